"""Training pairs through a patch shard on the GPU: a shard written on the
host is appended to a trainer's device patch set (dft_cnn_trainer_add_patches),
read back bit for bit, written out again, and evaluated by the network
against the numpy oracle.  Forward outputs are FP32 with BF16 parameters
against the oracle's FP64 on the same BF16-rounded parameters, so the mean
squared error agrees to a relative 1e-4."""
import numpy as np
import pytest

from oracle import cnn_oracle as O

pytestmark = pytest.mark.gpu


def test_trainer_round_trips_a_patch_shard(tmp_path):
    import dft_amd
    P, count = 8, 5
    rng = np.random.default_rng(11)
    xin = rng.standard_normal((count, 2, P ** 3)).astype(np.float32)
    tgt = rng.standard_normal((count, P ** 3)).astype(np.float32)
    path = tmp_path / "pairs.dftp"
    dft_amd.save_patch_shard(path, xin, tgt, target_scale=0, h=0.2)
    print(f"[shards] trainer patch {P}, capacity 6, shard of {count}", flush=True)

    tr = dft_amd.CnnTrainer(patch=P, capacity=6, seed=4)
    try:
        hdr = tr.load_patches(path)
        assert hdr["count"] == count and tr.count() == count
        a, b = tr.patches(0, count)
        assert a.tobytes() == xin.tobytes() and b.tobytes() == tgt.tobytes()
        # beyond the free capacity: refused, nothing appended
        with pytest.raises(dft_amd.DftError) as e:
            tr.add_patches(xin[:2], tgt[:2])
        assert e.value.code == -1 and tr.count() == count
        tr.add_patches(xin[:1], tgt[:1])
        assert tr.count() == count + 1
        out = tmp_path / "again.dftp"
        tr.save_patches(out, h=0.2)
        a, b, h2 = dft_amd.load_patch_shard(out)
        assert np.array_equal(a, np.concatenate([xin, xin[:1]])) and np.array_equal(b, np.concatenate([tgt, tgt[:1]]))
        assert h2["h"] == 0.2

        # the network evaluates the loaded pairs as the oracle does
        p = O.bf16_round(tr.params())
        o, _ = O.forward(p, xin.astype(np.float64).reshape(count, 2, P, P, P))
        want = float(np.mean((o - tgt.astype(np.float64).reshape(count, P, P, P)) ** 2))
        got = tr.eval(0, count)
        assert abs(got - want) <= 1e-4 * want, (got, want)
        loss = tr.train(3, batch=2, lr=1e-3)
        assert np.all(np.isfinite(loss))
    finally:
        tr.close()
