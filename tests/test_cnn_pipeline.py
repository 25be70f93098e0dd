"""Host logic of the learned-preconditioner pipeline (dft_amd/cnn_pipeline.py)
and the molecule set it draws from (dft_amd/geometry.py): seed lists, the
geometry constraints of every training and held-out molecule, and the shard
plan that splits training from validation.  The harvest, training and
evaluation steps themselves run on the GPU."""
import itertools

import numpy as np
import pytest

import dft_amd
from dft_amd import cnn_pipeline as cp
from dft_amd.geometry import random_cluster, training_molecule


def test_parse_seeds():
    assert cp.parse_seeds("1-64") == list(range(1, 65))
    assert cp.parse_seeds("1001-1008") == list(range(1001, 1009))
    assert cp.parse_seeds("3, 1-4,9") == [3, 1, 2, 4, 9]
    assert cp.parse_seeds("") == []
    for bad in ("5-2", "-3", "a-b"):
        with pytest.raises(ValueError):
            cp.parse_seeds(bad)


@pytest.mark.parametrize("seed", list(range(1, 65)) + list(range(1001, 1009)))
def test_training_molecules_meet_the_set_definition(seed):
    m = np.array(training_molecule(seed))
    z, pos = m[:, 0], m[:, 1:]
    assert set(z) <= {1.0, 6.0, 7.0, 8.0}
    # 4..12 drawn atoms, plus one H when their electron count is odd
    assert 4 <= len(m) <= 13 and int(z.sum()) % 2 == 0
    if len(m) == 13:
        assert z[-1] == 1.0
    assert np.all(np.linalg.norm(pos, axis=1) <= 5.0)
    d = [np.linalg.norm(a - b) for a, b in itertools.combinations(pos, 2)]
    assert min(d) >= 2.0
    assert np.array_equal(m, np.array(training_molecule(seed)))  # a seed names one geometry


def test_training_set_varies():
    sizes = {len(training_molecule(s)) for s in range(1, 65)}
    assert len(sizes) >= 5
    assert not np.array_equal(np.array(training_molecule(1)), np.array(training_molecule(2)))


def test_cluster_generator_keeps_the_benchmark_cluster():
    """The 40-atom cluster of the 256^3 workload: seed 5, radius 7, 150-170 electrons."""
    m = np.array(random_cluster(5, 40, 7.0, ne_range=(150, 170)))
    assert m.shape == (40, 4) and int(m[:, 0].sum()) == 168


def _shard(path, count, P):
    rng = np.random.default_rng(count)
    dft_amd.save_patch_shard(path, rng.standard_normal((count, 2, P ** 3)).astype(np.float32),
                             rng.standard_normal((count, P ** 3)).astype(np.float32))


def test_shard_plan(tmp_path):
    with pytest.raises(FileNotFoundError):
        cp.shard_plan(tmp_path)
    for seed, count in ((3, 5), (1, 7), (12, 4)):
        _shard(cp.shard_path(tmp_path, seed), count, 4)
    plan = cp.shard_plan(tmp_path)
    assert [p.rsplit("/", 1)[1] for p in plan["paths"]] == ["seed00001.dftp", "seed00003.dftp", "seed00012.dftp"]
    assert plan["counts"] == [7, 5, 4] and plan["patch"] == 4
    assert (plan["total"], plan["train_count"], plan["val_count"]) == (16, 12, 4)
    _shard(cp.shard_path(tmp_path, 20), 2, 5)
    with pytest.raises(ValueError, match="patch edges"):
        cp.shard_plan(tmp_path)


def test_single_shard_trains_on_everything(tmp_path):
    _shard(cp.shard_path(tmp_path, 1), 6, 4)
    plan = cp.shard_plan(tmp_path)
    assert (plan["train_count"], plan["val_count"]) == (6, 0)


class _FakeTrainer:
    """Stands in for CnnTrainer (a device patch set): records calls, holds patches in host memory."""
    made = []

    def __init__(self, patch=32, capacity=1, seed=0, channels=16):
        self.patch, self.capacity, self.seed = patch, capacity, seed
        self.xin = np.zeros((0, 2, patch ** 3), np.float32)
        self.tgt = np.zeros((0, patch ** 3), np.float32)
        self.calls = []
        _FakeTrainer.made.append(self)

    def harvest(self, grid, atoms, **kw):
        n = min(self.capacity, 3 + len(atoms))
        rng = np.random.default_rng(kw["seed"])
        self.add_patches(rng.standard_normal((n, 2, self.patch ** 3)), rng.standard_normal((n, self.patch ** 3)))
        self.calls.append(("harvest", kw["seed"], kw["nstates"]))
        return n

    def add_patches(self, xin, tgt):
        self.xin = np.concatenate([self.xin, np.asarray(xin, np.float32).reshape(-1, 2, self.patch ** 3)])
        self.tgt = np.concatenate([self.tgt, np.asarray(tgt, np.float32).reshape(-1, self.patch ** 3)])
        assert len(self.tgt) <= self.capacity

    def count(self):
        return len(self.tgt)

    def save_patches(self, path, target_scale=0, h=0.0):
        dft_amd.save_patch_shard(path, self.xin, self.tgt, target_scale=target_scale, h=h)

    def load_patches(self, path):
        xin, tgt, hdr = dft_amd.load_patch_shard(path)
        self.add_patches(xin, tgt)
        return hdr

    def train(self, steps, batch=8, lr=1e-3, train_count=0):
        self.calls.append(("train", steps, train_count))
        return np.full(steps, 0.5)

    def eval(self, first, count):
        self.calls.append(("eval", first, count))
        return 0.25

    def model(self, base=0):
        return dft_amd.CnnModel(np.zeros(dft_amd.cnn_param_count(), np.float32), base=base)

    def close(self):
        pass


def test_harvest_resumes_and_train_holds_out_the_last_molecule(tmp_path, monkeypatch):
    monkeypatch.setattr(cp, "CnnTrainer", _FakeTrainer)
    monkeypatch.setattr(cp, "Grid", type("G", (), {"centered": staticmethod(lambda n, h: (n, h))}))
    _FakeTrainer.made.clear()
    m = cp.harvest(tmp_path, [3, 1], n=16, h=0.5, patch=4, target_patches=20)
    assert [r["seed"] for r in m["molecules"]] == [3, 1] and m["per_molecule"] == 10
    assert all(r["patches"] == min(10, 3 + len(training_molecule(r["seed"]))) for r in m["molecules"])
    # nstates: the occupied count plus 4
    assert [c[2] for t in _FakeTrainer.made for c in t.calls] == \
        [int(round(np.array(training_molecule(s))[:, 0].sum())) // 2 + 4 for s in (3, 1)]
    # a second call harvests only the new seed
    _FakeTrainer.made.clear()
    m = cp.harvest(tmp_path, [1, 3, 2], n=16, h=0.5, patch=4, target_patches=30)
    assert [c[1] for t in _FakeTrainer.made for c in t.calls] == [2]
    assert sorted(r["seed"] for r in m["molecules"]) == [1, 2, 3]
    assert cp.patch_shard_info(cp.shard_path(tmp_path, 1))["h"] == 0.5
    # training: every shard loaded, the last one (seed 3) held out
    plan = cp.shard_plan(tmp_path)
    _FakeTrainer.made.clear()
    _, log = cp.train(tmp_path, steps=1200, batch=2, chunk=500)
    t = _FakeTrainer.made[0]
    assert t.count() == plan["total"]
    assert [c for c in t.calls if c[0] == "train"] == [("train", 500, plan["train_count"])] * 2 + \
        [("train", 200, plan["train_count"])]
    assert ("eval", plan["train_count"], plan["val_count"]) in t.calls
    assert [r["step"] for r in log["curve"]] == [500, 1000, 1200]
    assert (tmp_path / "model.npz").exists() and (tmp_path / "train_log.json").exists()


def test_evaluate_compares_diagonal_and_learned(tmp_path, monkeypatch):
    dft_amd.CnnModel(np.zeros(dft_amd.cnn_param_count(), np.float32)).save(tmp_path / "model.npz")
    runs = []

    def fake_scf(grid, atoms, precond=None, cnn=None, **kw):
        learned = precond == dft_amd.PRECOND_CNN
        assert (cnn is not None) == learned
        runs.append(precond)
        return dict(converged=True, happly_columns=300 if learned else 400, davidson_iterations=10,
                    seconds_total=2.0 if learned else 3.0, cnn_applied=7 if learned else 0, cnn_fallbacks=1,
                    total=-1.0 + (1e-9 if learned else 0.0))

    monkeypatch.setattr(cp, "scf_run", fake_scf)
    monkeypatch.setattr(cp, "scf_result_json", lambda g, a, r: "{}")
    monkeypatch.setattr(cp, "Grid", type("G", (), {"centered": staticmethod(lambda n, h: (n, h))}))
    s = cp.evaluate(tmp_path, [1001, 1002], n=16, h=0.5)
    assert runs == [dft_amd.PRECOND_DIAG, dft_amd.PRECOND_CNN] * 2
    assert s["mean_happly_gain"] == pytest.approx(4 / 3)
    assert s["molecules"][0]["happly"] == [400, 300] and s["molecules"][0]["time_gain"] == pytest.approx(1.5)
    assert s["max_abs_energy_difference"] == pytest.approx(1e-9)
    assert (tmp_path / "eval.json").exists() and (tmp_path / "eval_seed01001_learned.json").exists()


def test_harvest_refuses_to_resume_with_other_settings(tmp_path, monkeypatch):
    """A resumed harvest must use the grid, spacing, patch edge, functional and base it started with;
    shards of different spacings or targets never train together."""
    monkeypatch.setattr(cp, "CnnTrainer", _FakeTrainer)
    monkeypatch.setattr(cp, "Grid", type("G", (), {"centered": staticmethod(lambda n, h: (n, h))}))
    cp.harvest(tmp_path, [1], n=16, h=0.5, patch=4, target_patches=10)
    for kw in (dict(h=0.4), dict(n=20), dict(base="diagonal")):
        args = dict(n=16, h=0.5, patch=4, target_patches=10)
        args.update(kw)
        with pytest.raises(ValueError, match="harvested with"):
            cp.harvest(tmp_path, [2], **args)
    # a shard of another spacing dropped into the directory
    xin = np.zeros((2, 2, 64), np.float32)
    dft_amd.save_patch_shard(cp.shard_path(tmp_path, 7), xin, np.zeros((2, 64), np.float32), target_scale=2, h=0.3)
    with pytest.raises(ValueError, match="spacings"):
        cp.shard_plan(tmp_path)


def test_kinetic_base_harvest_trains_a_base_one_model(tmp_path, monkeypatch):
    """base "kinetic": shards carry target_scale 2 and the trained model adds to the kinetic correction;
    evaluation compares it with the exact kinetic correction."""
    monkeypatch.setattr(cp, "CnnTrainer", _FakeTrainer)
    monkeypatch.setattr(cp, "Grid", type("G", (), {"centered": staticmethod(lambda n, h: (n, h))}))
    cp.harvest(tmp_path, [1, 2], n=16, h=0.5, patch=4, target_patches=20, base="kinetic")
    assert cp.patch_shard_info(cp.shard_path(tmp_path, 1))["target_scale"] == 2
    model, _ = cp.train(tmp_path, steps=10, batch=2, chunk=10)
    assert model.base == dft_amd.CNN_BASE_KINETIC
    runs = []

    def fake_scf(grid, atoms, precond=None, cnn=None, **kw):
        runs.append(precond)
        return dict(converged=True, happly_columns=100, davidson_iterations=10, seconds_total=1.0, cnn_applied=0,
                    cnn_fallbacks=0, total=-1.0)

    monkeypatch.setattr(cp, "scf_run", fake_scf)
    monkeypatch.setattr(cp, "scf_result_json", lambda g, a, r: "{}")
    cp.evaluate(tmp_path, [1001], n=16, h=0.5)
    assert runs == [dft_amd.PRECOND_KINETIC_EXACT, dft_amd.PRECOND_CNN]
