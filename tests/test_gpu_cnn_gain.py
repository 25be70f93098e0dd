"""End to end on the GPU: harvest training pairs from the method's own SCF
runs (projected CG on the correction equation), train the network with the
HIP forward/backward/Adam kernels, and measure what the trained network gains
over the diagonal correction on held-out molecules.

The gain is reported, not assumed.  The assertions are:
- training lowers the loss on the training patches;
- the learned Davidson converges to the same eigenvalues as the diagonal one
  (the per-column diagonal fallback guarantees progress);
- the network was used.
The measured ratio of H applications, diagonal / learned, is printed and
recorded in the DESIGN document."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def cluster(seed, natoms, radius=3.0):
    from bench import random_cluster
    return np.array(random_cluster(seed, natoms, radius), float)


def test_train_and_gain(capsys):
    import dft_amd
    g = dft_amd.Grid.centered(48, 0.35)
    tr = dft_amd.CnnTrainer(patch=16, capacity=3000, seed=3)
    for s in (1, 2, 3):
        m = cluster(s, 3)
        ne = int(m[:, 0].sum())
        tr.harvest(g, m, nstates=ne // 2 + 2, max_iterations=4, poisson_fixed=200, davidson_tol=1e-5, max_columns=6,
                   patches_per_column=2, every=3, cg_tol=1e-4, cg_max_iter=150, seed=s)
    n = tr.count()
    assert n > 100
    before = tr.eval(0, min(n, 400))
    tr.train(1500, batch=16, lr=1e-3)
    after = tr.eval(0, min(n, 400))
    assert np.isfinite(after) and after < 0.9 * before, (before, after)
    model = tr.model()
    tr.close()

    dev = dft_amd.Device(0)
    ratios = []
    for s in (101, 102):
        m = cluster(s, 3)
        ne = int(m[:, 0].sum())
        v = dev.zeros((g.size,))
        dev.nuclear_potential(g, m, v)
        vh = v.download()
        k = ne // 2 + 2
        rd = dft_amd.davidson_run(g, vh, k, nconv=ne // 2, tol=1e-6, max_iter=800)
        rc = dft_amd.davidson_run(g, vh, k, nconv=ne // 2, tol=1e-6, max_iter=800, cnn=model)
        assert np.max(np.abs(rc["eigenvalues"][: ne // 2] - rd["eigenvalues"][: ne // 2])) < 1e-8
        assert rc["cnn_applied"] > 0
        ratios.append(rd["happly_columns"] / rc["happly_columns"])
        with capsys.disabled():
            print(f"\n[cnn gain] molecule {s}: H applies diagonal {rd['happly_columns']} learned "
                  f"{rc['happly_columns']} (learned corrections {rc['cnn_applied']}, fallbacks "
                  f"{rc['cnn_fallbacks']}); gain {ratios[-1]:.2f}x; loss {before:.3e} -> {after:.3e}")
    dev.close()


def test_train_and_gain_on_the_kinetic_base(capsys):
    """The model that adds to the exact kinetic correction (base 1, harvested with target_scale 2):
    training lowers the held-in loss; on held-out molecules the learned Davidson reaches the kinetic
    correction's eigenvalues with the network applied.  The ratio of H applications, kinetic / learned,
    is printed (measured below 1 in this build: the network's local correction converges, but more
    slowly than the kinetic correction alone, so the learned Davidson gets a larger iteration cap)."""
    import dft_amd
    g = dft_amd.Grid.centered(48, 0.35)
    tr = dft_amd.CnnTrainer(patch=16, capacity=3000, seed=5)
    for s in (1, 2, 3):
        m = cluster(s, 3)
        ne = int(m[:, 0].sum())
        tr.harvest(g, m, nstates=ne // 2 + 2, max_iterations=4, poisson_fixed=200, davidson_tol=1e-5, max_columns=6,
                   patches_per_column=4, every=3, cg_tol=1e-4, cg_max_iter=150, seed=s, target_scale=2)
    n = tr.count()
    assert n > 100
    before = tr.eval(0, min(n, 400))
    tr.train(1500, batch=16, lr=1e-3)
    after = tr.eval(0, min(n, 400))
    assert np.isfinite(after) and after < before, (before, after)
    model = tr.model(base=dft_amd.CNN_BASE_KINETIC)
    tr.close()
    dev = dft_amd.Device(0)
    for s in (101, 102):
        m = cluster(s, 3)
        ne = int(m[:, 0].sum())
        v = dev.zeros((g.size,))
        dev.nuclear_potential(g, m, v)
        vh = v.download()
        k = ne // 2 + 2
        rk = dft_amd.davidson_run(g, vh, k, nconv=ne // 2, tol=1e-6, max_iter=800, precond=dft_amd.PRECOND_KINETIC_EXACT)
        rc = dft_amd.davidson_run(g, vh, k, nconv=ne // 2, tol=1e-6, max_iter=6000, cnn=model)
        assert np.max(np.abs(rc["eigenvalues"][: ne // 2] - rk["eigenvalues"][: ne // 2])) < 1e-8
        assert rc["cnn_applied"] > 0
        with capsys.disabled():
            print(f"\n[cnn gain, kinetic base] molecule {s}: H applies kinetic {rk['happly_columns']} learned "
                  f"{rc['happly_columns']} (learned {rc['cnn_applied']}, fallbacks {rc['cnn_fallbacks']}); "
                  f"ratio {rk['happly_columns'] / rc['happly_columns']:.2f}; loss {before:.3e} -> {after:.3e}")
    dev.close()
