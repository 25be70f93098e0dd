"""GPU parity of the learned-preconditioner kernels (cnn.hip, harvest.hip)
against the numpy oracle (oracle/cnn_oracle.py), through the C-ABI.

Tolerances: the kernels compute in FP32 (BF16 parameters, FP32 activations
and accumulation) and the oracle in FP64 on the same BF16-rounded
parameters, so forward outputs are compared normwise within 3e-5 and
gradients within 3e-4 (sums of ~1e5 FP32 products); bf16 rounding and the
harvest maps are exact or at FP64/FP32 rounding."""
import numpy as np
import pytest

from oracle import cnn_oracle as O

pytestmark = pytest.mark.gpu

C = 16


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def dev():
    import dft_amd
    d = dft_amd.Device(0)
    yield d
    d.close()


def upload_params(dev, p):
    w32 = dev.array(p.astype(np.float32), dtype=np.float32)
    wbf = dev.array(shape=(p.size,), dtype=np.uint16)
    dev.cnn_to_bf16(w32, wbf)
    return w32, wbf


@pytest.mark.parametrize("dims,nb", [((37, 10, 9), 2), ((16, 16, 16), 3), ((5, 7, 3), 1)])
def test_cnn_forward_matches_oracle(dev, dims, nb):
    nx, ny, nz = dims
    rng = np.random.default_rng(nx)
    p = O.he_params(C, seed=nx)
    lay, n = O.layout(C)
    for _, co, _, b in lay:
        p[b:b + co] = rng.uniform(-0.1, 0.1, co)  # non-zero biases
    x = rng.standard_normal((nb, 2, nz, ny, nx)).astype(np.float32)
    w32, wbf = upload_params(dev, p)
    assert np.array_equal(wbf.download(), O.bf16_bits(p))
    xin = dev.array(x.reshape(nb, -1), dtype=np.float32)
    out = dev.array(shape=(nb, nx * ny * nz), dtype=np.float32)
    for training in (1, 0):
        act = dev.array(shape=(dev.cnn_act_count(dims, nb, training),), dtype=np.float32)
        dev.cnn_forward(dims, nb, wbf, xin, out, act, training)
        ref, zs = O.forward(O.bf16_round(p), x, C)
        assert rel(out.download().reshape(ref.shape), ref) <= 3e-5
        if training:
            a = act.download()
            off = (n + 63) // 64 * 64
            layer = C * nx * ny * nz * nb
            for l in range(3):
                z = a[off + l * layer: off + (l + 1) * layer].reshape(zs[l].shape)
                assert rel(z, zs[l]) <= 3e-5


@pytest.mark.parametrize("dims,nb", [((20, 9, 6), 2), ((16, 16, 16), 1)])
def test_cnn_backward_matches_oracle(dev, dims, nb):
    nx, ny, nz = dims
    rng = np.random.default_rng(7 + nx)
    p = O.he_params(C, seed=11)
    lay, n = O.layout(C)
    for _, co, _, b in lay:
        p[b:b + co] = rng.uniform(-0.1, 0.1, co)
    x = rng.standard_normal((nb, 2, nz, ny, nx)).astype(np.float32)
    t = rng.standard_normal((nb, nz, ny, nx)).astype(np.float32)
    w32, wbf = upload_params(dev, p)
    xin = dev.array(x.reshape(nb, -1), dtype=np.float32)
    tg = dev.array(t.reshape(nb, -1), dtype=np.float32)
    out = dev.array(shape=(nb, nx * ny * nz), dtype=np.float32)
    act = dev.array(shape=(dev.cnn_act_count(dims, nb, 1),), dtype=np.float32)
    scratch = dev.array(shape=(dev.cnn_scratch_count(dims, nb),), dtype=np.float32)
    grad = dev.array(shape=(n,), dtype=np.float32)
    dev.cnn_forward(dims, nb, wbf, xin, out, act, 1)
    loss = dev.cnn_backward(dims, nb, wbf, xin, act, out, tg, grad, scratch)
    ref_loss, ref_g = O.loss_and_grad(O.bf16_round(p), x, t, C)
    assert abs(loss - ref_loss) <= 3e-5 * ref_loss
    g = grad.download()
    # per layer, normwise
    for ci, co, w, b in lay:
        assert rel(g[w:b], ref_g[w:b]) <= 3e-4, (ci, co)
        assert rel(g[b:b + co], ref_g[b:b + co]) <= 3e-4, (ci, co)


def test_cnn_adam_matches_oracle(dev):
    rng = np.random.default_rng(3)
    n = O.layout(C)[1]
    w = rng.standard_normal(n).astype(np.float32)
    g = rng.standard_normal(n).astype(np.float32)
    m = (0.1 * rng.standard_normal(n)).astype(np.float32)
    v = (0.01 * rng.random(n)).astype(np.float32)
    dw, dg, dm, dv = (dev.array(a, dtype=np.float32) for a in (w, g, m, v))
    wbf = dev.array(shape=(n,), dtype=np.uint16)
    dev.cnn_adam(dw, wbf, dg, dm, dv, step=3, lr=1e-3)
    # the kernel's hyper-parameters are FP32: beta2 = 0.999f = 0.999000013 (1 - beta2 differs by 1.3e-5)
    f = lambda x: float(np.float32(x))
    rw, rm, rv = O.adam_step(w, g, m, v, 3, lr=f(1e-3), b1=f(0.9), b2=f(0.999), eps=f(1e-8))
    assert rel(dm.download(), rm) <= 1e-6
    assert rel(dv.download(), rv) <= 1e-6
    w1 = dw.download()
    assert np.max(np.abs(w1 - rw)) <= 1e-6 * np.max(np.abs(rw))
    assert np.array_equal(wbf.download(), O.bf16_bits(w1))


def test_preconditioner_maps(dev):
    import dft_amd
    g = dft_amd.Grid.centered(12, 0.4)
    n = g.size
    rng = np.random.default_rng(5)
    R = rng.standard_normal((3, n))
    V = rng.standard_normal(n)
    theta = np.array([-1.5, -0.25, 0.3])
    dR, dV, dth = dev.array(R), dev.array(V), dev.array(theta)
    sc = dev.zeros((3,))
    dev.column_maxabs(dR, n, 3, sc)
    assert np.array_equal(sc.download(), np.max(np.abs(R), axis=1))
    xin = dev.array(shape=(3, 2, n), dtype=np.float32)
    dev.cnn_prepare(g, dV, dR, 3, dth, sc, xin)
    h = (g.dx, g.dy, g.dz)
    got = xin.download()
    for b in range(3):
        ref, s, h2 = O.prepare(R[b], V, theta[b], h)
        assert rel(got[b], ref.astype(np.float32)) <= 1e-7
    out = dev.array(rng.standard_normal((3, n)).astype(np.float32), dtype=np.float32)
    T = dev.zeros((3, n))
    dev.cnn_finish(g, 3, out, sc, T)
    h2 = 1.0 / sum(1.0 / x ** 2 for x in h)
    ref = out.download().astype(np.float64) * np.max(np.abs(R), axis=1)[:, None] * h2
    assert rel(T.download(), ref) <= 1e-15


def test_harvest_kernels(dev):
    import dft_amd
    g = dft_amd.Grid.centered(16, 0.3)
    n = g.size
    rng = np.random.default_rng(9)
    X, Y = rng.standard_normal((4, n)), rng.standard_normal((4, n))
    dX, dY = dev.array(X), dev.array(Y)
    d = dev.zeros((4,))
    dev.column_dots(dX, dY, n, 4, d)
    assert rel(d.download(), np.sum(X * Y, axis=1)) <= 1e-13
    a, b = rng.standard_normal(4), rng.standard_normal(4)
    dev.cols_axpby(dev.array(a), dX, dev.array(b), dY, n, 4)
    assert rel(dY.download(), a[:, None] * X + b[:, None] * Y) <= 1e-15
    # patch energies and extraction on an 8^3 tiling of 16^3
    P = 8
    E = dev.zeros((4, 8))
    dev.patch_energy(g, P, dX, 4, E)
    Xg = X.reshape(4, 16, 16, 16)
    ref = np.array([[np.sum(Xg[j, bz * P:(bz + 1) * P, by * P:(by + 1) * P, bx * P:(bx + 1) * P] ** 2)
                     for bz in range(2) for by in range(2) for bx in range(2)] for j in range(4)])
    assert rel(E.download(), ref) <= 1e-13
    V = rng.standard_normal(n)
    Cm = rng.standard_normal((4, n))
    theta = rng.standard_normal(4)
    desc = np.array([[0, 0, 0, 0], [3, 8, 3, 5], [1, 2, 8, 8]])
    sc, cs = np.max(np.abs(X), axis=1), np.max(np.abs(Cm), axis=1)
    xin = dev.array(shape=(3, 2, P ** 3), dtype=np.float32)
    tgt = dev.array(shape=(3, P ** 3), dtype=np.float32)
    dev.extract_patches(g, P, desc, 4, dev.array(V), dX, dev.array(Cm), dev.array(theta), dev.array(sc),
                        dev.array(cs), xin, tgt)
    h2 = 1.0 / (3.0 / 0.3 ** 2)
    Vg, Cg = V.reshape(16, 16, 16), Cm.reshape(4, 16, 16, 16)
    gi, gt = xin.download(), tgt.download()
    for p, (j, x0, y0, z0) in enumerate(desc):
        sl = (slice(z0, z0 + P), slice(y0, y0 + P), slice(x0, x0 + P))
        assert rel(gi[p, 0], (Xg[j][sl] / sc[j]).ravel().astype(np.float32)) <= 1e-7
        assert rel(gi[p, 1], ((Vg[sl] - theta[j]) * h2).ravel().astype(np.float32)) <= 1e-7
        assert rel(gt[p], (Cg[j][sl] / cs[j]).ravel().astype(np.float32)) <= 1e-7
    with pytest.raises(dft_amd.DftError):
        dev.extract_patches(g, P, np.array([[0, 9, 0, 0]]), 4, dev.array(V), dX, dev.array(Cm), dev.array(theta),
                            dev.array(sc), dev.array(cs), xin, tgt)
    with pytest.raises(dft_amd.DftError):
        dev.extract_patches(g, P, np.array([[4, 0, 0, 0]]), 4, dev.array(V), dX, dev.array(Cm), dev.array(theta),
                            dev.array(sc), dev.array(cs), xin, tgt)


def _ramp(l, b, lo_in, hi_in, taper):
    w = np.ones(b)
    idx = np.arange(b)
    if lo_in:
        w = np.minimum(w, (idx + 1) / (taper + 1))
    if hi_in:
        w = np.minimum(w, (b - idx) / (taper + 1))
    return w


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_cnn_box_maps(dt):
    """dft_cnn_prepare_box / dft_cnn_add_box against numpy: the inputs cut from a sub-box, and
    t[box] += out s h2 w with the face ramps (a box touching the grid's x faces has no ramp there)."""
    import dft_amd
    from oracle.oracle import Grid
    og = Grid(20, 18, 16, -2.0, 2.1, -1.9, 1.6, -1.7, 1.4)
    g = dft_amd.Grid.make(*og.n, *og.bounds)
    dev = dft_amd.Device(0)
    rng = np.random.default_rng(4)
    n, nc = og.size, 3
    V = rng.standard_normal(n)
    R = rng.standard_normal((nc, n))
    th = rng.standard_normal(nc)
    sc = np.abs(R).max(axis=1)
    box = [0, 3, 2, 20, 11, 9]  # x spans the grid; y, z inside
    bx, by, bz = box[3:]
    nb = bx * by * bz
    xin = dev.array(shape=(nc, 2, nb), dtype=np.float32)
    dev.cnn_prepare_box(g, dev.array(V, dtype=dt), dev.array(R, dtype=dt), nc, dev.array(th), dev.array(sc), box, xin)
    Vb = V.astype(dt).reshape(og.shape)[2:2 + bz, 3:3 + by, 0:bx].ravel()
    Rb = R.astype(dt).reshape(nc, *og.shape)[:, 2:2 + bz, 3:3 + by, 0:bx].reshape(nc, -1)
    h2 = 1.0 / (1 / og.dx ** 2 + 1 / og.dy ** 2 + 1 / og.dz ** 2)
    got = xin.download()
    np.testing.assert_allclose(got[:, 0], Rb / sc[:, None], rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(got[:, 1], (Vb[None, :] - th[:, None]) * h2, rtol=1e-6, atol=1e-6)
    out = rng.standard_normal((nc, nb)).astype(np.float32)
    T0 = rng.standard_normal((nc, n))
    t = dev.array(T0, dtype=dt)
    dev.cnn_add_box(g, nc, dev.array(out, dtype=np.float32), dev.array(sc), box, 3, t)
    w = (_ramp(0, bz, True, True, 3)[:, None, None] * _ramp(0, by, True, True, 3)[None, :, None] *
         _ramp(0, bx, False, False, 3)[None, None, :]).ravel()
    ref = T0.astype(dt).astype(np.float64).reshape(nc, *og.shape).copy()
    add = (out.astype(np.float64) * sc[:, None] * h2 * w[None, :]).reshape(nc, bz, by, bx)
    ref[:, 2:2 + bz, 3:3 + by, 0:bx] += add
    tol = 1e-12 if dt == np.float64 else 1e-5
    np.testing.assert_allclose(t.download().astype(np.float64), ref.reshape(nc, n), rtol=tol, atol=tol)
    dev.close()


def test_zero_network_on_the_kinetic_base():
    """A base-1 model with all-zero parameters adds nothing to the exact kinetic correction: the
    Davidson reaches the kinetic correction's eigenvalues, with the network applied and the diagonal
    fallback used only for the columns whose residual did not go down."""
    import dft_amd
    from oracle.oracle import Grid, Oracle
    og = Grid.centered(40, 0.3)
    g = dft_amd.Grid.make(*og.n, *og.bounds)
    nuc = [[6, 0.05, 0.02, -0.03], [1, 2.0, 0.1, 0.0]]
    v = Oracle().nuclear_potential(og, nuc)
    zero = dft_amd.CnnModel(np.zeros(dft_amd.cnn_param_count(), np.float32), base=dft_amd.CNN_BASE_KINETIC)
    rk = dft_amd.davidson_run(g, v, 6, nconv=4, tol=1e-7, max_iter=400, precond=dft_amd.PRECOND_KINETIC_EXACT)
    rz = dft_amd.davidson_run(g, v, 6, nconv=4, tol=1e-7, max_iter=400, cnn=zero)
    print("zero-net", rk["happly_columns"], rz["happly_columns"], rz["cnn_applied"], rz["cnn_fallbacks"])
    assert np.max(np.abs(rz["eigenvalues"][:4] - rk["eigenvalues"][:4])) < 1e-8
    assert rz["cnn_applied"] > 0


@pytest.mark.parametrize("target_scale", [0, 2])
def test_harvest_cg_true_residual(target_scale):
    """The projected CG of the harvest stops when its recurrence's residual is below cg_tol; the
    residual recomputed from the solution afterwards, |b - A c| / |b| with
    A = (I - QQ^T)(H - theta)(I - QQ^T) and b = -(I - QQ^T) r, is within 2 cg_tol of it for every
    converged column (the two differ by FP64 rounding drift, orders of magnitude below cg_tol here).
    Both the diagonal-corrected harvest and the kinetic-base one (target_scale 2)."""
    import dft_amd
    g = dft_amd.Grid.centered(40, 0.35)
    tr = dft_amd.CnnTrainer(patch=16, capacity=600, seed=11)
    nuc = [[6, 0.05, 0.02, -0.03], [1, 2.0, 0.1, 0.0], [1, -1.1, 1.6, 0.2]]
    cg_tol = 1e-6
    tr.harvest(g, nuc, nstates=6, max_iterations=3, poisson_fixed=200, davidson_tol=1e-5, max_columns=4,
               patches_per_column=2, every=2, cg_tol=cg_tol, cg_max_iter=400, seed=3, multiplicity=1,
               target_scale=target_scale)
    st = tr.harvest_stats()
    tr.close()
    print("harvest stats", st)
    assert st["solves"] > 0 and st["solves"] > st["failed"] and st["patches"] > 0
    assert 0.0 < st["max_true_residual"] <= 2.0 * cg_tol
