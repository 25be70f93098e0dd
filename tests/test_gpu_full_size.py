"""Kernels at the workloads' full grid shapes: the benzene grid (160^3,
h = 0.2) against the FP64 oracle directly, and the 40-atom cluster's grid
(256^3, h = 0.18) through properties that do not depend on size.

- H psi: at 160^3 the FP64 block (a 4-orbital group plus a 2-orbital one) is
  compared with the oracle at relative 1e-12, the FP32 block at 2e-6 (inputs
  rounded to 2^-24, ~8 stencil terms).  At 256^3, Dirichlet sine modes are
  eigenvectors of the stencil with a closed-form eigenvalue, shifted by a
  constant potential: FP64 within 1e-12 of the mode's peak value; FP32 within
  2e-4 (the lowest mode's eigenvalue is ~1/70 of the stencil's terms, which
  scales the rounding of the input); the 64-orbital FP32 block takes the
  prefetching path of large blocks.
- Density: sum over the grid of sin^2 along an axis is (n + 1) / 2 for every
  mode, so normalised modes integrate to the electron count (1e-12).
- Poisson: a boundary sine-mode source has the Jacobi fixed point
  -4 pi rho / mu in closed form (tests/fullsize_forms.py, checked against the
  oracle's converged Jacobi on small grids), relative 1e-12 in FP64 and 1e-4
  in FP32 (three length-256 transforms each way).
- XC: PZ81 and PBE96 FP64 on a benzene-shaped open-shell density whose
  corners fall below the 1e-20 floor, against the oracle at relative 1e-12;
  the energy within the rounding bound of the oracle's own sequential sum of
  4.1M same-sign terms, (N - 1) 2^-53 relative."""
import numpy as np
import pytest

from fullsize_forms import boundary_sine_mode, dirichlet_sine_modes

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def dev():
    import dft_amd
    d = dft_amd.Device(0)
    yield d
    d.close()


def _grids(n, h):
    import dft_amd
    from oracle.oracle import Grid
    return Grid.centered(n, h), dft_amd.Grid.centered(n, h)


def test_hpsi_benzene_grid_against_oracle(dev, oracle):
    og, g = _grids(160, 0.2)
    print(f"[full size] H psi 160^3 x 6 vs oracle", flush=True)
    rng = np.random.default_rng(21)
    v = rng.standard_normal(og.size)
    psi = rng.standard_normal((6, og.size))
    ref = np.stack([oracle.hpsi(og, v, psi[o]) for o in range(6)])
    out = dev.array(shape=(6, og.size))
    dev.hpsi(g, dev.array(v), dev.array(psi), out)
    e64 = rel(out.download(), ref)
    o32 = dev.array(shape=(6, og.size), dtype=np.float32)
    dev.hpsi(g, dev.array(v, dtype=np.float32), dev.array(psi, dtype=np.float32), o32)
    e32 = rel(o32.download(), ref)
    print(f"[full size] H psi 160^3: FP64 rel {e64:.2e}, FP32 rel {e32:.2e}", flush=True)
    assert e64 <= 1e-12 and e32 <= 2e-6


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_hpsi_cluster_grid_sine_modes_and_density(dev, prec):
    import dft_amd
    og, g = _grids(256, 0.18)
    n, h = og.n, (og.dx, og.dy, og.dz)
    modes = [(1, 1, 1), (2, 7, 3), (40, 5, 129), (256, 255, 128)]
    print(f"[full size] H psi 256^3 sine modes {prec}", flush=True)
    psi, lam = dirichlet_sine_modes(n, h, modes)
    c = -0.37
    dt = np.float64 if prec == "f64" else np.float32
    out = dev.array(shape=psi.shape, dtype=dt)
    dpsi = dev.array(psi, dtype=dt)
    dev.hpsi(g, dev.array(np.full(og.size, c), dtype=dt), dpsi, out)
    got = out.download().astype(np.float64)
    errs = [rel(got[o], (lam[o] + c) * psi[o]) for o in range(len(modes))]
    print(f"[full size] H psi 256^3 {prec} per-mode rel {['%.1e' % e for e in errs]}", flush=True)
    assert max(errs) <= (1e-12 if prec == "f64" else 2e-4)
    # density of normalised modes (sum psi^2 dV = 1 needs the factor prod_a 2 / (n_a + 1) / h_a)
    dv = og.dx * og.dy * og.dz
    norm = np.prod([2.0 / (na + 1) for na in n]) / dv
    occ = np.array([2.0, 1.0, 1.0, 0.5]) * norm
    dens = dev.array(shape=(og.size,), dtype=dt)
    dev.density(dpsi, occ, dens)
    charge = dev.reduce(dft_amd.RED_SUM, dens) * dv
    print(f"[full size] density 256^3 {prec}: charge {charge!r} (4.5)", flush=True)
    assert abs(charge - 4.5) <= (1e-12 if prec == "f64" else 1e-5) * 4.5


def test_hpsi_cluster_grid_large_fp32_block(dev):
    """64 orbitals in FP32: the block size of the prefetching kernel path."""
    og, g = _grids(256, 0.18)
    n, h = og.n, (og.dx, og.dy, og.dz)
    modes = [(1 + (3 * o) % 11, 1 + (5 * o) % 13, 1 + (7 * o) % 17) for o in range(64)]
    print(f"[full size] H psi 256^3 x 64 FP32", flush=True)
    psi = np.empty((64, og.size), dtype=np.float32)
    lam = np.empty(64)
    for o in range(64):
        p, l = dirichlet_sine_modes(n, h, [modes[o]])
        psi[o], lam[o] = p[0], l[0]
    # the FP32 input itself is the reference's operand: (lam + c) psi32 differs from H psi32 by the
    # input's rounding times the stencil's norm, the same order as the bound below
    dpsi = dev.array(psi, dtype=np.float32)
    c = 0.25
    out = dev.array(shape=psi.shape, dtype=np.float32)
    dev.hpsi(g, dev.array(np.full(og.size, c), dtype=np.float32), dpsi, out)
    got = out.download()
    worst = max(rel(got[o], (lam[o] + c) * psi[o].astype(np.float64)) for o in range(64))
    print(f"[full size] H psi 256^3 x 64 FP32 worst per-mode rel {worst:.2e}", flush=True)
    assert worst <= 2e-4


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_poisson_direct_cluster_grid_sine_mode(dev, prec):
    og, g = _grids(256, 0.18)
    rho, want = boundary_sine_mode(og.n, (og.dx, og.dy, og.dz))
    dt = np.float64 if prec == "f64" else np.float32
    v = dev.array(np.full(og.size, 3.0), dtype=dt)  # every point is written
    dev.poisson_direct(g, dev.array(rho, dtype=dt), v)
    e = rel(v.download(), want)
    print(f"[full size] Poisson direct 256^3 {prec} rel {e:.2e}", flush=True)
    assert e <= (1e-12 if prec == "f64" else 1e-4)


@pytest.mark.parametrize("fid", [1, 2])
def test_xc_benzene_grid_against_oracle(dev, oracle, fid):
    from dft_amd.geometry import benzene
    og, g = _grids(160, 0.2)
    xs, ys, zs = og.coords()
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    na = np.zeros(og.shape)
    for z, ax, ay, az in benzene():
        r2 = (X - ax) ** 2 + (Y - ay) ** 2 + (Z - az) ** 2
        na += 0.05 * z * (np.exp(-0.6 * r2) + 0.1 * np.exp(-2.0 * np.sqrt(r2)))
    nb = 0.8 * na * (1.0 + 0.1 * np.sin(0.7 * X) * np.cos(0.3 * Y))  # zeta in [0.06, 0.17]: no empty channel
    na, nb = na.ravel(), nb.ravel()
    assert (na + nb < 1e-20).any()  # the floor branch is taken in the corners
    print(f"[full size] XC 160^3 functional {fid} vs oracle", flush=True)
    ref_a = oracle.xc_potential(og, fid, na, nb, 0)
    ref_b = oracle.xc_potential(og, fid, na, nb, 1)
    ref_e = oracle.xc_energy(og, fid, na, nb)
    va, vb = dev.array(shape=(og.size,)), dev.array(shape=(og.size,))
    e = dev.xc(g, fid, dev.array(na), dev.array(nb), va, vb)
    ea, eb = rel(va.download(), ref_a), rel(vb.download(), ref_b)
    print(f"[full size] XC 160^3 functional {fid}: rel {ea:.2e} {eb:.2e}, energy {abs(e - ref_e) / abs(ref_e):.2e}",
          flush=True)
    assert ea <= 1e-12 and eb <= 1e-12
    # the energy against the exactly rounded sum of the oracle's per-point terms (math.fsum): the
    # oracle's own sum in index order is only within (N - 1) 2^-53 of it (4.5e-10 relative at 160^3);
    # the device's tree of FP64 partial sums holds 1e-13, and each term is within the terms' own
    # device/libm difference (the potential's 1e-12 bar) of the oracle's
    import math
    exact = math.fsum(oracle.xc_energy_terms(og, fid, na, nb))
    print(f"[full size] XC 160^3 energy vs fsum {abs(e - exact) / abs(exact):.2e}, oracle's own sum "
          f"{abs(ref_e - exact) / abs(exact):.2e}", flush=True)
    assert abs(e - exact) <= 1e-13 * abs(exact)
    assert abs(ref_e - exact) <= (og.size - 1) * 2.0 ** -53 * abs(exact)


def test_davidson_products_benzene_shape_exact(dev):
    """The Davidson's dense products at the timed workload's shape: N = 160^3
    rows, a basis of 87 columns, a block of 29.  Operands in {-1, 0, 1} (and
    small integer coefficients) make every product and every partial sum an
    integer below 2^24, exact in FP32, BF16 and FP64 alike, so each mode must
    return the integer result bit for bit: the Gram in FP64, FP32 and
    split-BF16 (V^T V and V^T W), the block update Y = V C (and Y += V C), and
    the Ritz residual R = W Q - V Q diag(theta), whose squared column norms
    agree at the FP64 bar."""
    import dft_amd
    n, m, k = 160 ** 3, 87, 29
    print(f"[full size] Davidson products {n} x {m} x {k}", flush=True)
    rng = np.random.default_rng(87)
    V8 = rng.integers(-1, 2, size=(m, n), dtype=np.int8)
    W8 = rng.integers(-1, 2, size=(m, n), dtype=np.int8)
    Cm = rng.integers(-2, 3, size=(k, m)).astype(np.float64)  # column-major m x k
    th = rng.integers(-2, 3, size=k).astype(np.float64)
    V32, W32 = V8.astype(np.float32), W8.astype(np.float32)
    ref_vv = (V32 @ V32.T).astype(np.float64)  # integers below 2^24: exact in FP32 arithmetic
    ref_vw = (V32 @ W32[:k].T).astype(np.float64)
    ref_y = (Cm.astype(np.float32) @ V32).astype(np.float64)
    ref_r = (Cm.astype(np.float32) @ W32 - th.astype(np.float32)[:, None] * (Cm.astype(np.float32) @ V32)).astype(np.float64)
    assert np.max(np.abs(ref_vv)) < 2 ** 24 and np.max(np.abs(ref_r)) < 2 ** 24
    dC, dth = dev.array(Cm), dev.array(th)
    for dt, modes in ((np.float64, (dft_amd.GRAM_F64,)), (np.float32, (dft_amd.GRAM_F32, dft_amd.GRAM_SPLIT_BF16, dft_amd.GRAM_SPLIT6_BF16))):
        dV, dW = dev.array(V8.astype(dt), dtype=dt), dev.array(W8.astype(dt), dtype=dt)
        for mode in modes:
            out = dev.zeros((m, m))
            dev.gram(mode, dV, dV, out, n, m, m)
            assert np.array_equal(out.download().T, ref_vv), ("V^T V", mode)
            out = dev.zeros((k, m))
            dev.gram(mode, dV, dW, out, n, m, k)
            assert np.array_equal(out.download().T, ref_vw), ("V^T W", mode)
        y = dev.array(np.full((k, n), np.nan), dtype=dt)  # beta = 0 never reads Y
        dev.block_update(dV, dC, y, n, m, k, alpha=1.0, beta=0.0)
        assert np.array_equal(y.download().astype(np.float64), ref_y), ("V C", dt)
        dev.block_update(dV, dC, y, n, m, k, alpha=1.0, beta=1.0)
        assert np.array_equal(y.download().astype(np.float64), 2.0 * ref_y), ("V C + Y", dt)
        r = dev.array(np.full((k, n), np.nan), dtype=dt)
        rn = dev.zeros((k,))
        dev.ritz_residual(dV, dW, dC, dth, r, rn, n, m, k)
        got = r.download().astype(np.float64)
        assert np.array_equal(got, ref_r), ("ritz residual", dt)
        want = np.sum(ref_r ** 2, axis=1)
        assert np.max(np.abs(rn.download() - want)) <= 1e-12 * np.max(want), ("residual norms", dt)
        print(f"[full size] Davidson products {np.dtype(dt).name}: exact", flush=True)
        del dV, dW, y, r
