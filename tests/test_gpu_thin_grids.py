"""The smallest and thinnest grids the reference's CartesianGrid allows (two
points per axis and up), against the FP64 oracle: H psi on blocks whose
orbital count exercises every group size (and the empty block, which writes
nothing), and the direct Poisson solve down to grids with one interior point
or none (boundary values only).  Bars as for the other grid kernels: FP64
relative 1e-12, FP32 2e-6 for H psi (inputs rounded to 2^-24)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(2, 2, 2), (3, 2, 2), (2, 2, 37), (130, 2, 3), (2, 70, 2), (65, 17, 2)]


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def dev():
    import dft_amd
    d = dft_amd.Device(0)
    yield d
    d.close()


def _grids(shape):
    import dft_amd
    from oracle.oracle import Grid
    og = Grid(*shape, -1.3, 1.1, -0.9, 1.4, -1.2, 0.8)
    return og, dft_amd.Grid.make(*og.n, *og.bounds)


@pytest.mark.parametrize("shape", SHAPES)
def test_hpsi_thin_grids(dev, oracle, shape):
    og, g = _grids(shape)
    print(f"[thin] H psi {shape}", flush=True)
    rng = np.random.default_rng(og.size)
    v = rng.standard_normal(og.size)
    for nb in (1, 3, 5, 6):
        psi = rng.standard_normal((nb, og.size))
        ref = np.stack([oracle.hpsi(og, v, psi[o]) for o in range(nb)])
        out = dev.array(shape=(nb, og.size))
        dev.hpsi(g, dev.array(v), dev.array(psi), out)
        assert rel(out.download(), ref) <= 1e-12, (shape, nb)
        o32 = dev.array(shape=(nb, og.size), dtype=np.float32)
        dev.hpsi(g, dev.array(v, dtype=np.float32), dev.array(psi, dtype=np.float32), o32)
        assert rel(o32.download(), ref) <= 2e-6, (shape, nb)
    # an empty block is a no-op
    sentinel = dev.array(np.full(og.size, 5.0))
    dev.hpsi(g, dev.array(v), dev.array(np.zeros(og.size)), sentinel, nb=0)
    assert np.all(sentinel.download() == 5.0)


@pytest.mark.parametrize("shape", [(2, 2, 2), (3, 3, 3), (3, 5, 4), (40, 3, 3), (2, 9, 9)])
def test_poisson_direct_thin_grids(dev, oracle, shape):
    og, g = _grids(shape)
    print(f"[thin] Poisson direct {shape}", flush=True)
    rng = np.random.default_rng(7)
    rho = np.abs(rng.standard_normal(og.size))
    ref, _, md = oracle.poisson(og, rho, v0=np.zeros(og.size), max_sweeps=400000, tol=1e-15)
    assert md < 1e-15
    v = dev.array(np.full(og.size, 9.0))  # every point is written
    dev.poisson_direct(g, dev.array(rho), v)
    assert rel(v.download(), ref) <= 1e-12, shape


@pytest.mark.parametrize("shape", [(2, 2, 2), (3, 2, 2), (2, 3, 9), (5, 4, 3)])
@pytest.mark.parametrize("fid", [1, 2])
def test_xc_thin_grids(dev, oracle, shape, fid):
    """One-sided derivatives at every face and the radius-2 divergence of the
    GGA flux on grids two and three points thick, open shell."""
    og, g = _grids(shape)
    print(f"[thin] XC {fid} {shape}", flush=True)
    rng = np.random.default_rng(og.size + fid)
    na = 0.2 + rng.random(og.size)
    nb = 0.1 + 0.5 * rng.random(og.size)
    ref_a = oracle.xc_potential(og, fid, na, nb, 0)
    ref_b = oracle.xc_potential(og, fid, na, nb, 1)
    ref_e = oracle.xc_energy(og, fid, na, nb)
    va, vb = dev.array(shape=(og.size,)), dev.array(shape=(og.size,))
    e = dev.xc(g, fid, dev.array(na), dev.array(nb), va, vb)
    assert rel(va.download(), ref_a) <= 1e-12 and rel(vb.download(), ref_b) <= 1e-12, shape
    assert abs(e - ref_e) <= 1e-12 * abs(ref_e)


@pytest.mark.parametrize("shape", [(2, 2, 2), (3, 3, 3), (3, 5, 4), (40, 3, 3), (2, 9, 9), (70, 3, 4)])
def test_poisson_reference_rule_thin_grids(dev, oracle, shape):
    """The reference's Jacobi (zero start, max|dV| < 1e-8, cap 10000) on thin
    grids: the same potential and the same sweep count as the oracle."""
    og, g = _grids(shape)
    print(f"[thin] Jacobi {shape}", flush=True)
    rho = np.abs(np.random.default_rng(3).standard_normal(og.size))
    ref, rsw, _ = oracle.poisson(og, rho)
    v = dev.array(np.full(og.size, 9.0))
    sw, _ = dev.poisson(g, dev.array(rho), v)
    assert sw == rsw, (shape, sw, rsw)
    assert rel(v.download(), ref) <= 1e-12, shape
