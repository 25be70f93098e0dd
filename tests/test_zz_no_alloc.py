"""dft_xc_noalloc (the allocation-free XC entry point) against dft_xc through the C-ABI: the same
kernels, so every potential value and the energy must be bit-identical, for PZ81 and PBE96 (both
fluxes), FP64 and FP32, on two calls in a row (the second on a smaller grid), and a grid whose launch
needs more partials than the fixed array holds is rejected on the host."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import dft_amd
    d = dft_amd.Device(0)
    yield d
    d.close()


def _density(n, h, shift):
    from oracle.oracle import Grid
    og = Grid.centered(n, h)
    xs, ys, zs = og.coords()
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    r = np.sqrt((X - shift) ** 2 + Y ** 2 + Z ** 2)
    return og, (1.8 * np.exp(-2.0 * r)).ravel(), (1.2 * np.exp(-1.6 * r)).ravel()


def _noalloc(dev, g, functional, na, nb, va, vb):
    e = C.c_double()
    rc = dev.L.dft_xc_noalloc(dev.ctx, C.byref(g), C.c_int(na.prec), C.c_int(functional), C.c_void_p(na.ptr),
                              C.c_void_p(nb.ptr), C.c_void_p(va.ptr), C.c_void_p(vb.ptr), C.byref(e))
    return rc, e.value


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("functional", ["pz81", "pbe96", "pbe96_consistent"])
def test_noalloc_matches_dft_xc_twice(dev, dtype, functional):
    import dft_amd
    fid = {"pz81": dft_amd.PZ81, "pbe96": dft_amd.PBE96,
           "pbe96_consistent": dft_amd.PBE96 | dft_amd.XC_FLAG_FLUX_CONSISTENT}[functional]
    for n, h, shift in ((72, 0.25, 0.3), (40, 0.3, -0.2)):  # second call: a smaller grid
        print(f"[noalloc] {functional} {np.dtype(dtype).name} n={n}", flush=True)
        og, a, b = _density(n, h, shift)
        g = dft_amd.Grid.make(*og.n, *og.bounds)
        na, nb = dev.array(a, dtype=dtype), dev.array(b, dtype=dtype)
        va, vb = dev.array(shape=(og.size,), dtype=dtype), dev.array(shape=(og.size,), dtype=dtype)
        wa, wb = dev.array(shape=(og.size,), dtype=dtype), dev.array(shape=(og.size,), dtype=dtype)
        e_ref = dev.xc(g, fid, na, nb, va, vb)
        rc, e = _noalloc(dev, g, fid, na, nb, wa, wb)
        assert rc == 0
        assert e == e_ref
        assert np.array_equal(wa.download(), va.download())
        assert np.array_equal(wb.download(), vb.download())


def test_noalloc_rejects_grids_beyond_its_array(dev):
    import dft_amd
    # 4096 x 4096 x 2 points: 128 x 512 FP64 GGA tiles = 65536 partials > 16384 (nothing is launched)
    g = dft_amd.Grid.make(4096, 4096, 2, -1.0, 1.0, -1.0, 1.0, -0.5, 0.5)
    na = dev.array(shape=(8,))
    rc, _ = _noalloc(dev, g, dft_amd.PBE96, na, na, na, na)
    assert rc == -1
    assert "beyond the fixed scratch" in dev.L.dft_last_error(dev.ctx).decode()
