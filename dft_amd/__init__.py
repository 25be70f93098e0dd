"""dft_amd -- MI355X-native 3D real-space Kohn-Sham hot path.

The product is native: libdft_hip.so (gfx950 kernels behind the C-ABI in
include/dft_hip.h) and libdft_core.so (the C++ layer keeping the reference's
solver interfaces, include/dft_molecule.h and include/dft_core.h).  This
module loads them through ctypes for tests and benchmarks; it raises if the
native libraries are missing -- there is no Python or CPU fallback.

Also here: the file formats of include/dft_io.h (save_field / load_field,
save_patch_shard / load_patch_shard, scf_result_json / save_scf_result);
dft_amd.run (an SCF from a run description or a result document);
dft_amd.cnn_pipeline (harvest, training and evaluation of the learned
preconditioner on the seeded molecule set); dft_amd.geometry (the
workloads' molecules).
"""
import ctypes as C
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")
HIP_SO = os.path.join(LIBDIR, "libdft_hip.so")
CORE_SO = os.path.join(LIBDIR, "libdft_core.so")

F64, F32 = 0, 1
GRAM_F64, GRAM_F32, GRAM_SPLIT_BF16, GRAM_SPLIT6_BF16 = 0, 1, 2, 3
PZ81, PBE96 = 1, 2
RED_SUM, RED_DOT, RED_MAXABS, RED_DIFF2 = 0, 1, 2, 3


class NativeMissing(RuntimeError):
    pass


class DftError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"status {code}: {msg}")
        self.code = code


class Grid(C.Structure):
    """dft_grid: CartesianGrid as plain data (CartesianGrid.cpp:5-48)."""

    _fields_ = [("nx", C.c_int64), ("ny", C.c_int64), ("nz", C.c_int64)] + [
        (k, C.c_double) for k in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax", "dx", "dy", "dz")
    ]

    @classmethod
    def make(cls, nx, ny, nz, xmin, xmax, ymin, ymax, zmin, zmax):
        g = cls()
        rc = hip_lib().dft_grid_init(C.byref(g), nx, ny, nz, *[C.c_double(v) for v in (xmin, xmax, ymin, ymax, zmin, zmax)])
        if rc != 0:
            raise DftError(rc, "invalid grid")
        return g

    @classmethod
    def centered(cls, n, h):
        half = (n - 1) * h / 2.0
        return cls.make(n, n, n, -half, half, -half, half, -half, half)

    @property
    def size(self):
        return self.nx * self.ny * self.nz

    @property
    def bounds(self):
        return (self.xmin, self.xmax, self.ymin, self.ymax, self.zmin, self.zmax)


_hip = None
_core = None


def _load(path):
    if not os.path.exists(path):
        raise NativeMissing(f"{path} is not built; run __graft_entry__.build() (dft_amd/build.py)")
    return C.CDLL(path)


def hip_lib():
    global _hip
    if _hip is None:
        L = _load(HIP_SO)
        L.dft_last_error.restype = C.c_char_p
        L.dft_stream.restype = C.c_void_p
        _hip = L
    return _hip


def core_lib():
    global _core
    if _core is None:
        hip_lib()
        L = _load(CORE_SO)
        _core = L
    return _core


_vp = C.c_void_p
_dp = C.POINTER(C.c_double)


def _prec_of(dtype):
    dt = np.dtype(dtype)
    if dt == np.float64:
        return F64
    if dt == np.float32:
        return F32
    raise ValueError(f"unsupported dtype {dt}")


class DeviceArray:
    """A device allocation owned by a Device, with shape/dtype bookkeeping."""

    def __init__(self, dev, shape, dtype):
        self.dev = dev
        self.shape = tuple(int(s) for s in np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        dev._chk(dev.L.dft_alloc(dev.ctx, C.c_size_t(max(self.nbytes, 1)), C.byref(p)))
        self.ptr = p.value
        dev._live.add(self)

    @property
    def prec(self):
        return _prec_of(self.dtype)

    def col(self, j, ld=None):
        """Device address of column j of a column-major block stored as (ncols, ld)."""
        ld = ld if ld is not None else self.shape[-1]
        return self.ptr + j * ld * self.dtype.itemsize

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes <= self.nbytes
        self.dev._chk(self.dev.L.dft_upload(self.dev.ctx, _vp(self.ptr), a.ctypes.data_as(_vp), C.c_size_t(a.nbytes)))
        return self

    def download(self):
        out = np.empty(self.shape, dtype=self.dtype)
        self.dev._chk(self.dev.L.dft_download(self.dev.ctx, out.ctypes.data_as(_vp), _vp(self.ptr), C.c_size_t(self.nbytes)))
        return out

    def free(self):
        if self.ptr and self.dev.ctx:
            self.dev.L.dft_free(self.dev.ctx, _vp(self.ptr))
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Device:
    """One dft_ctx (a HIP stream and its workspaces) on one GPU."""

    def __init__(self, device=0):
        self.L = hip_lib()
        ctx = C.c_void_p()
        rc = self.L.dft_ctx_create(C.c_int(device), C.byref(ctx))
        if rc != 0:
            raise DftError(rc, "dft_ctx_create failed (no GPU?)")
        self.ctx = ctx
        self._live = weakref.WeakSet()

    def close(self):
        """Free every array still alive on this context, then destroy it."""
        if self.ctx:
            for a in list(self._live):
                a.free()
            self.L.dft_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != 0:
            raise DftError(rc, self.L.dft_last_error(self.ctx).decode())

    def array(self, host=None, shape=None, dtype=np.float64):
        if host is not None:
            host = np.asarray(host)
            a = DeviceArray(self, host.shape, dtype if dtype is not None else host.dtype)
            return a.upload(host)
        return DeviceArray(self, shape, dtype)

    def zeros(self, shape, dtype=np.float64):
        a = DeviceArray(self, shape, dtype)
        self._chk(self.L.dft_zero(self.ctx, _vp(a.ptr), C.c_size_t(a.nbytes)))
        return a

    def sync(self):
        self._chk(self.L.dft_sync(self.ctx))

    def clamp_min(self, y, lo=0.0, n=None):
        """y = max(y, lo) elementwise (NaN -> lo)."""
        self._chk(self.L.dft_clamp_min(self.ctx, C.c_int(y.prec), C.c_int64(n or int(np.prod(y.shape))), C.c_double(lo), _vp(y.ptr)))

    def copy(self, dst, src, nbytes=None):
        """Device-to-device copy on the context stream."""
        nb = src.nbytes if nbytes is None else int(nbytes)
        self._chk(self.L.dft_copy(self.ctx, _vp(dst.ptr), _vp(src.ptr), C.c_size_t(nb)))

    # ---- kernels -------------------------------------------------------
    def hpsi(self, g, veff, psi, out, nb=None, ld=None, ldo=None):
        n = g.size
        nb = nb if nb is not None else (psi.shape[0] if len(psi.shape) > 1 else 1)
        self._chk(self.L.dft_hpsi(self.ctx, C.byref(g), C.c_int(psi.prec), _vp(veff.ptr), _vp(psi.ptr),
                                  C.c_int64(ld or n), C.c_int(nb), _vp(out.ptr), C.c_int64(ldo or n)))

    def kinetic_energy(self, g, psi, electrons, nb=None):
        nb = nb if nb is not None else (psi.shape[0] if len(psi.shape) > 1 else 1)
        el = np.ascontiguousarray(electrons, dtype=np.int32)
        t = C.c_double()
        self._chk(self.L.dft_kinetic_energy(self.ctx, C.byref(g), C.c_int(psi.prec), _vp(psi.ptr), C.c_int64(g.size),
                                            C.c_int(nb), el.ctypes.data_as(C.POINTER(C.c_int)), C.byref(t)))
        return t.value

    def poisson(self, g, rho, v, max_sweeps=10000, tol=1e-8, zero_start=True, batch=64):
        sw = C.c_int64()
        md = C.c_double()
        self._chk(self.L.dft_poisson(self.ctx, C.byref(g), C.c_int(rho.prec), _vp(rho.ptr), _vp(v.ptr),
                                     C.c_int64(max_sweeps), C.c_double(tol), C.c_int(1 if zero_start else 0),
                                     C.c_int(batch), C.byref(sw), C.byref(md)))
        return sw.value, md.value

    def poisson_chebyshev(self, g, rho, v, max_sweeps=100000, tol=1e-8, zero_start=True, batch=64):
        sw, md = C.c_int64(), C.c_double()
        self._chk(self.L.dft_poisson_chebyshev(self.ctx, C.byref(g), C.c_int(rho.prec), _vp(rho.ptr), _vp(v.ptr),
                                               C.c_int64(max_sweeps), C.c_double(tol), C.c_int(1 if zero_start else 0),
                                               C.c_int(batch), C.byref(sw), C.byref(md)))
        return sw.value, md.value

    def poisson_direct(self, g, rho, v):
        """V_H at the Jacobi fixed point by sine transforms (boundary values set by the reference's rule)."""
        self._chk(self.L.dft_poisson_direct(self.ctx, C.byref(g), C.c_int(rho.prec), _vp(rho.ptr), _vp(v.ptr)))

    def hartree_energy(self, g, rho, vh):
        e = C.c_double()
        self._chk(self.L.dft_hartree_energy(self.ctx, C.byref(g), C.c_int(rho.prec), _vp(rho.ptr), _vp(vh.ptr), C.byref(e)))
        return e.value

    def xc(self, g, functional, na, nb, va, vb=None, energy=True):
        e = C.c_double()
        self._chk(self.L.dft_xc(self.ctx, C.byref(g), C.c_int(na.prec), C.c_int(functional), _vp(na.ptr), _vp(nb.ptr),
                                _vp(va.ptr), _vp(vb.ptr if vb is not None else None),
                                C.byref(e) if energy else None))
        return e.value if energy else None

    def xc_points(self, functional, pts):
        pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 4)
        din = self.array(pts)
        dout = self.array(shape=(pts.shape[0], 5))
        self._chk(self.L.dft_xc_points(self.ctx, C.c_int(functional), _vp(din.ptr), C.c_int64(pts.shape[0]), _vp(dout.ptr)))
        return dout.download()

    def density(self, psi, occ, out, accumulate=False, nb=None, n=None):
        nb = nb if nb is not None else psi.shape[0]
        n = n if n is not None else psi.shape[-1]
        occ = np.ascontiguousarray(occ, dtype=np.float64)
        self._chk(self.L.dft_density(self.ctx, C.c_int(psi.prec), _vp(psi.ptr), C.c_int64(n), C.c_int(nb),
                                     occ.ctypes.data_as(_dp), C.c_int64(n), _vp(out.ptr), C.c_int(1 if accumulate else 0)))

    def reduce(self, op, a, b=None, n=None):
        r = C.c_double()
        n = n if n is not None else int(np.prod(a.shape))
        self._chk(self.L.dft_reduce(self.ctx, C.c_int(op), C.c_int(a.prec), _vp(a.ptr), _vp(b.ptr if b else None),
                                    C.c_int64(n), C.byref(r)))
        return r.value

    def nuclear_potential(self, g, nuclei, out):
        nuc = np.ascontiguousarray(nuclei, dtype=np.float64).reshape(-1, 4)
        self._chk(self.L.dft_nuclear_potential(self.ctx, C.byref(g), C.c_int(out.prec), C.c_int(nuc.shape[0]),
                                               nuc.ctypes.data_as(_dp), _vp(out.ptr)))

    RADIAL_FIELDS = ("kinetic", "external", "hartree", "exchangeCorrelation", "total", "densityDifference",
                     "energyDifference", "maxKSResidual", "iterations", "converged", "status", "mixing")

    def radial_scf(self, Z, functional, points=2000, rmax=30.0, densities=False):
        """solveAtom (AtomicDFT.cpp:6-41) for every (Z[j], functional[j]) at once, one workgroup
        per atom.  Returns a list of dicts with the energy terms, the stop quantities, the iteration
        count and the occupied orbitals' eigenvalues (alpha subshells, then beta), and with
        densities=True the output spin densities as an array (njobs, 2, points)."""
        Z = np.ascontiguousarray(np.atleast_1d(Z), dtype=np.int32)
        fn = np.ascontiguousarray(np.broadcast_to(np.atleast_1d(functional), Z.shape), dtype=np.int32)
        res = np.zeros((Z.size, 26))
        dens = np.zeros((Z.size, 2, points)) if densities else None
        rc = self.L.dft_radial_scf(self.ctx, C.c_int(Z.size), Z.ctypes.data_as(C.POINTER(C.c_int)),
                                   fn.ctypes.data_as(C.POINTER(C.c_int)), C.c_int64(points), C.c_double(rmax),
                                   res.ctypes.data_as(_dp), dens.ctypes.data_as(_dp) if densities else None)
        self._chk(rc)
        out = []
        for j in range(Z.size):
            d = {k: float(res[j, q]) for q, k in enumerate(self.RADIAL_FIELDS)}
            norb = radial_orbital_count(int(Z[j]))
            d["eigenvalues"] = [float(x) for x in res[j, 12:12 + norb]]
            out.append(d)
        return (out, dens) if densities else out

    def gram(self, mode, a, b, out, n, ma, mb, alpha=1.0, beta=0.0, lda=None, ldb=None, ldo=None):
        self._chk(self.L.dft_gram(self.ctx, C.c_int(mode), C.c_int64(n), C.c_int(ma), C.c_int(mb), C.c_double(alpha),
                                  _vp(a.ptr), C.c_int64(lda or n), _vp(b.ptr), C.c_int64(ldb or n), C.c_double(beta),
                                  _vp(out.ptr), C.c_int64(ldo or ma)))

    def block_update(self, a, cmat, y, n, m, k, alpha=1.0, beta=0.0, lda=None, ldc=None, ldy=None):
        self._chk(self.L.dft_block_update(self.ctx, C.c_int(a.prec), C.c_int64(n), C.c_int(m), C.c_int(k),
                                          C.c_double(alpha), _vp(a.ptr), C.c_int64(lda or n), _vp(cmat.ptr),
                                          C.c_int64(ldc or m), C.c_double(beta), _vp(y.ptr), C.c_int64(ldy or n)))

    def ritz_residual(self, v, w, q, theta, r, rn2, n, m, k, ldq=None):
        """r[:, j] = w q[:, j] - theta[j] v q[:, j] and rn2[j] = ||r[:, j]||^2, j < k."""
        self._chk(self.L.dft_ritz_residual(self.ctx, C.c_int(v.prec), C.c_int64(n), C.c_int(m), C.c_int(k), _vp(v.ptr),
                                           C.c_int64(n), _vp(w.ptr), C.c_int64(n), _vp(q.ptr), C.c_int64(ldq or m),
                                           _vp(theta.ptr), _vp(r.ptr), C.c_int64(n), _vp(rn2.ptr)))

    def syev(self, a, m):
        w = self.array(shape=(m,))
        self._chk(self.L.dft_syev(self.ctx, C.c_int(m), _vp(a.ptr), C.c_int64(m), _vp(w.ptr)))
        return w

    def syev_lowest(self, a, m, k, w=None, z=None):
        """k lowest eigenpairs of the symmetric m x m device matrix a: (w[:k], z[:, :k])."""
        w = w if w is not None else self.array(shape=(m,))
        z = z if z is not None else self.array(shape=(m, m))
        self._chk(self.L.dft_syev_lowest(self.ctx, C.c_int(m), C.c_int(k), _vp(a.ptr), C.c_int64(m), _vp(z.ptr),
                                         C.c_int64(m), _vp(w.ptr)))
        return w, z

    def chol_inv(self, a, m):
        self._chk(self.L.dft_chol_inv(self.ctx, C.c_int(m), _vp(a.ptr), C.c_int64(m)))

    def seed_vectors(self, g, out, first, count):
        self._chk(self.L.dft_seed_vectors(self.ctx, C.byref(g), C.c_int(out.prec), C.c_int(first), C.c_int(count),
                                          _vp(out.ptr), C.c_int64(g.size)))

    def precond_kinetic_cheb(self, g, r, t, nb, sigma=1.0, sweeps=8):
        tmp = DeviceArray(self, (nb, g.size), r.dtype)
        self._chk(self.L.dft_precond_kinetic_cheb(self.ctx, C.byref(g), C.c_int(r.prec), C.c_double(sigma), C.c_int(sweeps),
                                                  _vp(r.ptr), C.c_int64(g.size), C.c_int(nb), _vp(t.ptr),
                                                  C.c_int64(g.size), _vp(tmp.ptr)))
        tmp.free()

    def precond_kinetic_exact_shifted(self, g, r, t, nb, theta, sigma_min=0.5, scale=0.35):
        """t[:, j] = (-1/2 Lap + max(sigma_min, -scale theta[j]))^-1 r[:, j] (theta: device FP64)."""
        self._chk(self.L.dft_precond_kinetic_exact_shifted(self.ctx, C.byref(g), C.c_int(r.prec), C.c_double(sigma_min),
                                                           C.c_double(scale), _vp(theta.ptr), _vp(r.ptr),
                                                           C.c_int64(g.size), C.c_int(nb), _vp(t.ptr), C.c_int64(g.size)))

    def precond_kinetic_exact(self, g, r, t, nb, sigma=1.5):
        self._chk(self.L.dft_precond_kinetic_exact(self.ctx, C.byref(g), C.c_int(r.prec), C.c_double(sigma), _vp(r.ptr),
                                                   C.c_int64(g.size), C.c_int(nb), _vp(t.ptr), C.c_int64(g.size)))

    def precond_kinetic(self, g, r, t, nb, sigma=1.0, omega=2.0 / 3.0, sweeps=4):
        tmp = DeviceArray(self, (nb, g.size), r.dtype)
        self._chk(self.L.dft_precond_kinetic(self.ctx, C.byref(g), C.c_int(r.prec), C.c_double(sigma), C.c_double(omega),
                                             C.c_int(sweeps), _vp(r.ptr), C.c_int64(g.size), C.c_int(nb), _vp(t.ptr),
                                             C.c_int64(g.size), _vp(tmp.ptr)))
        tmp.free()

    def precond_diag(self, g, veff, r, eps, t, nb):
        self._chk(self.L.dft_precond_diag(self.ctx, C.byref(g), C.c_int(r.prec), _vp(veff.ptr), _vp(r.ptr),
                                          C.c_int64(g.size), C.c_int(nb), _vp(eps.ptr), _vp(t.ptr), C.c_int64(g.size)))

    # ---- learned preconditioner kernels (cnn.hip, harvest.hip) ----------
    def cnn_act_count(self, dims, nbatch, training, channels=16):
        n = C.c_int64()
        self._chk(self.L.dft_cnn_act_count(C.c_int(channels), C.c_int64(dims[0]), C.c_int64(dims[1]),
                                           C.c_int64(dims[2]), C.c_int(nbatch), C.c_int(training), C.byref(n)))
        return n.value

    def cnn_scratch_count(self, dims, nbatch, channels=16):
        n = C.c_int64()
        self._chk(self.L.dft_cnn_scratch_count(C.c_int(channels), C.c_int64(dims[0]), C.c_int64(dims[1]),
                                               C.c_int64(dims[2]), C.c_int(nbatch), C.byref(n)))
        return n.value

    def cnn_to_bf16(self, w32, wbf):
        self._chk(self.L.dft_cnn_to_bf16(self.ctx, C.c_int64(int(np.prod(w32.shape))), _vp(w32.ptr), _vp(wbf.ptr)))

    def cnn_forward(self, dims, nbatch, wbf, xin, out, act, training, channels=16):
        """dims = (nx, ny, nz); xin [nbatch][2][N], out [nbatch][N] (FP32 device arrays)."""
        self._chk(self.L.dft_cnn_forward(self.ctx, C.c_int(channels), C.c_int64(dims[0]), C.c_int64(dims[1]),
                                         C.c_int64(dims[2]), C.c_int(nbatch), _vp(wbf.ptr), _vp(xin.ptr), _vp(out.ptr),
                                         _vp(act.ptr), C.c_int(training)))

    def cnn_backward(self, dims, nbatch, wbf, xin, act, out, target, grad, scratch, channels=16):
        loss = C.c_double()
        self._chk(self.L.dft_cnn_backward(self.ctx, C.c_int(channels), C.c_int64(dims[0]), C.c_int64(dims[1]),
                                          C.c_int64(dims[2]), C.c_int(nbatch), _vp(wbf.ptr), _vp(xin.ptr), _vp(act.ptr),
                                          _vp(out.ptr), _vp(target.ptr), _vp(grad.ptr), _vp(scratch.ptr), C.byref(loss)))
        return loss.value

    def cnn_adam(self, w32, wbf, grad, m, v, step, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self._chk(self.L.dft_cnn_adam(self.ctx, C.c_int64(int(np.prod(w32.shape))), _vp(w32.ptr), _vp(wbf.ptr),
                                      _vp(grad.ptr), _vp(m.ptr), _vp(v.ptr), C.c_int(step), C.c_float(lr),
                                      C.c_float(beta1), C.c_float(beta2), C.c_float(eps)))

    def column_maxabs(self, x, n, ncols, out):
        self._chk(self.L.dft_column_maxabs(self.ctx, C.c_int(x.prec), C.c_int64(n), C.c_int(ncols), _vp(x.ptr),
                                           C.c_int64(n), _vp(out.ptr)))

    def cnn_prepare(self, g, veff, r, ncols, theta, scale, xin):
        self._chk(self.L.dft_cnn_prepare(self.ctx, C.byref(g), C.c_int(r.prec), _vp(veff.ptr), _vp(r.ptr),
                                         C.c_int64(g.size), C.c_int(ncols), _vp(theta.ptr), _vp(scale.ptr), _vp(xin.ptr)))

    def cnn_prepare_box(self, g, veff, r, ncols, theta, scale, box, xin):
        b = np.ascontiguousarray(box, dtype=np.int64)
        self._chk(self.L.dft_cnn_prepare_box(self.ctx, C.byref(g), C.c_int(r.prec), _vp(veff.ptr), _vp(r.ptr),
                                             C.c_int64(g.size), C.c_int(ncols), _vp(theta.ptr), _vp(scale.ptr),
                                             b.ctypes.data_as(C.POINTER(C.c_int64)), _vp(xin.ptr)))

    def cnn_add_box(self, g, ncols, out, scale, box, taper, t):
        b = np.ascontiguousarray(box, dtype=np.int64)
        self._chk(self.L.dft_cnn_add_box(self.ctx, C.byref(g), C.c_int(t.prec), C.c_int(ncols), _vp(out.ptr),
                                         _vp(scale.ptr), b.ctypes.data_as(C.POINTER(C.c_int64)), C.c_int(taper),
                                         _vp(t.ptr), C.c_int64(g.size)))

    def cnn_finish(self, g, ncols, out, scale, t):
        self._chk(self.L.dft_cnn_finish(self.ctx, C.byref(g), C.c_int(t.prec), C.c_int(ncols), _vp(out.ptr),
                                        _vp(scale.ptr), _vp(t.ptr), C.c_int64(g.size)))

    def column_dots(self, x, y, n, ncols, out):
        self._chk(self.L.dft_column_dots(self.ctx, C.c_int(x.prec), C.c_int64(n), C.c_int(ncols), _vp(x.ptr),
                                         C.c_int64(n), _vp(y.ptr), C.c_int64(n), _vp(out.ptr)))

    def cols_axpby(self, a, x, b, y, n, ncols):
        self._chk(self.L.dft_cols_axpby(self.ctx, C.c_int(x.prec), C.c_int64(n), C.c_int(ncols), _vp(a.ptr),
                                        _vp(x.ptr), C.c_int64(n), _vp(b.ptr), _vp(y.ptr), C.c_int64(n)))

    def patch_energy(self, g, patch, r, ncols, out):
        self._chk(self.L.dft_patch_energy(self.ctx, C.byref(g), C.c_int(r.prec), C.c_int(patch), _vp(r.ptr),
                                          C.c_int64(g.size), C.c_int(ncols), _vp(out.ptr)))

    def extract_patches(self, g, patch, desc, ncols, veff, r, cmat, theta, scale, cscale, xin, tgt):
        d = np.ascontiguousarray(desc, dtype=np.int32).reshape(-1, 4)
        self._chk(self.L.dft_extract_patches(self.ctx, C.byref(g), C.c_int(r.prec), C.c_int(patch), C.c_int(d.shape[0]),
                                             d.ctypes.data_as(C.POINTER(C.c_int)), C.c_int(ncols), _vp(veff.ptr),
                                             _vp(r.ptr), C.c_int64(g.size), _vp(cmat.ptr), C.c_int64(g.size),
                                             _vp(theta.ptr), _vp(scale.ptr), _vp(cscale.ptr), _vp(xin.ptr),
                                             _vp(tgt.ptr)))

    # ---- timing on the context stream ---------------------------------
    def event(self):
        e = C.c_void_p()
        self._chk(self.L.dft_event_create(self.ctx, C.byref(e)))
        return e

    def record(self, ev):
        self._chk(self.L.dft_event_record(self.ctx, ev))

    def elapsed_ms(self, a, b):
        ms = C.c_float()
        self._chk(self.L.dft_event_elapsed_ms(self.ctx, a, b, C.byref(ms)))
        return ms.value


# ---------------------------------------------------------------- core layer
SCF_FP64, SCF_MIXED = 0, 1
STOP_REFERENCE, STOP_DE = 0, 1
INIT_BARE, INIT_HYDROGENIC = 0, 1
PRECOND_DIAG, PRECOND_CNN, PRECOND_KINETIC, PRECOND_KINETIC_CHEB, PRECOND_KINETIC_EXACT = 0, 1, 2, 3, 4
POISSON_JACOBI, POISSON_CHEBYSHEV, POISSON_DIRECT = 0, 1, 2
MIX_LINEAR, MIX_PULAY = 0, 1
XC_FLUX_REFERENCE, XC_FLUX_CONSISTENT = 0, 1
XC_FLAG_FLUX_CONSISTENT = 0x100  # OR-ed into the functional id of Device.xc


class ScfConfig(C.Structure):
    """dft_scf_config (include/dft_core.h)."""

    _fields_ = [("functional", C.c_int), ("charge", C.c_int), ("multiplicity", C.c_int), ("precision", C.c_int),
                ("nstates", C.c_int), ("max_subspace", C.c_int), ("max_iterations", C.c_int),
                ("poisson_fixed", C.c_int), ("poisson_batch", C.c_int), ("poisson_method", C.c_int),
                ("poisson_tol", C.c_double), ("stop_rule", C.c_int),
                ("energy_tol", C.c_double), ("init_guess", C.c_int), ("davidson_tol", C.c_double),
                ("davidson_max_iter", C.c_int), ("refine_iterations", C.c_int), ("precond", C.c_int),
                ("cnn", C.c_void_p), ("verbose", C.c_int), ("precond_sweeps", C.c_int),
                ("precond_shift", C.c_double), ("mixer", C.c_int), ("mix_depth", C.c_int), ("mix_beta", C.c_double),
                ("smearing", C.c_double), ("ortho_gram", C.c_int), ("mix_kerker", C.c_double),
                ("precond_shift_scale", C.c_double), ("xc_flux", C.c_int)]


class ScfReport(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("first_de_iteration", C.c_int),
                ("refine_iterations", C.c_int), ("energy", C.c_double * 6), ("energy_difference", C.c_double),
                ("density_difference", C.c_double), ("ks_residual", C.c_double), ("seconds_total", C.c_double),
                ("seconds_to_first_de", C.c_double), ("seconds_poisson", C.c_double), ("seconds_xc", C.c_double),
                ("seconds_davidson", C.c_double), ("seconds_other", C.c_double), ("happly_columns", C.c_int64),
                ("poisson_sweeps", C.c_int64), ("davidson_iterations", C.c_int64), ("cnn_applied", C.c_int64),
                ("cnn_fallbacks", C.c_int64), ("block_passes", C.c_int64), ("block_fallbacks", C.c_int64),
                ("fermi_level", C.c_double * 2), ("smearing_energy", C.c_double)]


def _core_err(rc):
    L = core_lib()
    L.dft_core_last_error.restype = C.c_char_p
    raise DftError(rc, L.dft_core_last_error().decode())


def radial_config(Z):
    """getAtomicConfiguration (ElectronicConfiguration.cpp:50-145): the subshells of Z as
    (n, l, alpha electrons, beta electrons) tuples."""
    L = hip_lib()
    a = [(C.c_int * 7)() for _ in range(4)]
    k = L.dft_radial_config(C.c_int(Z), *a)
    if k < 0:
        raise ValueError(f"no configuration for Z = {Z}")
    return [(a[0][i], a[1][i], a[2][i], a[3][i]) for i in range(k)]


def radial_orbital_count(Z):
    """Occupied radial orbitals of Z: one per subshell and spin with electrons."""
    return sum((al > 0) + (be > 0) for _, _, al, be in radial_config(Z))


def scf_run(grid, nuclei, functional=PZ81, charge=0, multiplicity=1, precision=SCF_FP64, nstates=0, max_subspace=0,
            max_iterations=250, poisson_fixed=0, poisson_batch=64, poisson_method=POISSON_JACOBI, poisson_tol=0.0,
            stop_rule=STOP_REFERENCE, energy_tol=1e-5,
            init_guess=INIT_BARE, davidson_tol=1e-7, davidson_max_iter=0, refine_iterations=0,
            precond=PRECOND_DIAG, cnn=None, verbose=False, want_density=False, precond_sweeps=0, precond_shift=0.0,
            mixer=MIX_LINEAR, mix_depth=0, mix_beta=0.0, smearing=0.0, ortho_gram=0, mix_kerker=0.0,
            precond_shift_scale=0.0, xc_flux=XC_FLUX_REFERENCE):
    """Run the device-resident SCF; returns a dict of results.  smearing: Fermi-Dirac kT in Ha (0: whole
    orbitals, the reference's rule); with smearing the total is the free energy E - TS.  xc_flux:
    XC_FLUX_REFERENCE (the reference's GGA potential) or XC_FLUX_CONSISTENT (the derivative of E_xc)."""
    L = core_lib()
    nuc = np.ascontiguousarray(nuclei, dtype=np.float64).reshape(-1, 4)
    cfg = ScfConfig(functional, charge, multiplicity, precision, nstates, max_subspace, max_iterations,
                    poisson_fixed, poisson_batch, poisson_method, poisson_tol, stop_rule, energy_tol, init_guess, davidson_tol,
                    davidson_max_iter, refine_iterations, precond, _model_ptr(cnn), 1 if verbose else 0,
                    precond_sweeps, precond_shift, mixer, mix_depth, mix_beta, smearing, ortho_gram, mix_kerker,
                    precond_shift_scale, xc_flux)
    rep = ScfReport()
    ne = int(round(nuc[:, 0].sum())) - charge
    eig = np.zeros(ne + 2)
    cap = max_iterations + max(refine_iterations, 0) + 1
    trace = np.zeros((cap, 5))
    rho = np.zeros(2 * grid.size) if want_density else None
    rc = L.dft_scf_run(C.byref(grid), C.c_int(nuc.shape[0]), nuc.ctypes.data_as(_dp), C.byref(cfg), C.byref(rep),
                       eig.ctypes.data_as(_dp), trace.ctypes.data_as(_dp), C.c_int(cap),
                       None if rho is None else rho.ctypes.data_as(_dp))
    if rc != 0:
        _core_err(rc)
    it = rep.iterations
    out = {k: getattr(rep, k) for k, _ in ScfReport._fields_ if k not in ("energy", "fermi_level")}
    out["fermi_level"] = list(rep.fermi_level)
    out["energy"] = list(rep.energy)
    out["total"] = rep.energy[5]
    out["converged"] = bool(rep.converged)
    out["trace"] = trace[:it].copy()
    na = (ne + multiplicity - 1) // 2
    nb = ne - na
    out["eigenvalues"] = eig[: na + nb].copy()
    out["n_alpha"], out["n_beta"] = na, nb
    out["config"], out["report"] = cfg, rep  # for scf_result_json
    if rho is not None:
        out["rho_alpha"], out["rho_beta"] = rho[: grid.size], rho[grid.size:]
    return out


def fermi_occupations(eps, nelec, kT):
    """Fermi-Dirac occupations (each in [0, 1]) of one spin channel summing to nelec; returns (occ, mu, entropy)."""
    L = core_lib()
    e = np.ascontiguousarray(eps, dtype=np.float64)
    occ = np.zeros(e.size)
    mu, S = C.c_double(), C.c_double()
    rc = L.dft_fermi_occupations(e.ctypes.data_as(_dp), C.c_int(e.size), C.c_double(nelec), C.c_double(kT),
                                 occ.ctypes.data_as(_dp), C.byref(mu), C.byref(S))
    if rc != 0:
        _core_err(rc)
    return occ, mu.value, S.value


def davidson_run(grid, veff, nstates, nconv=None, prec=F64, gram_mode=GRAM_F64, max_subspace=0, tol=1e-7,
                 max_iter=0, want_psi=False, cnn=None, precond=None):
    """Block Davidson on a fixed potential; cnn: a CnnModel for the learned correction."""
    L = core_lib()
    v = np.ascontiguousarray(veff, dtype=np.float64)
    nconv = nconv or nstates
    eig = np.zeros(nstates)
    res = np.zeros(nstates)
    its = C.c_int()
    stats = np.zeros(4, dtype=np.int64)
    psi = np.zeros((nconv, grid.size)) if want_psi else None
    rc = L.dft_davidson_run(C.byref(grid), C.c_int(prec), C.c_int(gram_mode), v.ctypes.data_as(_dp), C.c_int(nstates),
                            C.c_int(nconv), C.c_int(max_subspace), C.c_double(tol), C.c_int(max_iter),
                            C.c_int(precond if precond is not None else (PRECOND_CNN if cnn is not None else PRECOND_DIAG)),
                            _model_ptr(cnn), eig.ctypes.data_as(_dp), None if psi is None else psi.ctypes.data_as(_dp),
                            res.ctypes.data_as(_dp), C.byref(its), stats.ctypes.data_as(C.POINTER(C.c_int64)))
    if rc != 0:
        _core_err(rc)
    return dict(eigenvalues=eig, residuals=res, iterations=its.value, psi=psi, happly_columns=int(stats[0]),
                restarts=int(stats[1]), cnn_applied=int(stats[2]), cnn_fallbacks=int(stats[3]))


# ---------------------------------------------------------------- learned preconditioner
CNN_CHANNELS = 16


def cnn_param_count(channels=CNN_CHANNELS):
    n = C.c_int64()
    if hip_lib().dft_cnn_param_count(C.c_int(channels), C.byref(n)) != 0:
        raise ValueError("bad channel count")
    return n.value


class _CnnModelC(C.Structure):
    _fields_ = [("channels", C.c_int), ("count", C.c_int64), ("params", C.POINTER(C.c_float)), ("base", C.c_int)]


CNN_BASE_NONE, CNN_BASE_KINETIC = 0, 1


class CnnModel:
    """Trained network parameters (FP32 master copy, kernel layout); save/load as .npz.  base:
    CNN_BASE_NONE (the output is the correction) or CNN_BASE_KINETIC (the output is added to the
    exact kinetic correction over the molecule's box; trained on harvests with target_scale 2)."""

    def __init__(self, params, channels=CNN_CHANNELS, base=CNN_BASE_NONE):
        self.params = np.ascontiguousarray(params, dtype=np.float32)
        if self.params.size != cnn_param_count(channels):
            raise ValueError("parameter count does not match the network")
        if base not in (CNN_BASE_NONE, CNN_BASE_KINETIC):
            raise ValueError("base must be CNN_BASE_NONE or CNN_BASE_KINETIC")
        self.channels, self.base = channels, base
        self._c = _CnnModelC(channels, self.params.size, self.params.ctypes.data_as(C.POINTER(C.c_float)), base)

    def save(self, path):
        np.savez(path, params=self.params, channels=np.int64(self.channels), base=np.int64(self.base))

    @classmethod
    def load(cls, path):
        with np.load(path, allow_pickle=False) as z:
            return cls(z["params"], int(z["channels"]), int(z["base"]) if "base" in z.files else CNN_BASE_NONE)


def _model_ptr(m):
    if m is None:
        return None
    if not isinstance(m, CnnModel):
        raise TypeError("cnn must be a CnnModel")
    return C.cast(C.pointer(m._c), C.c_void_p)


class HarvestConfig(C.Structure):
    """dft_cnn_harvest_config (include/dft_core.h)."""

    _fields_ = [("patch", C.c_int), ("max_columns", C.c_int), ("patches_per_column", C.c_int), ("every", C.c_int),
                ("cg_tol", C.c_double), ("cg_max_iter", C.c_int), ("seed", C.c_uint64), ("target_scale", C.c_int)]


class HarvestStats(C.Structure):
    _fields_ = [("solves", C.c_int64), ("failed", C.c_int64), ("cg_iterations", C.c_int64), ("patches", C.c_int64),
                ("max_true_residual", C.c_double)]


class CnnTrainer:
    """Device patch set + network + Adam state (dft_cnn_trainer)."""

    def __init__(self, patch=32, capacity=4096, seed=1, channels=CNN_CHANNELS):
        self.L = core_lib()
        self.L.dft_cnn_trainer_create.argtypes = [C.c_int, C.c_int, C.c_int64, C.c_uint64, C.POINTER(C.c_void_p)]
        self.L.dft_cnn_trainer_destroy.argtypes = [C.c_void_p]
        h = C.c_void_p()
        rc = self.L.dft_cnn_trainer_create(channels, patch, capacity, seed, C.byref(h))
        if rc != 0:
            _core_err(rc)
        self.h = h
        self.patch, self.channels = patch, channels
        self.count_params = cnn_param_count(channels)

    def close(self):
        if self.h:
            self.L.dft_cnn_trainer_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def harvest(self, grid, nuclei, functional=PZ81, nstates=0, max_iterations=30, poisson_fixed=0, energy_tol=1e-5,
                init_guess=INIT_HYDROGENIC, davidson_tol=1e-6, max_columns=8, patches_per_column=4, every=3,
                poisson_method=POISSON_JACOBI,
                cg_tol=1e-6, cg_max_iter=400, seed=7, multiplicity=1, target_scale=0):
        nuc = np.ascontiguousarray(nuclei, dtype=np.float64).reshape(-1, 4)
        cfg = ScfConfig(functional, 0, multiplicity, SCF_FP64, nstates, 0, max_iterations, poisson_fixed, 64,
                        poisson_method, 0.0, STOP_DE,
                        energy_tol, init_guess, davidson_tol, 0, 0, PRECOND_DIAG, None, 0)
        hc = HarvestConfig(self.patch, max_columns, patches_per_column, every, cg_tol, cg_max_iter, seed, target_scale)
        rep = ScfReport()
        added = C.c_int64()
        rc = self.L.dft_cnn_trainer_harvest(self.h, C.byref(grid), C.c_int(nuc.shape[0]), nuc.ctypes.data_as(_dp),
                                            C.byref(cfg), C.byref(hc), C.byref(rep), C.byref(added))
        if rc != 0:
            _core_err(rc)
        return added.value

    def count(self):
        n = C.c_int64()
        self.L.dft_cnn_trainer_count(self.h, C.byref(n))
        return n.value

    def harvest_stats(self):
        """Totals over this trainer's harvests (dft_cnn_trainer_harvest_stats)."""
        st = HarvestStats()
        rc = self.L.dft_cnn_trainer_harvest_stats(self.h, C.byref(st))
        if rc != 0:
            _core_err(rc)
        return {k: getattr(st, k) for k, _ in HarvestStats._fields_}

    def train(self, steps, batch=8, lr=1e-3, train_count=0):
        loss = np.zeros(max(steps, 1))
        rc = self.L.dft_cnn_trainer_train(self.h, C.c_int(steps), C.c_int(batch), C.c_float(lr), C.c_int64(train_count),
                                          loss.ctypes.data_as(_dp))
        if rc != 0:
            _core_err(rc)
        return loss[:steps]

    def eval(self, first, count):
        loss = C.c_double()
        rc = self.L.dft_cnn_trainer_eval(self.h, C.c_int64(first), C.c_int64(count), C.byref(loss))
        if rc != 0:
            _core_err(rc)
        return loss.value

    def params(self):
        p = np.zeros(self.count_params, dtype=np.float32)
        rc = self.L.dft_cnn_trainer_get_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float)), C.c_int64(p.size))
        if rc != 0:
            _core_err(rc)
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        rc = self.L.dft_cnn_trainer_set_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float)), C.c_int64(p.size))
        if rc != 0:
            _core_err(rc)

    def model(self, base=CNN_BASE_NONE):
        return CnnModel(self.params(), self.channels, base)

    def patches(self, first, count):
        P3 = self.patch ** 3
        xin = np.zeros((count, 2, P3), dtype=np.float32)
        tgt = np.zeros((count, P3), dtype=np.float32)
        rc = self.L.dft_cnn_trainer_get_patches(self.h, C.c_int64(first), C.c_int64(count),
                                                xin.ctypes.data_as(C.POINTER(C.c_float)),
                                                tgt.ctypes.data_as(C.POINTER(C.c_float)))
        if rc != 0:
            _core_err(rc)
        return xin, tgt

    def add_patches(self, xin, tgt):
        """Append training pairs (count x 2 x P^3 inputs, count x P^3 targets) after those held."""
        P3 = self.patch ** 3
        xin = np.ascontiguousarray(xin, dtype=np.float32).reshape(-1, 2, P3)
        tgt = np.ascontiguousarray(tgt, dtype=np.float32).reshape(-1, P3)
        if xin.shape[0] != tgt.shape[0]:
            raise ValueError("inputs and targets hold different patch counts")
        rc = self.L.dft_cnn_trainer_add_patches(self.h, C.c_int64(xin.shape[0]), xin.ctypes.data_as(C.POINTER(C.c_float)),
                                                tgt.ctypes.data_as(C.POINTER(C.c_float)))
        if rc != 0:
            _core_err(rc)

    def save_patches(self, path, first=0, count=None, target_scale=0, h=0.0):
        """Write patches [first, first + count) as a patch shard (include/dft_io.h)."""
        count = self.count() - first if count is None else count
        xin, tgt = self.patches(first, count)
        save_patch_shard(path, xin, tgt, target_scale=target_scale, h=h)

    def load_patches(self, path):
        """Append every patch of a shard; returns the shard's header."""
        xin, tgt, hdr = load_patch_shard(path)
        if hdr["patch"] != self.patch:
            raise ValueError(f"shard patch edge {hdr['patch']} differs from the trainer's {self.patch}")
        self.add_patches(xin, tgt)
        return hdr


class ScfFields(C.Structure):
    _fields_ = [(k, C.POINTER(C.c_double)) for k in ("vh", "vxc_alpha", "vxc_beta", "rho_alpha", "rho_beta", "vh_out")]


def scf_step(grid, nuclei, rho_alpha, rho_beta, functional=PZ81, multiplicity=1, precision=SCF_FP64, nstates=0,
             poisson_fixed=0, davidson_tol=1e-7):
    """One SCF iteration on given input densities; returns fields and scalars."""
    L = core_lib()
    nuc = np.ascontiguousarray(nuclei, dtype=np.float64).reshape(-1, 4)
    n = grid.size
    ra = np.ascontiguousarray(rho_alpha, dtype=np.float64)
    rb = np.ascontiguousarray(rho_beta, dtype=np.float64)
    cfg = ScfConfig(functional, 0, multiplicity, precision, nstates, 0, 1, poisson_fixed, 64, POISSON_JACOBI, 0.0,
                    STOP_DE, 0.0,
                    INIT_BARE, davidson_tol, 0, 0, PRECOND_DIAG, None, 0)
    rep = ScfReport()
    out = {k: np.zeros(n) for k in ("vh", "vxc_alpha", "vxc_beta", "rho_alpha", "rho_beta", "vh_out")}
    f = ScfFields(*[out[k].ctypes.data_as(_dp) for k in ("vh", "vxc_alpha", "vxc_beta", "rho_alpha", "rho_beta", "vh_out")])
    ne = int(round(nuc[:, 0].sum()))
    eig = np.zeros(ne + 2)
    rc = L.dft_scf_step(C.byref(grid), C.c_int(nuc.shape[0]), nuc.ctypes.data_as(_dp), C.byref(cfg),
                        ra.ctypes.data_as(_dp), rb.ctypes.data_as(_dp), C.byref(rep), eig.ctypes.data_as(_dp), C.byref(f))
    if rc != 0:
        _core_err(rc)
    out["energy"] = np.array(rep.energy)
    out["eigenvalues"] = eig[:ne]
    out["dRho"] = rep.density_difference
    out["ks"] = rep.ks_residual
    return out


# ---------------------------------------------------------------- on-disk formats (include/dft_io.h)
class FieldHeader(C.Structure):
    _fields_ = [("dtype", C.c_int), ("ncols", C.c_int64), ("grid", Grid), ("data_bytes", C.c_uint64),
                ("crc", C.c_uint32), ("name", C.c_char * 64)]


class PatchShardHeader(C.Structure):
    _fields_ = [("channels", C.c_int), ("target_scale", C.c_int), ("patch", C.c_int64), ("count", C.c_int64),
                ("h", C.c_double), ("crc_in", C.c_uint32), ("crc_tgt", C.c_uint32)]


def _path(p):
    return os.fsencode(os.fspath(p))


def crc32(data, crc=0):
    """CRC-32 (zlib's) of a bytes-like object or array, continuing from crc."""
    L = core_lib()
    L.dft_crc32.restype = C.c_uint32
    buf = np.ascontiguousarray(np.frombuffer(memoryview(data).cast("B"), dtype=np.uint8))
    return L.dft_crc32(C.c_uint32(crc), buf.ctypes.data_as(_vp), C.c_size_t(buf.size))


def _field_dict(h):
    return dict(dtype=np.float64 if h.dtype == F64 else np.float32, ncols=h.ncols,
                grid=Grid.from_buffer_copy(h.grid), name=h.name.decode(), crc=h.crc, data_bytes=h.data_bytes)


def save_field(path, grid, data, name=""):
    """Write a grid field (N values) or an orbital block ((ncols, N): one column per row, the device layout).
    data: a float64/float32 numpy array or a DeviceArray (downloaded first)."""
    if isinstance(data, DeviceArray):
        data = data.download()
    a = np.ascontiguousarray(data)
    prec = _prec_of(a.dtype)
    n = grid.size
    if a.size % n or a.size == 0:
        raise ValueError(f"{a.size} values are not whole columns of the {n}-point grid")
    rc = core_lib().dft_field_write(_path(path), C.byref(grid), C.c_int(prec), C.c_int64(a.size // n), name.encode(),
                                    a.ctypes.data_as(_vp))
    if rc != 0:
        _core_err(rc)


def field_info(path):
    """Header of a field file as a dict (dtype, ncols, grid, name, crc, data_bytes)."""
    h = FieldHeader()
    rc = core_lib().dft_field_read_header(_path(path), C.byref(h))
    if rc != 0:
        _core_err(rc)
    return _field_dict(h)


def load_field(path):
    """Read a field file (checksum verified); returns (grid, array, name): array is (N,) for one column,
    else (ncols, N)."""
    info = field_info(path)
    n = info["grid"].size
    shape = (n,) if info["ncols"] == 1 else (info["ncols"], n)
    out = np.empty(shape, dtype=info["dtype"])
    h = FieldHeader()
    rc = core_lib().dft_field_read(_path(path), C.byref(h), out.ctypes.data_as(_vp), C.c_uint64(out.nbytes))
    if rc != 0:
        _core_err(rc)
    info = _field_dict(h)
    return info["grid"], out, info["name"]


def save_patch_shard(path, xin, tgt, target_scale=0, h=0.0):
    """Write training pairs: xin (count, 2, P^3) and tgt (count, P^3), FP32."""
    xin = np.ascontiguousarray(xin, dtype=np.float32)
    tgt = np.ascontiguousarray(tgt, dtype=np.float32)
    if xin.ndim != 3 or tgt.ndim != 2:
        raise ValueError("inputs must be (count, 2, P^3) and targets (count, P^3)")
    count = tgt.shape[0]
    P = int(round(tgt.shape[1] ** (1.0 / 3.0)))
    if xin.shape != (count, 2, P ** 3) or tgt.shape != (count, P ** 3):
        raise ValueError("inputs must be (count, 2, P^3) and targets (count, P^3)")
    rc = core_lib().dft_patch_shard_write(_path(path), C.c_int64(P), C.c_int64(count), C.c_int(target_scale),
                                          C.c_double(h), xin.ctypes.data_as(C.POINTER(C.c_float)),
                                          tgt.ctypes.data_as(C.POINTER(C.c_float)))
    if rc != 0:
        _core_err(rc)


def patch_shard_info(path):
    h = PatchShardHeader()
    rc = core_lib().dft_patch_shard_read_header(_path(path), C.byref(h))
    if rc != 0:
        _core_err(rc)
    return {k: getattr(h, k) for k, _ in PatchShardHeader._fields_}


def load_patch_shard(path, first=0, count=None):
    """Read patches [first, first + count) of a shard; returns (xin, tgt, header)."""
    hdr = patch_shard_info(path)
    count = hdr["count"] - first if count is None else count
    P3 = hdr["patch"] ** 3
    xin = np.empty((count, 2, P3), dtype=np.float32)
    tgt = np.empty((count, P3), dtype=np.float32)
    rc = core_lib().dft_patch_shard_read(_path(path), C.c_int64(first), C.c_int64(count),
                                         xin.ctypes.data_as(C.POINTER(C.c_float)), tgt.ctypes.data_as(C.POINTER(C.c_float)))
    if rc != 0:
        _core_err(rc)
    return xin, tgt, hdr


def scf_result_json(grid, nuclei, result):
    """The result document (include/dft_io.h) of a scf_run result, as a string."""
    L = core_lib()
    nuc = np.ascontiguousarray(nuclei, dtype=np.float64).reshape(-1, 4)
    eig = np.ascontiguousarray(result["eigenvalues"], dtype=np.float64)
    na, nb = result["n_alpha"], result["n_beta"]
    if eig.size != na + nb:
        raise ValueError("eigenvalue count differs from n_alpha + n_beta")
    args = [C.byref(grid), C.c_int(nuc.shape[0]), nuc.ctypes.data_as(_dp), C.byref(result["config"]),
            C.byref(result["report"]), eig.ctypes.data_as(_dp), C.c_int(na), C.c_int(nb)]
    n = C.c_size_t()
    rc = L.dft_scf_result_json(*args, None, C.c_size_t(0), C.byref(n))
    if rc != 0:
        _core_err(rc)
    buf = C.create_string_buffer(n.value + 1)
    rc = L.dft_scf_result_json(*args, buf, C.c_size_t(n.value + 1), C.byref(n))
    if rc != 0:
        _core_err(rc)
    return buf.value.decode()


def save_scf_result(path, grid, nuclei, result, fields=True):
    """Write the result document to path; with fields, also the output spin densities (when the run
    kept them, want_density=True) as field files path + '.rho_alpha.dftf' / '.rho_beta.dftf'."""
    with open(path, "w") as f:
        f.write(scf_result_json(grid, nuclei, result))
    if fields:
        for k in ("rho_alpha", "rho_beta"):
            if result.get(k) is not None:
                save_field(f"{os.fspath(path)}.{k}.dftf", grid, result[k], name=k)
