"""ctypes access to the checker libraries (TEST INFRASTRUCTURE ONLY).

``Oracle`` wraps oracle/libks_oracle.so, the FP64 C restatement of the
reference's 3D Kohn-Sham path; ``Reference`` wraps oracle/_ref/libdftref.so,
the reference's own sources compiled where they lie (present only where the
reference checkout exists).  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py import this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libks_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libdftref.so")

_dp = C.POINTER(C.c_double)
_lp = C.POINTER(C.c_long)
_ip = C.POINTER(C.c_int)


def _d(a):
    return a.ctypes.data_as(_dp)


def build(with_ref=True):
    """Compile the restatement (and the reference shim when its sources exist)."""
    target = "all" if with_ref else "oracle"
    subprocess.run(["make", "-s", "-C", HERE, target], check=True)


class Grid:
    """Cubic-box grid descriptor mirroring CartesianGrid (CartesianGrid.cpp:5-48)."""

    def __init__(self, nx, ny, nz, xmin, xmax, ymin, ymax, zmin, zmax):
        self.n = (int(nx), int(ny), int(nz))
        self.bounds = (float(xmin), float(xmax), float(ymin), float(ymax), float(zmin), float(zmax))
        self.dx = (xmax - xmin) / (nx - 1)
        self.dy = (ymax - ymin) / (ny - 1)
        self.dz = (zmax - zmin) / (nz - 1)

    @classmethod
    def centered(cls, n, h):
        """n^3 points, spacing h, bounds +-(n-1)h/2 (box centre between points for even n)."""
        half = (n - 1) * h / 2.0
        return cls(n, n, n, -half, half, -half, half, -half, half)

    @property
    def size(self):
        return self.n[0] * self.n[1] * self.n[2]

    @property
    def shape(self):  # numpy (z, y, x) order, x fastest
        return (self.n[2], self.n[1], self.n[0])

    def coords(self):
        xs = self.bounds[0] + np.arange(self.n[0]) * self.dx
        ys = self.bounds[2] + np.arange(self.n[1]) * self.dy
        zs = self.bounds[4] + np.arange(self.n[2]) * self.dz
        return xs, ys, zs


class _OgGrid(C.Structure):
    _fields_ = [("nx", C.c_long), ("ny", C.c_long), ("nz", C.c_long)] + [
        (k, C.c_double) for k in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax", "dx", "dy", "dz")
    ]


class _OgScfConfig(C.Structure):
    _fields_ = [("functional", C.c_int), ("charge", C.c_int), ("multiplicity", C.c_int),
                ("max_iterations", C.c_int), ("poisson_fixed", C.c_int), ("stop_rule", C.c_int),
                ("energy_tol", C.c_double), ("init_guess", C.c_int), ("max_happly", C.c_long)]


class _OgScfResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("energy", C.c_double * 6),
                ("energy_difference", C.c_double), ("density_difference", C.c_double),
                ("ks_residual", C.c_double), ("eig_alpha0", C.c_double), ("eig_beta0", C.c_double),
                ("first_de_iteration", C.c_int)]


class OracleError(RuntimeError):
    pass


def _f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None and a.size != n:
        raise ValueError("size mismatch")
    return a


class Oracle:
    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            build(with_ref=False)
        L = C.CDLL(path)
        self.L = L
        L.og_last_error.restype = C.c_char_p
        L.og_hartree_energy.restype = C.c_double
        L.og_xc_energy.restype = C.c_double
        L.og_kinetic_energy.restype = C.c_double
        L.og_external_energy.restype = C.c_double

    def _chk(self, rc):
        if rc != 0:
            raise OracleError(f"oracle rc={rc}: {self.L.og_last_error().decode()}")

    def _g(self, g):
        og = _OgGrid()
        self._chk(self.L.og_grid_init(C.byref(og), C.c_long(g.n[0]), C.c_long(g.n[1]), C.c_long(g.n[2]),
                                      *[C.c_double(b) for b in g.bounds]))
        return og

    def hpsi(self, g, v, psi):
        out = np.empty(g.size)
        self._chk(self.L.og_hpsi(C.byref(self._g(g)), _d(_f64(v, g.size)), _d(_f64(psi, g.size)), _d(out)))
        return out

    def poisson(self, g, rho, v0=None, max_sweeps=10000, tol=1e-8):
        """Reference call when v0 is None; otherwise warm start from v0."""
        v = np.zeros(g.size) if v0 is None else _f64(v0, g.size).copy()
        sw = C.c_long()
        md = C.c_double()
        self._chk(self.L.og_poisson(C.byref(self._g(g)), _d(_f64(rho, g.size)), _d(v), C.c_long(max_sweeps),
                                    C.c_double(tol), C.c_int(1 if v0 is None else 0), C.byref(sw), C.byref(md)))
        return v, sw.value, md.value

    def hartree_energy(self, g, rho, vh):
        return self.L.og_hartree_energy(C.byref(self._g(g)), _d(_f64(rho, g.size)), _d(_f64(vh, g.size)))

    def xc_points(self, fid, inp):
        inp = _f64(inp).reshape(-1, 4)
        out = np.empty((inp.shape[0], 5))
        for p in range(inp.shape[0]):
            self._chk(self.L.og_xc_point(C.c_int(fid), _d(inp[p]), _d(out[p])))
        return out

    def xc_potential(self, g, fid, na, nb, spin, consistent=False):
        out = np.empty(g.size)
        self._chk(self.L.og_xc_potential_flux(C.byref(self._g(g)), C.c_int(fid), _d(_f64(na, g.size)),
                                              _d(_f64(nb, g.size)), C.c_int(spin), C.c_int(int(consistent)),
                                              _d(out)))
        return out

    def xc_energy_terms(self, g, fid, na, nb):
        """The per-point summands n eps dV of xc_energy."""
        out = np.empty(g.size)
        self._chk(self.L.og_xc_energy_terms(C.byref(self._g(g)), C.c_int(fid), _d(_f64(na, g.size)),
                                            _d(_f64(nb, g.size)), _d(out)))
        return out

    def xc_energy(self, g, fid, na, nb):
        return self.L.og_xc_energy(C.byref(self._g(g)), C.c_int(fid), _d(_f64(na, g.size)), _d(_f64(nb, g.size)))

    def spin_density(self, psi, electrons):
        psi = _f64(psi)
        norb, n = psi.shape
        out = np.empty(n)
        el = np.ascontiguousarray(electrons, dtype=np.int32)
        self._chk(self.L.og_spin_density(C.c_long(n), C.c_int(norb), _d(psi), el.ctypes.data_as(_ip), _d(out)))
        return out

    def electron_density(self, psi, electrons, spins):
        psi = _f64(psi)
        norb, n = psi.shape
        out = np.empty(n)
        el = np.ascontiguousarray(electrons, dtype=np.int32)
        sp = np.ascontiguousarray(spins, dtype=np.int32)
        self._chk(self.L.og_electron_density(C.c_long(n), C.c_int(norb), _d(psi), el.ctypes.data_as(_ip),
                                             sp.ctypes.data_as(_ip), _d(out)))
        return out

    def integrate_density(self, rho, dx, dy, dz):
        r = _f64(rho)
        e = C.c_double()
        self._chk(self.L.og_integrate_density(C.c_long(r.size), _d(r), C.c_double(dx), C.c_double(dy), C.c_double(dz),
                                              C.byref(e)))
        return e.value

    def ks_residual(self, g, va, vb, psi, eig, spins):
        """calculateMaximumMolecularKSResidual (MolecularSCFHelpers.cpp:119-183); spins[o] 0 alpha, 1 beta."""
        psi = _f64(psi).reshape(-1, g.size)
        sp = np.ascontiguousarray(spins, dtype=np.int32)
        out = C.c_double()
        self._chk(self.L.og_ks_residual(C.byref(self._g(g)), _d(_f64(va, g.size)), _d(_f64(vb, g.size)),
                                        C.c_int(psi.shape[0]), _d(psi), _d(_f64(eig, psi.shape[0])),
                                        sp.ctypes.data_as(_ip), C.byref(out)))
        return out.value

    def kinetic_energy(self, g, psi, electrons):
        psi = _f64(psi).reshape(-1, g.size)
        el = np.ascontiguousarray(electrons, dtype=np.int32)
        return self.L.og_kinetic_energy(C.byref(self._g(g)), C.c_int(psi.shape[0]), _d(psi), el.ctypes.data_as(_ip))

    def external_energy(self, g, rho, vnuc):
        return self.L.og_external_energy(C.byref(self._g(g)), _d(_f64(rho, g.size)), _d(_f64(vnuc, g.size)))

    def nuclear_potential(self, g, nuclei):
        nuc = _f64(nuclei).reshape(-1, 4)
        out = np.empty(g.size)
        self._chk(self.L.og_nuclear_potential(C.byref(self._g(g)), C.c_int(nuc.shape[0]), _d(nuc), _d(out)))
        return out

    def nuclear_repulsion(self, nuclei):
        nuc = _f64(nuclei).reshape(-1, 4)
        e = C.c_double()
        self._chk(self.L.og_nuclear_repulsion(C.c_int(nuc.shape[0]), _d(nuc), C.byref(e)))
        return e.value

    def davidson(self, g, v, norb, init=None, max_happly=5000):
        psi = np.empty((norb, g.size))
        eig = np.empty(norb)
        hap = np.zeros(norb, dtype=np.int64)
        occ = np.ones(norb, dtype=np.int32)
        ip = None if init is None else _d(_f64(init).reshape(norb, g.size))
        self._chk(self.L.og_davidson(C.byref(self._g(g)), _d(_f64(v, g.size)), C.c_int(norb), occ.ctypes.data_as(_ip),
                                     ip, C.c_long(max_happly), _d(psi), _d(eig), hap.ctypes.data_as(_lp)))
        return eig, psi, hap

    def scf(self, g, nuclei, functional=1, charge=0, multiplicity=1, max_iterations=250, poisson_fixed=0,
            stop_rule=0, energy_tol=1e-8, init_guess=0, max_happly=5000, want_density=False):
        nuc = _f64(nuclei).reshape(-1, 4)
        cfg = _OgScfConfig(functional, charge, multiplicity, max_iterations, poisson_fixed, stop_rule,
                           energy_tol, init_guess, max_happly)
        res = _OgScfResult()
        trace = np.zeros((max_iterations, 5))
        ne = int(round(nuc[:, 0].sum())) - charge
        eig = np.zeros(ne + 2)
        rho = np.zeros(2 * g.size) if want_density else None
        self._chk(self.L.og_scf(C.byref(self._g(g)), C.c_int(nuc.shape[0]), _d(nuc), C.byref(cfg), C.byref(res),
                                _d(trace), _d(eig), None if rho is None else _d(rho)))
        it = res.iterations
        out = dict(iterations=it, converged=bool(res.converged), energy=list(res.energy),
                   total=res.energy[5], dE=res.energy_difference, dRho=res.density_difference,
                   ks=res.ks_residual, eig_alpha0=res.eig_alpha0, eig_beta0=res.eig_beta0,
                   first_de_iteration=res.first_de_iteration, trace=trace[:it].copy(), eigenvalues=eig)
        if rho is not None:
            out["rho_alpha"], out["rho_beta"] = rho[: g.size], rho[g.size:]
        return out


def _scf_step(self, g, nuclei, functional, multiplicity, ra, rb):
    nuc = _f64(nuclei).reshape(-1, 4)
    n = g.size
    out = {k: np.empty(n) for k in ("vh", "vxa", "vxb", "ora", "orb", "ovh")}
    ne = int(round(nuc[:, 0].sum()))
    eig = np.empty(ne + 2)
    E = np.empty(6)
    dd, ks = C.c_double(), C.c_double()
    self._chk(self.L.og_scf_step(C.byref(self._g(g)), C.c_int(nuc.shape[0]), _d(nuc), C.c_int(functional),
                                 C.c_int(multiplicity), _d(_f64(ra, n)), _d(_f64(rb, n)), _d(out["vh"]), _d(out["vxa"]),
                                 _d(out["vxb"]), _d(out["ora"]), _d(out["orb"]), _d(out["ovh"]), _d(eig), _d(E),
                                 C.byref(dd), C.byref(ks)))
    out.update(eigenvalues=eig[:ne], energy=E, dRho=dd.value, ks=ks.value)
    return out


Oracle.scf_step = _scf_step


class Reference:
    """The reference's compiled 3D path (only where oracle/_ref was built)."""

    def __init__(self, path=REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = C.CDLL(path)
        self.L = L
        L.rf_last_error.restype = C.c_char_p

    @staticmethod
    def available():
        return os.path.exists(REF_SO)

    def _chk(self, rc):
        if rc != 0:
            raise OracleError(f"reference rc={rc}: {self.L.rf_last_error().decode()}")

    @staticmethod
    def _g(g):
        return (C.c_long * 3)(*g.n), (C.c_double * 6)(*g.bounds)

    def hpsi(self, g, v, psi):
        out = np.empty(g.size)
        self._chk(self.L.rf_hpsi(*self._g(g), _d(_f64(v, g.size)), _d(_f64(psi, g.size)), _d(out)))
        return out

    def poisson(self, g, rho):
        out = np.empty(g.size)
        self._chk(self.L.rf_poisson(*self._g(g), _d(_f64(rho, g.size)), _d(out)))
        return out

    def hartree_energy(self, g, rho, vh):
        e = C.c_double()
        self._chk(self.L.rf_hartree_energy(*self._g(g), _d(_f64(rho, g.size)), _d(_f64(vh, g.size)), C.byref(e)))
        return e.value

    def xc_points(self, fid, inp):
        inp = _f64(inp).reshape(-1, 4)
        out = np.empty((inp.shape[0], 5))
        self._chk(self.L.rf_xc_points(C.c_int(fid), C.c_long(inp.shape[0]), _d(inp), _d(out)))
        return out

    def xc_potential(self, g, fid, na, nb, spin):
        out = np.empty(g.size)
        self._chk(self.L.rf_xc_potential(*self._g(g), C.c_int(fid), _d(_f64(na, g.size)), _d(_f64(nb, g.size)),
                                         C.c_int(spin), _d(out)))
        return out

    def xc_energy(self, g, fid, na, nb):
        e = C.c_double()
        self._chk(self.L.rf_xc_energy(*self._g(g), C.c_int(fid), _d(_f64(na, g.size)), _d(_f64(nb, g.size)), C.byref(e)))
        return e.value

    def spin_density(self, psi, electrons):
        psi = _f64(psi)
        norb, n = psi.shape
        out = np.empty(n)
        el = np.ascontiguousarray(electrons, dtype=np.int32)
        self._chk(self.L.rf_spin_density(C.c_long(n), C.c_int(norb), _d(psi), el.ctypes.data_as(_ip), _d(out)))
        return out

    def electron_density(self, psi, electrons, spins):
        psi = _f64(psi)
        norb, n = psi.shape
        out = np.empty(n)
        el = np.ascontiguousarray(electrons, dtype=np.int32)
        sp = np.ascontiguousarray(spins, dtype=np.int32)
        self._chk(self.L.rf_electron_density(C.c_long(n), C.c_int(norb), _d(psi), el.ctypes.data_as(_ip),
                                             sp.ctypes.data_as(_ip), _d(out)))
        return out

    def integrate_density(self, rho, dx, dy, dz):
        r = _f64(rho)
        e = C.c_double()
        self._chk(self.L.rf_integrate_density(C.c_long(r.size), _d(r), C.c_double(dx), C.c_double(dy), C.c_double(dz),
                                              C.byref(e)))
        return e.value

    def ks_residual(self, g, va, vb, psi, eig, spins):
        psi = _f64(psi).reshape(-1, g.size)
        sp = np.ascontiguousarray(spins, dtype=np.int32)
        out = C.c_double()
        self._chk(self.L.rf_ks_residual(*self._g(g), _d(_f64(va, g.size)), _d(_f64(vb, g.size)), C.c_int(psi.shape[0]),
                                        _d(psi), _d(_f64(eig, psi.shape[0])), sp.ctypes.data_as(_ip), C.byref(out)))
        return out.value

    def kinetic_energy(self, g, psi, electrons):
        psi = _f64(psi).reshape(-1, g.size)
        el = np.ascontiguousarray(electrons, dtype=np.int32)
        e = C.c_double()
        self._chk(self.L.rf_kinetic(*self._g(g), C.c_int(psi.shape[0]), _d(psi), el.ctypes.data_as(_ip), C.byref(e)))
        return e.value

    def external_energy(self, g, rho, vnuc):
        e = C.c_double()
        self._chk(self.L.rf_external(*self._g(g), _d(_f64(rho, g.size)), _d(_f64(vnuc, g.size)), C.byref(e)))
        return e.value

    def nuclear_potential(self, g, nuclei):
        nuc = _f64(nuclei).reshape(-1, 4)
        out = np.empty(g.size)
        self._chk(self.L.rf_nuclear_potential(*self._g(g), C.c_int(nuc.shape[0]), _d(nuc), _d(out)))
        return out

    def nuclear_repulsion(self, nuclei):
        nuc = _f64(nuclei).reshape(-1, 4)
        e = C.c_double()
        self._chk(self.L.rf_nuclear_repulsion(C.c_int(nuc.shape[0]), _d(nuc), C.byref(e)))
        return e.value

    def davidson(self, g, v, norb, init=None, max_happly=5000):
        psi = np.empty((norb, g.size))
        eig = np.empty(norb)
        ip = None if init is None else _d(_f64(init).reshape(norb, g.size))
        self._chk(self.L.rf_davidson(*self._g(g), _d(_f64(v, g.size)), C.c_int(norb), ip, C.c_long(max_happly),
                                     _d(psi), _d(eig)))
        return eig, psi

    def scf_step(self, g, nuclei, functional, multiplicity, ra, rb):
        """executeMolecularSCFIteration on given densities, the same outputs as Oracle.scf_step."""
        nuc = _f64(nuclei).reshape(-1, 4)
        n = g.size
        out = {k: np.empty(n) for k in ("vh", "vxa", "vxb", "ora", "orb", "ovh")}
        ne = int(round(nuc[:, 0].sum()))
        eig = np.empty(ne + 2)
        E = np.empty(6)
        dd, ks = C.c_double(), C.c_double()
        self._chk(self.L.rf_scf_step(*self._g(g), C.c_int(nuc.shape[0]), _d(nuc), C.c_int(functional),
                                     C.c_int(multiplicity), _d(_f64(ra, n)), _d(_f64(rb, n)), _d(out["vh"]),
                                     _d(out["vxa"]), _d(out["vxb"]), _d(out["ora"]), _d(out["orb"]), _d(out["ovh"]),
                                     _d(eig), _d(E), C.byref(dd), C.byref(ks)))
        out.update(eigenvalues=eig[:ne], energy=E, dRho=dd.value, ks=ks.value)
        return out

    def scf(self, g, nuclei, functional=1, charge=0, multiplicity=1):
        nuc = _f64(nuclei).reshape(-1, 4)
        out = np.zeros(11)
        ne = int(round(nuc[:, 0].sum())) - charge
        eig = np.zeros(ne + 2)
        self._chk(self.L.rf_scf(*self._g(g), C.c_int(nuc.shape[0]), _d(nuc), C.c_int(charge), C.c_int(multiplicity),
                                C.c_int(functional), _d(out), _d(eig), C.c_int(len(eig)), C.c_int(1)))
        keys = ["iterations", "converged", "T", "Eext", "EH", "Exc", "Enn", "total", "dE", "dRho", "ks"]
        r = dict(zip(keys, out.tolist()))
        r["iterations"] = int(r["iterations"])
        r["eigenvalues"] = eig
        return r


# ---- the radial (atomic) stack: oracle/radial_oracle.c and the reference's solveAtom
RADIAL_KEYS = ("kinetic", "external", "hartree", "exchangeCorrelation", "total", "densityDifference",
               "energyDifference", "maxKSResidual")


def _radial_scf(self, Z, functional=1, points=2000, rmax=30.0, densities=False):
    """og_radial_scf: the atom's SCF as the reference's solveAtom (AtomicDFT.cpp:6-41)."""
    res = np.zeros(8)
    eig = np.zeros(14)
    no = C.c_int()
    na = np.zeros(points) if densities else None
    nb = np.zeros(points) if densities else None
    self.L.og_radial_scf.restype = C.c_int
    rc = self.L.og_radial_scf(C.c_int(Z), C.c_int(functional), C.c_long(points), C.c_double(rmax), _d(res), _d(eig),
                              C.byref(no), None if na is None else _d(na), None if nb is None else _d(nb))
    if rc < -1000 or rc == -1001:
        raise OracleError(f"og_radial_scf rc={rc}")
    out = dict(zip(RADIAL_KEYS, res.tolist()))
    out.update(iterations=abs(rc), converged=rc > 0, eigenvalues=eig[:no.value].tolist())
    if densities:
        out["alpha"], out["beta"] = na, nb
    return out


def _radial_orbital(self, V, l, index, rmax=30.0):
    """og_radial_orbital: eigenvalue `index` of the radial H(l) with potential V, and its vector."""
    V = _f64(V)
    u = np.zeros(V.size)
    e = C.c_double()
    self._chk(self.L.og_radial_orbital(C.c_long(V.size), C.c_double(rmax), _d(V), C.c_int(l), C.c_long(index),
                                       C.byref(e), _d(u)))
    return e.value, u


def _radial_hartree(self, rho, rmax=30.0):
    rho = _f64(rho)
    out = np.zeros(rho.size)
    self._chk(self.L.og_radial_hartree(C.c_long(rho.size), C.c_double(rmax), _d(rho), _d(out)))
    return out


Oracle.radial_scf = _radial_scf
Oracle.radial_orbital = _radial_orbital
Oracle.radial_hartree = _radial_hartree


def _ref_radial_scf(self, Z, functional=1, points=2000, rmax=30.0):
    """rf_radial_scf: the reference's own solveAtom(RadialGrid(points, rmax), getAtomicConfiguration(Z), f)."""
    out = np.zeros(10)
    eig = np.zeros(14)
    no = C.c_int()
    self._chk(self.L.rf_radial_scf(C.c_int(Z), C.c_int(functional), C.c_long(points), C.c_double(rmax), _d(out),
                                   _d(eig), C.c_int(14), C.byref(no)))
    keys = ["iterations", "converged", "kinetic", "external", "hartree", "exchangeCorrelation", "total",
            "energyDifference", "densityDifference", "maxKSResidual"]
    r = dict(zip(keys, out.tolist()))
    r["iterations"] = int(r["iterations"])
    r["converged"] = bool(r["converged"])
    r["eigenvalues"] = eig[:no.value].tolist()
    return r


Reference.radial_scf = _ref_radial_scf
