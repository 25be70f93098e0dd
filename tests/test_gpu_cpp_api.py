"""The reference-shaped C++ API (include/dft_molecule.h) used from C++: a
program written against the header (tests/cpp/molecule_api_check.cpp) is
compiled with g++ by the build step, linked to libdft_core.so, run on the GPU, and its outputs
are compared with the FP64 oracle (the same bars as the kernel tests)."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


def test_reference_cpp_api_against_oracle(tmp_path, oracle):
    """The program runs while this process holds a device context that has
    just run a Jacobi solve on the same GPU; the child's own context and
    solves must be unaffected and match the oracle."""
    import dft_amd
    from oracle.oracle import Grid
    exe = os.path.join(ROOT, "tests", "cpp", "bin", "molecule_api_check")  # built by __graft_entry__.build()
    assert os.path.exists(exe), "tests/cpp/bin/molecule_api_check is not built"
    dev = dft_amd.Device(0)
    gs = dft_amd.Grid.centered(20, 0.4)
    dev.poisson(gs, dev.array(np.full(gs.size, 1e-3)), dev.zeros((gs.size,)), max_sweeps=300, tol=0.0,
                zero_start=True, batch=100)
    dev.sync()
    og = Grid(21, 18, 16, -4.0, 4.2, -3.9, 3.6, -3.3, 3.5)
    n = og.size
    rng = np.random.default_rng(11)
    xs, ys, zs = og.coords()
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    r2 = (X - 0.1) ** 2 + (Y + 0.2) ** 2 + Z ** 2
    data = dict(v=rng.standard_normal(n), psi=rng.standard_normal(n), rho=(np.exp(-r2) * 0.8).ravel(),
                na=(0.6 * np.exp(-0.9 * r2)).ravel(), nb=(0.4 * np.exp(-1.1 * r2)).ravel())
    np.array([*og.n, *og.bounds], dtype=np.float64).tofile(tmp_path / "grid.bin")
    for k, a in data.items():
        np.asarray(a, dtype=np.float64).tofile(tmp_path / f"{k}.bin")
    out = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, timeout=120)
    dev.close()
    assert out.returncode == 0, out.stderr
    rd = lambda k: np.fromfile(tmp_path / f"{k}.out", dtype=np.float64)
    assert rel(rd("hpsi"), oracle.hpsi(og, data["v"], data["psi"])) <= 1e-12
    vh_ref, _, _ = oracle.poisson(og, data["rho"])
    assert rel(rd("vh"), vh_ref) <= 1e-12
    assert rel(rd("vxc_pz"), oracle.xc_potential(og, 1, data["na"], data["nb"], 0)) <= 1e-12
    assert rel(rd("vxc_pbe"), oracle.xc_potential(og, 2, data["na"], data["nb"], 1)) <= 1e-12
    nuc = [[8, 0.01, 0.02, 0.03], [1, 1.43, 1.11, 0.0], [1, -1.43, 1.11, 0.0]]
    vn = oracle.nuclear_potential(og, nuc)
    assert rel(rd("vnuc"), vn) <= 1e-14
    s = rd("scalars")
    assert abs(s[0] - oracle.hartree_energy(og, data["rho"], vh_ref)) <= 1e-12 * abs(s[0])
    assert abs(s[1] - oracle.xc_energy(og, 2, data["na"], data["nb"])) <= 1e-12 * abs(s[1])
    eig_ref, psi_ref, _ = oracle.davidson(og, vn, 3)
    assert np.max(np.abs(s[3:6] - eig_ref)) <= 1e-8
    dens_ref = 2.0 * np.sum(psi_ref ** 2, axis=0)
    assert rel(rd("density"), dens_ref) <= 1e-6
    scf = rd("scf")
    assert scf[2] == 1.0 and abs(scf[0] - (-2.8564865005239386)) <= 1e-7
