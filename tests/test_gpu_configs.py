"""The benchmark definition's small configurations on the GPU, against the reference.

C1: He at the centre of a 48^3 grid, h = 0.25, PZ81, Davidson block 4.  Under the reference's rules
(bare-nucleus start, reference Jacobi Poisson, the dE/dRho/KS triple stop) the reference's own
solveMolecularSelfConsistentField gives E = -2.967196211849 Ha and e_1s = -0.826151374116 Ha (the
compiled reference, SelfConsistentFieldMolecule.cpp:50-561); from the hydrogenic 1s start with the
|dE| < 1e-8 stop the same state is reached.

C2: Ne, 96^3, h = 0.2, PBE96, block 8: the kernels at full size against the oracle on identical inputs
(H psi, 200 Jacobi sweeps from zero, V_xc), and one SCF step (MolecularSCFIteration.cpp:12-252, the
reference Poisson rule) against the reference-written fixture tests/golden/c2_step.npz.

Also the Kohn-Sham residual of the reference-shaped C++ API (MolecularSCFHelpers.cpp:119-183) against
the oracle on identical orbitals.
"""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
C2_FIXTURE = os.path.join(ROOT, "tests", "golden", "c2_step.npz")
C1_E, C1_EPS = -2.967196211849, -0.826151374116


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def dgrid(og):
    import dft_amd
    return dft_amd.Grid.make(*og.n, *og.bounds)


@pytest.fixture(scope="module")
def dev():
    import dft_amd
    d = dft_amd.Device(0)
    yield d
    d.close()


# ------------------------------------------------------------------ C1
def test_c1_helium_reference_rules():
    import dft_amd
    from oracle.oracle import Grid
    og = Grid.centered(48, 0.25)
    print("[c1] He 48^3 reference rules", flush=True)
    r = dft_amd.scf_run(dgrid(og), [[2.0, 0.0, 0.0, 0.0]], functional=dft_amd.PZ81, nstates=4,
                        precision=dft_amd.SCF_FP64, poisson_method=dft_amd.POISSON_JACOBI, poisson_fixed=0,
                        stop_rule=dft_amd.STOP_REFERENCE, init_guess=dft_amd.INIT_BARE)
    print("c1 reference rules", r["total"], r["eigenvalues"], r["iterations"])
    assert r["converged"]
    assert abs(r["total"] - C1_E) <= 1e-7
    # 5e-8: both runs stop under the reference's Jacobi rule, max|dV| < 1e-8 per sweep with V_H still
    # moving (HartreePotentialMolecule.cpp:312-322); one sweep more or less at the last step moves the
    # eigenvalue by up to that much, and the two Davidson methods reach the last step by different paths
    assert abs(r["eigenvalues"][0] - C1_EPS) <= 5e-8


def test_c1_helium_hydrogenic_start_de_stop():
    import dft_amd
    from oracle.oracle import Grid
    og = Grid.centered(48, 0.25)
    print("[c1] He 48^3 hydrogenic start, |dE| < 1e-8", flush=True)
    r = dft_amd.scf_run(dgrid(og), [[2.0, 0.0, 0.0, 0.0]], functional=dft_amd.PZ81, nstates=4,
                        precision=dft_amd.SCF_FP64, poisson_method=dft_amd.POISSON_JACOBI, poisson_fixed=0,
                        stop_rule=dft_amd.STOP_DE, energy_tol=1e-8, init_guess=dft_amd.INIT_HYDROGENIC)
    print("c1 hydrogenic", r["total"], r["eigenvalues"], r["iterations"])
    assert r["converged"] and abs(r["energy_difference"]) < 1e-8
    assert abs(r["total"] - C1_E) <= 1e-7
    # the |dE| stop leaves the density where the energy (second order in the density error) is within
    # 1e-8 but the eigenvalue (first order) is not: measured 5.3e-5 from the reference-rule value at
    # 13 steps, where the reference's own stop also requires the density change and the KS residual
    assert abs(r["eigenvalues"][0] - C1_EPS) <= 1e-4


# ------------------------------------------------------------------ C2
def _c2_inputs():
    from oracle.oracle import Grid
    og = Grid.centered(96, 0.2)
    xs, ys, zs = og.coords()
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2).ravel()
    n = np.exp(-2.0 * 9.6 * r) * 2.0 + (r * r) * np.exp(-2.0 * 2.8 * r) * 8.0  # as tests/golden/make_c2_step.py
    n *= 5.0 / (n.sum() * og.dx * og.dy * og.dz)
    return og, n


def test_c2_kernels_full_size(dev, oracle):
    """H psi (FP64, bitwise equal to the oracle's reference-order arithmetic), 200 Jacobi sweeps from zero
    (the workload's sweep count; bitwise) and the PBE96 potential (relative 1e-12) at 96^3."""
    import dft_amd
    og, n = _c2_inputs()
    g = dgrid(og)
    vnuc = oracle.nuclear_potential(og, [[10.0, 0.0, 0.0, 0.0]])
    rng = np.random.default_rng(96)
    psi = rng.standard_normal(og.size)
    print("[c2] H psi", flush=True)
    out = dev.array(shape=(1, og.size))
    dev.hpsi(g, dev.array(vnuc), dev.array(psi.reshape(1, -1)), out)
    assert np.array_equal(out.download().ravel(), oracle.hpsi(og, vnuc, psi))
    print("[c2] Jacobi 200 sweeps", flush=True)
    rho = 2.0 * n
    vh = dev.zeros((og.size,))
    sw, md = dev.poisson(g, dev.array(rho), vh, max_sweeps=200, tol=0.0, zero_start=True, batch=64)
    ref, sw_ref, md_ref = oracle.poisson(og, rho, v0=np.zeros(og.size), max_sweeps=200, tol=0.0)
    assert sw == 200 and np.array_equal(vh.download(), ref)
    print("[c2] PBE96 V_xc", flush=True)
    va, vb = dev.array(shape=(og.size,)), dev.array(shape=(og.size,))
    e = dev.xc(g, dft_amd.PBE96, dev.array(n), dev.array(n), va, vb)
    assert rel(va.download(), oracle.xc_potential(og, 2, n, n, 0)) <= 1e-12
    e_ref = oracle.xc_energy(og, 2, n, n)
    assert abs(e - e_ref) <= 1e-12 * abs(e_ref)


@pytest.mark.skipif(not os.path.exists(C2_FIXTURE), reason="tests/golden/c2_step.npz not written")
def test_c2_scf_step_against_reference():
    """One SCF step at 96^3 on the fixture's input density, reference Poisson rule (10 000 sweeps from
    zero at this size), GPU FP64 against the reference's executeMolecularSCFIteration: V_H and V_xc at
    4096 sampled points to relative 1e-12; occupied eigenvalues to 1e-8 Ha; output densities and output
    V_H (first order in the orbitals, each side converged to residual 1e-7) to relative 1e-6; energy
    terms to 1e-5 Ha (first order in the orbital error on a non-self-consistent input); dRho to 1e-6
    relative."""
    import dft_amd
    fx = np.load(C2_FIXTURE)
    og, n = _c2_inputs()
    assert abs(n.sum() - float(fx["input_sum"])) <= 1e-12 * float(fx["input_sum"])
    print("[c2] SCF step 96^3", flush=True)
    got = dft_amd.scf_step(dgrid(og), fx["nuclei"], n, n, functional=int(fx["functional"]), multiplicity=1,
                           nstates=8)
    idx = fx["idx"]
    for k_got, k_ref, tol in (("vh", "vh", 1e-12), ("vxc_alpha", "vxa", 1e-12), ("vxc_beta", "vxb", 1e-12),
                              ("rho_alpha", "ora", 1e-6), ("rho_beta", "orb", 1e-6), ("vh_out", "ovh", 1e-6)):
        e = rel(got[k_got][idx], fx[f"{k_ref}_at"])
        print(k_got, e)
        assert e <= tol, k_got
    de = np.max(np.abs(got["eigenvalues"] - fx["eigenvalues"]))
    print("eig", de, "E", got["energy"] - fx["energy"], "dRho", got["dRho"], float(fx["dRho"]))
    assert de <= 1e-8
    assert np.max(np.abs(got["energy"] - fx["energy"])) <= 1e-5
    assert abs(got["dRho"] - float(fx["dRho"])) <= 1e-6 * float(fx["dRho"])
    assert got["ks"] < 1e-7 and float(fx["ks"]) < 1e-7


# ------------------------------------------------------------------ KS residual
def test_ks_residual_cpp_api_against_oracle(tmp_path, oracle):
    """calculateMaximumMolecularKSResidual through the C++ API on the GPU against the oracle on the same
    orbitals, eigenvalues and potentials (both spins; near-eigenvectors, so the residual is a
    cancellation): relative 1e-12."""
    from oracle.oracle import Grid
    exe = os.path.join(ROOT, "tests", "cpp", "bin", "ks_residual_check")  # built by __graft_entry__.build()
    assert os.path.exists(exe), "tests/cpp/bin/ks_residual_check is not built"
    og = Grid(23, 21, 19, -3.3, 3.3, -3.0, 3.0, -2.7, 2.7)
    nx, ny, nz = og.n
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    rng = np.random.default_rng(5)
    modes, eig = [], []
    for (p, q, r) in ((1, 1, 1), (1, 2, 1), (2, 2, 3), (3, 1, 2)):
        u = (np.sin(np.pi * p * (x + 1) / (nx + 1)) * np.sin(np.pi * q * (y + 1) / (ny + 1)) *
             np.sin(np.pi * r * (z + 1) / (nz + 1))).ravel()
        u = u + 1e-4 * rng.standard_normal(u.size)
        modes.append(u / np.sqrt(np.sum(u * u) * og.dx * og.dy * og.dz))
        eig.append(sum((1 - np.cos(np.pi * m / (nn + 1))) / d ** 2
                       for m, nn, d in ((p, nx, og.dx), (q, ny, og.dy), (r, nz, og.dz))))
    psi = np.array(modes)
    va = 0.2 + 1e-3 * rng.standard_normal(og.size)
    vb = -0.1 + 1e-3 * rng.standard_normal(og.size)
    spins = np.array([0, 1, 0, 1], dtype=np.int32)
    eig = np.array(eig) + np.where(spins == 0, 0.2, -0.1)
    np.array([*og.n, *og.bounds], dtype=np.float64).tofile(tmp_path / "grid.bin")
    np.concatenate([va, vb]).astype(np.float64).tofile(tmp_path / "potentials.bin")
    np.concatenate([[psi.shape[0]], eig, spins.astype(float), psi.ravel()]).astype(np.float64).tofile(
        tmp_path / "orbitals.bin")
    print("[ks residual api]", flush=True)
    out = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    got = float(np.fromfile(tmp_path / "ks.out", dtype=np.float64)[0])
    ref = oracle.ks_residual(og, va, vb, psi, eig, spins)
    print("ks", got, ref)
    assert ref > 1e-4
    assert abs(got - ref) <= 1e-12 * ref
