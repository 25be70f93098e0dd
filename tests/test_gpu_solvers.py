"""GPU block Davidson and device-resident SCF against the oracle and the
reference's known answers.  Bars: occupied eigenvalues within 1e-8 Ha and
total energies within 1e-7 Ha of the FP64 CPU path (reference rules)."""
import numpy as np
import pytest

from conftest import grid_from

pytestmark = pytest.mark.gpu


def dgrid(og):
    import dft_amd
    return dft_amd.Grid.make(*og.n, *og.bounds)


def test_davidson_golden_eigenvalues(goldens):
    import dft_amd
    og = grid_from(goldens["dav_grid"])
    r = dft_amd.davidson_run(dgrid(og), goldens["dav_v"], nstates=6, nconv=4, want_psi=True)
    print("davidson", r["eigenvalues"][:4], goldens["dav_eig"], r["iterations"])
    assert np.max(np.abs(r["eigenvalues"][:4] - goldens["dav_eig"])) <= 1e-8
    assert np.all(r["residuals"][:4] < 1e-7)
    dv = og.dx * og.dy * og.dz
    psi = r["psi"]
    np.testing.assert_allclose(psi @ psi.T * dv, np.eye(4), atol=1e-9)
    # same subspace as the reference's orbitals: projector distance
    ref = goldens["dav_psi"]
    Pg = psi.T @ psi * dv
    Pr = ref.T @ ref * dv
    assert np.max(np.abs(Pg - Pr)) <= 1e-6


def test_davidson_vs_dense(oracle):
    import dft_amd
    from oracle.oracle import Grid
    og = Grid.centered(6, 0.6)
    v = oracle.nuclear_potential(og, [[2, 0.05, 0.1, -0.02]])
    eye = np.eye(og.size)
    H = np.stack([oracle.hpsi(og, v, eye[i]) for i in range(og.size)], 1)
    w = np.linalg.eigvalsh(H)
    r = dft_amd.davidson_run(dgrid(og), v, nstates=5, nconv=3)
    assert np.max(np.abs(r["eigenvalues"][:3] - w[:3])) <= 1e-9


@pytest.mark.parametrize("gram", ["f32", "split"])
def test_davidson_mixed_precision(goldens, gram):
    import dft_amd
    og = grid_from(goldens["dav_grid"])
    mode = dft_amd.GRAM_F32 if gram == "f32" else dft_amd.GRAM_SPLIT_BF16
    r = dft_amd.davidson_run(dgrid(og), goldens["dav_v"], nstates=6, nconv=4, prec=dft_amd.F32, gram_mode=mode,
                             tol=5e-5)
    err = np.max(np.abs(r["eigenvalues"][:4] - goldens["dav_eig"]))
    print("mixed davidson", gram, err, r["iterations"])
    # FP32 storage with split-BF16 products (~2^-17 relative per product) on
    # N = 12^3 points: eigenvalues to 1e-5 relative; the FP64 refinement pass
    # of the mixed SCF removes the rest
    assert err <= 1e-5 * np.max(np.abs(goldens["dav_eig"]))


@pytest.mark.parametrize("name,total,steps", [("h2", -2.8564865005, 29), ("h2o", -101.39828181, 87)])
def test_scf_main_cpp_known_answers(goldens, name, total, steps):
    import dft_amd
    og = grid_from(goldens[f"scf_{name}_grid"])
    r = dft_amd.scf_run(dgrid(og), goldens[f"scf_{name}_nuc"], functional=int(goldens[f"scf_{name}_fid"]))
    ref = goldens[f"scf_{name}_result"]
    print(name, r["total"], ref[6], r["iterations"], steps)
    assert r["converged"]
    assert abs(r["total"] - ref[6]) <= 1e-7
    assert abs(r["total"] - total) <= 1e-7 + 5e-11 * abs(total)
    # Two SCF runs that each satisfy the reference's stop rule (dE < 1e-11,
    # dRho < 1e-9, KS < 1e-7) differ by up to ~1e-7 in eigenvalues (the stop
    # rule bounds the step, not the distance to the fixed point); the 1e-8
    # eigenvalue bar is asserted on identical inputs in test_scf_step_parity.
    ne = len(r["eigenvalues"])
    assert np.max(np.abs(r["eigenvalues"] - goldens[f"scf_{name}_eig"][:ne])) <= 1e-7


def test_scf_he_pbe_golden(goldens):
    import dft_amd
    og = grid_from(goldens["scf_he_pbe_grid"])
    r = dft_amd.scf_run(dgrid(og), goldens["scf_he_pbe_nuc"], functional=2)
    ref = goldens["scf_he_pbe_result"]
    print("he pbe", r["total"], ref[6], r["energy"], ref[1:6])
    assert r["converged"]
    np.testing.assert_allclose(r["energy"], ref[1:7], atol=1e-7)
    assert abs(r["eigenvalues"][0] - goldens["scf_he_pbe_eig"][0]) <= 1e-8


@pytest.mark.parametrize("case", ["h2o_pz81", "ne_pbe_open", "n_pbe_open_sc"])
def test_scf_step_parity(oracle, case):
    """One SCF iteration on identical input densities, GPU vs the FP64 CPU
    path: V_H and V_xc to relative 1e-12, occupied eigenvalues to 1e-8 Ha,
    energy components to 1e-6 Ha, output densities to the Davidson tolerance;
    the total energy to 1e-7 Ha at a self-consistent input (closed-shell H2O,
    PZ81; the open-shell N quartet, PBE96, at the density of its converged
    SCF)."""
    import dft_amd
    from oracle.oracle import Grid
    if case == "h2o_pz81":
        og = Grid(9, 9, 9, -8.1, 7.9, -8.1, 7.9, -8.1, 7.9)
        nuc, fid, mult = [[8, 0, 0, 0], [1, 1.430, 0, 1.108], [1, -1.430, 0, 1.108]], 1, 1
        base = oracle.scf(og, nuc, fid, want_density=True)  # self-consistent input density
        ra, rb = base["rho_alpha"], base["rho_beta"]
    elif case == "n_pbe_open_sc":
        og = Grid.centered(20, 0.45)
        nuc, fid, mult = [[7, 0.05, -0.02, 0.03]], 2, 4
        sc = dft_amd.scf_run(dgrid(og), nuc, functional=fid, multiplicity=mult, stop_rule=dft_amd.STOP_DE,
                             energy_tol=1e-10, max_iterations=200, want_density=True)
        assert sc["converged"]
        ra, rb = sc["rho_alpha"], sc["rho_beta"]
    else:
        og = Grid.centered(20, 0.45)
        nuc, fid, mult = [[7, 0.05, -0.02, 0.03]], 2, 4  # N atom, quartet: 5 alpha + 2 beta
        xs, ys, zs = og.coords()
        Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
        r = np.sqrt((X - 0.05) ** 2 + (Y + 0.02) ** 2 + (Z - 0.03) ** 2).ravel()
        dv = og.dx * og.dy * og.dz
        ra = np.exp(-1.6 * r) + 0.3 * np.exp(-0.7 * r)
        ra *= 5.0 / (ra.sum() * dv)
        rb = np.exp(-2.2 * r)
        rb *= 2.0 / (rb.sum() * dv)
    ref = oracle.scf_step(og, nuc, fid, mult, ra, rb)
    got = dft_amd.scf_step(dgrid(og), nuc, ra, rb, functional=fid, multiplicity=mult)

    def rel(a, b):
        return np.max(np.abs(a - b)) / np.max(np.abs(b))
    print(case, "vh", rel(got["vh"], ref["vh"]), "vxa", rel(got["vxc_alpha"], ref["vxa"]),
          "vxb", rel(got["vxc_beta"], ref["vxb"]), "eig", np.max(np.abs(got["eigenvalues"] - ref["eigenvalues"])),
          "E", got["energy"] - ref["energy"])
    assert rel(got["vh"], ref["vh"]) <= 1e-12
    assert rel(got["vxc_alpha"], ref["vxa"]) <= 1e-12 and rel(got["vxc_beta"], ref["vxb"]) <= 1e-12
    assert np.max(np.abs(got["eigenvalues"] - ref["eigenvalues"])) <= 1e-8
    # At a self-consistent input the total energy is second order in the
    # orbital error (orbitals converged to residual 1e-7 on both sides); on a
    # non-self-consistent input it is first order, like the components.
    if case in ("h2o_pz81", "n_pbe_open_sc"):
        assert abs(got["energy"][5] - ref["energy"][5]) <= 1e-7
    assert np.max(np.abs(got["energy"] - ref["energy"])) <= 1e-6
    assert rel(got["rho_alpha"], ref["ora"]) <= 1e-6 and rel(got["rho_beta"], ref["orb"]) <= 1e-6
    assert got["ks"] < 1e-7 and ref["ks"] < 1e-7


@pytest.mark.parametrize("kind", ["kinetic", "kinetic_cheb", "kinetic_exact"])
def test_davidson_kinetic_correction(goldens, kind):
    """The kinetic smoothing corrections (damped-Jacobi or Chebyshev-combined
    steps) reach the reference's eigenvalues in fewer iterations than the
    diagonal one."""
    import dft_amd
    og = grid_from(goldens["dav_grid"])
    pc = {"kinetic": dft_amd.PRECOND_KINETIC, "kinetic_cheb": dft_amd.PRECOND_KINETIC_CHEB,
          "kinetic_exact": dft_amd.PRECOND_KINETIC_EXACT}[kind]
    rd = dft_amd.davidson_run(dgrid(og), goldens["dav_v"], nstates=6, nconv=4)
    rk = dft_amd.davidson_run(dgrid(og), goldens["dav_v"], nstates=6, nconv=4, precond=pc)
    assert np.max(np.abs(rk["eigenvalues"][:4] - goldens["dav_eig"])) <= 1e-8
    assert rk["iterations"] <= rd["iterations"], (rk["iterations"], rd["iterations"])


@pytest.mark.parametrize("case", ["f64_dense", "f32_many"])
def test_davidson_many_states(oracle, case):
    """Many states for the grid (the seed block is numerically rank deficient
    in FP32 and gets completed with random directions): eigenvalues match
    dense eigh (FP64) or the FP64 solve (FP32 storage, split-BF16 products)."""
    import dft_amd
    from oracle.oracle import Grid
    if case == "f64_dense":
        og = Grid.centered(8, 0.6)
        v = oracle.nuclear_potential(og, [[3, 0.05, 0.1, -0.02]])
        eye = np.eye(og.size)
        H = np.stack([oracle.hpsi(og, v, eye[i]) for i in range(og.size)], 1)
        w = np.linalg.eigvalsh(H)
        r = dft_amd.davidson_run(dgrid(og), v, nstates=50, nconv=40, max_subspace=150, tol=1e-8, max_iter=600)
        assert np.max(np.abs(r["eigenvalues"][:40] - w[:40])) <= 1e-8
    else:
        og = Grid.centered(24, 0.6)
        v = oracle.nuclear_potential(og, [[6, 0.05, 0.1, -0.02], [8, 1.3, -0.4, 0.2]])
        ref = dft_amd.davidson_run(dgrid(og), v, nstates=96, nconv=84, tol=1e-7, max_iter=600,
                                   precond=dft_amd.PRECOND_KINETIC)
        # FP32 storage with split-BF16 orthogonalisation of 96+ columns: the
        # residual floor is ~1e-3 here (the SCF's FP32 phase tolerates it and
        # refines in FP64); eigenvalues carry the square of it
        r = dft_amd.davidson_run(dgrid(og), v, nstates=96, nconv=84, prec=dft_amd.F32, gram_mode=dft_amd.GRAM_SPLIT_BF16,
                                 tol=3e-3, max_iter=300, precond=dft_amd.PRECOND_KINETIC)
        err = np.max(np.abs(r["eigenvalues"][:84] - ref["eigenvalues"][:84]))
        print("f32 many states", err, r["iterations"], np.max(r["residuals"][:84]))
        assert err <= 3e-5 * np.max(np.abs(ref["eigenvalues"][:84])), err


@pytest.mark.parametrize("method", ["chebyshev", "direct"])
def test_scf_fast_poisson_matches_converged_jacobi(oracle, method):
    """SCF with Chebyshev-accelerated or direct (sine-transform) Poisson
    converges to the SCF whose Jacobi Poisson is run to full convergence every
    call (5000 warm sweeps on 16^3), both under the reference's triple stop
    criterion (dE < 1e-11, dRho < 1e-9, KS < 1e-7); also in mixed precision
    with the kinetic correction (FP32 phase, then FP64 refinement to the same
    criterion)."""
    import dft_amd
    from oracle.oracle import Grid
    og = Grid.centered(16, 0.42)
    nuc = [[1, -0.71, 0.02, 0.0], [1, 0.69, 0.0, 0.03]]
    ref = oracle.scf(og, nuc, 1, poisson_fixed=5000, stop_rule=0, init_guess=1)
    assert ref["converged"]
    pm = dft_amd.POISSON_CHEBYSHEV if method == "chebyshev" else dft_amd.POISSON_DIRECT
    common = dict(functional=dft_amd.PZ81, poisson_method=pm, poisson_tol=1e-13,
                  stop_rule=dft_amd.STOP_REFERENCE, init_guess=dft_amd.INIT_HYDROGENIC, davidson_tol=1e-9,
                  max_iterations=250)
    r = dft_amd.scf_run(dgrid(og), nuc, **common)
    print(method, "scf", r["total"], ref["total"], r["iterations"], ref["iterations"])
    # the total-energy bar of the FP64 path (1e-7 Ha); both runs stop at
    # dRho < 1e-9, which leaves ~1e-8 Ha in this non-variational expression
    assert r["converged"] and abs(r["total"] - ref["total"]) <= 1e-7
    m = dft_amd.scf_run(dgrid(og), nuc, precision=dft_amd.SCF_MIXED, refine_iterations=250,
                        precond=dft_amd.PRECOND_KINETIC_CHEB if method == "chebyshev" else dft_amd.PRECOND_KINETIC_EXACT,
                        **common)
    print(method, "scf mixed", m["total"], m["iterations"], m["refine_iterations"])
    assert m["converged"] and abs(m["total"] - ref["total"]) <= 1e-7


def test_scf_pulay_mixing_matches_linear(oracle):
    """Pulay (DIIS) density mixing converges to the same SCF as the reference's
    lagged linear mixing (the converged-Jacobi oracle SCF, reference stop rule),
    closed and open shell, FP64 and mixed precision."""
    import dft_amd
    from oracle.oracle import Grid
    og = Grid.centered(16, 0.42)
    nuc = [[1, -0.71, 0.02, 0.0], [1, 0.69, 0.0, 0.03]]
    ref = oracle.scf(og, nuc, 1, poisson_fixed=5000, stop_rule=0, init_guess=1)
    common = dict(functional=dft_amd.PZ81, poisson_method=dft_amd.POISSON_DIRECT, stop_rule=dft_amd.STOP_REFERENCE,
                  init_guess=dft_amd.INIT_HYDROGENIC, davidson_tol=1e-9, max_iterations=250, mixer=dft_amd.MIX_PULAY)
    r = dft_amd.scf_run(dgrid(og), nuc, **common)
    print("pulay", r["total"], ref["total"], r["iterations"], ref["iterations"])
    assert r["converged"] and abs(r["total"] - ref["total"]) <= 1e-7
    assert r["iterations"] <= ref["iterations"]
    m = dft_amd.scf_run(dgrid(og), nuc, precision=dft_amd.SCF_MIXED, refine_iterations=250,
                        precond=dft_amd.PRECOND_KINETIC_EXACT, **common)
    assert m["converged"] and abs(m["total"] - ref["total"]) <= 1e-7
    # open shell (doublet H3 chain): both spin channels mixed as one vector
    h3 = [[1, -1.61, 0.02, 0.0], [1, 0.05, 0.0, 0.03], [1, 1.66, -0.02, 0.01]]
    lin = dft_amd.scf_run(dgrid(og), h3, multiplicity=2, **dict(common, mixer=dft_amd.MIX_LINEAR))
    pul = dft_amd.scf_run(dgrid(og), h3, multiplicity=2, **common)
    print("doublet", lin["total"], pul["total"], lin["iterations"], pul["iterations"])
    assert lin["converged"] and pul["converged"] and abs(lin["total"] - pul["total"]) <= 1e-7
    # Kerker-screened update (q0 = 1/bohr): the same fixed points, closed and open shell
    kc = dft_amd.scf_run(dgrid(og), nuc, mix_kerker=1.0, **common)
    ko = dft_amd.scf_run(dgrid(og), h3, multiplicity=2, mix_kerker=1.0, **common)
    print("kerker", kc["total"], kc["iterations"], ko["total"], ko["iterations"])
    assert kc["converged"] and abs(kc["total"] - ref["total"]) <= 1e-7
    assert ko["converged"] and abs(ko["total"] - lin["total"]) <= 1e-7


@pytest.mark.parametrize("prec", ["f64", "f32_split"])
def test_davidson_free_box_full_size(prec):
    """Block Davidson at the 40-atom cluster's shapes (256^3 grid, h = 0.18,
    96 states, subspace up to 288 columns: basis offsets beyond 2^31 values)
    with V_eff = 0, where the 7-point operator's spectrum is known in closed
    form: sum over the axes of (1 - cos(pi p / (n + 1))) / h^2, p = 1..n.
    The lowest 80 eigenvalues (degenerate clusters included) must match."""
    import time
    import dft_amd
    n, h, ns, nc = 256, 0.18, 96, 80
    g = dft_amd.Grid.centered(n, h)
    l1 = (1.0 - np.cos(np.pi * np.arange(1, 13) / (n + 1))) / h ** 2
    exact = np.sort((l1[:, None, None] + l1[None, :, None] + l1[None, None, :]).ravel())[:nc]
    kw = dict(prec=dft_amd.F64, gram_mode=dft_amd.GRAM_F64, tol=1e-7) if prec == "f64" else \
        dict(prec=dft_amd.F32, gram_mode=dft_amd.GRAM_SPLIT_BF16, tol=1e-4)
    t = time.perf_counter()
    r = dft_amd.davidson_run(g, np.zeros(g.size), nstates=ns, nconv=nc, max_subspace=3 * ns,
                             precond=dft_amd.PRECOND_KINETIC_EXACT, **kw)
    err = np.max(np.abs(r["eigenvalues"][:nc] - exact))
    print(prec, "iterations", r["iterations"], "seconds", round(time.perf_counter() - t, 2), "max error", err)
    assert err <= (1e-10 if prec == "f64" else 2e-6)


def test_scf_smearing():
    """Fermi-Dirac smearing (no reference counterpart).  A wide-gap molecule
    (H2) with small kT gives the whole-orbital SCF's energy, since every
    occupation is 0 or 1 to rounding.  A spin-unpolarised carbon atom (a
    partly filled, degenerate 2p level: no whole-orbital ground state)
    converges with fractional occupations: -TS lies between 0 and the bound
    -kT 2 k ln 2 of k half-filled states per spin, and the mixed-precision run
    reaches the FP64 free energy.  The occupation solver itself is pinned on
    the CPU (degenerate levels, symmetric spectra)."""
    import dft_amd
    g = dft_amd.Grid.centered(16, 0.42)
    nuc = [[1, -0.71, 0.02, 0.0], [1, 0.69, 0.0, 0.03]]
    common = dict(functional=dft_amd.PZ81, poisson_method=dft_amd.POISSON_DIRECT, stop_rule=dft_amd.STOP_DE,
                  energy_tol=1e-9, init_guess=dft_amd.INIT_HYDROGENIC, davidson_tol=1e-9, max_iterations=200,
                  mixer=dft_amd.MIX_PULAY)
    tight = dict(common, stop_rule=dft_amd.STOP_REFERENCE)  # both runs at the reference's tight stop rule
    r0 = dft_amd.scf_run(g, nuc, nstates=4, **tight)
    r1 = dft_amd.scf_run(g, nuc, nstates=4, smearing=1e-3, **tight)
    print("h2", r0["total"], r1["total"], r1["smearing_energy"], r1["fermi_level"])
    assert r0["converged"] and r1["converged"]
    assert abs(r1["total"] - r0["total"]) <= 1e-9 and abs(r1["smearing_energy"]) <= 1e-100
    assert r1["fermi_level"][0] > r0["eigenvalues"][0]
    gc = dft_amd.Grid.centered(40, 0.3)
    kT = 0.01
    c = [[6, 0.0, 0.0, 0.0]]
    ce = dict(common, energy_tol=1e-8, nstates=8, smearing=kT)
    f64 = dft_amd.scf_run(gc, c, **ce)
    print("carbon", f64["total"], f64["smearing_energy"], f64["iterations"], f64["fermi_level"])
    assert f64["converged"]
    assert -kT * 2 * 8 * np.log(2.0) < f64["smearing_energy"] < 0.0
    mx = dft_amd.scf_run(gc, c, precision=dft_amd.SCF_MIXED, refine_iterations=100,
                         precond=dft_amd.PRECOND_KINETIC_EXACT, **ce)
    print("carbon mixed", mx["total"], mx["iterations"], mx["refine_iterations"])
    assert mx["converged"] and abs(mx["total"] - f64["total"]) <= 1e-6


@pytest.mark.parametrize("init", ["bare", "hydrogenic"])
def test_scf_mixed_neon_finds_the_fp64_state(init):
    """A Ne atom (PBE96), whose 1s level the box-centred seed vectors barely
    overlap: the mixed-precision SCF with the default Davidson settings
    (exact kinetic correction, |dE|-following tolerances) must reach the same
    state and energy as the FP64 SCF with the reference's diagonal
    correction, from either starting guess."""
    import dft_amd
    g = dft_amd.Grid.centered(48, 0.25)
    ne = [[10, 0.0, 0.0, 0.0]]
    kw = dict(functional=dft_amd.PBE96, nstates=8, poisson_method=dft_amd.POISSON_DIRECT, stop_rule=dft_amd.STOP_DE,
              energy_tol=1e-6, init_guess=dft_amd.INIT_BARE if init == "bare" else dft_amd.INIT_HYDROGENIC,
              davidson_tol=1e-7, max_iterations=200)
    ref = dft_amd.scf_run(g, ne, **kw)
    mx = dft_amd.scf_run(g, ne, precision=dft_amd.SCF_MIXED, refine_iterations=50,
                         precond=dft_amd.PRECOND_KINETIC_EXACT, **kw)
    print(init, "fp64", ref["total"], ref["iterations"], "mixed", mx["total"], mx["iterations"], mx["eigenvalues"][:2])
    assert ref["converged"] and mx["converged"]
    assert abs(mx["total"] - ref["total"]) <= 1e-5
    assert abs(mx["eigenvalues"][0] - ref["eigenvalues"][0]) <= 1e-3


@pytest.mark.parametrize("seed", [107, 108])
def test_scf_smearing_converges_random_radicals(seed):
    """Seeded random H/C/N/O molecules within 3 bohr (the benchmark's random
    geometry rule) that never converge with whole-orbital occupations: with
    Fermi-Dirac smearing (kT = 0.01) and Pulay mixing both precisions converge,
    to the same free energy within 1e-5 Ha."""
    import sys, os
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    import dft_amd
    rng = np.random.default_rng(seed)
    atoms = bench.random_cluster(seed, int(rng.integers(3, 7)), 3.0)
    ne = int(sum(a[0] for a in atoms))
    g = dft_amd.Grid.centered(96, 0.25)
    kw = dict(functional=dft_amd.PZ81, nstates=ne // 2 + 10, poisson_method=dft_amd.POISSON_DIRECT,
              stop_rule=dft_amd.STOP_DE, energy_tol=1e-5, init_guess=dft_amd.INIT_HYDROGENIC, davidson_tol=1e-7,
              max_iterations=120, smearing=0.01, mixer=dft_amd.MIX_PULAY)
    ref = dft_amd.scf_run(g, atoms, **kw)
    mx = dft_amd.scf_run(g, atoms, precision=dft_amd.SCF_MIXED, refine_iterations=50,
                         precond=dft_amd.PRECOND_KINETIC_EXACT, **kw)
    print(seed, len(atoms), ne, ref["total"], ref["iterations"], mx["total"], mx["iterations"])
    assert ref["converged"] and mx["converged"]
    assert abs(mx["total"] - ref["total"]) <= 1e-5
