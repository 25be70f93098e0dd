/*
 * ks_oracle.h -- FP64 CPU restatement of the reference's 3D real-space
 * Kohn-Sham path (willandru/DFT, src/dft/*Molecule*.cpp, PZ81.cpp, PBE96.cpp).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP path is
 * compared against.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load it; the product library never links or calls it.
 *
 * Every function restates the reference's arithmetic in the same operation
 * order (built with -O2 -ffp-contract=off, like the reference's ISO C++20
 * build), so results agree bitwise with the compiled reference on the same
 * libm.  Citations are reference file:line.
 *
 * Errors: functions return 0 on success, -1 for an invalid argument (the
 * reference's std::invalid_argument), -2 for a numerical failure
 * (std::runtime_error).  og_last_error() holds the message.
 */
#ifndef KS_ORACLE_H
#define KS_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* CartesianGrid.cpp:5-159 -- spacing (max-min)/(n-1), x fastest. */
typedef struct {
    long nx, ny, nz;
    double xmin, xmax, ymin, ymax, zmin, zmax;
    double dx, dy, dz;
} og_grid;

int og_grid_init(og_grid* g, long nx, long ny, long nz, double xmin, double xmax,
                 double ymin, double ymax, double zmin, double zmax);
const char* og_last_error(void);

/* KohnShamHamiltonianMolecule.cpp:7-183 */
int og_hpsi(const og_grid* g, const double* veff, const double* psi, double* out);

/* HartreePotentialMolecule.cpp:37-95 boundary rule; 99-324 Jacobi.
 * zero_start=1: the reference call (boundary rule, interior 0).
 * zero_start=0: warm start, interior taken from v on entry (boundary still
 * set by the rule).  Stops after a sweep whose max|dV| < tol or after
 * max_sweeps sweeps; v holds the latest iterate. */
int og_poisson(const og_grid* g, const double* rho, double* v, long max_sweeps,
               double tol, int zero_start, long* sweeps_done, double* last_maxdiff);
/* HartreePotentialMolecule.cpp:326-379 */
double og_hartree_energy(const og_grid* g, const double* rho, const double* vh);

/* XCFunctional.h:5-21 XCInput -> XCResult: in = {na, nb, |grad na|, |grad nb|},
 * out = {eps_xc, v_a, v_b, c_a, c_b}.  id 1 = PZ81 (PZ81.cpp:437-501),
 * id 2 = PBE96 (PBE96.cpp:814-966). */
int og_xc_point(int functional, const double in[4], double out[5]);

/* ExchangeCorrelationMolecule.cpp:576-747 */
int og_xc_potential(const og_grid* g, int functional, const double* na,
                    const double* nb, int spin, double* v);
/* As og_xc_potential; consistent = 1 replaces the flux c_sigma grad n_sigma
 * by c_sigma grad n_sigma / |grad n_sigma|, the functional derivative's flux
 * (build option; no counterpart in the reference). */
int og_xc_potential_flux(const og_grid* g, int functional, const double* na,
                         const double* nb, int spin, int consistent, double* v);
/* The per-point terms n eps_xc dV of og_xc_energy (0 at or below the floor), so that a test can sum
 * them exactly (math.fsum) instead of in the reference's sequential order. */
int og_xc_energy_terms(const og_grid* g, int functional, const double* na, const double* nb, double* terms);
/* ExchangeCorrelationMolecule.cpp:749-842 */
double og_xc_energy(const og_grid* g, int functional, const double* na, const double* nb);

/* ElectronDensityMolecule.cpp:5-76 (orbitals stored column-major, N each). */
int og_spin_density(long n, int norb, const double* psi, const int* electrons, double* out);
/* ElectronDensityMolecule.cpp:78-114: n_alpha + n_beta; spins[o] 0 alpha, 1 beta. */
int og_electron_density(long n, int norb, const double* psi, const int* electrons, const int* spins, double* out);
/* ElectronDensityMolecule.cpp:116-140: sum(rho) dx dy dz. */
int og_integrate_density(long n, const double* rho, double dx, double dy, double dz, double* electrons);

/* MolecularSCFHelpers.cpp:119-183: max over the orbitals of ||H psi - eps psi|| sqrt(dV), H built
 * on va for spins[o] == 0 and on vb otherwise; orbitals column-major, N each. */
int og_ks_residual(const og_grid* g, const double* va, const double* vb, int norb, const double* psi,
                   const double* eig, const int* spins, double* out);
/* TotalEnergyMolecule.cpp:13-199 (Neumann-like stencil) */
double og_kinetic_energy(const og_grid* g, int norb, const double* psi, const int* electrons);
/* TotalEnergyMolecule.cpp:201-235 */
double og_external_energy(const og_grid* g, const double* rho, const double* vnuc);

/* NuclearPotential.cpp:6-76; nuclei = {Z, x, y, z} * natoms. */
int og_nuclear_potential(const og_grid* g, int natoms, const double* nuclei, double* v);
/* NuclearRepulsion.cpp:6-54 */
int og_nuclear_repulsion(int natoms, const double* nuclei, double* e);

/* EigenvalueSolverMolecule.cpp:666-1300: single-vector Davidson, orbitals
 * solved one after another with deflation.  init may be NULL (seed vectors).
 * psi_out: norb columns of N; eig_out: norb values; happly_out (optional):
 * H applications per orbital. */
int og_davidson(const og_grid* g, const double* veff, int norb, const int* occupations,
                const double* init, long max_happly, double* psi_out, double* eig_out,
                long* happly_out);

/* SCF driver: SelfConsistentFieldMolecule.cpp:50-561, MolecularSCFIteration.cpp:12-252,
 * MolecularSCFHelpers.cpp:27-249,515-579, with the knobs a configuration needs. */
typedef struct {
    int functional;        /* 1 PZ81, 2 PBE96 */
    int charge;
    int multiplicity;
    int max_iterations;    /* reference: 250 */
    int poisson_fixed;     /* 0: reference rule (tol 1e-8, cap 10000, zero start);
                              k>0: exactly k sweeps, warm-started from the last V_H */
    int stop_rule;         /* 0: reference triple criterion; 1: |dE| < energy_tol only */
    double energy_tol;     /* used by stop_rule 1 */
    int init_guess;        /* 0: bare-nucleus orbitals (reference); 1: hydrogenic 1s density */
    long max_happly;       /* Davidson cap per orbital, reference 5000 */
} og_scf_config;

typedef struct {
    int iterations;
    int converged;
    double energy[6];      /* kinetic, external, hartree, xc, nuclear, total */
    double energy_difference;
    double density_difference;
    double ks_residual;
    double eig_alpha0;     /* lowest alpha eigenvalue */
    double eig_beta0;
    int first_de_iteration; /* first step with |dE| < energy_tol (0 if none) */
} og_scf_result;

/* trace (optional): per step {E_total, dE, dRho, KS, mixing}; capacity max_iterations.
 * eig_out (optional): n_alpha + n_beta final eigenvalues.
 * rho_out (optional): 2 * N, final output alpha then beta densities. */
int og_scf(const og_grid* g, int natoms, const double* nuclei, const og_scf_config* cfg,
           og_scf_result* res, double* trace, double* eig_out, double* rho_out);

/* One SCF iteration on given input spin densities (MolecularSCFIteration.cpp:12-252)
 * with the reference Poisson rule and seed-vector Davidson starts.  Outputs
 * (all N unless noted): vh, vxa, vxb, ora, orb, ovh; eig (n_alpha + n_beta);
 * E[6]; dd (effective density difference); ks (max KS residual). */
int og_scf_step(const og_grid* g, int natoms, const double* nuclei, int functional, int multiplicity,
                const double* ra, const double* rb, double* vh, double* vxa, double* vxb, double* ora,
                double* orb, double* ovh, double* eig, double* E, double* dd, double* ks);

#ifdef __cplusplus
}
#endif

/* ---- radial (atomic) stack, oracle/radial_oracle.c ------------------------
 * Grid r_i = (i + 1) rmax / (points + 1), i < points (RadialGrid.cpp:5-20). */
/* ElectronicConfiguration.cpp:16-145: subshells (n, l, alpha, beta) of Z = 1..30; returns their count */
int og_radial_config(int Z, int* n, int* l, int* alpha, int* beta);
/* EigenvalueSolver.cpp:314-403 solveOrbital: eigenvalue `index` of the tridiagonal H(l) with
 * potential v, and its eigenvector u normalised to dr sum u^2 = 1 */
int og_radial_orbital(long points, double rmax, const double* v, int l, long index, double* eig, double* u);
/* HartreePotential.cpp:8-86 */
int og_radial_hartree(long points, double rmax, const double* rho, double* vh);
/* ExchangeCorrelation.cpp:167-279 (spin 0 or 1) and 281-370 */
int og_radial_xc_potential(long points, double rmax, int functional, const double* na, const double* nb, int spin,
                           double* v);
double og_radial_xc_energy(long points, double rmax, int functional, const double* na, const double* nb);
/* AtomicDFT.cpp:6-41 / SelfConsistentField.cpp: the atom's SCF.  res[8] = {kinetic, external,
 * hartree, xc, total, dRho, |dE|, KS residual} of the last iteration; eig: orbital eigenvalues,
 * alpha then beta (<= 14); the output spin densities if na_out / nb_out are given.  Returns the
 * iteration count (negated if not converged in 250), < -1000 on an error. */
int og_radial_scf(int Z, int functional, long points, double rmax, double* res, double* eig, int* norb_out,
                  double* na_out, double* nb_out);
#endif
