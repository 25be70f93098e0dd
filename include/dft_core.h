/*
 * dft_core.h -- C-ABI of the device-resident solver layer (libdft_core.so).
 *
 * The reference's drop-in C++ interface (applyMolecularKohnShamHamiltonian,
 * solveMolecularOrbitals, calculateHartreePotential, ..., the SCF driver)
 * is declared in dft_molecule.h.  The entry points here run the same SCF
 * (SelfConsistentFieldMolecule.cpp:50-561, MolecularSCFIteration.cpp:12-252)
 * with every field resident on the GPU, plus the configuration knobs a
 * workload needs (Poisson mode, stop rule, initial guess, precision mode,
 * Davidson block size, preconditioner).  Host arrays only cross the
 * boundary at entry (geometry) and exit (results).
 * Status codes: 0 ok, -1 invalid argument, -2 numerical failure or
 * non-convergence, -3 GPU error; dft_core_last_error() holds the message.
 */
#ifndef DFT_CORE_H
#define DFT_CORE_H

#include <stdint.h>

#include "dft_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { DFT_SCF_FP64 = 0, DFT_SCF_MIXED = 1 };
enum { DFT_STOP_REFERENCE = 0, DFT_STOP_DE = 1 };
enum { DFT_INIT_BARE = 0, DFT_INIT_HYDROGENIC = 1 };
/* DIAG: the reference's diagonal correction; CNN: the learned correction with
 * the diagonal one as fallback; KINETIC: 4 damped-Jacobi steps on
 * (-1/2 Lap + 1) t = r (a smoothing correction, no reference counterpart). */
enum { DFT_PRECOND_DIAG = 0, DFT_PRECOND_CNN = 1, DFT_PRECOND_KINETIC = 2, DFT_PRECOND_KINETIC_CHEB = 3,
       DFT_PRECOND_KINETIC_EXACT = 4 };
enum { DFT_POISSON_JACOBI = 0, DFT_POISSON_CHEBYSHEV = 1, DFT_POISSON_DIRECT = 2 };
enum { DFT_MIX_LINEAR = 0, DFT_MIX_PULAY = 1 };

typedef struct {
    int functional;        /* DFT_XC_PZ81 / DFT_XC_PBE96 */
    int charge;
    int multiplicity;      /* 1: closed shell (beta copies alpha, MolecularSCFIteration.cpp:77-90,127-133) */
    int precision;         /* DFT_SCF_FP64 or DFT_SCF_MIXED */
    int nstates;           /* Davidson block per spin (>= occupied); 0: occupied count */
    int max_subspace;      /* 0: max(3 * block, block + 16) */
    int max_iterations;    /* SCF steps; 0: 250 (DFTConstants.h:15) */
    int poisson_fixed;     /* 0: reference rule (tol 1e-8, cap 10000, zero start); k > 0: k warm sweeps */
    int poisson_batch;     /* sweeps between host checks; 0: 64 */
    int poisson_method;    /* DFT_POISSON_JACOBI (the reference's sweeps), DFT_POISSON_CHEBYSHEV or DFT_POISSON_DIRECT */
    double poisson_tol;    /* Chebyshev stop: max|dV| per sweep (0: 1e-8 FP64; FP32 uses >= 1e-5) */
    int stop_rule;         /* DFT_STOP_REFERENCE (dE, dRho, KS triple) or DFT_STOP_DE */
    double energy_tol;     /* |dE| threshold for DFT_STOP_DE */
    int init_guess;        /* DFT_INIT_BARE (reference) or DFT_INIT_HYDROGENIC */
    double davidson_tol;   /* residual threshold; 0: 1e-7 (EigenvalueSolverMolecule.cpp:19) */
    int davidson_max_iter; /* 0: 200 */
    int refine_iterations; /* mixed mode: max FP64 refinement steps after the FP32 phase stops */
    int precond;           /* DFT_PRECOND_DIAG, DFT_PRECOND_KINETIC(_CHEB, _EXACT) or DFT_PRECOND_CNN */
    const void* cnn;       /* trained network (const dft_cnn_model*) for DFT_PRECOND_CNN */
    int verbose;           /* 1: one line per SCF step on stderr */
    int precond_sweeps;    /* DFT_PRECOND_KINETIC(_CHEB): smoothing steps per correction; 0: 4 (8 for _CHEB) */
    double precond_shift;  /* DFT_PRECOND_KINETIC*: sigma in (-1/2 Lap + sigma)^-1; 0: 1.0 (_CHEB 1.5, _EXACT see
                              precond_shift_scale) */
    int mixer;             /* DFT_MIX_LINEAR (the reference's lagged linear mixing) or DFT_MIX_PULAY */
    int mix_depth;         /* DFT_MIX_PULAY: history length; 0: 6 */
    double mix_beta;       /* DFT_MIX_PULAY: residual weight; 0: 0.3 */
    double smearing;       /* Fermi-Dirac width kT in Ha; 0: the reference's whole-orbital (aufbau) occupations.
                              > 0: fractional occupations over the nstates block less its top 4 states (kept as
                              the eigensolver's unoccupied buffer when nstates exceeds the occupied count by 8 or more),
                              and the energy reported and tested by the stop rule is the free energy E - TS */
    int ortho_gram;        /* mixed mode: Gram products of the FP32 orthonormalisation (block Gram-Schmidt, CholQR):
                              0 or DFT_GRAM_SPLIT6_BF16: FP32-level products as six bf16 MFMA products;
                              DFT_GRAM_F32: FP32 products on the FP32 MFMA; DFT_GRAM_SPLIT_BF16: as the
                              Rayleigh-Ritz products */
    double mix_kerker;     /* DFT_MIX_PULAY: Kerker screening wave number q0 (1/bohr) of the update; 0: none */
    double precond_shift_scale; /* DFT_PRECOND_KINETIC_EXACT: per-state shifts sigma_j = max(precond_shift,
                                   -scale theta_j); 0: scale 0.35 with precond_shift 0.5 by default; < 0: one shift
                                   (precond_shift, default 2.0) for every state */
    int xc_flux;           /* DFT_XC_FLUX_REFERENCE (0): the GGA potential's flux c_sigma grad n_sigma as
                              ExchangeCorrelationMolecule.cpp:264-274 builds it; DFT_XC_FLUX_CONSISTENT_SCF (1):
                              c_sigma grad n_sigma / |grad n_sigma|, the derivative of the E_xc the SCF reports */
} dft_scf_config;

enum { DFT_XC_FLUX_REFERENCE = 0, DFT_XC_FLUX_CONSISTENT_SCF = 1 };

typedef struct {
    int iterations;            /* SCF steps taken (all phases) */
    int converged;
    int first_de_iteration;    /* first step with |dE| < energy_tol */
    int refine_iterations;     /* FP64 steps after the mixed phase */
    double energy[6];          /* kinetic, external, hartree, xc, nuclear, total (EnergyComponents, DFTData.h:47-54) */
    double energy_difference;
    double density_difference;
    double ks_residual;
    double seconds_total;      /* wall time of the SCF loop, GPU synchronised */
    double seconds_to_first_de;
    double seconds_poisson, seconds_xc, seconds_davidson, seconds_other;
    int64_t happly_columns;    /* orbital H applications */
    int64_t poisson_sweeps;
    int64_t davidson_iterations;
    int64_t cnn_applied, cnn_fallbacks;
    int64_t block_passes;     /* Davidson block Gram-Schmidt passes (one or two per new block) */
    int64_t block_fallbacks;  /* blocks that needed the rank-revealing orthonormalisation */
    double fermi_level[2];    /* smearing > 0: chemical potential of the alpha and beta channels */
    double smearing_energy;   /* smearing > 0: -TS, included in energy[5] */
} dft_scf_report;

const char* dft_core_last_error(void);

/* Run the SCF.  nuclei: {Z, x, y, z} * natoms.  Optional outputs (NULL to
 * skip): eig (n_alpha + n_beta occupied eigenvalues), trace (5 doubles per
 * step: E, dE, dRho, KS, mixing; capacity max_iterations + refine), rho
 * (2*N: output alpha then beta density). */
int dft_scf_run(const dft_grid* g, int natoms, const double* nuclei, const dft_scf_config* cfg, dft_scf_report* rep,
                double* eig, double* trace, int trace_cap, double* rho);

/* Fields of one SCF step (all FP64 host arrays of N values; NULL to skip). */
typedef struct {
    double* vh;         /* V_H of the input density */
    double* vxc_alpha;  /* V_xc of the input densities */
    double* vxc_beta;
    double* rho_alpha;  /* output spin densities */
    double* rho_beta;
    double* vh_out;     /* V_H of the output density */
} dft_scf_fields;

/* One SCF iteration (MolecularSCFIteration.cpp:12-252) on given input spin
 * densities (host, N each), orbitals started from seed vectors; cfg selects
 * functional, multiplicity, precision, Poisson mode and Davidson block.
 * rep receives energies, dRho and the KS residual of the step; eig the
 * occupied eigenvalues (alpha then beta). */
int dft_scf_step(const dft_grid* g, int natoms, const double* nuclei, const dft_scf_config* cfg,
                 const double* rho_alpha_in, const double* rho_beta_in, dft_scf_report* rep, double* eig,
                 dft_scf_fields* fields);

/* Block Davidson on a given potential (tests and benchmarks).  veff_host: N
 * FP64 values.  prec/gram_mode select the arithmetic.  eig: nstates values;
 * psi (optional): nconv orbitals, N each, normalised sum psi^2 dV = 1 like
 * the reference's MolecularOrbital::psi; resid (optional): nstates residuals.
 * precond: DFT_PRECOND_DIAG / CNN / KINETIC / KINETIC_CHEB / KINETIC_EXACT; cnn: the trained network
 * (dft_cnn_model) for DFT_PRECOND_CNN.  stats (optional, 4 values): H-applied columns,
 * restarts, learned corrections applied, diagonal fallbacks. */
/* Fermi-Dirac occupations of one spin channel: occ[i] = 1 / (1 + exp((eps[i]
 * - mu) / kT)) with mu such that sum occ = nelec (bisection to rounding),
 * and the entropy S = -sum [occ ln occ + (1 - occ) ln(1 - occ)].  kT > 0,
 * 0 <= nelec <= n.  No reference counterpart (whole orbitals only there). */
int dft_fermi_occupations(const double* eps, int n, double nelec, double kT, double* occ, double* mu, double* entropy);

int dft_davidson_run(const dft_grid* g, int prec, int gram_mode, const double* veff_host, int nstates, int nconv,
                     int max_subspace, double tol, int max_iter, int precond, const void* cnn, double* eig,
                     double* psi, double* resid, int* iterations, int64_t* stats);

/* ---- learned preconditioner: model, training-data harvest, training ----
 * No reference counterpart (the reference preconditions with the diagonal
 * correction only, EigenvalueSolverMolecule.cpp:483-555).  The network maps
 * {r / max|r|, (V_eff - theta) h2} to c / max|c| (h2 = 1 / sum(1/h^2); only
 * the direction of a correction column matters to Davidson), where c solves
 * the projected correction equation
 * (I - QQ^T)(H - theta)(I - QQ^T) c = -r, Q = the Ritz vectors up to and
 * including the state's own. */
typedef struct {
    int channels;          /* 16 */
    int64_t count;         /* dft_cnn_param_count(channels) */
    const float* params;   /* FP32 master parameters (host) */
    int base;              /* 0: the network's output is the correction; 1: it is added, scaled by max|r| h2, to the
                              exact kinetic correction (-1/2 Lap + sigma_j)^-1 r over the molecule's box (harvested
                              with target_scale 2) */
} dft_cnn_model;

typedef struct {
    int patch;             /* patch edge (32) */
    int max_columns;       /* residual columns harvested per Davidson correction step */
    int patches_per_column;
    int every;             /* harvest at every k-th correction step of the solve */
    double cg_tol;         /* relative residual of the projected CG */
    int cg_max_iter;
    uint64_t seed;         /* patch jitter */
    int target_scale;      /* 0: target c / max|c|;  1: c / (max|r| h2), the input's scale;
                              2: the Davidson runs with the exact kinetic correction t_kin (per-state shifts) and the
                              target is (beta c - t_kin) / (max|r| h2), beta = <t_kin, c> / <c, c>: what the network
                              must add to t_kin (a model with base 1) */
} dft_cnn_harvest_config;

typedef struct dft_cnn_trainer dft_cnn_trainer;

/* A trainer owns a context, a device patch set of `capacity` patches and the
 * network (parameters initialised from `seed`, He-uniform). */
int dft_cnn_trainer_create(int channels, int patch, int64_t capacity, uint64_t seed, dft_cnn_trainer** out);
void dft_cnn_trainer_destroy(dft_cnn_trainer* t);
/* Run an FP64 SCF (cfg) and harvest training patches from its Davidson
 * corrections; *added = patches added (stops adding when full). */
int dft_cnn_trainer_harvest(dft_cnn_trainer* t, const dft_grid* g, int natoms, const double* nuclei,
                            const dft_scf_config* cfg, const dft_cnn_harvest_config* hc, dft_scf_report* rep,
                            int64_t* added);
int dft_cnn_trainer_count(dft_cnn_trainer* t, int64_t* patches);
/* Totals over the trainer's harvests: columns solved by the projected CG,
 * columns that failed (no convergence or negative curvature), CG iterations,
 * patches added, and the largest true relative residual |b - A c| / |b| of a
 * converged column, recomputed from c after the CG (the CG stops on its
 * recurrence's residual <= cg_tol). */
typedef struct {
    int64_t solves;
    int64_t failed;
    int64_t cg_iterations;
    int64_t patches;
    double max_true_residual;
} dft_cnn_harvest_stats;
int dft_cnn_trainer_harvest_stats(dft_cnn_trainer* t, dft_cnn_harvest_stats* out);
/* Adam steps on random batches of the first `train_count` patches
 * (0: all); loss (optional) receives one value per step. */
int dft_cnn_trainer_train(dft_cnn_trainer* t, int steps, int batch, float lr, int64_t train_count, double* loss);
/* Mean loss over patches [first, first + count) in batches. */
int dft_cnn_trainer_eval(dft_cnn_trainer* t, int64_t first, int64_t count, double* loss);
int dft_cnn_trainer_get_params(dft_cnn_trainer* t, float* params, int64_t count);
int dft_cnn_trainer_set_params(dft_cnn_trainer* t, const float* params, int64_t count);
/* Copy patches [first, first + count) to the host: in (count x 2 x P^3), tgt (count x P^3). */
int dft_cnn_trainer_get_patches(dft_cnn_trainer* t, int64_t first, int64_t count, float* in, float* tgt);
/* Append count patches from the host (same layout) after the patches held,
 * e.g. from a patch shard (dft_io.h); -1 when they exceed the free capacity. */
int dft_cnn_trainer_add_patches(dft_cnn_trainer* t, int64_t count, const float* in, const float* tgt);

#ifdef __cplusplus
}
#endif
#endif
