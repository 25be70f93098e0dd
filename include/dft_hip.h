/*
 * dft_hip.h -- C-ABI of the gfx950 kernel library (libdft_hip.so).
 *
 * The reference (willandru/DFT) has no FFI: its 3D Kohn-Sham hot path is a
 * set of free C++ functions on std::vector<double>.  Each entry point below
 * replaces the loop nest of one of those functions; the C++ layer in
 * include/dft_molecule.h keeps the reference's own names and signatures on
 * top of these.  Conventions:
 *   - every function returns an int status: 0 ok, -1 invalid argument,
 *     -2 numerical failure / non-convergence, -3 HIP error;
 *     dft_last_error(ctx) holds the message;
 *   - field/orbital pointers are DEVICE pointers from dft_alloc (or any
 *     device allocation on the context's device); host pointers are named
 *     *_host;
 *   - fields are nx*ny*nz values with x fastest (index (k*ny+j)*nx+i,
 *     CartesianGrid.cpp:145-159); an orbital block is column-major, column
 *     o at psi + o*ld, ld >= nx*ny*nz;
 *   - all work is queued on the context's stream; functions that return a
 *     host scalar synchronise that stream;
 *   - one context per host thread; calls on one context are not re-entrant.
 */
#ifndef DFT_HIP_H
#define DFT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dft_ctx dft_ctx;

enum { DFT_OK = 0, DFT_EINVAL = -1, DFT_ENUMERIC = -2, DFT_EHIP = -3 };
/* Arithmetic/storage precision of a field or orbital block. */
enum { DFT_F64 = 0, DFT_F32 = 1 };
/* Gram product modes (dft_gram). */
enum { DFT_GRAM_F64 = 0, DFT_GRAM_F32 = 1, DFT_GRAM_SPLIT_BF16 = 2, DFT_GRAM_SPLIT6_BF16 = 3 };
/* XCFunctional subclasses (XCFunctional.h:15-21, PZ81.h, PBE96.h). */
enum { DFT_XC_PZ81 = 1, DFT_XC_PBE96 = 2 };
/* Flag OR-ed into a functional id for dft_xc: the GGA flux is
 * c_sigma grad n_sigma / |grad n_sigma| (the functional derivative of E_xc)
 * instead of the reference's c_sigma grad n_sigma
 * (ExchangeCorrelationMolecule.cpp:264-274 with c_sigma = d f / d|grad n_sigma|,
 * PBE96.cpp:928-957).  No counterpart in the reference; PZ81 is unaffected. */
enum { DFT_XC_FLUX_CONSISTENT = 0x100 };
/* Reductions (dft_reduce). */
enum { DFT_RED_SUM = 0, DFT_RED_DOT = 1, DFT_RED_MAXABS = 2, DFT_RED_DIFF2 = 3 };

/* CartesianGrid (CartesianGrid.h:5-61) as plain data; dft_grid_init sets
 * dx = (xmax-xmin)/(nx-1) exactly as CartesianGrid.cpp:45-47. */
typedef struct {
    int64_t nx, ny, nz;
    double xmin, xmax, ymin, ymax, zmin, zmax;
    double dx, dy, dz;
} dft_grid;

int dft_grid_init(dft_grid* g, int64_t nx, int64_t ny, int64_t nz, double xmin, double xmax, double ymin,
                  double ymax, double zmin, double zmax);

/* ---- context, memory, stream ---------------------------------------- */
int dft_ctx_create(int device, dft_ctx** out);
int dft_ctx_destroy(dft_ctx* ctx);
const char* dft_last_error(const dft_ctx* ctx);
int dft_sync(dft_ctx* ctx);
/* the context's hipStream_t, for callers that interleave their own work */
void* dft_stream(dft_ctx* ctx);
int dft_alloc(dft_ctx* ctx, size_t bytes, void** dev_ptr);
int dft_free(dft_ctx* ctx, void* dev_ptr);
int dft_upload(dft_ctx* ctx, void* dev_dst, const void* host_src, size_t bytes);
int dft_download(dft_ctx* ctx, void* host_dst, const void* dev_src, size_t bytes);
int dft_copy(dft_ctx* ctx, void* dev_dst, const void* dev_src, size_t bytes);
int dft_zero(dft_ctx* ctx, void* dev_ptr, size_t bytes);
/* y[i] = (prec_dst) x[i], n elements */
int dft_convert(dft_ctx* ctx, const void* x, int prec_src, void* y, int prec_dst, int64_t n);
/* Timing on the context stream: events are opaque. */
int dft_event_create(dft_ctx* ctx, void** ev);
int dft_event_record(dft_ctx* ctx, void* ev);
int dft_event_elapsed_ms(dft_ctx* ctx, void* start, void* stop, float* ms);
int dft_event_destroy(dft_ctx* ctx, void* ev);

/* ---- (a) H psi ---------------------------------------------------------
 * Replaces applyMolecularKohnShamHamiltonian (KohnShamHamiltonianMolecule.cpp:7-183):
 * out[:,o] = -1/2 Lap_h psi[:,o] + veff .* psi[:,o] for nb orbitals, zero
 * ghosts outside the box.  All three arrays in precision prec. */
int dft_hpsi(dft_ctx* ctx, const dft_grid* g, int prec, const void* veff, const void* psi, int64_t ld, int nb,
             void* out, int64_t ldo);

/* Kinetic energy sum_o e_o <psi_o| -1/2 Lap_Neumann |psi_o> dV with the
 * reference's face rule (TotalEnergyMolecule.cpp:13-199).  electrons_host:
 * nb ints.  Result in FP64 on the host. */
int dft_kinetic_energy(dft_ctx* ctx, const dft_grid* g, int prec, const void* psi, int64_t ld, int nb,
                       const int* electrons_host, double* t_host);

/* The same sum with fractional weights (occ_host: nb doubles, electrons per
 * orbital), for smeared occupations.  No reference counterpart: the
 * reference occupies whole orbitals only. */
int dft_kinetic_energy_occ(dft_ctx* ctx, const dft_grid* g, int prec, const void* psi, int64_t ld, int nb,
                           const double* occ_host, double* t_host);

/* Davidson diagonal preconditioner (EigenvalueSolverMolecule.cpp:483-555):
 * t[:,o] = r[:,o] / (eps_o - (sum 1/h^2 + veff)), or r where |denominator| < 1e-10
 * or not finite.  eps_dev: nb FP64 values on the device. */
int dft_precond_diag(dft_ctx* ctx, const dft_grid* g, int prec, const void* veff, const void* r, int64_t ldr,
                     int nb, const double* eps_dev, void* t, int64_t ldt);
/* Kinetic smoothing correction (no reference counterpart; DFT_PRECOND_KINETIC
 * in dft_core.h): t = r / D, D = sum(1/h^2) + sigma, then `sweeps` damped
 * Jacobi steps t <- t + (omega / D)(r - (-1/2 Lap + sigma) t), each a fused
 * stencil pass.  tmp: nb columns of N (leading dimension N) of scratch. */
int dft_precond_kinetic(dft_ctx* ctx, const dft_grid* g, int prec, double sigma, double omega, int sweeps,
                        const void* r, int64_t ldr, int nb, void* t, int64_t ldt, void* tmp);
/* The same correction with the steps combined by Chebyshev semi-iteration
 * (D is the centre of the spectrum of -1/2 Lap + sigma, so the Jacobi
 * iteration's spectral radius rho = sum cos(pi/(n+1))/h^2 / D is known):
 * t_1 = r / D, then `sweeps` fused stencil passes of the three-term
 * recurrence.  tmp: nb columns of N of scratch. */
int dft_precond_kinetic_cheb(dft_ctx* ctx, const dft_grid* g, int prec, double sigma, int sweeps, const void* r,
                             int64_t ldr, int nb, void* t, int64_t ldt, void* tmp);
/* The limit of both: t = (-1/2 Lap + sigma)^{-1} r (zero ghosts) applied
 * exactly by sine transforms along x, y, z, as BF16 GEMMs with FP32
 * accumulation on the matrix cores (a correction direction needs ~2 digits).
 * No tmp argument: the context's workspace holds the BF16 blocks. */
int dft_precond_kinetic_exact(dft_ctx* ctx, const dft_grid* g, int prec, double sigma, const void* r, int64_t ldr,
                              int nb, void* t, int64_t ldt);
/* The same with a shift per column: sigma_j = max(sigma_min, -scale theta[j])
 * (theta: nb device FP64 Ritz values), so each state's correction solves
 * (-1/2 Lap + sigma_j) t = r with sigma_j following its own eigenvalue. */
int dft_precond_kinetic_exact_shifted(dft_ctx* ctx, const dft_grid* g, int prec, double sigma_min, double scale,
                                      const double* theta, const void* r, int64_t ldr, int nb, void* t, int64_t ldt);
/* One damped-Jacobi step of the kinetic smoothing correction (no reference
 * counterpart; see dft_core.h DFT_PRECOND_KINETIC): first = 1: t = cf r;
 * otherwise t = t + cf (r - kt), per column. */
int dft_smooth_step(dft_ctx* ctx, int prec, int64_t n, int nb, double cf, const void* r, int64_t ldr, const void* kt,
                    int64_t ldk, void* t, int64_t ldt, int first);
/* x[0..n) = value */
int dft_fill(dft_ctx* ctx, int prec, int64_t n, double value, void* x);

/* ---- (b) Poisson ---------------------------------------------------------
 * Replaces calculateHartreePotential (HartreePotentialMolecule.cpp:99-324).
 * Boundary values follow the reference rule (:37-95: 0 on a face, else
 * Q/|r|).  zero_start=1 is the reference call (interior starts at 0);
 * zero_start=0 warm-starts from v on entry.  Sweeps stop after the first
 * sweep whose max|dV| over the interior is < tol (tol<=0: never), or after
 * max_sweeps; v holds that sweep's iterate (the exact reference stop, found
 * by replaying a batch that overshoots).  sweeps_per_launch >=1 groups
 * launches between host checks.  prec F32 runs the sweeps in FP32 on FP32
 * rho/v. */
int dft_poisson(dft_ctx* ctx, const dft_grid* g, int prec, const void* rho, void* v, int64_t max_sweeps, double tol,
                int zero_start, int sweeps_per_launch, int64_t* sweeps_done, double* last_maxdiff);
/* The same discrete problem (boundary rule of HartreePotentialMolecule.cpp:37-95,
 * Jacobi splitting of :295-322) solved by Chebyshev semi-iteration on the
 * Jacobi sweeps: O(n) sweeps to a given accuracy instead of the reference's
 * O(n^2).  Not the reference's iterate sequence; its fixed point.  Stops
 * after the batch (sweeps between host checks) in which max|dV| < tol. */
int dft_poisson_chebyshev(dft_ctx* ctx, const dft_grid* g, int prec, const void* rho, void* v, int64_t max_sweeps,
                          double tol, int zero_start, int batch, int64_t* sweeps_done, double* last_maxdiff);
/* The same fixed point solved directly: interior equations Lap_h V = -4 pi rho
 * with the boundary values of the reference's rule, diagonalised by 1D
 * discrete sine transforms (dense FP64 GEMMs along each axis, the operator's
 * eigenvalues divided out).  The converged limit of dft_poisson's sweeps to
 * rounding; FP64 internally for either precision of rho/v. */
int dft_poisson_direct(dft_ctx* ctx, const dft_grid* g, int prec, const void* rho, void* v);
/* Hartree energy 1/2 sum rho V dV (HartreePotentialMolecule.cpp:326-379). */
int dft_hartree_energy(dft_ctx* ctx, const dft_grid* g, int prec, const void* rho, const void* vh, double* e_host);

/* ---- (c) exchange-correlation -------------------------------------------
 * Replaces calculateMolecularSpinExchangeCorrelationPotential
 * (ExchangeCorrelationMolecule.cpp:576-747) for both spins in one call
 * (functional may carry DFT_XC_FLUX_CONSISTENT):
 * va/vb receive v_alpha/v_beta (vb may be NULL to compute alpha only, the
 * reference's closed-shell call).  exc_host (may be NULL) receives the
 * energy of calculateMolecularExchangeCorrelationEnergy (:749-842). */
int dft_xc(dft_ctx* ctx, const dft_grid* g, int prec, int functional, const void* na, const void* nb, void* va,
           void* vb, double* exc_host);
/* dft_xc without any allocation: the same kernels and results, with the
 * energy partials in a fixed device array of 16384 doubles (one per GPU,
 * held until the process ends) instead of context memory.  No counterpart in
 * the reference; it is for callers that cannot allocate (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Grids whose
 * launch needs more than 16383 workgroup partials are rejected with
 * DFT_EINVAL.  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (several contexts or threads) race; each call overwrites
 * every partial it reads, so nothing needs clearing between calls. */
int dft_xc_noalloc(dft_ctx* ctx, const dft_grid* g, int prec, int functional, const void* na, const void* nb,
                   void* va, void* vb, double* exc_host);
/* Pointwise XCFunctional::evaluate for npts inputs {na, nb, |grad na|,
 * |grad nb|} (FP64, device) -> {eps, v_a, v_b, c_a, c_b}. */
int dft_xc_points(dft_ctx* ctx, int functional, const double* in_dev, int64_t npts, double* out_dev);

/* ---- (d) density and reductions ------------------------------------------
 * calculateMolecularSpinDensity (ElectronDensityMolecule.cpp:5-76):
 * n = (accumulate ? n : 0) + sum_o occ_o psi_o^2; occ_host: nb doubles. */
int dft_density(dft_ctx* ctx, int prec, const void* psi, int64_t ld, int nb, const double* occ_host, int64_t n_pts,
                void* n_out, int accumulate);
/* FP64 reductions over n elements of a (and b): SUM a, DOT a.b, MAXABS a,
 * DIFF2 sum (a-b)^2.  Result on the host. */
int dft_reduce(dft_ctx* ctx, int op, int prec, const void* a, const void* b, int64_t n, double* result_host);
/* Elementwise: y = alpha*x + beta*y */
int dft_axpby(dft_ctx* ctx, int prec, int64_t n, double alpha, const void* x, double beta, void* y);
/* Elementwise y = max(y, lo) (a NaN becomes lo): keeps an extrapolated
 * (Pulay-mixed) density non-negative.  No reference counterpart: the
 * reference's linear mixing of non-negative densities needs none. */
int dft_clamp_min(dft_ctx* ctx, int prec, int64_t n, double lo, void* y);
/* out = a + b + c (c may be NULL) -- V_eff assembly (MolecularSCFIteration.cpp:92-115) */
int dft_add3(dft_ctx* ctx, int prec, int64_t n, const void* a, const void* b, const void* c, void* out);
/* NuclearPotential::calculate (NuclearPotential.cpp:6-76): nuclei_host =
 * {Z,x,y,z} per atom.  Returns -2 if a nucleus sits exactly on a grid point. */
int dft_nuclear_potential(dft_ctx* ctx, const dft_grid* g, int prec, int natoms, const double* nuclei_host, void* v);

/* ---- radial (atomic) stack ------------------------------------------------
 * The reference's one-centre solver (AtomicDFT.cpp, SelfConsistentField.cpp,
 * EigenvalueSolver.cpp, HartreePotential.cpp, ExchangeCorrelation.cpp,
 * ElectronDensity.cpp, TotalEnergy.cpp) on the radial grid
 * r_i = (i + 1) rmax / (points + 1), i < points (RadialGrid.cpp:5-20). */
/* getAtomicConfiguration (ElectronicConfiguration.cpp:50-145): the subshells
 * (n, l, alpha, beta electrons; at most 7) of Z = 1..30; returns their count
 * or -1. */
int dft_radial_config(int Z, int* n, int* l, int* alpha, int* beta);
/* solveAtom (AtomicDFT.cpp:6-41) for njobs atoms at once, one workgroup
 * each: Z[j] in 1..30, functional[j] DFT_XC_PZ81 or DFT_XC_PBE96, 2 <= points
 * <= 2048 (the reference's default: 2000 points, rmax 30).  results: njobs
 * rows of 26 doubles {kinetic, external, hartree, xc, total, dRho, |dE|, KS
 * residual, iterations, converged (0/1), status, final mixing factor,
 * eigenvalues of the occupied orbitals (alpha subshells, then beta; 14
 * slots)}.  densities (optional, host): njobs x 2 x points, the last output
 * alpha and beta densities.  The SCF stops as the reference's does (dRho <
 * 1e-9, |dE| < 1e-11, KS < 1e-7, at most 250 iterations).  Returns -2 when a
 * job met a singular tridiagonal pivot (its status is nonzero). */
int dft_radial_scf(dft_ctx* ctx, int njobs, const int* Z, const int* functional, int64_t points, double rmax,
                   double* results, double* densities);

/* ---- (e) Gram / Rayleigh-Ritz products ------------------------------------
 * out (device, FP64, column-major ma x mb, leading dim ldo) =
 *   alpha * A^T B  (+ beta * out), A: N x ma (lda), B: N x mb (ldb).
 * DFT_GRAM_F64: FP64 operands on FP64 MFMA.  DFT_GRAM_F32: FP32 operands,
 * FP32 FMA.  DFT_GRAM_SPLIT_BF16: FP32 operands split x = hi + lo (bf16),
 * A^T B ~ hi.hi + hi.lo + lo.hi on bf16 MFMA with FP32 accumulation,
 * partial sums combined in FP64.  DFT_GRAM_SPLIT6_BF16: FP32 operands split
 * exactly into three bf16 pieces, the six products of order <= 2^-18 on bf16
 * MFMA: FP32-product accuracy (as DFT_GRAM_F32) at the bf16 rate; outputs up
 * to 96 x 32 run as DFT_GRAM_F32. */
int dft_gram(dft_ctx* ctx, int mode, int64_t n, int ma, int mb, double alpha, const void* a, int64_t lda,
             const void* b, int64_t ldb, double beta, double* out, int64_t ldo);
/* Block update Y = alpha * A C + beta * Y: A N x m (prec), C m x k FP64
 * device column-major (ldc), Y N x k (prec).  Y may overlap A (an in-place
 * update of the last columns of a basis block). */
int dft_block_update(dft_ctx* ctx, int prec, int64_t n, int m, int k, double alpha, const void* a, int64_t lda,
                     const double* c, int64_t ldc, double beta, void* y, int64_t ldy);
/* 1 when dft_block_update with this precision and shape runs as one kernel
 * pass that may overwrite A in place (Y = the first k columns of A) without
 * a staging copy; 0 when an overlapping Y would be staged through scratch
 * memory (n k values).  A helper for callers that choose between an
 * in-place update and separate buffers; no reference counterpart. */
int dft_block_update_in_place(int prec, int m, int k);
/* Per-column scaling y[:,j] = x[:,j] * s[j] (s device FP64, nb values);
 * and residual r[:,j] = hx[:,j] - theta[j] x[:,j]. */
int dft_scale_columns(dft_ctx* ctx, int prec, int64_t n, int nb, const void* x, int64_t ldx, const double* s,
                      void* y, int64_t ldy);
int dft_residual(dft_ctx* ctx, int prec, int64_t n, int nb, const void* hx, int64_t ldh, const void* x, int64_t ldx,
                 const double* theta, void* r, int64_t ldr);
/* Davidson residuals from the basis: r[:,j] = w q[:,j] - theta[j] v q[:,j]
 * (w = H v, N x m; q m x k device FP64, ldq >= m; theta device FP64) and
 * rn2[j] = ||r[:,j]||^2 (device FP64, k values), in one pass over v and w
 * when 2m x k coefficients fit the kernel's LDS budget (two passes and a
 * norm pass otherwise).  Replaces the residual and norm loop of the
 * reference's Davidson (src/dft/EigenvalueSolverMolecule.cpp:151-175, 893-903) for a
 * block, without forming the Ritz vectors. */
int dft_ritz_residual(dft_ctx* ctx, int prec, int64_t n, int m, int k, const void* v, int64_t ldv, const void* w,
                      int64_t ldw, const double* q, int64_t ldq, const double* theta, void* r, int64_t ldr,
                      double* rn2);
/* Column norms^2 (FP64 device out, nb values) */
int dft_column_norms2(dft_ctx* ctx, int prec, int64_t n, int nb, const void* x, int64_t ldx, double* out);

/* ---- dense FP64 subspace algebra on the device (m <= 1024) --------------- */
/* Symmetric eigensolve of an m x m FP64 device matrix (column-major, ld):
 * eigenvalues ascending into w (device), eigenvectors overwrite a. */
int dft_syev(dft_ctx* ctx, int m, double* a, int64_t lda, double* w);
/* The k lowest eigenpairs of a symmetric m x m FP64 device matrix (lower
 * triangle read, a unchanged): eigenvalues ascending in w (room for m
 * values), eigenvectors in the first k columns of z (ldz >= m, z != a).
 * One launch, no host synchronisation, when m <= 96 and
 * m (m + 1) + 2 k m <= 19000 (Householder tridiagonalisation, Sturm
 * multisection, inverse iteration, all in LDS); rocSOLVER dsyevd otherwise.
 * The Rayleigh-Ritz step of the block Davidson; replaces the reference's
 * Jacobi `diagonalize` (src/dft/EigenvalueSolverMolecule.cpp:290-449), which
 * returns the lowest pair of the projected matrix. */
int dft_syev_lowest(dft_ctx* ctx, int m, int k, const double* a, int64_t lda, double* z, int64_t ldz, double* w);
/* Cholesky S = L L^T of an SPD m x m device matrix, then a <- L^{-T}
 * (upper triangular, zeros below): for S = T^T T, T * a has orthonormal
 * columns (CholQR).  Returns -2 if S is not positive definite. */
int dft_chol_inv(dft_ctx* ctx, int m, double* a, int64_t lda);

/* Small FP64 device transpose: dst(j,i) = src(i,j), src rows x cols. */
int dft_transpose(dft_ctx* ctx, int rows, int cols, const double* src, int64_t lds, double* dst, int64_t ldd);

/* ---- setup fields -------------------------------------------------------- */
/* Davidson seed vectors buildInitialVector (EigenvalueSolverMolecule.cpp:177-283)
 * for indices first_index .. first_index+count-1 into out columns. */
int dft_seed_vectors(dft_ctx* ctx, const dft_grid* g, int prec, int first_index, int count, void* out, int64_t ld);
/* count columns of deterministic pseudo-random values in [-1, 1) (splitmix64
 * of seed, column and point): completes a seed block that is numerically
 * rank-deficient (no reference counterpart). */
int dft_random_vectors(dft_ctx* ctx, int prec, int64_t n, int count, uint64_t seed, void* out, int64_t ld);
/* Initial density guess: sum over atoms of Z * hydrogenic 1s density
 * (Z^3/pi) exp(-2 Z r), normalised on the grid to `electrons`. */
int dft_hydrogenic_density(dft_ctx* ctx, const dft_grid* g, int prec, int natoms, const double* nuclei_host,
                           double electrons, void* out);

/* ---- learned preconditioner: 3D CNN (BF16 weights, FP32 accumulation) ----
 * No reference counterpart: the network augments the reference's
 * davidsonCorrection (EigenvalueSolverMolecule.cpp:483-555) at its call site
 * (:1042-1047), with the diagonal correction as the fallback.
 * Network: 4 layers of 3x3x3 convolutions, channels 2 -> C -> C -> C -> 1,
 * SiLU after layers 1-3, zero padding.  Parameters: one array, for each
 * layer W[cout][cin][27] (tap = kz*9 + ky*3 + kx; the output at v reads the
 * input at v + (kx-1, ky-1, kz-1)) followed by b[cout]; FP32 master copy
 * plus a BF16 copy that the kernels use.  Activations are FP32,
 * [item][channel][z][y][x].  Only C = 16 is built. */
int dft_cnn_param_count(int channels, int64_t* count);
/* Floats of the activation buffer: training = 1 keeps the three hidden
 * pre-activations for backward, 0 uses two ping-pong layers. */
int dft_cnn_act_count(int channels, int64_t nx, int64_t ny, int64_t nz, int nbatch, int training, int64_t* count);
/* Forward over nbatch items of nx*ny*nz: in [nbatch][2][N], out [nbatch][N]. */
int dft_cnn_forward(dft_ctx* ctx, int channels, int64_t nx, int64_t ny, int64_t nz, int nbatch,
                    const uint16_t* w_bf16, const float* in, float* out, float* act, int training);
/* Mean-squared-error loss L = mean((out - target)^2) and its gradient with
 * respect to the BF16 parameters, written (FP32) to grad.  act must come
 * from dft_cnn_forward with training = 1.  loss_host may be NULL. */
int dft_cnn_backward(dft_ctx* ctx, int channels, int64_t nx, int64_t ny, int64_t nz, int nbatch,
                     const uint16_t* w_bf16, const float* in, const float* act, const float* out,
                     const float* target, float* grad, float* scratch, double* loss_host);
int dft_cnn_scratch_count(int channels, int64_t nx, int64_t ny, int64_t nz, int nbatch, int64_t* count);
/* Adam step (bias-corrected) on the FP32 master weights; refreshes the BF16 copy. */
int dft_cnn_adam(dft_ctx* ctx, int64_t count, float* w, uint16_t* w_bf16, const float* grad, float* m, float* v,
                 int step, float lr, float beta1, float beta2, float eps);
int dft_cnn_to_bf16(dft_ctx* ctx, int64_t count, const float* w, uint16_t* w_bf16);
/* Preconditioner maps.  s_b = scale_dev[b] (max |r_b|), h2 = 1 / sum(1/h^2):
 * prepare: in[b] = {r_b / s_b, (V_eff - theta_b) h2};  finish: T_b = out_b s_b h2
 * (a positive per-column scale: Davidson uses only each column's direction). */
int dft_column_maxabs(dft_ctx* ctx, int prec, int64_t n, int ncols, const void* x, int64_t ldx, double* out_dev);
int dft_cnn_prepare(dft_ctx* ctx, const dft_grid* g, int prec, const void* veff, const void* r, int64_t ldr, int ncols,
                    const double* theta_dev, const double* scale_dev, float* in);
int dft_cnn_finish(dft_ctx* ctx, const dft_grid* g, int prec, int ncols, const float* out, const double* scale_dev,
                   void* t, int64_t ldt);
/* The same two maps on a sub-box of the grid, box = {x0, y0, z0, bx, by, bz}
 * (host array, inside the grid): in [ncols][2][bx by bz] from the box, and
 * t[box] += out s h2 w, with w a linear ramp over `taper` points at each box
 * face that lies inside the grid (the learned correction added to a base
 * correction over the region the network covers). */
int dft_cnn_prepare_box(dft_ctx* ctx, const dft_grid* g, int prec, const void* veff, const void* r, int64_t ldr,
                        int ncols, const double* theta_dev, const double* scale_dev, const int64_t* box, float* in);
int dft_cnn_add_box(dft_ctx* ctx, const dft_grid* g, int prec, int ncols, const float* out, const double* scale_dev,
                    const int64_t* box, int taper, void* t, int64_t ldt);

/* ---- training-data harvest (projected CG on (H - theta) c = -r) ---- */
/* out_j = x_j . y_j (FP64, device) */
int dft_column_dots(dft_ctx* ctx, int prec, int64_t n, int nb, const void* x, int64_t ldx, const void* y, int64_t ldy,
                    double* out_dev);
/* y_j = a_j x_j + b_j y_j with a, b device arrays of nb doubles */
int dft_cols_axpby(dft_ctx* ctx, int prec, int64_t n, int nb, const double* a_dev, const void* x, int64_t ldx,
                   const double* b_dev, void* y, int64_t ldy);
/* sum r^2 of each column over each aligned patch^3 block:
 * out_dev[col * nblocks + block], block = bx + nbx*(by + nby*bz), nb* = n* / patch */
int dft_patch_energy(dft_ctx* ctx, const dft_grid* g, int prec, int patch, const void* r, int64_t ldr, int ncols,
                     double* out_dev);
/* Cut npatch training patches; desc (host) = {column, x0, y0, z0} per patch,
 * each inside the grid and column < ncols.  in [npatch][2][patch^3] =
 * {r/s, (V_eff - theta) h2}, tgt [npatch][patch^3] = c / cs, with s and cs
 * the columns' max |r| and max |c| (scale_dev, cscale_dev). */
int dft_extract_patches(dft_ctx* ctx, const dft_grid* g, int prec, int patch, int npatch, const int* desc, int ncols,
                        const void* veff, const void* r, int64_t ldr, const void* c, int64_t ldc,
                        const double* theta_dev, const double* scale_dev, const double* cscale_dev, float* in,
                        float* tgt);

#ifdef __cplusplus
}
#endif
#endif
