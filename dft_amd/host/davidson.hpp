// davidson.hpp -- block Davidson eigensolver on device-resident orbital
// blocks.  Replaces the reference's one-orbital-at-a-time Davidson
// (EigenvalueSolverMolecule.cpp:666-1300) with a block method whose dense
// products are the MFMA Gram kernels:
//   basis V (N x m, orthonormal), W = H V, G = V^T W (FP64, m x m)
//   Rayleigh-Ritz: G = Q diag(theta) Q^T (FP64 on the device)
//   R = W Q_k - V Q_k theta in one pass over [W | V], norms in the same pass;
//   the Ritz block X = V Q_k (and HX = W Q_k) is formed only where it is
//   read: before a restart, for a preconditioner that reads X, and at exit
//   restart to V = [X, P] when the new block would exceed the subspace limit;
//   T = P(R) for unconverged columns, written straight after V's columns;
//   T orthonormalised against V and itself in place by Pythagorean block
//   Gram-Schmidt: G = [V T]^T T in one pass, S' = T^T T - (V^T T)^T (V^T T)
//   = R^T R, T <- [V T] [-(V^T T) R^-1 ; R^-1] in one pass (a second pass
//   only when the first lost accuracy).
// Restart keeps momentum (LOBPCG-style thick restart): the new basis is
// [X, P] with P the previous Ritz block orthogonalised against X, formed in
// the coefficient space of V (no extra H applications).  Corrections are
// computed only for the unconverged states that must converge.
// Vectors are kept at unit 2-norm; with psi = u / sqrt(dV) the reference's
// residual ||H psi - eps psi|| sqrt(dV) equals ||H u - eps u||.
#pragma once

#include <cstdint>
#include <vector>

#include "gpu.hpp"

namespace dftc {

// A correction operator T = P(R) on the residual columns first..first+ncols-1
// of the block.  R, T and theta_dev point at column `first`; X is the whole
// Ritz block (columns 0..first+ncols-1 are current).
struct Preconditioner {
    virtual ~Preconditioner() = default;
    // Writes the T columns; returns how many of them used the fallback.
    virtual int apply(dft_ctx* c, const dft_grid& g, int prec, const void* veff, const void* X, int64_t ldx,
                      const void* R, int64_t ldr, int first, int ncols, const double* theta_dev, void* T,
                      int64_t ldt) = 0;
    // Residual norms of every block column before and after an iteration.
    virtual void feedback(const std::vector<double>& /*before*/, const std::vector<double>& /*after*/) {}
    // Called when a solve starts (new operator).
    virtual void reset() {}
    // True when apply() reads the Ritz block X (then it is formed every iteration).
    virtual bool needs_ritz_vectors() const { return false; }
    virtual int64_t applied() const { return 0; }
    virtual int64_t fallbacks() const { return 0; }
};

// davidsonCorrection (EigenvalueSolverMolecule.cpp:483-555)
struct DiagonalPreconditioner : Preconditioner {
    int apply(dft_ctx* c, const dft_grid& g, int prec, const void* veff, const void* /*X*/, int64_t /*ldx*/,
              const void* R, int64_t ldr, int /*first*/, int ncols, const double* theta_dev, void* T,
              int64_t ldt) override {
        check(c, dft_precond_diag(c, &g, prec, veff, R, ldr, ncols, theta_dev, T, ldt), "dft_precond_diag");
        return 0;
    }
};

// Kinetic smoothing correction: t ~ (-1/2 Lap + sigma)^{-1} r by `sweeps`
// damped-Jacobi steps started from the diagonal solution r / D,
// D = sum(1/h^2) + sigma.  A polynomial low-pass filter on the residual: the
// high-frequency components (kinetic energy far above the occupied
// spectrum) are damped relative to the smooth ones, which the diagonal
// correction (a near-constant scaling on these grids) cannot do.  No
// reference counterpart; the diagonal correction stays the reference's choice.
// Mode kChebyshev: the same stencil steps combined by Chebyshev
// semi-iteration (dft_precond_kinetic_cheb), a much closer approximation of
// (-1/2 Lap + sigma)^{-1} r for the same number of passes.  Mode kExact: that
// inverse itself, by sine transforms on the matrix cores
// (dft_precond_kinetic_exact).
class KineticPreconditioner : public Preconditioner {
   public:
    enum Mode { kJacobi = 0, kChebyshev = 1, kExact = 2 };
    // kExact with scale > 0: per-state shifts sigma_j = max(sigma, -scale theta_j)
    KineticPreconditioner(int sweeps = 4, double sigma = 1.0, double omega = 2.0 / 3.0, Mode mode = kJacobi,
                          double scale = 0.0)
        : sweeps_(sweeps), sigma_(sigma), omega_(omega), scale_(scale), mode_(mode) {}
    int apply(dft_ctx* c, const dft_grid& g, int prec, const void* veff, const void* X, int64_t ldx, const void* R,
              int64_t ldr, int first, int ncols, const double* theta_dev, void* T, int64_t ldt) override;

   private:
    int sweeps_;
    double sigma_, omega_, scale_;
    Mode mode_;
    Buf kt_;  // scratch columns for the ping-pong steps
    int64_t kt_cols_ = 0;
    int sig_prec_ = -1;
};

struct DavidsonOptions {
    int nconv = 1;         // leading states that must reach tol
    double tol = 1.0e-7;   // EigenvalueSolverMolecule.cpp:19
    int max_iter = 200;
    int gram_mode = DFT_GRAM_F64;
    // Gram mode of the orthonormalisation products (Pythagorean block
    // Gram-Schmidt, CholQR) of FP32 blocks; -1: gram_mode
    int ortho_gram_mode = -1;
};

struct DavidsonStats {
    int iterations = 0;
    int restarts = 0;
    int64_t happly_columns = 0;
    double max_residual = 0.0;
    bool converged = false;
    int block_passes = 0;     // Pythagorean Gram-Schmidt passes over new blocks
    int block_fallbacks = 0;  // blocks orthonormalised by the rank-revealing path
};

class BlockDavidson {
   public:
    BlockDavidson(dft_ctx* c, const dft_grid& g, int prec, int nstates, int max_subspace);
    // Solve for the lowest nstates of H = -1/2 Lap + veff.  When `warm` is
    // true the current X is the starting block, otherwise seed vectors
    // (buildInitialVector) are used.
    DavidsonStats solve(const void* veff, bool warm, const DavidsonOptions& opt, Preconditioner& P);

    void* X() const { return X_.get(); }          // N x k Ritz vectors (unit 2-norm)
    int64_t ld() const { return n_; }
    int nstates() const { return k_; }
    int prec() const { return prec_; }
    const std::vector<double>& eigenvalues() const { return theta_h_; }
    const std::vector<double>& residuals() const { return rn_h_; }
    // Replace X with a block given in another precision (refinement hand-over).
    void load_X(const void* src, int src_prec, int ncols);
    // Overwrite columns first..k-1 of X with the seed vectors of those indices.
    void seed_columns(int first);

   private:
    void hpsi(const void* veff, const void* in, void* out, int ncols);
    void gram(int ma, int mb, const void* A, const void* B, double* out, int64_t ldo, double beta = 0.0);
    // the orthonormalisation products (mode ortho_mode_ for FP32 blocks)
    void ogram(int ma, int mb, const void* A, const void* B, double* out, int64_t ldo);
    // Orthonormalise `cnt` columns at src into dst (CholQR twice, with an
    // eigenvalue filter if the block is rank deficient); returns kept count.
    int orthonormalise(void* src, int cnt, void* dst, void* tmp);
    // Thick restart: V <- [X, P], W <- [HX, HP] with at most pmax columns of
    // P; returns the new basis size.
    // Thick restart.  restart_coefficients: the current Ritz coefficients (host
    // copy of Q, ld mmax) and P = (I - QQ^T) Qprev orthonormalised (m x kept);
    // restart_in_place: [V | W][:, :k+kept] = [V | W] [Q_k | P] in one in-place
    // pass each, X = V[:, :k]; restart: from X and HX formed by the caller.
    int restart_coefficients(int m, int mprev, int pmax, std::vector<double>& Q, std::vector<double>& Pc);
    int restart_in_place(int m, int kept, const std::vector<double>& Q, const std::vector<double>& Pc);
    int restart(int m, int kept, const std::vector<double>& Pc);
    // The nt correction columns at V[:, m:m+nt] made orthonormal and
    // orthogonal to V[:, :m], in place: Pythagorean block Gram-Schmidt (one
    // Gram pass [V T]^T T and one update per pass, a second pass only when
    // the first lost accuracy); the rank-revealing path when the block is
    // (numerically) dependent.  Returns the number of columns now at
    // V[:, m:...], 0 when nothing independent remains.
    int append_block(int m, int nt, DavidsonStats& st);

    dft_ctx* c_;
    dft_grid g_;
    int prec_, k_, mmax_;
    int64_t n_;
    size_t es_;
    Buf V_, W_, X_, HX_, R_, T_, PV_;
    Buf G_, Q_, P_, S_, theta_, rn_, Qprev_, Cb_, Cr_;
    std::vector<double> theta_h_, rn_h_;
    int gram_mode_ = DFT_GRAM_F64;
    int ortho_mode_ = DFT_GRAM_F64;
};

}  // namespace dftc
