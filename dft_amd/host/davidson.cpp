// davidson.cpp -- block Davidson (see davidson.hpp).
#include "davidson.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

namespace dftc {

BlockDavidson::BlockDavidson(dft_ctx* c, const dft_grid& g, int prec, int nstates, int max_subspace)
    : c_(c), g_(g), prec_(prec), k_(nstates), mmax_(max_subspace) {
    if (k_ <= 0) throw std::invalid_argument("El numero de orbitales debe ser mayor que cero.");
    if (mmax_ < 2 * k_) mmax_ = 2 * k_;
    n_ = g.nx * g.ny * g.nz;
    es_ = esize(prec_);
    V_ = Buf(c, (size_t)n_ * mmax_ * es_);
    W_ = Buf(c, (size_t)n_ * mmax_ * es_);
    X_ = Buf(c, (size_t)n_ * k_ * es_);
    HX_ = Buf(c, (size_t)n_ * k_ * es_);
    R_ = Buf(c, (size_t)n_ * k_ * es_);
    T_ = Buf(c, (size_t)n_ * k_ * es_);
    PV_ = Buf(c, (size_t)n_ * k_ * es_);
    G_ = Buf(c, sizeof(double) * (size_t)mmax_ * mmax_);
    Q_ = Buf(c, sizeof(double) * (size_t)mmax_ * mmax_);
    P_ = Buf(c, sizeof(double) * (size_t)mmax_ * k_);
    S_ = Buf(c, sizeof(double) * (size_t)k_ * k_);
    theta_ = Buf(c, sizeof(double) * (size_t)mmax_);
    rn_ = Buf(c, sizeof(double) * (size_t)k_);
    Qprev_ = Buf(c, sizeof(double) * (size_t)mmax_ * k_);
    Cb_ = Buf(c, sizeof(double) * (size_t)mmax_ * k_);
    Cr_ = Buf(c, sizeof(double) * (size_t)mmax_ * 2 * k_);
    G_.zero();
}

void BlockDavidson::hpsi(const void* veff, const void* in, void* out, int ncols) {
    check(c_, dft_hpsi(c_, &g_, prec_, veff, in, n_, ncols, out, n_), "dft_hpsi");
}

void BlockDavidson::gram(int ma, int mb, const void* A, const void* B, double* out, int64_t ldo, double beta) {
    const int mode = prec_ == DFT_F64 ? DFT_GRAM_F64 : gram_mode_;
    check(c_, dft_gram(c_, mode, n_, ma, mb, 1.0, A, n_, B, n_, beta, out, ldo), "dft_gram");
}

void BlockDavidson::ogram(int ma, int mb, const void* A, const void* B, double* out, int64_t ldo) {
    const int mode = prec_ == DFT_F64 ? DFT_GRAM_F64 : ortho_mode_;
    check(c_, dft_gram(c_, mode, n_, ma, mb, 1.0, A, n_, B, n_, 0.0, out, ldo), "dft_gram");
}

void BlockDavidson::load_X(const void* src, int src_prec, int ncols) {
    ncols = std::min(ncols, k_);
    check(c_, dft_convert(c_, src, src_prec, X_.get(), prec_, n_ * (int64_t)ncols), "dft_convert");
}

void BlockDavidson::seed_columns(int first) {
    if (first >= k_) return;
    check(c_, dft_seed_vectors(c_, &g_, prec_, first, k_ - first, static_cast<char*>(X_.get()) + (size_t)first * n_ * es_, n_),
          "dft_seed_vectors");
}

int BlockDavidson::orthonormalise(void* src, int cnt, void* dst, void* tmp) {
    // two CholQR passes: src -> tmp -> dst
    void* in = src;
    for (int pass = 0; pass < 2; ++pass) {
        void* out = pass == 0 ? tmp : dst;
        double* S = S_.as<double>();
        ogram(cnt, cnt, in, in, S, cnt);
        const int rc = dft_chol_inv(c_, cnt, S, cnt);
        if (rc == DFT_OK) {
            check(c_, dft_block_update(c_, prec_, n_, cnt, cnt, 1.0, in, n_, S, cnt, 0.0, out, n_), "dft_block_update");
        } else if (rc == DFT_ENUMERIC) {
            // rank-deficient block: keep eigen-directions above a relative floor
            ogram(cnt, cnt, in, in, S, cnt);
            check(c_, dft_syev(c_, cnt, S, cnt, theta_.as<double>() /* scratch */), "dft_syev");
            std::vector<double> lam(cnt), U((size_t)cnt * cnt);
            check(c_, dft_download(c_, lam.data(), theta_.get(), sizeof(double) * cnt), "dft_download");
            check(c_, dft_download(c_, U.data(), S, sizeof(double) * cnt * cnt), "dft_download");
            const double lmax = std::max(lam.back(), 0.0);
            const double floor = (prec_ == DFT_F64 ? 1e-24 : 1e-12) * std::max(lmax, 1e-300);
            int first = 0;
            while (first < cnt && !(lam[first] > floor)) ++first;
            const int kept = cnt - first;
            if (kept == 0) return 0;
            std::vector<double> C((size_t)cnt * kept);
            for (int j = 0; j < kept; ++j) {
                const double s = 1.0 / std::sqrt(lam[first + j]);
                for (int i = 0; i < cnt; ++i) C[(size_t)j * cnt + i] = U[(size_t)(first + j) * cnt + i] * s;
            }
            check(c_, dft_upload(c_, S, C.data(), sizeof(double) * C.size()), "dft_upload");
            check(c_, dft_block_update(c_, prec_, n_, cnt, kept, 1.0, in, n_, S, cnt, 0.0, out, n_), "dft_block_update");
            cnt = kept;
        } else {
            check(c_, rc, "dft_chol_inv");
        }
        in = out;
    }
    return cnt;
}

int BlockDavidson::append_block(int m, int nt, DavidsonStats& st) {
    const size_t colb = (size_t)n_ * es_;
    char* V = static_cast<char*>(V_.get());
    char* T = V + (size_t)m * colb;
    const int mt = m + nt;
    double* G1 = P_.as<double>();  // (m + nt) x nt, ld m + nt (P_ holds mmax x k)
    // accuracy of one Gram entry relative to the column norms: FP64; FP32 products;
    // split-BF16 (operands kept to ~2^-17, which for orbitals concentrated on a few
    // grid points is also the entry's relative accuracy)
    const double eps_g = prec_ == DFT_F64 ? 2.0e-16 : ortho_mode_ == DFT_GRAM_SPLIT_BF16 ? 1.0e-5 : 1.0e-7;
    std::vector<double> g((size_t)mt * nt), L((size_t)nt * nt), Ri((size_t)nt * nt), C((size_t)mt * nt);
    for (int pass = 0; pass < 2; ++pass) {
        ogram(mt, nt, V, T, G1, mt);
        check(c_, dft_download(c_, g.data(), G1, sizeof(double) * g.size()), "dft_download");
        auto G = [&](int i, int j) { return g[(size_t)j * mt + i]; };  // rows 0..m-1: V^T T; m..: T^T T
        // S' = T^T T - (V^T T)^T (V^T T), the Gram matrix of the projected block;
        // column norms^2 before (sn) and after (sp) the projection
        std::vector<double> sn(nt), sp(nt);
        bool ok = true;
        for (int j = 0; j < nt; ++j) {
            sn[j] = G(m + j, j);
            ok = ok && sn[j] > 0.0 && std::isfinite(sn[j]);
        }
        for (int j = 0; j < nt && ok; ++j)
            for (int i = j; i < nt; ++i) {
                double s = 0.5 * (G(m + i, j) + G(m + j, i));
                for (int l = 0; l < m; ++l) s -= G(l, i) * G(l, j);
                L[(size_t)j * nt + i] = s;  // lower triangle, column-major
                if (i == j) sp[j] = s;
            }
        // Cholesky S' = L L^T; a pivot not clearly above the Gram noise of its own
        // column (entries are accurate to eps_g times the column norms) means dependence
        for (int j = 0; j < nt && ok; ++j) {
            double d = L[(size_t)j * nt + j];
            for (int l = 0; l < j; ++l) d -= L[(size_t)l * nt + j] * L[(size_t)l * nt + j];
            if (!(d > 64.0 * eps_g * sn[j])) {
                ok = false;
                break;
            }
            d = std::sqrt(d);
            L[(size_t)j * nt + j] = d;
            for (int i = j + 1; i < nt; ++i) {
                double s = L[(size_t)j * nt + i];
                for (int l = 0; l < j; ++l) s -= L[(size_t)l * nt + i] * L[(size_t)l * nt + j];
                L[(size_t)j * nt + i] = s / d;
            }
        }
        if (!ok) break;
        // R = L^T; Ri = R^-1 (upper triangular), column by column; rs2 = ||D R^-1||_F^2
        // with D = diag(column norms before the projection): ~nt for an independent
        // block that kept its norms
        std::fill(Ri.begin(), Ri.end(), 0.0);
        double rs2 = 0.0;
        for (int j = 0; j < nt; ++j) {
            Ri[(size_t)j * nt + j] = 1.0 / L[(size_t)j * nt + j];
            for (int i = j - 1; i >= 0; --i) {  // (R Ri)[i][j] = 0: sum_{l=i..j} R[i][l] Ri[l][j]
                double s = 0.0;
                for (int l = i + 1; l <= j; ++l) s += L[(size_t)i * nt + l] * Ri[(size_t)j * nt + l];
                Ri[(size_t)j * nt + i] = -s / L[(size_t)i * nt + i];
            }
            for (int i = 0; i <= j; ++i) rs2 += sn[i] * Ri[(size_t)j * nt + i] * Ri[(size_t)j * nt + i];
        }
        // C = [-(V^T T) Ri ; Ri]
        for (int j = 0; j < nt; ++j) {
            for (int l = 0; l < m; ++l) {
                double s = 0.0;
                for (int q = 0; q <= j; ++q) s += G(l, q) * Ri[(size_t)j * nt + q];
                C[(size_t)j * mt + l] = -s;
            }
            for (int i = 0; i < nt; ++i) C[(size_t)j * mt + m + i] = i <= j ? Ri[(size_t)j * nt + i] : 0.0;
        }
        check(c_, dft_upload(c_, Cb_.get(), C.data(), sizeof(double) * C.size()), "dft_upload");
        check(c_, dft_block_update(c_, prec_, n_, mt, nt, 1.0, V, n_, Cb_.as<double>(), mt, 0.0, T, n_),
              "dft_block_update");
        ++st.block_passes;
        // the block's loss of orthogonality after a pass is ~ eps_g ||D R^-1||^2: one pass
        // is enough when every column kept at least half its norm and the block is well
        // conditioned; otherwise a second pass on the (now nearly orthonormal) block
        bool single = pass == 0 && rs2 <= 4.0 * nt;
        for (int j = 0; j < nt && single; ++j) single = sp[j] >= 0.25 * sn[j];
        if (single || pass == 1) return nt;
    }
    // dependent block: copy it out, project twice, then the rank-revealing orthonormalisation
    ++st.block_fallbacks;
    check(c_, dft_copy(c_, T_.get(), T, colb * nt), "dft_copy");
    for (int pass = 0; pass < 2; ++pass) {
        ogram(m, nt, V, T_.get(), P_.as<double>(), mmax_);
        check(c_, dft_block_update(c_, prec_, n_, m, nt, -1.0, V, n_, P_.as<double>(), mmax_, 1.0, T_.get(), n_),
              "dft_block_update");
    }
    return orthonormalise(T_.get(), nt, T, R_.get());
}

int BlockDavidson::restart_coefficients(int m, int mprev, int pmax, std::vector<double>& Q, std::vector<double>& Pc) {
    // Coefficients (in the current basis V, m columns) of the Ritz block X
    // (= Q[:, :k]) and of the previous Ritz block (Qprev padded with zeros).
    const int k = k_;
    Q.assign((size_t)mmax_ * k, 0.0);
    check(c_, dft_download(c_, Q.data(), Q_.get(), sizeof(double) * (size_t)mmax_ * k), "dft_download");
    Pc.assign((size_t)m * k, 0.0);
    if (mprev <= 0 || pmax <= 0) return 0;
    std::vector<double> Qp((size_t)mmax_ * k, 0.0);
    check(c_, dft_download(c_, Qp.data(), Qprev_.get(), sizeof(double) * (size_t)mmax_ * k), "dft_download");
    auto q = [&](int i, int j) -> double& { return Q[(size_t)j * mmax_ + i]; };
    auto p = [&](int i, int j) -> double& { return Qp[(size_t)j * mmax_ + i]; };
    for (int j = 0; j < k; ++j)
        for (int i = mprev; i < m; ++i) p(i, j) = 0.0;
    // P = (I - Q Q^T) Qprev, twice, then modified Gram-Schmidt among its columns
    int kept = 0;
    for (int j = 0; j < k; ++j) {
        std::vector<double> v(m);
        for (int i = 0; i < m; ++i) v[i] = p(i, j);
        for (int pass = 0; pass < 2; ++pass) {
            for (int a = 0; a < k; ++a) {
                double d = 0;
                for (int i = 0; i < m; ++i) d += q(i, a) * v[i];
                for (int i = 0; i < m; ++i) v[i] -= d * q(i, a);
            }
            for (int a = 0; a < kept; ++a) {
                double d = 0;
                for (int i = 0; i < m; ++i) d += Pc[(size_t)a * m + i] * v[i];
                for (int i = 0; i < m; ++i) v[i] -= d * Pc[(size_t)a * m + i];
            }
        }
        if (kept >= pmax) break;
        double nn = 0;
        for (double x : v) nn += x * x;
        nn = std::sqrt(nn);
        if (nn < 1e-8) continue;  // previous block already inside span(X)
        for (int i = 0; i < m; ++i) Pc[(size_t)kept * m + i] = v[i] / nn;
        ++kept;
    }
    return kept;
}

int BlockDavidson::restart_in_place(int m, int kept, const std::vector<double>& Q, const std::vector<double>& Pc) {
    const int k = k_, kn = k + kept;
    // [Q_k | P], m x (k + kept), column-major with ld m
    std::vector<double> C((size_t)m * kn);
    for (int j = 0; j < k; ++j)
        for (int i = 0; i < m; ++i) C[(size_t)j * m + i] = Q[(size_t)j * mmax_ + i];
    std::copy(Pc.begin(), Pc.begin() + (size_t)m * kept, C.begin() + (size_t)m * k);
    check(c_, dft_upload(c_, Cr_.get(), C.data(), sizeof(double) * C.size()), "dft_upload");
    // one pass each over V and W; every output point is written by the wave that has read all of its inputs
    check(c_, dft_block_update(c_, prec_, n_, m, kn, 1.0, V_.get(), n_, Cr_.as<double>(), m, 0.0, V_.get(), n_),
          "dft_block_update");
    check(c_, dft_block_update(c_, prec_, n_, m, kn, 1.0, W_.get(), n_, Cr_.as<double>(), m, 0.0, W_.get(), n_),
          "dft_block_update");
    // the Ritz block is now V's first k columns; it is copied to X_ only if
    // the iteration ends before the next Rayleigh-Ritz step
    gram(kn, kn, V_.get(), W_.get(), G_.as<double>(), mmax_);
    return kn;
}

int BlockDavidson::restart(int m, int kept, const std::vector<double>& Pc) {
    const int k = k_;
    const size_t colb = (size_t)n_ * es_;
    char* V = static_cast<char*>(V_.get());
    char* W = static_cast<char*>(W_.get());
    double* S = P_.as<double>();  // mmax x k scratch for the coefficients
    if (kept > 0) {
        check(c_, dft_upload(c_, S, Pc.data(), sizeof(double) * (size_t)m * kept), "dft_upload");
        // P in PV_, HP in T_ (both free at this point; R_ still holds the residuals)
        check(c_, dft_block_update(c_, prec_, n_, m, kept, 1.0, V, n_, S, m, 0.0, PV_.get(), n_), "dft_block_update");
        check(c_, dft_block_update(c_, prec_, n_, m, kept, 1.0, W, n_, S, m, 0.0, T_.get(), n_), "dft_block_update");
    }
    check(c_, dft_copy(c_, V, X_.get(), colb * k), "dft_copy");
    check(c_, dft_copy(c_, W, HX_.get(), colb * k), "dft_copy");
    if (kept > 0) {
        check(c_, dft_copy(c_, V + (size_t)k * colb, PV_.get(), colb * kept), "dft_copy");
        check(c_, dft_copy(c_, W + (size_t)k * colb, T_.get(), colb * kept), "dft_copy");
    }
    const int mn = k + kept;
    gram(mn, mn, V, W, G_.as<double>(), mmax_);
    return mn;
}

int KineticPreconditioner::apply(dft_ctx* c, const dft_grid& g, int prec, const void* /*veff*/, const void* /*X*/,
                                 int64_t /*ldx*/, const void* R, int64_t ldr, int /*first*/, int ncols,
                                 const double* theta_dev, void* T, int64_t ldt) {
    if (ncols <= 0) return 0;
    if (mode_ == kExact) {
        if (scale_ > 0.0)
            check(c, dft_precond_kinetic_exact_shifted(c, &g, prec, sigma_, scale_, theta_dev, R, ldr, ncols, T, ldt),
                  "dft_precond_kinetic_exact_shifted");
        else
            check(c, dft_precond_kinetic_exact(c, &g, prec, sigma_, R, ldr, ncols, T, ldt), "dft_precond_kinetic_exact");
        return 0;
    }
    const int64_t n = g.nx * g.ny * g.nz;
    if (kt_cols_ < ncols || sig_prec_ != prec) {
        kt_ = Buf(c, (size_t)n * ncols * esize(prec));
        kt_cols_ = ncols;
        sig_prec_ = prec;
    }
    if (mode_ == kChebyshev)
        check(c, dft_precond_kinetic_cheb(c, &g, prec, sigma_, sweeps_, R, ldr, ncols, T, ldt, kt_.get()),
              "dft_precond_kinetic_cheb");
    else
        check(c, dft_precond_kinetic(c, &g, prec, sigma_, omega_, sweeps_, R, ldr, ncols, T, ldt, kt_.get()),
              "dft_precond_kinetic");
    return 0;
}

DavidsonStats BlockDavidson::solve(const void* veff, bool warm, const DavidsonOptions& opt, Preconditioner& P) {
    DavidsonStats st;
    gram_mode_ = opt.gram_mode;
    ortho_mode_ = opt.ortho_gram_mode >= 0 ? opt.ortho_gram_mode : opt.gram_mode;
    const int k = k_;
    const size_t colb = (size_t)n_ * es_;
    char* V = static_cast<char*>(V_.get());
    char* W = static_cast<char*>(W_.get());
    double* G = G_.as<double>();
    double* Q = Q_.as<double>();
    double* Pm = P_.as<double>();
    double* theta = theta_.as<double>();
    if (!warm) check(c_, dft_seed_vectors(c_, &g_, prec_, 0, k, X_.get(), n_), "dft_seed_vectors");
    int m = orthonormalise(X_.get(), k, V, T_.get());
    if (m < k) {
        // The starting block is numerically rank deficient (wide seed Gaussians
        // on large boxes, or FP32 storage): complete it with pseudo-random
        // directions orthogonalised against the kept ones.
        const int need = k - m;
        const size_t colb0 = (size_t)n_ * es_;
        check(c_, dft_random_vectors(c_, prec_, n_, need, 0x5eedULL + (uint64_t)m, T_.get(), n_), "dft_random_vectors");
        if (m > 0)
            for (int pass = 0; pass < 2; ++pass) {
                ogram(m, need, V, T_.get(), P_.as<double>(), mmax_);
                check(c_, dft_block_update(c_, prec_, n_, m, need, -1.0, V, n_, P_.as<double>(), mmax_, 1.0, T_.get(), n_),
                      "dft_block_update");
            }
        m += orthonormalise(T_.get(), need, V + (size_t)m * colb0, R_.get());
    }
    if (m < k) throw std::runtime_error("No se pudo construir el vector inicial.");
    hpsi(veff, V, W, k);
    st.happly_columns += k;
    gram(k, k, V, W, G, mmax_);
    const int nconv = std::min(std::max(opt.nconv, 1), k);
    theta_h_.assign(k, 0.0);
    rn_h_.assign(k, INFINITY);
    std::vector<double> prev_rn;
    P.reset();
    int mprev = 0;  // basis size when Qprev was taken (0: none yet)
    const int nact = nconv;
    // X_ holds V Q_k of the latest Rayleigh-Ritz step only when x_fresh; the
    // Ritz block is V's first k columns when x_in_v (after an in-place restart)
    bool x_fresh = false, x_in_v = false;
    int m_rr = m;
    auto ritz_block = [&](bool with_hx) {
        check(c_, dft_block_update(c_, prec_, n_, m_rr, k, 1.0, V, n_, Q, mmax_, 0.0, X_.get(), n_), "dft_block_update");
        if (with_hx)
            check(c_, dft_block_update(c_, prec_, n_, m_rr, k, 1.0, W, n_, Q, mmax_, 0.0, HX_.get(), n_),
                  "dft_block_update");
        x_fresh = true;
    };
    const bool px = P.needs_ritz_vectors();
    for (int it = 1; it <= opt.max_iter; ++it) {
        st.iterations = it;
        if (int rc = dft_syev_lowest(c_, m, k, G, mmax_, Q, mmax_, theta)) {
            // say what the projected matrix looked like
            std::vector<double> gh((size_t)mmax_ * m);
            check(c_, dft_download(c_, gh.data(), G, sizeof(double) * gh.size()), "dft_download");
            int bad = 0;
            double big = 0.0, asym = 0.0;
            for (int j = 0; j < m; ++j)
                for (int i = 0; i < m; ++i) {
                    const double x = gh[(size_t)j * mmax_ + i];
                    if (!std::isfinite(x)) ++bad;
                    else big = std::max(big, std::fabs(x));
                    if (i > j) asym = std::max(asym, std::fabs(x - gh[(size_t)i * mmax_ + j]));
                }
            throw std::runtime_error("Davidson: Rayleigh-Ritz eigensolve failed (iteration " + std::to_string(it) +
                                     ", subspace " + std::to_string(m) + ", " + std::to_string(bad) +
                                     " non-finite entries, max |G| " + std::to_string(big) + ", max asymmetry " +
                                     std::to_string(asym) + "): " + std::string(dft_last_error(c_)));
        }
        m_rr = m;
        x_fresh = x_in_v = false;
        if (px) ritz_block(false);
        check(c_, dft_ritz_residual(c_, prec_, n_, m, k, V, n_, W, n_, Q, mmax_, theta, R_.get(), n_, rn_.as<double>()),
              "dft_ritz_residual");
        check(c_, dft_download(c_, theta_h_.data(), theta, sizeof(double) * k), "dft_download");
        check(c_, dft_download(c_, rn_h_.data(), rn_.get(), sizeof(double) * k), "dft_download");
        double maxres = 0.0;
        for (int j = 0; j < k; ++j) {
            rn_h_[j] = std::sqrt(std::max(rn_h_[j], 0.0));
            if (!std::isfinite(rn_h_[j])) throw std::runtime_error("Davidson: non-finite residual");
            if (j < nconv) maxres = std::max(maxres, rn_h_[j]);
        }
        if (!prev_rn.empty()) P.feedback(prev_rn, rn_h_);
        prev_rn = rn_h_;
        st.max_residual = maxres;
        if (maxres < opt.tol) {
            st.converged = true;
            break;
        }
        if (it == opt.max_iter) break;
        // corrections for the unconverged states that must converge, in contiguous runs
        std::vector<std::pair<int, int>> runs;
        int nt = 0;
        for (int j = 0; j < nact;) {
            if (rn_h_[j] < opt.tol) {
                ++j;
                continue;
            }
            int j1 = j;
            while (j1 < nact && rn_h_[j1] >= opt.tol) ++j1;
            runs.push_back({j, j1});
            nt += j1 - j;
            j = j1;
        }
        if (m + nt > mmax_) {  // thick restart on [X, P] before the new block is formed
            std::vector<double> Qh, Pc;
            const int kept = restart_coefficients(m, mprev, std::max(0, mmax_ - k - nt), Qh, Pc);
            if (dft_block_update_in_place(prec_, m, k + kept)) {
                // X_, if fresh, already holds the same V Q_k for the correction below
                m = restart_in_place(m, kept, Qh, Pc);
                x_in_v = true;
            } else {
                ritz_block(true);
                m = restart(m, kept, Pc);
            }
            ++st.restarts;
            mprev = 0;
        } else {
            // remember this iteration's Ritz coefficients for the next restart
            check(c_, dft_copy(c_, Qprev_.get(), Q, sizeof(double) * (size_t)mmax_ * k), "dft_copy");
            mprev = m;
        }
        if (m + nt > mmax_) {  // a subspace limit below 3 k: only the lowest states' corrections fit
            int room = mmax_ - m;
            for (auto& r : runs) {
                r.second = std::min(r.second, r.first + std::max(room, 0));
                room -= r.second - r.first;
            }
            nt = mmax_ - m;
        }
        if (nt <= 0) break;
        // the corrections go straight to V[:, m : m + nt]
        int off = 0;
        for (const auto& r : runs) {
            if (r.second <= r.first) continue;
            P.apply(c_, g_, prec_, veff, X_.get(), n_, static_cast<char*>(R_.get()) + r.first * colb, n_, r.first,
                    r.second - r.first, theta + r.first, V + (size_t)(m + off) * colb, n_);
            off += r.second - r.first;
        }
        const int kept = append_block(m, nt, st);
        if (kept == 0) break;  // no new direction: stagnation
        hpsi(veff, V + (size_t)m * colb, W + (size_t)m * colb, kept);
        st.happly_columns += kept;
        gram(m + kept, kept, V, W + (size_t)m * colb, G + (size_t)m * mmax_, mmax_);
        check(c_, dft_transpose(c_, m, kept, G + (size_t)m * mmax_, mmax_, G + m, mmax_), "dft_transpose");
        m += kept;
    }
    // the Ritz block of the last Rayleigh-Ritz step (V's first m_rr columns
    // are unchanged unless a restart formed X_ itself or left it in V)
    if (x_in_v) check(c_, dft_copy(c_, X_.get(), V_.get(), (size_t)n_ * es_ * k), "dft_copy");
    else if (!x_fresh) ritz_block(false);
    return st;
}

}  // namespace dftc
