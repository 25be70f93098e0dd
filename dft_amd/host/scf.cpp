// scf.cpp -- device-resident SCF driver (see scf.hpp).
#include "scf.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <limits>
#include <stdexcept>
#include <string>

namespace dftc {

namespace {

using Clock = std::chrono::steady_clock;
double secs(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double>(b - a).count(); }

// DFTConstants.h:14-19 and SelfConsistentFieldMolecule.cpp:27-31
constexpr double kMixing0 = 0.25, kMinMix = 0.10, kMaxMix = 0.50, kMixUp = 1.10, kMixDown = 0.75, kOsc = 1.05;
constexpr double kEnergyTol = 1.0e-11, kDensityTol = 1.0e-9, kKsTol = 1.0e-7;
// mixed mode: where the FP32 phase hands over to FP64 (|dE| below the floor
// or the stop tolerance, whichever is larger).  Near the hand-over an FP32
// step costs a few Davidson iterations and an FP64 step some twenty, so the
// FP32 phase runs on while the density still settles: on the 40-atom
// cluster at 256^3 (PBE96, smeared) a floor of 3e-4 handed over at dRho
// 2.8e-5 and took 4 FP64 steps; 1e-5 hands over at dRho 8.9e-6 after three
// more FP32 steps and takes 3.  (Run on to |dE| < 1e-7 the deeper FP32 phase
// costs FP64 steps instead: its last Pulay entries hold FP32 noise, which
// the refinement drops when it stalls; see below.)
constexpr double kF32EnergyFloor = 1.0e-5, kF32DensityFloor = 1.0e-5;
// ... or, below this relative density change, once three FP32 steps in a row
// have not lowered it by 10% (the FP32 phase has reached its own noise floor)
constexpr double kF32StallDensity = 1.0e-3;
// grid points between the nuclei's bounding box and the faces of the learned correction's box
constexpr int kCnnBoxMargin = 16;

}  // namespace

double fermi_occupations(const double* eps, int n, double nelec, double kT, double* occ) {
    auto fill = [&](double mu) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) {
            const double x = (eps[i] - mu) / kT;
            // both branches stay finite: exp of a non-positive argument
            occ[i] = x > 0.0 ? std::exp(-x) / (1.0 + std::exp(-x)) : 1.0 / (1.0 + std::exp(x));
            s += occ[i];
        }
        return s;
    };
    const double lo0 = *std::min_element(eps, eps + n), hi0 = *std::max_element(eps, eps + n);
    double lo = lo0 - 50.0 * kT - 1.0, hi = hi0 + 50.0 * kT + 1.0;
    if (nelec >= n) {
        std::fill(occ, occ + n, 1.0);
        return hi;
    }
    if (nelec <= 0.0) {
        std::fill(occ, occ + n, 0.0);
        return lo;
    }
    // The electron count rises monotonically with mu but, in floating point,
    // is flat across a gap much wider than kT: take the midpoint of the
    // smallest mu with count >= nelec and the largest with count <= nelec
    // (equal to rounding without a gap; mid-gap, with whole orbitals, across one).
    auto bisect = [&](bool upper) {
        double a = lo, b = hi;
        for (int it = 0; it < 200 && b - a > 1e-15 * std::max(1.0, std::fabs(a) + std::fabs(b)); ++it) {
            const double mid = 0.5 * (a + b);
            const double cnt = fill(mid);
            ((upper ? cnt <= nelec : cnt < nelec) ? a : b) = mid;
        }
        return 0.5 * (a + b);
    };
    const double mu = 0.5 * (bisect(false) + bisect(true));
    fill(mu);
    return mu;
}

double fermi_entropy(const double* occ, int n) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) {
        const double o = occ[i];
        if (o > 0.0 && o < 1.0) s -= o * std::log(o) + (1.0 - o) * std::log(1.0 - o);
    }
    return s;
}

// Pulay (DIIS) density mixing (P. Pulay, Chem. Phys. Lett. 73, 393 (1980)):
// with the last `depth` input densities rho_i and residuals R_i = rho_out,i -
// rho_i, minimise |sum c_i R_i| subject to sum c_i = 1 and take
// rho_next = sum c_i (rho_i + beta R_i).  Spin densities are mixed together
// (one vector of both channels for open shells).  The history is one block
// [rho_0 .. rho_{M-1} | R_0 .. R_{M-1}] (leading dimension L = n or 2 n), so
// the new row of the DIIS matrix is one Gram pass and rho_next one block
// update.  No reference counterpart: the reference's lagged linear mixing
// (SelfConsistentFieldMolecule.cpp:447-484) stays the default.
class PulayMixer {
   public:
    // q0 > 0: Kerker screening of the update, rho_next = sum c_i rho_i + beta K(sum c_i R_i)
    // with K = -Lap (-Lap + q0^2)^-1 = 1 - q0^2 (-Lap + q0^2)^-1 (G. P. Kerker, Phys. Rev. B
    // 23, 3082 (1981)), the inverse by the sine-transform kinetic solve: it damps the
    // long-wavelength components of the residual, where a metallic system's density
    // response is largest.  The DIIS coefficients still minimise the unscreened residual.
    PulayMixer(dft_ctx* c, const dft_grid& g, int prec, int64_t n, bool closed, int depth, double beta, double q0)
        : c_(c), g_(g), prec_(prec), n_(n), closed_(closed), depth_(depth), beta_(beta), q0_(q0) {
        L_ = closed ? n : 2 * n;
        hist_ = Buf(c, (size_t)2 * depth * L_ * esize(prec));
        hist_.zero();  // unused columns enter the update with coefficient 0
        coef_ = Buf(c, sizeof(double) * (size_t)4 * depth);
        dots_ = Buf(c, sizeof(double) * (size_t)depth);
        A_.assign((size_t)depth * depth, 0.0);
        if (q0_ > 0.0) {
            rbar_ = Buf(c, (size_t)L_ * esize(prec));
            kr_ = Buf(c, (size_t)L_ * esize(prec));
        }
    }
    // Continue another mixer's history in this mixer's precision (the FP32 steps' history at the
    // hand-over of a mixed-precision run): densities and residuals converted, the DIIS matrix kept,
    // so the first FP64 steps extrapolate over the FP32 steps instead of starting from one entry.
    void take_history(const PulayMixer& o) {
        if (o.depth_ != depth_ || o.L_ != L_) return;
        check(c_, dft_convert(c_, o.hist_.get(), o.prec_, hist_.get(), prec_, (int64_t)2 * depth_ * L_), "dft_convert");
        A_ = o.A_;
        count_ = o.count_;
        head_ = o.head_;
        inherited_.assign(depth_, false);
        for (int i = 0; i < count_; ++i) inherited_[i] = true;
    }

    // Forget the entries taken over from another mixer that this one has not
    // yet overwritten: the kept entries move to slots 0.. in slot order (a
    // kept slot only ever moves down, onto a dropped or already-moved one)
    // and the DIIS matrix follows.  Returns the number of entries dropped.
    int drop_inherited() {
        if (inherited_.empty()) return 0;
        std::vector<int> keep;
        for (int i = 0; i < count_; ++i)
            if (!inherited_[i]) keep.push_back(i);
        const int dropped = count_ - (int)keep.size();
        if (dropped == 0) return 0;
        const size_t cb = (size_t)L_ * esize(prec_);
        char* h = static_cast<char*>(hist_.get());
        std::vector<double> A2((size_t)depth_ * depth_, 0.0);
        for (int t = 0; t < (int)keep.size(); ++t) {
            const int s = keep[t];
            if (s != t) {
                check(c_, dft_copy(c_, h + (size_t)t * cb, h + (size_t)s * cb, cb), "dft_copy");
                check(c_, dft_copy(c_, h + (size_t)(depth_ + t) * cb, h + (size_t)(depth_ + s) * cb, cb), "dft_copy");
            }
            for (int u = 0; u < (int)keep.size(); ++u) A2[(size_t)t * depth_ + u] = A_[(size_t)s * depth_ + keep[u]];
        }
        A_ = A2;
        count_ = (int)keep.size();
        head_ = count_ % depth_;
        inherited_.assign(depth_, false);
        return dropped;
    }

    void mix(void* ra, void* rb, const void* ora, const void* orb) {
        const size_t es = esize(prec_), cb = (size_t)L_ * es;
        char* h = static_cast<char*>(hist_.get());
        char* rin = h + (size_t)head_ * cb;
        char* res = h + (size_t)(depth_ + head_) * cb;
        const size_t fb = (size_t)n_ * es;
        check(c_, dft_copy(c_, rin, ra, fb), "dft_copy");
        check(c_, dft_copy(c_, res, ora, fb), "dft_copy");
        check(c_, dft_axpby(c_, prec_, n_, -1.0, ra, 1.0, res), "dft_axpby");
        if (!closed_) {
            check(c_, dft_copy(c_, rin + fb, rb, fb), "dft_copy");
            check(c_, dft_copy(c_, res + fb, orb, fb), "dft_copy");
            check(c_, dft_axpby(c_, prec_, n_, -1.0, rb, 1.0, res + fb), "dft_axpby");
        }
        count_ = std::min(count_ + 1, depth_);
        if (!inherited_.empty()) inherited_[head_] = false;
        // new row: <R_i, R_head> for every stored i (slots 0..count-1 hold the history)
        const int mode = prec_ == DFT_F64 ? DFT_GRAM_F64 : DFT_GRAM_F32;
        check(c_, dft_gram(c_, mode, L_, count_, 1, 1.0, h + (size_t)depth_ * cb, L_, res, L_, 0.0,
                           dots_.as<double>(), count_),
              "dft_gram");
        std::vector<double> d(count_);
        check(c_, dft_download(c_, d.data(), dots_.get(), sizeof(double) * count_), "dft_download");
        for (int i = 0; i < count_; ++i) A_[(size_t)i * depth_ + head_] = A_[(size_t)head_ * depth_ + i] = d[i];
        // bordered system [A 1; 1^T 0] [c; lambda] = [0; 1], scaled and lightly regularised
        const int m = count_ + 1;
        std::vector<double> M((size_t)m * m, 0.0), b(m, 0.0);
        double scale = 0.0;
        for (int i = 0; i < count_; ++i) scale = std::max(scale, A_[(size_t)i * depth_ + i]);
        scale = scale > 0.0 ? 1.0 / scale : 1.0;
        for (int i = 0; i < count_; ++i) {
            for (int j = 0; j < count_; ++j) M[(size_t)i * m + j] = A_[(size_t)i * depth_ + j] * scale;
            M[(size_t)i * m + i] *= 1.0 + 1e-10;
            M[(size_t)i * m + count_] = M[(size_t)count_ * m + i] = 1.0;
        }
        b[count_] = 1.0;
        std::vector<double> cvec;
        if (!solve(M, b, m, cvec)) {  // singular history: restart it from the newest entry
            cvec.assign(count_ + 1, 0.0);
            cvec[head_] = 1.0;
        }
        if (q0_ > 0.0) {
            // [rho_bar | R_bar] = history * [c 0 ; 0 c], then rho_next = rho_bar + beta K(R_bar)
            std::vector<double> C2((size_t)2 * depth_ * 2, 0.0);
            for (int i = 0; i < count_; ++i) {
                C2[i] = cvec[i];
                C2[(size_t)2 * depth_ + depth_ + i] = cvec[i];
            }
            check(c_, dft_upload(c_, coef_k_(), C2.data(), sizeof(double) * C2.size()), "dft_upload");
            const int ch = closed_ ? 1 : 2;
            for (int s = 0; s < ch; ++s) {
                char* dst = static_cast<char*>(s == 0 ? ra : rb);
                check(c_, dft_block_update(c_, prec_, n_, 2 * depth_, 1, 1.0, h + s * fb, L_, coef_k_(), 2 * depth_, 0.0,
                                           dst, n_),
                      "dft_block_update");
                check(c_, dft_block_update(c_, prec_, n_, 2 * depth_, 1, 1.0, h + s * fb, L_, coef_k_() + 2 * depth_,
                                           2 * depth_, 0.0, static_cast<char*>(rbar_.get()) + s * fb, n_),
                      "dft_block_update");
            }
            // (-1/2 Lap + q0^2/2)^-1 R = 2 (-Lap + q0^2)^-1 R, so K R = R - (q0^2 / 2) kin(R)
            check(c_, dft_precond_kinetic_exact(c_, &g_, prec_, 0.5 * q0_ * q0_, rbar_.get(), n_, ch, kr_.get(), n_),
                  "dft_precond_kinetic_exact");
            for (int s = 0; s < ch; ++s) {
                void* dst = s == 0 ? ra : rb;
                check(c_, dft_axpby(c_, prec_, n_, beta_, static_cast<char*>(rbar_.get()) + s * fb, 1.0, dst), "dft_axpby");
                check(c_, dft_axpby(c_, prec_, n_, -0.5 * beta_ * q0_ * q0_, static_cast<char*>(kr_.get()) + s * fb, 1.0,
                                    dst),
                      "dft_axpby");
                check(c_, dft_clamp_min(c_, prec_, n_, 0.0, dst), "dft_clamp_min");
            }
            if (closed_) check(c_, dft_copy(c_, rb, ra, fb), "dft_copy");
            head_ = (head_ + 1) % depth_;
            return;
        }
        // rho_next = sum c_i (rho_i + beta R_i), one pass over the history block
        std::vector<double> C((size_t)2 * depth_, 0.0);
        for (int i = 0; i < count_; ++i) {
            C[i] = cvec[i];
            C[depth_ + i] = beta_ * cvec[i];
        }
        check(c_, dft_upload(c_, coef_.get(), C.data(), sizeof(double) * C.size()), "dft_upload");
        check(c_, dft_block_update(c_, prec_, n_, 2 * depth_, 1, 1.0, h, L_, coef_.as<double>(), 2 * depth_, 0.0, ra, n_),
              "dft_block_update");
        // the extrapolation can take the density below zero in its tails
        check(c_, dft_clamp_min(c_, prec_, n_, 0.0, ra), "dft_clamp_min");
        if (closed_) check(c_, dft_copy(c_, rb, ra, fb), "dft_copy");
        else {
            check(c_, dft_block_update(c_, prec_, n_, 2 * depth_, 1, 1.0, h + fb, L_, coef_.as<double>(), 2 * depth_, 0.0,
                                       rb, n_),
                  "dft_block_update");
            check(c_, dft_clamp_min(c_, prec_, n_, 0.0, rb), "dft_clamp_min");
        }
        head_ = (head_ + 1) % depth_;
    }

   private:
    // Gaussian elimination with partial pivoting; false when singular
    static bool solve(std::vector<double> M, std::vector<double> b, int m, std::vector<double>& x) {
        for (int col = 0; col < m; ++col) {
            int piv = col;
            for (int r = col + 1; r < m; ++r)
                if (std::fabs(M[(size_t)r * m + col]) > std::fabs(M[(size_t)piv * m + col])) piv = r;
            if (!(std::fabs(M[(size_t)piv * m + col]) > 1e-300)) return false;
            if (piv != col) {
                for (int j = 0; j < m; ++j) std::swap(M[(size_t)piv * m + j], M[(size_t)col * m + j]);
                std::swap(b[piv], b[col]);
            }
            for (int r = col + 1; r < m; ++r) {
                const double f = M[(size_t)r * m + col] / M[(size_t)col * m + col];
                for (int j = col; j < m; ++j) M[(size_t)r * m + j] -= f * M[(size_t)col * m + j];
                b[r] -= f * b[col];
            }
        }
        x.assign(m, 0.0);
        for (int r = m - 1; r >= 0; --r) {
            double s = b[r];
            for (int j = r + 1; j < m; ++j) s -= M[(size_t)r * m + j] * x[j];
            x[r] = s / M[(size_t)r * m + r];
            if (!std::isfinite(x[r])) return false;
        }
        return true;
    }

    double* coef_k_() { return coef_.as<double>(); }

    dft_ctx* c_;
    dft_grid g_;
    int prec_;
    int64_t n_, L_ = 0;
    bool closed_;
    int depth_;
    double beta_, q0_;
    Buf hist_, coef_, dots_, rbar_, kr_;
    std::vector<double> A_;
    int count_ = 0, head_ = 0;
    std::vector<bool> inherited_;  // slots still holding another mixer's entries
};

struct ScfDriver::Fields {
    int prec;
    int64_t n;
    Buf vnuc, ra, rb, rho, vh, vhl, vxa, vxb, va, vb, ora, orb, orho, ovh, ta, tb;
    std::unique_ptr<BlockDavidson> da, db;
    Fields(dft_ctx* c, const dft_grid& g, int p, int k, int mmax, bool closed) : prec(p) {
        n = g.nx * g.ny * g.nz;
        const size_t b = (size_t)n * esize(p);
        for (Buf* f : {&vnuc, &ra, &rb, &rho, &vh, &vhl, &vxa, &vxb, &va, &vb, &ora, &orb, &orho, &ovh, &ta, &tb}) *f = Buf(c, b);
        vhl.zero();
        da.reset(new BlockDavidson(c, g, p, k, mmax));
        if (!closed) db.reset(new BlockDavidson(c, g, p, k, mmax));
    }
};

void nuclei_box(const dft_grid& g, int natoms, const double* nuclei, int margin, int64_t box[6]) {
    const int64_t n[3] = {g.nx, g.ny, g.nz};
    const double lo[3] = {g.xmin, g.ymin, g.zmin}, h[3] = {g.dx, g.dy, g.dz};
    for (int a = 0; a < 3; ++a) {
        double mn = 1e300, mx = -1e300;
        for (int i = 0; i < natoms; ++i) {
            mn = std::min(mn, nuclei[4 * i + 1 + a]);
            mx = std::max(mx, nuclei[4 * i + 1 + a]);
        }
        int64_t i0 = natoms ? (int64_t)std::floor((mn - lo[a]) / h[a]) - margin : 0;
        int64_t i1 = natoms ? (int64_t)std::ceil((mx - lo[a]) / h[a]) + margin + 1 : n[a];
        i0 = std::max<int64_t>(0, std::min(i0, n[a] - 1));
        i1 = std::max<int64_t>(i0 + 1, std::min(i1, n[a]));
        box[a] = i0;
        box[3 + a] = i1 - i0;
    }
}

std::unique_ptr<Preconditioner> make_preconditioner(dft_ctx* c, int kind, const void* model, const dft_grid& g,
                                                    int sweeps, double shift, double shift_scale, const int64_t* box) {
    switch (kind) {
        case DFT_PRECOND_DIAG:
            return std::unique_ptr<Preconditioner>(new DiagonalPreconditioner());
        case DFT_PRECOND_KINETIC:
            if (sweeps < 0 || sweeps > 64) throw std::invalid_argument("precond_sweeps must be in 0..64");
            if (!(shift >= 0.0) || !std::isfinite(shift)) throw std::invalid_argument("precond_shift must be >= 0");
            return std::unique_ptr<Preconditioner>(
                new KineticPreconditioner(sweeps > 0 ? sweeps : 4, shift > 0.0 ? shift : 1.0));
        case DFT_PRECOND_KINETIC_CHEB:
            if (sweeps < 0 || sweeps > 64) throw std::invalid_argument("precond_sweeps must be in 0..64");
            if (!(shift >= 0.0) || !std::isfinite(shift)) throw std::invalid_argument("precond_shift must be >= 0");
            return std::unique_ptr<Preconditioner>(
                new KineticPreconditioner(sweeps > 0 ? sweeps : 8, shift > 0.0 ? shift : 1.5, 2.0 / 3.0,
                                          KineticPreconditioner::kChebyshev));
        case DFT_PRECOND_KINETIC_EXACT:
            if (!(shift >= 0.0) || !std::isfinite(shift)) throw std::invalid_argument("precond_shift must be >= 0");
        {
            if (!std::isfinite(shift_scale)) throw std::invalid_argument("precond_shift_scale must be finite");
            // per-state shifts by default: sigma_j = max(0.5, 0.35 |theta_j|) (the best of a scan on the benzene
            // workload: 98 Davidson iterations against 124 for the best single shift, 2.0)
            const double scale = shift_scale > 0.0 ? shift_scale : shift_scale < 0.0 ? 0.0 : 0.35;
            return std::unique_ptr<Preconditioner>(new KineticPreconditioner(
                0, shift > 0.0 ? shift : scale > 0.0 ? 0.5 : 2.0, 2.0 / 3.0, KineticPreconditioner::kExact, scale));
        }
        case DFT_PRECOND_CNN: {
            auto p = make_cnn_preconditioner(c, model, g, box);
            if (!p) throw std::invalid_argument("learned preconditioner requested without a model");
            return p;
        }
        default:
            throw std::invalid_argument("unknown preconditioner");
    }
}

ScfDriver::ScfDriver(dft_ctx* c, const dft_grid& g, int natoms, const double* nuclei, const dft_scf_config& cfg)
    : c_(c), g_(g), nuc_(nuclei, nuclei + 4 * (size_t)natoms), cfg_(cfg) {}

void ScfDriver::run(dft_scf_report& rep, ScfOutputs& out) {
    rep = dft_scf_report{};
    const int natoms = (int)(nuc_.size() / 4);
    // SelfConsistentFieldMolecule.cpp:58-127
    if (natoms == 0) throw std::invalid_argument("La molecula debe contener al menos un nucleo.");
    if (cfg_.multiplicity < 1) throw std::invalid_argument("La multiplicidad electronica debe ser al menos 1.");
    if (cfg_.functional != DFT_XC_PZ81 && cfg_.functional != DFT_XC_PBE96) throw std::invalid_argument("unknown functional");
    if (cfg_.xc_flux != DFT_XC_FLUX_REFERENCE && cfg_.xc_flux != DFT_XC_FLUX_CONSISTENT_SCF)
        throw std::invalid_argument("xc_flux must be DFT_XC_FLUX_REFERENCE or DFT_XC_FLUX_CONSISTENT_SCF");
    int zsum = 0;
    for (int a = 0; a < natoms; ++a) zsum += (int)nuc_[4 * a];
    const int ne = zsum - cfg_.charge;
    if (ne < 0) throw std::invalid_argument("La carga molecular produce un numero negativo de electrones.");
    if (ne == 0) throw std::invalid_argument("El SCF molecular requiere al menos un electron.");
    const int sd = cfg_.multiplicity - 1;
    if (sd > ne) throw std::invalid_argument("La multiplicidad electronica es incompatible con el numero de electrones.");
    if ((ne - sd) % 2 != 0)
        throw std::invalid_argument("La multiplicidad electronica es incompatible con la paridad del numero de electrones.");
    const int nal = (ne + sd) / 2, nbe = (ne - sd) / 2;
    const bool closed = cfg_.multiplicity == 1;
    if (nbe <= 0) throw std::invalid_argument("El numero de orbitales debe ser mayor que cero.");
    const int nocc_max = std::max(nal, nbe);
    const int k = std::max(cfg_.nstates > 0 ? cfg_.nstates : nocc_max, nocc_max);
    const double kT = cfg_.smearing;
    const bool smear = kT > 0.0;
    if (!(kT >= 0.0) || !std::isfinite(kT)) throw std::invalid_argument("smearing must be >= 0");
    if (cfg_.ortho_gram != 0 && cfg_.ortho_gram != DFT_GRAM_F32 && cfg_.ortho_gram != DFT_GRAM_SPLIT_BF16 &&
        cfg_.ortho_gram != DFT_GRAM_SPLIT6_BF16)
        throw std::invalid_argument("ortho_gram must be 0, DFT_GRAM_F32, DFT_GRAM_SPLIT_BF16 or DFT_GRAM_SPLIT6_BF16");
    if (smear && k < nocc_max + 2)
        throw std::invalid_argument("smearing needs empty states: nstates must exceed the occupied count by 2 or more");
    const int mmax = cfg_.max_subspace > 0 ? cfg_.max_subspace : std::max(3 * k, k + 16);
    const int maxit = cfg_.max_iterations > 0 ? cfg_.max_iterations : 250;
    const int batch = cfg_.poisson_batch > 0 ? cfg_.poisson_batch : 64;
    const double dv = g_.dx * g_.dy * g_.dz;
    const int64_t n = g_.nx * g_.ny * g_.nz;
    double enn = 0.0;  // NuclearRepulsion.cpp:6-54
    for (int a = 0; a < natoms; ++a)
        for (int b = a + 1; b < natoms; ++b) {
            const double dx = nuc_[4 * a + 1] - nuc_[4 * b + 1], dy = nuc_[4 * a + 2] - nuc_[4 * b + 2],
                         dz = nuc_[4 * a + 3] - nuc_[4 * b + 3];
            const double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 == 0.0) throw std::runtime_error("NuclearRepulsion is undefined for coincident nuclei.");
            enn += nuc_[4 * a] * nuc_[4 * b] / std::sqrt(d2);
        }

    std::unique_ptr<Preconditioner> own_precond;
    if (!ext_precond_) {
        int64_t box[6];
        nuclei_box(g_, natoms, nuc_.data(), kCnnBoxMargin, box);
        own_precond = make_preconditioner(c_, cfg_.precond, cfg_.cnn, g_, cfg_.precond_sweeps, cfg_.precond_shift,
                                          cfg_.precond_shift_scale, box);
    }
    Preconditioner* precond = ext_precond_ ? ext_precond_ : own_precond.get();

    const bool mixed = cfg_.precision == DFT_SCF_MIXED;
    std::unique_ptr<Fields> F(new Fields(c_, g_, mixed ? DFT_F32 : DFT_F64, k, mmax, closed));
    const auto t0 = Clock::now();
    double t_pois = 0, t_xc = 0, t_dav = 0;
    auto tick = [&](double& acc, Clock::time_point& t) {
        check(c_, dft_sync(c_), "dft_sync");
        const auto now = Clock::now();
        acc += secs(t, now);
        t = now;
    };

    std::vector<double> occ_a(nal, 1.0 / dv), occ_b(nbe, 1.0 / dv);  // unit-norm vectors -> psi^2 = u^2/dV
    std::vector<int> el_a(nal, 1), el_b(nbe, 1);
    // smearing: per-channel occupations of the block's states (0..1), the states
    // the density and the energy take (up to the last with occ > 1e-14) and the
    // states the next Davidson must converge (occ > 1e-6)
    struct Occ {
        std::vector<double> o, w;  // occupations; o / dV for the density kernel
        int used = 0, conv = 0;
        double mu = 0.0, entropy = 0.0;
    };
    Occ sm_a, sm_b;
    sm_a.conv = nal;
    sm_b.conv = nbe;
    // The top 4 states of the block (when it has 8 or more beyond the
    // electrons) are Davidson's buffer: never occupied, never required to
    // converge, so the occupied set is always made of converged states.
    const int kocc = k - 4 >= nocc_max + 4 ? k - 4 : k;
    auto occupy = [&](Occ& q, const std::vector<double>& eps, int nel) {
        q.o.assign(k, 0.0);
        q.mu = fermi_occupations(eps.data(), kocc, (double)nel, kT, q.o.data());
        q.entropy = fermi_entropy(q.o.data(), kocc);
        q.used = q.conv = 0;
        for (int i = 0; i < kocc; ++i) {
            if (q.o[i] > 1e-14) q.used = i + 1;
            if (q.o[i] > 1e-6) q.conv = i + 1;
        }
        q.conv = std::max(q.conv, nel);
        q.used = std::max(q.used, nel);
        q.w.resize(q.used);
        for (int i = 0; i < q.used; ++i) q.w[i] = q.o[i] / dv;
    };
    DavidsonOptions dop;
    dop.tol = cfg_.davidson_tol > 0 ? cfg_.davidson_tol : 1.0e-7;
    dop.max_iter = cfg_.davidson_max_iter > 0 ? cfg_.davidson_max_iter : 200;
    bool have_prev = false;
    const double etol = cfg_.energy_tol > 0 ? cfg_.energy_tol : 1.0e-5;

    // Under the |dE| stop rule the orbitals need only be as accurate as the
    // SCF is settled: tol = clamp(min(0.1 dRho_prev, 0.3 sqrt(max(etol,
    // 0.1 |dE_prev|))), floor, 0.1) (smeared runs: below), the floor being the
    // configured tolerance (FP64) or the split-BF16 Rayleigh-Ritz floor (FP32
    // storage, ~3e-4).  The reference stop rule keeps the fixed tolerance of
    // EigenvalueSolverMolecule.cpp:19.
    double dd_prev = std::numeric_limits<double>::infinity();
    double de_prev = std::numeric_limits<double>::infinity();  // |dE| of the last step
    auto davidson = [&](Fields& f, BlockDavidson& D, const void* veff, int nocc) {
        DavidsonOptions o = dop;
        o.nconv = nocc;
        if (f.prec == DFT_F32) {
            o.gram_mode = DFT_GRAM_SPLIT_BF16;
            // FP32-product Grams: six bf16 products (0, the default) or the FP32 MFMA
            o.ortho_gram_mode = cfg_.ortho_gram == 0 ? DFT_GRAM_SPLIT6_BF16 : cfg_.ortho_gram;
            o.tol = std::max(o.tol, 3.0e-4);
            // a warm start needs few iterations; a cold start (seed vectors) keeps the full budget
            if (have_prev) o.max_iter = std::min(o.max_iter, 60);
        }
        // |dE| < etol only means convergence if the orbitals are converged
        // well beyond it: energy errors go with the square of the residual,
        // so the tolerance never exceeds 0.3 sqrt(etol) (else a loose solve
        // that leaves the orbitals unchanged gives dE = 0 by itself)
        if (cfg_.stop_rule == DFT_STOP_DE) {
            // With smeared occupations a residual r turns orbitals within ~kT of
            // each other by ~r / gap while their occupations differ by ~gap / 4kT:
            // the density moves by ~r / 4kT, so r must stay below ~0.4 kT of dRho
            const double dd_tol = smear ? std::min(0.1, 0.4 * kT) * dd_prev : 0.1 * dd_prev;
            // Whole-orbital runs: while the energy still moves by |dE| the orbitals need
            // only be converged to ~0.3 sqrt(0.1 |dE|) (capped at 0.1): early steps,
            // whose density changes by O(1), take a few iterations each.  Smeared runs
            // keep the 1e-3 / 0.3 sqrt(etol) caps: a loose early solve there can settle
            // the fractional occupations on another stationary point.
            // The first (cold) solve keeps the tight caps too: starting from the seed
            // vectors a loose solve can stop before a deep state (a 1s level the
            // seeds barely overlap) enters the lowest Ritz values, and the warm
            // solves that follow would not bring it back.
            const bool loose = !smear && have_prev;
            const double cap = loose ? 1.0e-1 : 1.0e-3;
            const double dE_ref = loose ? std::max(etol, 0.1 * de_prev) : etol;
            o.tol = std::max(o.tol, std::min({cap, dd_tol, 0.3 * std::sqrt(dE_ref)}));
        }
        DavidsonStats st = D.solve(veff, have_prev, o, *precond);
        rep.happly_columns += st.happly_columns;
        rep.block_passes += st.block_passes;
        rep.block_fallbacks += st.block_fallbacks;
        rep.davidson_iterations += st.iterations;
        if (!st.converged && f.prec == DFT_F64 && st.max_residual > 1e3 * o.tol)
            throw std::runtime_error("El orbital molecular no convergio. residual = " + std::to_string(st.max_residual));
        return st;
    };

    int cheb_cold = 1;  // the first Chebyshev solve starts from zero
    auto poisson = [&](Fields& f, const void* rho, void* v) {
        int64_t sw = 0;
        double md = 0;
        if (cfg_.poisson_method == DFT_POISSON_DIRECT) {
            check(c_, dft_poisson_direct(c_, &g_, f.prec, rho, v), "dft_poisson_direct");
        } else if (cfg_.poisson_method == DFT_POISSON_CHEBYSHEV) {
            // warm start from the last V_H; converged to tol every call
            const double tol = std::max(cfg_.poisson_tol > 0 ? cfg_.poisson_tol : 1.0e-8, f.prec == DFT_F64 ? 0.0 : 1.0e-5);
            // O(n log(1/tol)) sweeps suffice; the cap only guards against a non-converging setup
            const int64_t cap = 200 * std::max({g_.nx, g_.ny, g_.nz});
            check(c_, dft_copy(c_, v, f.vhl.get(), (size_t)n * esize(f.prec)), "dft_copy");
            check(c_, dft_poisson_chebyshev(c_, &g_, f.prec, rho, v, cap, tol, cheb_cold, batch, &sw, &md),
                  "dft_poisson_chebyshev");
            cheb_cold = 0;
            check(c_, dft_copy(c_, f.vhl.get(), v, (size_t)n * esize(f.prec)), "dft_copy");
        } else if (cfg_.poisson_fixed > 0) {
            check(c_, dft_copy(c_, v, f.vhl.get(), (size_t)n * esize(f.prec)), "dft_copy");
            check(c_, dft_poisson(c_, &g_, f.prec, rho, v, cfg_.poisson_fixed, 0.0, 0, batch, &sw, &md), "dft_poisson");
            check(c_, dft_copy(c_, f.vhl.get(), v, (size_t)n * esize(f.prec)), "dft_copy");
        } else {
            // HartreePotentialMolecule.cpp:14-15; FP32 cannot resolve 1e-8 on V ~ 1
            const double tol = f.prec == DFT_F64 ? 1.0e-8 : 1.0e-6;
            check(c_, dft_poisson(c_, &g_, f.prec, rho, v, 10000, tol, 1, batch, &sw, &md), "dft_poisson");
        }
        rep.poisson_sweeps += sw;
    };

    auto ddiff = [&](int p, const void* oldd, const void* newd) {  // MolecularSCFHelpers.cpp:46-95
        double d2 = 0, n2 = 0;
        check(c_, dft_reduce(c_, DFT_RED_DIFF2, p, newd, oldd, n, &d2), "dft_reduce");
        check(c_, dft_reduce(c_, DFT_RED_DOT, p, newd, newd, n, &n2), "dft_reduce");
        const double dn = std::sqrt(d2 * dv), nn = std::sqrt(n2 * dv);
        return nn <= 1.0e-14 ? dn : dn / nn;
    };

    // ---- initial densities (MolecularSCFHelpers.cpp:515-579) ----
    {
        Fields& f = *F;
        check(c_, dft_nuclear_potential(c_, &g_, f.prec, natoms, nuc_.data(), f.vnuc.get()), "dft_nuclear_potential");
        if (out.init_rho_a && out.init_rho_b) {
            Buf tmp(c_, (size_t)n * 8);
            tmp.upload(out.init_rho_a, (size_t)n * 8);
            check(c_, dft_convert(c_, tmp.get(), DFT_F64, f.ra.get(), f.prec, n), "dft_convert");
            tmp.upload(out.init_rho_b, (size_t)n * 8);
            check(c_, dft_convert(c_, tmp.get(), DFT_F64, f.rb.get(), f.prec, n), "dft_convert");
        } else if (cfg_.init_guess == DFT_INIT_HYDROGENIC) {
            check(c_, dft_hydrogenic_density(c_, &g_, f.prec, natoms, nuc_.data(), nal, f.ra.get()), "dft_hydrogenic_density");
            check(c_, dft_hydrogenic_density(c_, &g_, f.prec, natoms, nuc_.data(), nbe, f.rb.get()), "dft_hydrogenic_density");
        } else {
            auto t = Clock::now();
            davidson(f, *f.da, f.vnuc.get(), nal);
            if (!closed) davidson(f, *f.db, f.vnuc.get(), nbe);
            tick(t_dav, t);
            check(c_, dft_density(c_, f.prec, f.da->X(), n, nal, occ_a.data(), n, f.ra.get(), 0), "dft_density");
            if (closed) check(c_, dft_copy(c_, f.rb.get(), f.ra.get(), (size_t)n * esize(f.prec)), "dft_copy");
            else check(c_, dft_density(c_, f.prec, f.db->X(), n, nbe, occ_b.data(), n, f.rb.get(), 0), "dft_density");
            have_prev = true;
        }
    }

    double prevE = std::numeric_limits<double>::infinity(), prevDD = prevE, mixing = kMixing0;
    if (cfg_.mixer != DFT_MIX_LINEAR && cfg_.mixer != DFT_MIX_PULAY) throw std::invalid_argument("unknown density mixer");
    if (cfg_.mix_depth < 0 || cfg_.mix_depth > 32) throw std::invalid_argument("mix_depth must be in 0..32");
    if (!(cfg_.mix_kerker >= 0.0) || !std::isfinite(cfg_.mix_kerker)) throw std::invalid_argument("mix_kerker must be >= 0");
    std::unique_ptr<PulayMixer> pulay;  // rebuilt for each precision phase, keeping the history
    int pulay_prec = -1;
    bool refining = false;
    double best_dd = std::numeric_limits<double>::infinity();  // FP32 phase: lowest dRho so far
    int stall = 0;
    int refine_steps = 0;
    // FP64 refinement: three steps in a row without a 10% lower dRho drop the
    // FP32 steps' entries from the Pulay history (near the FP32 noise floor
    // they hold noise at the size of the remaining residual)
    double best_dd64 = std::numeric_limits<double>::infinity();
    int stall64 = 0;
    bool dropped32 = false;
    for (int it = 1; it <= maxit + (mixed ? std::max(cfg_.refine_iterations, 0) : 0); ++it) {
        if (!refining && it > maxit) break;
        Fields& f = *F;
        const int p = f.prec;
        const size_t fb = (size_t)n * esize(p);
        const double imix = mixing;
        auto t = Clock::now();
        // MolecularSCFIteration.cpp:48-66
        check(c_, dft_add3(c_, p, n, f.ra.get(), f.rb.get(), nullptr, f.rho.get()), "dft_add3");
        poisson(f, f.rho.get(), f.vh.get());
        tick(t_pois, t);
        // :68-90
        check(c_, dft_xc(c_, &g_, p, cfg_.functional | (cfg_.xc_flux ? DFT_XC_FLUX_CONSISTENT : 0), f.ra.get(),
                         f.rb.get(), f.vxa.get(), closed ? nullptr : f.vxb.get(), nullptr), "dft_xc");
        tick(t_xc, t);
        // a non-finite point anywhere in V_eff would poison every Davidson product: name the field
        {
            double sr = 0, sh = 0, sx = 0;
            check(c_, dft_reduce(c_, DFT_RED_SUM, p, f.rho.get(), nullptr, n, &sr), "dft_reduce");
            check(c_, dft_reduce(c_, DFT_RED_SUM, p, f.vh.get(), nullptr, n, &sh), "dft_reduce");
            check(c_, dft_reduce(c_, DFT_RED_SUM, p, f.vxa.get(), nullptr, n, &sx), "dft_reduce");
            if (!std::isfinite(sr) || !std::isfinite(sh) || !std::isfinite(sx))
                throw std::runtime_error("SCF step " + std::to_string(it) + ": non-finite input field (sum rho " +
                                         std::to_string(sr) + ", sum V_H " + std::to_string(sh) + ", sum V_xc " +
                                         std::to_string(sx) + ")");
        }
        // :92-115
        check(c_, dft_add3(c_, p, n, f.vnuc.get(), f.vh.get(), f.vxa.get(), f.va.get()), "dft_add3");
        if (!closed) check(c_, dft_add3(c_, p, n, f.vnuc.get(), f.vh.get(), f.vxb.get(), f.vb.get()), "dft_add3");
        // :117-144
        DavidsonStats sa = davidson(f, *f.da, f.va.get(), smear ? sm_a.conv : nal), sb;
        if (!closed) sb = davidson(f, *f.db, f.vb.get(), smear ? sm_b.conv : nbe);
        have_prev = true;
        if (smear) {
            occupy(sm_a, f.da->eigenvalues(), nal);
            if (!closed) occupy(sm_b, f.db->eigenvalues(), nbe);
        }
        tick(t_dav, t);
        // :163-193
        if (smear) check(c_, dft_density(c_, p, f.da->X(), n, sm_a.used, sm_a.w.data(), n, f.ora.get(), 0), "dft_density");
        else check(c_, dft_density(c_, p, f.da->X(), n, nal, occ_a.data(), n, f.ora.get(), 0), "dft_density");
        if (closed) check(c_, dft_copy(c_, f.orb.get(), f.ora.get(), fb), "dft_copy");
        else if (smear) check(c_, dft_density(c_, p, f.db->X(), n, sm_b.used, sm_b.w.data(), n, f.orb.get(), 0), "dft_density");
        else check(c_, dft_density(c_, p, f.db->X(), n, nbe, occ_b.data(), n, f.orb.get(), 0), "dft_density");
        check(c_, dft_add3(c_, p, n, f.ora.get(), f.orb.get(), nullptr, f.orho.get()), "dft_add3");
        poisson(f, f.orho.get(), f.ovh.get());
        tick(t_pois, t);
        // :195-205 -> TotalEnergyMolecule.cpp:237-333
        double E[6];
        {
            double ta = 0, tb = 0;
            if (smear) check(c_, dft_kinetic_energy_occ(c_, &g_, p, f.da->X(), n, sm_a.used, sm_a.o.data(), &ta), "dft_kinetic_energy");
            else check(c_, dft_kinetic_energy(c_, &g_, p, f.da->X(), n, nal, el_a.data(), &ta), "dft_kinetic_energy");
            if (closed) tb = ta;
            else if (smear) check(c_, dft_kinetic_energy_occ(c_, &g_, p, f.db->X(), n, sm_b.used, sm_b.o.data(), &tb), "dft_kinetic_energy");
            else check(c_, dft_kinetic_energy(c_, &g_, p, f.db->X(), n, nbe, el_b.data(), &tb), "dft_kinetic_energy");
            E[0] = (ta + tb) / dv;
            double s = 0;
            check(c_, dft_reduce(c_, DFT_RED_DOT, p, f.orho.get(), f.vnuc.get(), n, &s), "dft_reduce");
            E[1] = s * dv;
            check(c_, dft_hartree_energy(c_, &g_, p, f.orho.get(), f.ovh.get(), &E[2]), "dft_hartree_energy");
            check(c_, dft_xc(c_, &g_, p, cfg_.functional, f.ora.get(), f.orb.get(), f.ta.get(), f.tb.get(), &E[3]), "dft_xc");
            E[4] = enn;
            E[5] = E[0] + E[1] + E[2] + E[3] + E[4];
            if (smear) {
                // free energy E - TS, TS = kT (S_alpha + S_beta)
                rep.smearing_energy = -kT * (sm_a.entropy + (closed ? sm_a.entropy : sm_b.entropy));
                rep.fermi_level[0] = sm_a.mu;
                rep.fermi_level[1] = closed ? sm_a.mu : sm_b.mu;
                E[5] += rep.smearing_energy;
            }
        }
        // :207-241
        const double ddt = ddiff(p, f.rho.get(), f.orho.get());
        const double dds = std::max(ddiff(p, f.ra.get(), f.ora.get()), ddiff(p, f.rb.get(), f.orb.get()));
        const double dd = std::max(ddt, dds);
        // :243-249: max KS residual over occupied orbitals (Davidson's final residuals)
        double ks = 0.0;
        for (int j = 0; j < (smear ? std::min(sm_a.conv, k) : nal); ++j) ks = std::max(ks, f.da->residuals()[j]);
        if (!closed)
            for (int j = 0; j < (smear ? std::min(sm_b.conv, k) : nbe); ++j) ks = std::max(ks, f.db->residuals()[j]);
        // SelfConsistentFieldMolecule.cpp:297-400
        const double de = std::isfinite(prevE) ? std::fabs(E[5] - prevE) : std::numeric_limits<double>::infinity();
        rep.iterations = it;
        std::copy(E, E + 6, rep.energy);
        rep.energy_difference = de;
        rep.density_difference = dd;
        rep.ks_residual = ks;
        out.trace.insert(out.trace.end(), {E[5], de, dd, ks, imix});
        out.eig.assign(f.da->eigenvalues().begin(), f.da->eigenvalues().begin() + nal);
        const auto& eb = closed ? f.da->eigenvalues() : f.db->eigenvalues();
        out.eig.insert(out.eig.end(), eb.begin(), eb.begin() + nbe);
        tick(rep.seconds_other, t);
        if (!rep.first_de_iteration && de < etol) {
            rep.first_de_iteration = it;
            rep.seconds_to_first_de = secs(t0, Clock::now());
        }
        if (cfg_.verbose) {
            // HOMO and the lowest empty Ritz value of the alpha channel (approximate: only occupied states are converged)
            const auto& ea = f.da->eigenvalues();
            const double homo = ea[nal - 1], lumo = (int)ea.size() > nal ? ea[nal] : homo;
            std::fprintf(stderr,
                         "[scf] %3d %s E=%.12f dE=%.3e dRho=%.3e KS=%.3e mix=%.3f eps0=%.8f homo=%.5f lumo=%.5f dav=%d/%d "
                         "t=%.2fs\n",
                         it, p == DFT_F64 ? "f64" : "f32", E[5], de, dd, ks, imix, out.eig[0], homo, lumo, sa.iterations,
                         closed ? 0 : sb.iterations, secs(t0, Clock::now()));
        }
        bool conv = cfg_.stop_rule == DFT_STOP_DE ? de < etol : (de < kEnergyTol && dd < kDensityTol && ks < kKsTol);
        if (refining) ++refine_steps;
        // The FP32 phase stops when |dE| reaches the floor above or when dRho
        // stops falling (FP32 fields carry ~1e-4 Ha of energy noise and ~1e-6
        // of relative density noise) and hands over to the FP64 refinement,
        // which then applies the configured rule.
        if (mixed && !refining && f.prec == DFT_F32) {
            conv = cfg_.stop_rule == DFT_STOP_DE ? de < std::max(etol, kF32EnergyFloor) : dd < kF32DensityFloor;
            if (dd < 0.9 * best_dd) {
                best_dd = dd;
                stall = 0;
            } else if (++stall >= 3 && best_dd < kF32StallDensity) {
                conv = true;
            }
        }
        const bool hand_over = conv && mixed && !refining && cfg_.refine_iterations > 0;
        if (conv && !hand_over) {
            rep.converged = 1;
            break;
        }
        if (refining && !dropped32) {
            if (dd < 0.9 * best_dd64) {
                best_dd64 = dd;
                stall64 = 0;
            } else if (++stall64 >= 3 && pulay && pulay_prec == p) {
                pulay->drop_inherited();
                dropped32 = true;
            }
        }
        if (cfg_.mixer == DFT_MIX_PULAY) {
            if (!pulay || pulay_prec != p) {
                std::unique_ptr<PulayMixer> prev = std::move(pulay);
                pulay.reset(new PulayMixer(c_, g_, p, n, closed, cfg_.mix_depth > 0 ? cfg_.mix_depth : 6,
                                           cfg_.mix_beta > 0.0 ? cfg_.mix_beta : 0.3, cfg_.mix_kerker));
                if (prev) pulay->take_history(*prev);
                pulay_prec = p;
            }
            pulay->mix(f.ra.get(), f.rb.get(), f.ora.get(), f.orb.get());
        } else {
            // :447-484 lagged adaptive mixing
            if (std::isfinite(prevDD)) {
                if (dd > prevDD * kOsc) mixing = std::max(kMinMix, mixing * kMixDown);
                else if (dd < prevDD) mixing = std::min(kMaxMix, mixing * kMixUp);
            }
            check(c_, dft_axpby(c_, p, n, imix, f.ora.get(), 1.0 - imix, f.ra.get()), "dft_axpby");
            check(c_, dft_axpby(c_, p, n, imix, f.orb.get(), 1.0 - imix, f.rb.get()), "dft_axpby");
        }
        prevDD = dd;
        dd_prev = dd;
        de_prev = de;
        prevE = E[5];
        if (hand_over) {
            // FP64 refinement: carry densities, V_H warm start and orbitals over
            std::unique_ptr<Fields> G(new Fields(c_, g_, DFT_F64, k, mmax, closed));
            check(c_, dft_convert(c_, f.ra.get(), DFT_F32, G->ra.get(), DFT_F64, n), "dft_convert");
            check(c_, dft_convert(c_, f.rb.get(), DFT_F32, G->rb.get(), DFT_F64, n), "dft_convert");
            check(c_, dft_convert(c_, f.vhl.get(), DFT_F32, G->vhl.get(), DFT_F64, n), "dft_convert");
            check(c_, dft_nuclear_potential(c_, &g_, DFT_F64, natoms, nuc_.data(), G->vnuc.get()), "dft_nuclear_potential");
            G->da->load_X(f.da->X(), DFT_F32, k);
            if (!closed) G->db->load_X(f.db->X(), DFT_F32, k);
            F = std::move(G);
            refining = true;
            prevE = std::numeric_limits<double>::infinity();
            prevDD = prevE;
            dd_prev = 1.0e-3;
        }
    }
    check(c_, dft_sync(c_), "dft_sync");
    rep.seconds_total = secs(t0, Clock::now());
    rep.seconds_poisson = t_pois;
    rep.seconds_xc = t_xc;
    rep.seconds_davidson = t_dav;
    rep.cnn_applied = precond->applied();
    rep.cnn_fallbacks = precond->fallbacks();
    rep.refine_iterations = refine_steps;
    if (out.export_fields) {
        Fields& f = *F;
        Buf tmp(c_, (size_t)n * 8);
        auto put = [&](const void* src, double* dst) {
            if (!dst) return;
            check(c_, dft_convert(c_, src, f.prec, tmp.get(), DFT_F64, n), "dft_convert");
            tmp.download(dst, (size_t)n * 8);
        };
        const dft_scf_fields* e = out.export_fields;
        put(f.vh.get(), e->vh);
        put(f.vxa.get(), e->vxc_alpha);
        put(closed ? f.vxa.get() : f.vxb.get(), e->vxc_beta);
        put(f.ora.get(), e->rho_alpha);
        put(f.orb.get(), e->rho_beta);
        put(f.ovh.get(), e->vh_out);
    }
    if (out.want_fields) {
        Fields& f = *F;
        Buf tmp(c_, (size_t)n * 8 * std::max(nal, nbe));
        auto grab = [&](const void* src, int cols, std::vector<double>& dst, double scale) {
            check(c_, dft_convert(c_, src, f.prec, tmp.get(), DFT_F64, n * (int64_t)cols), "dft_convert");
            dst.assign((size_t)n * cols, 0.0);
            tmp.download(dst.data(), (size_t)n * cols * 8);
            if (scale != 1.0)
                for (double& x : dst) x *= scale;
        };
        const double s = 1.0 / std::sqrt(dv);
        grab(f.da->X(), nal, out.psi_a, s);
        if (closed) out.psi_b = out.psi_a; else grab(f.db->X(), nbe, out.psi_b, s);
        grab(f.va.get(), 1, out.veff_a, 1.0);
        if (closed) out.veff_b = out.veff_a; else grab(f.vb.get(), 1, out.veff_b, 1.0);
    }
    if (out.want_rho) {
        Fields& f = *F;
        out.rho.assign(2 * (size_t)n, 0.0);
        Buf tmp(c_, (size_t)n * 8);
        check(c_, dft_convert(c_, f.ora.get(), f.prec, tmp.get(), DFT_F64, n), "dft_convert");
        tmp.download(out.rho.data(), (size_t)n * 8);
        check(c_, dft_convert(c_, f.orb.get(), f.prec, tmp.get(), DFT_F64, n), "dft_convert");
        tmp.download(out.rho.data() + n, (size_t)n * 8);
    }
}

}  // namespace dftc
