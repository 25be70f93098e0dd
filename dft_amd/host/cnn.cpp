// cnn.cpp -- the learned Davidson preconditioner around the HIP CNN kernels
// (cnn.hip): inference with a per-column diagonal fallback, the projected-CG
// harvest of training pairs from the method's own SCF runs, and the trainer.
//
// The reference preconditions each Davidson correction with the diagonal of
// H only (davidsonCorrection, EigenvalueSolverMolecule.cpp:483-555, called at
// :1042-1047; its non-finite / tiny-denominator fallbacks are :1049-1104).
// Here that diagonal correction stays the fallback, per iteration: a column
// whose residual norm did not go down after a learned correction gets the
// diagonal one at the next iteration, and the learned one again after that.
//
// Two kinds of model (dft_cnn_model::base):
//   0: the network's output is the whole correction (c / (max|r| h2) or
//      c / max|c|), on the full grid;
//   1: the network's output, times max|r| h2, is added to the exact kinetic
//      correction t_kin = (-1/2 Lap + sigma_j)^-1 r (per-state shifts
//      sigma_j = max(0.5, 0.35 |theta_j|)) over a box around the nuclei,
//      ramped to zero over 4 points at box faces inside the grid; outside the
//      box the correction is t_kin.  The network then carries only what the
//      potential adds to the kinetic inverse, which is largest near nuclei.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include <functional>

#include "capi_util.hpp"
#include "davidson.hpp"
#include "dft_core.h"
#include "gpu.hpp"
#include "scf.hpp"

namespace dftc {

namespace {

constexpr int kChannels = 16;

int64_t param_count(int ch) {
    int64_t n = 0;
    if (dft_cnn_param_count(ch, &n) != DFT_OK) throw std::invalid_argument("dft_cnn_param_count: bad channel count");
    return n;
}

int64_t grid_n(const dft_grid& g) { return g.nx * g.ny * g.nz; }

}  // namespace

// He-uniform initial parameters in the kernel layout (W[cout][cin][27], b[cout]
// per layer; biases zero).
std::vector<float> cnn_initial_params(int ch, uint64_t seed) {
    std::vector<float> p((size_t)param_count(ch), 0.0f);
    std::mt19937_64 rng(seed);
    const int cin[4] = {2, ch, ch, ch}, cout[4] = {ch, ch, ch, 1};
    size_t off = 0;
    for (int l = 0; l < 4; ++l) {
        const double b = std::sqrt(6.0 / (cin[l] * 27.0)) * (l == 3 ? 0.1 : 1.0);
        std::uniform_real_distribution<double> u(-b, b);
        const size_t nw = (size_t)cout[l] * cin[l] * 27;
        for (size_t i = 0; i < nw; ++i) p[off + i] = (float)u(rng);
        off += nw + cout[l];
    }
    return p;
}

// ---- inference ---------------------------------------------------------------

class CnnPreconditioner final : public Preconditioner {
   public:
    CnnPreconditioner(dft_ctx* c, const dft_cnn_model& m, const dft_grid& g, const int64_t* box)
        : c_(c), g_(g), n_(grid_n(g)), base_(m.base == 1),
          kin_(0, 0.5, 2.0 / 3.0, KineticPreconditioner::kExact, 0.35) {
        if (m.channels != kChannels || !m.params || m.count != param_count(m.channels))
            throw std::invalid_argument("learned preconditioner: model does not match the 16-channel network");
        if (m.base != 0 && m.base != 1) throw std::invalid_argument("learned preconditioner: model base must be 0 or 1");
        const int64_t full[6] = {0, 0, 0, g.nx, g.ny, g.nz};
        for (int i = 0; i < 6; ++i) box_[i] = base_ && box ? box[i] : full[i];
        if (box_[0] < 0 || box_[1] < 0 || box_[2] < 0 || box_[3] < 1 || box_[4] < 1 || box_[5] < 1 ||
            box_[0] + box_[3] > g.nx || box_[1] + box_[4] > g.ny || box_[2] + box_[5] > g.nz)
            throw std::invalid_argument("learned preconditioner: box outside the grid");
        nb_ = box_[3] * box_[4] * box_[5];
        count_ = m.count;
        w32_ = Buf(c, sizeof(float) * count_);
        w32_.upload(m.params, sizeof(float) * count_);
        wbf_ = Buf(c, sizeof(uint16_t) * count_);
        check(c, dft_cnn_to_bf16(c, count_, w32_.as<float>(), wbf_.as<uint16_t>()), "dft_cnn_to_bf16");
        // columns per forward pass: activations of 2 x 16 channels plus input/output per column
        const double per_col = 4.0 * nb_ * (2 * kChannels + 3);
        chunk_ = (int)std::max(1.0, std::min(32.0, std::floor(kBudget / per_col)));
        int64_t act = 0;
        check(c, dft_cnn_act_count(kChannels, box_[3], box_[4], box_[5], chunk_, 0, &act), "dft_cnn_act_count");
        act_ = Buf(c, sizeof(float) * act);
        in_ = Buf(c, sizeof(float) * 2 * nb_ * chunk_);
        out_ = Buf(c, sizeof(float) * nb_ * chunk_);
    }

    int apply(dft_ctx* c, const dft_grid& g, int prec, const void* veff, const void* X, int64_t ldx, const void* R,
              int64_t ldr, int first, int ncols, const double* theta_dev, void* T, int64_t ldt) override {
        if (ncols <= 0) return 0;
        const size_t es = esize(prec);
        if ((int)mode_.size() < first + ncols) {
            mode_.resize(first + ncols, 0);
            fails_.resize(first + ncols, 0);
        }
        if ((int)scale_cap_ < ncols) {
            scale_ = Buf(c, sizeof(double) * ncols);
            scale_cap_ = ncols;
        }
        if (base_) kin_.apply(c, g, prec, veff, X, ldx, R, ldr, first, ncols, theta_dev, T, ldt);
        double* sc = scale_.as<double>();
        for (int j0 = 0; j0 < ncols; j0 += chunk_) {
            const int nc = std::min(chunk_, ncols - j0);
            const void* Rj = static_cast<const char*>(R) + (size_t)j0 * ldr * es;
            void* Tj = static_cast<char*>(T) + (size_t)j0 * ldt * es;
            check(c, dft_column_maxabs(c, prec, n_, nc, Rj, ldr, sc + j0), "dft_column_maxabs");
            if (base_) {
                check(c, dft_cnn_prepare_box(c, &g, prec, veff, Rj, ldr, nc, theta_dev + j0, sc + j0, box_,
                                             in_.as<float>()),
                      "dft_cnn_prepare_box");
                check(c, dft_cnn_forward(c, kChannels, box_[3], box_[4], box_[5], nc, wbf_.as<uint16_t>(),
                                         in_.as<float>(), out_.as<float>(), act_.as<float>(), 0),
                      "dft_cnn_forward");
                check(c, dft_cnn_add_box(c, &g, prec, nc, out_.as<float>(), sc + j0, box_, kTaper, Tj, ldt),
                      "dft_cnn_add_box");
            } else {
                check(c, dft_cnn_prepare(c, &g, prec, veff, Rj, ldr, nc, theta_dev + j0, sc + j0, in_.as<float>()),
                      "dft_cnn_prepare");
                check(c, dft_cnn_forward(c, kChannels, g.nx, g.ny, g.nz, nc, wbf_.as<uint16_t>(), in_.as<float>(),
                                         out_.as<float>(), act_.as<float>(), 0),
                      "dft_cnn_forward");
                check(c, dft_cnn_finish(c, &g, prec, nc, out_.as<float>(), sc + j0, Tj, ldt), "dft_cnn_finish");
            }
        }
        // diagonal correction for the columns in fallback, in contiguous runs
        int nf = 0;
        for (int j = 0; j < ncols;) {
            const int gj = first + j;
            const bool fb = mode_[gj] == kFallbackNext;
            if (!fb) {
                mode_[gj] = kLearned;
                ++j;
                continue;
            }
            int j1 = j;
            while (j1 < ncols && mode_[first + j1] == kFallbackNext) ++j1;
            DiagonalPreconditioner d;
            d.apply(c, g, prec, veff, X, ldx, static_cast<const char*>(R) + (size_t)j * ldr * es, ldr, gj, j1 - j,
                    theta_dev + j, static_cast<char*>(T) + (size_t)j * ldt * es, ldt);
            for (int q = j; q < j1; ++q) mode_[first + q] = kDiagonal;
            nf += j1 - j;
            j = j1;
        }
        applied_ += ncols - nf;
        fallbacks_ += nf;
        return nf;
    }

    void feedback(const std::vector<double>& before, const std::vector<double>& after) override {
        const size_t n = std::min({before.size(), after.size(), mode_.size()});
        for (size_t j = 0; j < n; ++j) {
            if (mode_[j] == kLearned) {
                const bool reduced = after[j] < before[j];
                if (!reduced) ++fails_[j];
                mode_[j] = reduced ? kNone : kFallbackNext;
            } else if (mode_[j] == kDiagonal) {
                mode_[j] = kNone;  // try the learned correction again
            }
        }
    }

    void reset() override {
        std::fill(mode_.begin(), mode_.end(), kNone);
        std::fill(fails_.begin(), fails_.end(), 0);
    }
    int64_t applied() const override { return applied_; }
    int64_t fallbacks() const override { return fallbacks_; }

   private:
    static constexpr double kBudget = 6.0e9;  // bytes of activations per forward pass
    static constexpr int kTaper = 4;
    enum : int { kNone = 0, kLearned = 1, kDiagonal = 2, kFallbackNext = 3 };
    dft_ctx* c_;
    dft_grid g_;
    int64_t n_, count_ = 0, nb_ = 0;
    int64_t box_[6];
    bool base_;
    KineticPreconditioner kin_;
    int chunk_ = 1;
    Buf w32_, wbf_, act_, in_, out_, scale_;
    int scale_cap_ = 0;
    std::vector<int> mode_, fails_;
    int64_t applied_ = 0, fallbacks_ = 0;
};

std::unique_ptr<Preconditioner> make_cnn_preconditioner(dft_ctx* c, const void* model, const dft_grid& g,
                                                        const int64_t* box) {
    if (!model) return nullptr;
    return std::unique_ptr<Preconditioner>(new CnnPreconditioner(c, *static_cast<const dft_cnn_model*>(model), g, box));
}

// ---- training-data harvest ---------------------------------------------------

struct PatchSet {
    int P = 32;
    int64_t P3 = 0, cap = 0, count = 0;
    Buf in, tgt;
    PatchSet(dft_ctx* c, int patch, int64_t capacity) : P(patch), P3((int64_t)patch * patch * patch), cap(capacity) {
        in = Buf(c, sizeof(float) * 2 * P3 * cap);
        tgt = Buf(c, sizeof(float) * P3 * cap);
    }
};

// The diagonal correction (Davidson runs as usual), plus, at every k-th
// correction step of a solve, the projected CG solve of
// (I - QQ^T)(H - theta_j)(I - QQ^T) c_j = -r_j (Q = Ritz vectors 0..j) for
// up to max_columns residuals, whose high-residual patches become training pairs.
class HarvestPreconditioner final : public Preconditioner {
   public:
    bool needs_ritz_vectors() const override { return true; }  // projects the CG targets against X
    HarvestPreconditioner(dft_ctx* c, const dft_grid& g, PatchSet& ps, const dft_cnn_harvest_config& hc)
        : c_(c), g_(g), n_(grid_n(g)), ps_(ps), hc_(hc), rng_(hc.seed),
          kin_(0, 0.5, 2.0 / 3.0, KineticPreconditioner::kExact, 0.35) {
        if (hc.patch != ps.P) throw std::invalid_argument("harvest: patch size differs from the trainer's");
        if (hc.target_scale < 0 || hc.target_scale > 2) throw std::invalid_argument("harvest: target_scale must be 0, 1 or 2");
        if (g.nx < ps.P || g.ny < ps.P || g.nz < ps.P) throw std::invalid_argument("harvest: grid smaller than a patch");
        if (hc.max_columns < 1 || hc.patches_per_column < 1 || hc.every < 1 || hc.cg_max_iter < 1 || !(hc.cg_tol > 0))
            throw std::invalid_argument("harvest: bad configuration");
        const size_t colb = sizeof(double) * n_ * hc.max_columns;
        Cc_ = Buf(c, colb);
        Res_ = Buf(c, colb);
        Pd_ = Buf(c, colb);
        Ap_ = Buf(c, colb);
        small_ = Buf(c, sizeof(double) * 8 * hc.max_columns);
    }

    int apply(dft_ctx* c, const dft_grid& g, int prec, const void* veff, const void* X, int64_t ldx, const void* R,
              int64_t ldr, int first, int ncols, const double* theta_dev, void* T, int64_t ldt) override {
        if (hc_.target_scale == 2) {
            kin_.apply(c, g, prec, veff, X, ldx, R, ldr, first, ncols, theta_dev, T, ldt);
        } else {
            DiagonalPreconditioner d;
            d.apply(c, g, prec, veff, X, ldx, R, ldr, first, ncols, theta_dev, T, ldt);
        }
        if (prec != DFT_F64 || ps_.count >= ps_.cap || ncols <= 0) return 0;
        if (calls_++ % hc_.every != 0) return 0;
        harvest(veff, X, ldx, R, ldr, first, std::min(ncols, hc_.max_columns), theta_dev, T, ldt);
        return 0;
    }
    void reset() override { calls_ = 0; }

    int64_t added = 0, solves = 0, failed = 0, cg_iterations = 0;
    // max over converged columns of |b - A c| / |b|, recomputed from c after the CG (b = -(I - QQ^T) r,
    // A = (I - QQ^T)(H - theta)(I - QQ^T)): the recurrence's residual can drift from the true one
    double max_true_residual = 0.0;

   private:
    void project(const void* X, int64_t ldx, int first, int nc, void* Y) {
        const int ma = first + nc;
        if ((int64_t)S_h_.size() < (int64_t)ma * nc) {
            S_h_.resize((size_t)ma * nc);
            S_ = Buf(c_, sizeof(double) * ma * nc);
        }
        check(c_, dft_gram(c_, DFT_GRAM_F64, n_, ma, nc, 1.0, X, ldx, Y, n_, 0.0, S_.as<double>(), ma), "dft_gram");
        S_.download(S_h_.data(), sizeof(double) * ma * nc);
        for (int j = 0; j < nc; ++j)
            for (int i = first + j + 1; i < ma; ++i) S_h_[(size_t)j * ma + i] = 0.0;
        S_.upload(S_h_.data(), sizeof(double) * ma * nc);
        check(c_, dft_block_update(c_, DFT_F64, n_, ma, nc, -1.0, X, ldx, S_.as<double>(), ma, 1.0, Y, n_),
              "dft_block_update");
    }

    void dots(const void* A, const void* B, int nc, std::vector<double>& out, int64_t lda = 0) {
        double* d = small_.as<double>() + 6 * hc_.max_columns;
        check(c_, dft_column_dots(c_, DFT_F64, n_, nc, A, lda ? lda : n_, B, n_, d), "dft_column_dots");
        out.resize(nc);
        check(c_, dft_download(c_, out.data(), d, sizeof(double) * nc), "dft_download");
    }

    void harvest(const void* veff, const void* X, int64_t ldx, const void* R, int64_t ldr, int first, int nc,
                 const double* theta_dev, const void* Tb, int64_t ldt) {
        const int mc = hc_.max_columns;
        double* a = small_.as<double>();
        double* b = a + mc;
        double* ma = a + 2 * mc;
        double* ones = a + 3 * mc;
        double* scale = a + 4 * mc;
        std::vector<double> h_ones(nc, 1.0), h_m1(nc, -1.0), h(nc);
        check(c_, dft_upload(c_, ones, h_ones.data(), sizeof(double) * nc), "dft_upload");
        check(c_, dft_upload(c_, ma, h_m1.data(), sizeof(double) * nc), "dft_upload");
        // Res = P(-R), C = 0, P = Res
        check(c_, dft_scale_columns(c_, DFT_F64, n_, nc, R, ldr, ma, Res_.get(), n_), "dft_scale_columns");
        project(X, ldx, first, nc, Res_.get());
        check(c_, dft_zero(c_, Cc_.get(), sizeof(double) * n_ * nc), "dft_zero");
        check(c_, dft_copy(c_, Pd_.get(), Res_.get(), sizeof(double) * n_ * nc), "dft_copy");
        std::vector<double> rr, b2, pap, rrn;
        dots(Res_.get(), Res_.get(), nc, rr);
        b2 = rr;
        std::vector<int> state(nc, 0);  // 0 active, 1 converged, 2 failed
        for (int j = 0; j < nc; ++j)
            if (!(rr[j] > 0.0)) state[j] = 2;
        int it = 0;
        for (; it < hc_.cg_max_iter; ++it) {
            if (std::none_of(state.begin(), state.end(), [](int s) { return s == 0; })) break;
            check(c_, dft_hpsi(c_, &g_, DFT_F64, veff, Pd_.get(), n_, nc, Ap_.get(), n_), "dft_hpsi");
            check(c_, dft_residual(c_, DFT_F64, n_, nc, Ap_.get(), n_, Pd_.get(), n_, theta_dev, Ap_.get(), n_),
                  "dft_residual");
            project(X, ldx, first, nc, Ap_.get());
            dots(Pd_.get(), Ap_.get(), nc, pap);
            for (int j = 0; j < nc; ++j) {
                h[j] = 0.0;
                if (state[j] != 0) continue;
                if (!(pap[j] > 0.0)) {  // not positive definite on this direction
                    state[j] = 2;
                    continue;
                }
                h[j] = rr[j] / pap[j];
            }
            check(c_, dft_upload(c_, a, h.data(), sizeof(double) * nc), "dft_upload");
            for (int j = 0; j < nc; ++j) h[j] = -h[j];
            check(c_, dft_upload(c_, b, h.data(), sizeof(double) * nc), "dft_upload");
            check(c_, dft_cols_axpby(c_, DFT_F64, n_, nc, a, Pd_.get(), n_, ones, Cc_.get(), n_), "dft_cols_axpby");
            check(c_, dft_cols_axpby(c_, DFT_F64, n_, nc, b, Ap_.get(), n_, ones, Res_.get(), n_), "dft_cols_axpby");
            dots(Res_.get(), Res_.get(), nc, rrn);
            for (int j = 0; j < nc; ++j) {
                h[j] = 0.0;
                if (state[j] != 0) continue;
                if (std::sqrt(rrn[j]) <= hc_.cg_tol * std::sqrt(b2[j])) state[j] = 1;
                h[j] = rrn[j] / rr[j];
                rr[j] = rrn[j];
            }
            check(c_, dft_upload(c_, b, h.data(), sizeof(double) * nc), "dft_upload");
            check(c_, dft_cols_axpby(c_, DFT_F64, n_, nc, ones, Res_.get(), n_, b, Pd_.get(), n_), "dft_cols_axpby");
        }
        cg_iterations += it;
        solves += nc;
        if (std::any_of(state.begin(), state.end(), [](int s) { return s == 1; })) {
            // true residual: Pd = b (recomputed), Ap = A c, Pd -= Ap
            check(c_, dft_scale_columns(c_, DFT_F64, n_, nc, R, ldr, ma, Pd_.get(), n_), "dft_scale_columns");
            project(X, ldx, first, nc, Pd_.get());
            check(c_, dft_hpsi(c_, &g_, DFT_F64, veff, Cc_.get(), n_, nc, Ap_.get(), n_), "dft_hpsi");
            check(c_, dft_residual(c_, DFT_F64, n_, nc, Ap_.get(), n_, Cc_.get(), n_, theta_dev, Ap_.get(), n_),
                  "dft_residual");
            project(X, ldx, first, nc, Ap_.get());
            check(c_, dft_cols_axpby(c_, DFT_F64, n_, nc, ma, Ap_.get(), n_, ones, Pd_.get(), n_), "dft_cols_axpby");
            std::vector<double> tr;
            dots(Pd_.get(), Pd_.get(), nc, tr);
            for (int j = 0; j < nc; ++j)
                if (state[j] == 1) max_true_residual = std::max(max_true_residual, std::sqrt(tr[j] / b2[j]));
        }
        if (hc_.target_scale == 2) {
            // what the network must add to t_kin (the base correction, in Tb) beyond the span of the
            // Ritz vectors, which Davidson projects out: with t = (I - QQ^T) t_kin (Q as in the CG),
            // the target is beta c - t, beta = <t, c> / <c, c> (the scale of c closest to t)
            double* t = Pd_.as<double>();
            for (int j = 0; j < nc; ++j)
                check(c_, dft_copy(c_, t + (size_t)j * n_, static_cast<const char*>(Tb) + (size_t)j * ldt * sizeof(double),
                                   sizeof(double) * n_), "dft_copy");
            project(X, ldx, first, nc, t);
            std::vector<double> tc, cc;
            dots(t, Cc_.get(), nc, tc);
            dots(Cc_.get(), Cc_.get(), nc, cc);
            std::vector<double> beta(nc), m1(nc, -1.0);
            for (int j = 0; j < nc; ++j) beta[j] = cc[j] > 0.0 ? tc[j] / cc[j] : 0.0;
            check(c_, dft_upload(c_, a, m1.data(), sizeof(double) * nc), "dft_upload");
            check(c_, dft_upload(c_, b, beta.data(), sizeof(double) * nc), "dft_upload");
            check(c_, dft_cols_axpby(c_, DFT_F64, n_, nc, a, t, n_, b, Cc_.get(), n_), "dft_cols_axpby");
        }
        // patches: the highest-residual aligned blocks of each converged column, jittered
        const int P = ps_.P;
        const int nbx = (int)(g_.nx / P), nby = (int)(g_.ny / P), nbz = (int)(g_.nz / P), nb = nbx * nby * nbz;
        if ((int64_t)E_h_.size() < (int64_t)nb * nc) {
            E_h_.resize((size_t)nb * nc);
            E_ = Buf(c_, sizeof(double) * nb * nc);
        }
        check(c_, dft_patch_energy(c_, &g_, DFT_F64, P, R, ldr, nc, E_.as<double>()), "dft_patch_energy");
        E_.download(E_h_.data(), sizeof(double) * nb * nc);
        std::vector<int> desc;
        std::uniform_int_distribution<int> jit(-P / 4, P / 4);
        for (int j = 0; j < nc; ++j) {
            if (state[j] != 1) {
                ++failed;
                continue;
            }
            std::vector<int> order(nb);
            std::iota(order.begin(), order.end(), 0);
            const double* e = E_h_.data() + (size_t)j * nb;
            std::sort(order.begin(), order.end(), [&](int p, int q) { return e[p] > e[q]; });
            // the highest-residual blocks, then as many blocks drawn at random from the rest (the
            // network also sees where the residual, and so the correction, is small)
            const int ntop = std::min(nb, (hc_.patches_per_column + 1) / 2);
            const int nrand = std::min(nb - ntop, hc_.patches_per_column - ntop);
            std::shuffle(order.begin() + ntop, order.end(), rng_);
            for (int k = 0; k < ntop + nrand; ++k) {
                if (ps_.count + (int64_t)desc.size() / 4 >= ps_.cap) break;
                const int blk = order[k];
                const int bx = blk % nbx, by = (blk / nbx) % nby, bz = blk / (nbx * nby);
                auto place = [&](int base, int n) { return std::min(std::max(base + jit(rng_), 0), n - P); };
                desc.insert(desc.end(), {j, place(bx * P, (int)g_.nx), place(by * P, (int)g_.ny), place(bz * P, (int)g_.nz)});
            }
        }
        const int np = (int)desc.size() / 4;
        if (np == 0) return;
        double* cscale = a + 5 * mc;
        check(c_, dft_column_maxabs(c_, DFT_F64, n_, nc, R, ldr, scale), "dft_column_maxabs");
        if (hc_.target_scale >= 1) {
            // cs = max|r| h2: the target keeps the input's scale (the correction operator stays linear)
            const double h2 = 1.0 / (1.0 / (g_.dx * g_.dx) + 1.0 / (g_.dy * g_.dy) + 1.0 / (g_.dz * g_.dz));
            std::vector<double> hs(nc);
            check(c_, dft_download(c_, hs.data(), scale, sizeof(double) * nc), "dft_download");
            for (double& x : hs) x *= h2;
            check(c_, dft_upload(c_, cscale, hs.data(), sizeof(double) * nc), "dft_upload");
        } else {
            check(c_, dft_column_maxabs(c_, DFT_F64, n_, nc, Cc_.get(), n_, cscale), "dft_column_maxabs");
        }
        for (int p0 = 0; p0 < np; p0 += 65535) {
            const int cnt = std::min(65535, np - p0);
            check(c_, dft_extract_patches(c_, &g_, DFT_F64, P, cnt, desc.data() + 4 * p0, nc, veff, R, ldr, Cc_.get(), n_,
                                          theta_dev, scale, cscale, ps_.in.as<float>() + 2 * ps_.P3 * ps_.count,
                                          ps_.tgt.as<float>() + ps_.P3 * ps_.count),
                  "dft_extract_patches");
            ps_.count += cnt;
            added += cnt;
        }
    }

    dft_ctx* c_;
    dft_grid g_;
    int64_t n_;
    PatchSet& ps_;
    dft_cnn_harvest_config hc_;
    std::mt19937_64 rng_;
    KineticPreconditioner kin_;
    int64_t calls_ = 0;
    Buf Cc_, Res_, Pd_, Ap_, small_, S_, E_;
    std::vector<double> S_h_, E_h_;
};

}  // namespace dftc

// ---- trainer (C-ABI object) ---------------------------------------------------

struct dft_cnn_trainer {
    dftc::Ctx ctx{0};
    int ch = 16;
    std::unique_ptr<dftc::PatchSet> ps;
    int64_t count = 0;  // parameters
    dftc::Buf w32, wbf, grad, m, v;
    int step = 0;
    std::mt19937_64 rng;
    dft_cnn_harvest_stats hstats{};
    // batch buffers
    int bcap = 0;
    dftc::Buf bin, btg, bout, act, scratch;

    void ensure_batch(int B) {
        if (B <= bcap) return;
        const int64_t P = ps->P, P3 = ps->P3;
        bin = dftc::Buf(ctx, sizeof(float) * 2 * P3 * B);
        btg = dftc::Buf(ctx, sizeof(float) * P3 * B);
        bout = dftc::Buf(ctx, sizeof(float) * P3 * B);
        int64_t na = 0, ns = 0;
        dftc::check(ctx, dft_cnn_act_count(ch, P, P, P, B, 1, &na), "dft_cnn_act_count");
        dftc::check(ctx, dft_cnn_scratch_count(ch, P, P, P, B, &ns), "dft_cnn_scratch_count");
        act = dftc::Buf(ctx, sizeof(float) * na);
        scratch = dftc::Buf(ctx, sizeof(float) * ns);
        bcap = B;
    }
    void gather(const std::vector<int64_t>& idx) {
        const int64_t P3 = ps->P3;
        for (size_t b = 0; b < idx.size(); ++b) {
            dftc::check(ctx, dft_copy(ctx, bin.as<float>() + 2 * P3 * b, ps->in.as<float>() + 2 * P3 * idx[b], sizeof(float) * 2 * P3),
                        "dft_copy");
            dftc::check(ctx, dft_copy(ctx, btg.as<float>() + P3 * b, ps->tgt.as<float>() + P3 * idx[b], sizeof(float) * P3),
                        "dft_copy");
        }
    }
};

namespace {
template <class F>
int guarded_cnn(F&& f) {
    return dftc::guarded_call(std::function<void()>(std::forward<F>(f)));
}
}  // namespace

extern "C" {

int dft_cnn_trainer_create(int channels, int patch, int64_t capacity, uint64_t seed, dft_cnn_trainer** out) {
    return guarded_cnn([&] {
        if (!out || channels != 16 || patch < 4 || capacity < 1) throw std::invalid_argument("dft_cnn_trainer_create: bad arguments");
        std::unique_ptr<dft_cnn_trainer> t(new dft_cnn_trainer());
        t->ch = channels;
        t->ps.reset(new dftc::PatchSet(t->ctx, patch, capacity));
        t->count = dftc::param_count(channels);
        std::vector<float> p = dftc::cnn_initial_params(channels, seed);
        t->w32 = dftc::Buf(t->ctx, sizeof(float) * t->count);
        t->wbf = dftc::Buf(t->ctx, sizeof(uint16_t) * t->count);
        t->grad = dftc::Buf(t->ctx, sizeof(float) * t->count);
        t->m = dftc::Buf(t->ctx, sizeof(float) * t->count);
        t->v = dftc::Buf(t->ctx, sizeof(float) * t->count);
        t->w32.upload(p.data(), sizeof(float) * t->count);
        t->m.zero();
        t->v.zero();
        dftc::check(t->ctx, dft_cnn_to_bf16(t->ctx, t->count, t->w32.as<float>(), t->wbf.as<uint16_t>()), "dft_cnn_to_bf16");
        t->rng.seed(seed ^ 0x9E3779B97F4A7C15ULL);
        *out = t.release();
    });
}

void dft_cnn_trainer_destroy(dft_cnn_trainer* t) { delete t; }

int dft_cnn_trainer_harvest(dft_cnn_trainer* t, const dft_grid* g, int natoms, const double* nuclei,
                            const dft_scf_config* cfg, const dft_cnn_harvest_config* hc, dft_scf_report* rep,
                            int64_t* added) {
    return guarded_cnn([&] {
        if (!t || !g || !cfg || !hc || !rep || (natoms > 0 && !nuclei)) throw std::invalid_argument("dft_cnn_trainer_harvest: null argument");
        if (cfg->precision != DFT_SCF_FP64) throw std::invalid_argument("dft_cnn_trainer_harvest: harvest runs the FP64 SCF");
        dftc::HarvestPreconditioner hp(t->ctx, *g, *t->ps, *hc);
        dftc::ScfDriver drv(t->ctx, *g, natoms, nuclei, *cfg);
        drv.set_preconditioner(&hp);
        dftc::ScfOutputs out;
        drv.run(*rep, out);
        if (added) *added = hp.added;
        t->hstats.solves += hp.solves;
        t->hstats.failed += hp.failed;
        t->hstats.cg_iterations += hp.cg_iterations;
        t->hstats.patches += hp.added;
        t->hstats.max_true_residual = std::max(t->hstats.max_true_residual, hp.max_true_residual);
    });
}

int dft_cnn_trainer_harvest_stats(dft_cnn_trainer* t, dft_cnn_harvest_stats* out) {
    if (!t || !out) return DFT_EINVAL;
    *out = t->hstats;
    return DFT_OK;
}

int dft_cnn_trainer_count(dft_cnn_trainer* t, int64_t* patches) {
    if (!t || !patches) return DFT_EINVAL;
    *patches = t->ps->count;
    return DFT_OK;
}

int dft_cnn_trainer_train(dft_cnn_trainer* t, int steps, int batch, float lr, int64_t train_count, double* loss) {
    return guarded_cnn([&] {
        if (!t || steps < 0 || batch < 1 || !(lr > 0)) throw std::invalid_argument("dft_cnn_trainer_train: bad arguments");
        const int64_t avail = train_count > 0 ? std::min(train_count, t->ps->count) : t->ps->count;
        if (avail < 1) throw std::invalid_argument("dft_cnn_trainer_train: no patches");
        t->ensure_batch(batch);
        const int64_t P = t->ps->P;
        std::uniform_int_distribution<int64_t> pick(0, avail - 1);
        std::vector<int64_t> idx(batch);
        for (int s = 0; s < steps; ++s) {
            for (int b = 0; b < batch; ++b) idx[b] = pick(t->rng);
            t->gather(idx);
            dftc::check(t->ctx, dft_cnn_forward(t->ctx, t->ch, P, P, P, batch, t->wbf.as<uint16_t>(), t->bin.as<float>(),
                                                t->bout.as<float>(), t->act.as<float>(), 1),
                        "dft_cnn_forward");
            double l = 0.0;
            dftc::check(t->ctx, dft_cnn_backward(t->ctx, t->ch, P, P, P, batch, t->wbf.as<uint16_t>(), t->bin.as<float>(),
                                                 t->act.as<float>(), t->bout.as<float>(), t->btg.as<float>(),
                                                 t->grad.as<float>(), t->scratch.as<float>(), loss ? &l : nullptr),
                        "dft_cnn_backward");
            ++t->step;
            dftc::check(t->ctx, dft_cnn_adam(t->ctx, t->count, t->w32.as<float>(), t->wbf.as<uint16_t>(), t->grad.as<float>(),
                                             t->m.as<float>(), t->v.as<float>(), t->step, lr, 0.9f, 0.999f, 1e-8f),
                        "dft_cnn_adam");
            if (loss) loss[s] = l;
        }
        t->ctx.sync();
    });
}

int dft_cnn_trainer_eval(dft_cnn_trainer* t, int64_t first, int64_t count, double* loss) {
    return guarded_cnn([&] {
        if (!t || !loss || first < 0 || count < 1 || first + count > t->ps->count) throw std::invalid_argument("dft_cnn_trainer_eval: bad range");
        const int B = 8;
        t->ensure_batch(B);
        const int64_t P = t->ps->P, P3 = t->ps->P3;
        std::vector<float> o((size_t)P3 * B), tg((size_t)P3 * B);
        double s = 0.0;
        for (int64_t p0 = first; p0 < first + count; p0 += B) {
            const int nb = (int)std::min<int64_t>(B, first + count - p0);
            std::vector<int64_t> idx(nb);
            for (int b = 0; b < nb; ++b) idx[b] = p0 + b;
            t->gather(idx);
            dftc::check(t->ctx, dft_cnn_forward(t->ctx, t->ch, P, P, P, nb, t->wbf.as<uint16_t>(), t->bin.as<float>(),
                                                t->bout.as<float>(), t->act.as<float>(), 0),
                        "dft_cnn_forward");
            t->bout.download(o.data(), sizeof(float) * P3 * nb);
            t->btg.download(tg.data(), sizeof(float) * P3 * nb);
            for (int64_t i = 0; i < P3 * nb; ++i) {
                const double e = (double)o[i] - (double)tg[i];
                s += e * e;
            }
        }
        *loss = s / ((double)count * P3);
    });
}

int dft_cnn_trainer_get_params(dft_cnn_trainer* t, float* params, int64_t count) {
    return guarded_cnn([&] {
        if (!t || !params || count != t->count) throw std::invalid_argument("dft_cnn_trainer_get_params: bad arguments");
        t->w32.download(params, sizeof(float) * count);
    });
}

int dft_cnn_trainer_set_params(dft_cnn_trainer* t, const float* params, int64_t count) {
    return guarded_cnn([&] {
        if (!t || !params || count != t->count) throw std::invalid_argument("dft_cnn_trainer_set_params: bad arguments");
        t->w32.upload(params, sizeof(float) * count);
        dftc::check(t->ctx, dft_cnn_to_bf16(t->ctx, t->count, t->w32.as<float>(), t->wbf.as<uint16_t>()), "dft_cnn_to_bf16");
        t->m.zero();
        t->v.zero();
        t->step = 0;
    });
}

int dft_cnn_trainer_get_patches(dft_cnn_trainer* t, int64_t first, int64_t count, float* in, float* tgt) {
    return guarded_cnn([&] {
        if (!t || first < 0 || count < 0 || first + count > t->ps->count) throw std::invalid_argument("dft_cnn_trainer_get_patches: bad range");
        const int64_t P3 = t->ps->P3;
        if (in) dftc::check(t->ctx, dft_download(t->ctx, in, t->ps->in.as<float>() + 2 * P3 * first, sizeof(float) * 2 * P3 * count), "dft_download");
        if (tgt) dftc::check(t->ctx, dft_download(t->ctx, tgt, t->ps->tgt.as<float>() + P3 * first, sizeof(float) * P3 * count), "dft_download");
    });
}

int dft_cnn_trainer_add_patches(dft_cnn_trainer* t, int64_t count, const float* in, const float* tgt) {
    return guarded_cnn([&] {
        if (!t || count < 0 || (count > 0 && (!in || !tgt))) throw std::invalid_argument("dft_cnn_trainer_add_patches: bad arguments");
        if (count > t->ps->cap - t->ps->count) throw std::invalid_argument("dft_cnn_trainer_add_patches: more patches than free capacity");
        if (count == 0) return;
        const int64_t P3 = t->ps->P3, first = t->ps->count;
        dftc::check(t->ctx, dft_upload(t->ctx, t->ps->in.as<float>() + 2 * P3 * first, in, sizeof(float) * 2 * P3 * count), "dft_upload");
        dftc::check(t->ctx, dft_upload(t->ctx, t->ps->tgt.as<float>() + P3 * first, tgt, sizeof(float) * P3 * count), "dft_upload");
        t->ps->count += count;
    });
}

}  // extern "C"
