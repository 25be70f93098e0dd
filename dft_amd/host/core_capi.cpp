// core_capi.cpp -- extern "C" entry points of libdft_core.so (dft_core.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>
#include <vector>

#include <functional>

#include "capi_util.hpp"
#include "davidson.hpp"
#include "dft_core.h"
#include "scf.hpp"

namespace {
thread_local std::string g_err;
}  // namespace

namespace dftc {
int guarded_call(const std::function<void()>& f) {
    try {
        f();
        return DFT_OK;
    } catch (const std::invalid_argument& e) {
        g_err = e.what();
        return DFT_EINVAL;
    } catch (const std::exception& e) {
        g_err = e.what();
        return DFT_ENUMERIC;
    }
}
}  // namespace dftc

namespace {
template <class F>
int guarded(F&& f) {
    return dftc::guarded_call(std::function<void()>(std::forward<F>(f)));
}
}  // namespace

extern "C" {

const char* dft_core_last_error(void) { return g_err.c_str(); }

int dft_fermi_occupations(const double* eps, int n, double nelec, double kT, double* occ, double* mu, double* entropy) {
    return guarded([&] {
        if (!eps || !occ || n <= 0 || !(kT > 0.0) || !(nelec >= 0.0) || nelec > n)
            throw std::invalid_argument("dft_fermi_occupations: bad arguments");
        const double m = dftc::fermi_occupations(eps, n, nelec, kT, occ);
        if (mu) *mu = m;
        if (entropy) *entropy = dftc::fermi_entropy(occ, n);
    });
}

int dft_scf_run(const dft_grid* g, int natoms, const double* nuclei, const dft_scf_config* cfg, dft_scf_report* rep,
                double* eig, double* trace, int trace_cap, double* rho) {
    return guarded([&] {
        if (!g || !cfg || !rep || (natoms > 0 && !nuclei)) throw std::invalid_argument("dft_scf_run: null argument");
        dftc::Ctx ctx(0);
        dftc::ScfDriver drv(ctx, *g, natoms, nuclei, *cfg);
        dftc::ScfOutputs out;
        out.want_rho = rho != nullptr;
        drv.run(*rep, out);
        if (eig) std::copy(out.eig.begin(), out.eig.end(), eig);
        if (trace && trace_cap > 0) {
            const size_t nv = std::min(out.trace.size(), (size_t)trace_cap * 5);
            std::copy(out.trace.begin(), out.trace.begin() + nv, trace);
        }
        if (rho) std::copy(out.rho.begin(), out.rho.end(), rho);
    });
}

int dft_scf_step(const dft_grid* g, int natoms, const double* nuclei, const dft_scf_config* cfg,
                 const double* ra, const double* rb, dft_scf_report* rep, double* eig, dft_scf_fields* fields) {
    return guarded([&] {
        if (!g || !cfg || !rep || !ra || !rb || (natoms > 0 && !nuclei)) throw std::invalid_argument("dft_scf_step: null argument");
        dftc::Ctx ctx(0);
        dft_scf_config c = *cfg;
        c.max_iterations = 1;
        c.refine_iterations = 0;
        c.stop_rule = DFT_STOP_REFERENCE;  // fixed Davidson tolerance; one step never meets the stop rule
        dftc::ScfDriver drv(ctx, *g, natoms, nuclei, c);
        dftc::ScfOutputs out;
        out.init_rho_a = ra;
        out.init_rho_b = rb;
        out.export_fields = fields;
        drv.run(*rep, out);
        if (eig) std::copy(out.eig.begin(), out.eig.end(), eig);
    });
}

int dft_davidson_run(const dft_grid* g, int prec, int gram_mode, const double* veff_host, int nstates, int nconv,
                     int max_subspace, double tol, int max_iter, int precond, const void* cnn, double* eig,
                     double* psi, double* resid, int* iterations, int64_t* stats) {
    return guarded([&] {
        if (!g || !veff_host || !eig || nstates <= 0) throw std::invalid_argument("dft_davidson_run: bad arguments");
        dftc::Ctx ctx(0);
        const int64_t n = g->nx * g->ny * g->nz;
        dftc::Buf v64(ctx, (size_t)n * 8), v(ctx, (size_t)n * dftc::esize(prec));
        v64.upload(veff_host, (size_t)n * 8);
        dftc::check(ctx, dft_convert(ctx, v64.get(), DFT_F64, v.get(), prec, n), "dft_convert");
        dftc::BlockDavidson D(ctx, *g, prec, nstates, max_subspace > 0 ? max_subspace : std::max(3 * nstates, nstates + 16));
        std::unique_ptr<dftc::Preconditioner> P = dftc::make_preconditioner(ctx, precond, cnn, *g);
        dftc::DavidsonOptions o;
        o.nconv = nconv > 0 ? nconv : nstates;
        o.tol = tol > 0 ? tol : 1.0e-7;
        o.max_iter = max_iter > 0 ? max_iter : 200;
        o.gram_mode = gram_mode;
        dftc::DavidsonStats st = D.solve(v.get(), false, o, *P);
        std::copy(D.eigenvalues().begin(), D.eigenvalues().end(), eig);
        if (resid) std::copy(D.residuals().begin(), D.residuals().end(), resid);
        if (iterations) *iterations = st.iterations;
        if (stats) {
            stats[0] = st.happly_columns;
            stats[1] = st.restarts;
            stats[2] = P->applied();
            stats[3] = P->fallbacks();
        }
        if (psi) {
            // reference normalisation: sum psi^2 dV = 1
            const int nc = std::min(o.nconv, nstates);
            dftc::Buf out(ctx, (size_t)n * nc * 8);
            dftc::check(ctx, dft_convert(ctx, D.X(), prec, out.get(), DFT_F64, n * (int64_t)nc), "dft_convert");
            out.download(psi, (size_t)n * nc * 8);
            const double s = 1.0 / std::sqrt(g->dx * g->dy * g->dz);
            for (int64_t i = 0; i < n * nc; ++i) psi[i] *= s;
        }
        if (!st.converged) throw std::runtime_error("Davidson did not converge");
    });
}

}  // extern "C"
