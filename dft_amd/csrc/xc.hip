// xc.hip -- (c) exchange-correlation potential and energy on the grid.
//
// Pass 1 (k_xc_point): per point with n > 1e-20, |grad n_sigma| by the
// reference's finite differences (central inside, one-sided on faces), the
// functional for both spins at once -> d f/d n_sigma into v_sigma, the
// gradient coefficient c_sigma into scratch, and n*eps*dV into a per-block
// energy partial.  PZ81 has c_sigma = 0, so its potential is final after
// pass 1.  Pass 2 (k_xc_div, PBE96) subtracts div(c_sigma grad n_sigma)
// built from the flux at i-1, i, i+1 exactly as ExchangeCorrelationMolecule
// .cpp:216-572 does (a radius-2 footprint in n, radius 1 in c).
#include "common.hpp"
#include "xc_functionals.hpp"

#include <algorithm>

using namespace dft;

namespace {

struct GridDev {
    int nx, ny, nz;
    double dx, dy, dz;
};

// calculateCartesianDerivative (ExchangeCorrelationMolecule.cpp:11-170)
template <typename T>
__device__ inline T fd(const T* __restrict__ f, int64_t p, int i, int n, int64_t s, T h) {
    if (i == 0) return (f[p + s] - f[p]) / h;
    if (i == n - 1) return (f[p] - f[p - s]) / h;
    return (f[p + s] - f[p - s]) / (T(2) * h);
}

template <typename T>
__device__ inline T gradmag(const T* __restrict__ f, int64_t p, int i, int j, int k, const GridDev& g) {
    const int64_t sy = g.nx, sz = (int64_t)g.nx * g.ny;
    const T gx = fd(f, p, i, g.nx, 1, (T)g.dx);
    const T gy = fd(f, p, j, g.ny, sy, (T)g.dy);
    const T gz = fd(f, p, k, g.nz, sz, (T)g.dz);
    return sqrt(gx * gx + gy * gy + gz * gz);
}

template <typename T>
__global__ __launch_bounds__(256) void k_xc_point(GridDev g, int functional, int consistent, const T* __restrict__ na,
                                                  const T* __restrict__ nb, T* __restrict__ va, T* __restrict__ vb,
                                                  T* __restrict__ ca, T* __restrict__ cb, double dv,
                                                  double* __restrict__ partials) {
    __shared__ double lds[4];
    const int64_t n = (int64_t)g.nx * g.ny * g.nz, plane = (int64_t)g.nx * g.ny;
    double e = 0.0;
    for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < n; p += (int64_t)gridDim.x * 256) {
        const T a = na[p], b = nb[p];
        const T tot = a + b;
        T oa = T(0), ob = T(0), oca = T(0), ocb = T(0), ga_keep = T(1), gb_keep = T(1);
        if (!(tot <= T(xc::kRhoFloor))) {
            const int i = (int)(p % g.nx), j = (int)((p / g.nx) % g.ny), k = (int)(p / plane);
            const T ga = functional == DFT_XC_PBE96 ? gradmag(na, p, i, j, k, g) : T(0);
            const T gb = functional == DFT_XC_PBE96 ? gradmag(nb, p, i, j, k, g) : T(0);
            ga_keep = ga;
            gb_keep = gb;
            // FP64 evaluation for small densities and for extreme relative gradients
            // (the GGA's t^6 terms overflow FP32 beyond |grad n| / n ~ 1e6)
            if (sizeof(T) == 4 && ((double)tot < xc::kF32Floor || (double)ga + (double)gb > 1.0e4 * (double)tot)) {
                const xc::Out<double> r = xc::evaluate<double>(functional, (double)a, (double)b, (double)ga, (double)gb);
                oa = (T)r.va; ob = (T)r.vb; oca = (T)r.ca; ocb = (T)r.cb;
                e += ((double)tot * r.eps) * dv;
            } else {
                const xc::Out<T> r = xc::evaluate<T>(functional, a, b, ga, gb);
                oa = r.va; ob = r.vb; oca = r.ca; ocb = r.cb;
                e += (double)(tot * r.eps) * dv;
            }
        }
        va[p] = oa;
        if (vb) vb[p] = ob;
        if (ca) {
            if (consistent) {  // flux c grad n / |grad n| (0 where the gradient vanishes)
                oca = ga_keep > T(0) ? oca / ga_keep : T(0);
                ocb = gb_keep > T(0) ? ocb / gb_keep : T(0);
            }
            ca[p] = oca;
            cb[p] = ocb;
        }
    }
    e = block_sum(e, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = e;
}

// flux c*dn/dx_a at the neighbour q (which has index iq along the axis)
template <typename T>
__device__ inline T flux(const T* __restrict__ c, const T* __restrict__ f, int64_t q, int iq, int n, int64_t s, T h) {
    return c[q] * fd(f, q, iq, n, s, h);
}

// calculateCartesianDivergence (ExchangeCorrelationMolecule.cpp:216-572)
template <typename T>
__device__ inline T div_axis(const T* __restrict__ c, const T* __restrict__ f, int64_t p, int i, int n, int64_t s,
                             T h, T fc) {
    if (i == 0) return (flux(c, f, p + s, i + 1, n, s, h) - fc) / h;
    if (i == n - 1) return (fc - flux(c, f, p - s, i - 1, n, s, h)) / h;
    const T fp = flux(c, f, p - s, i - 1, n, s, h);
    const T fn = flux(c, f, p + s, i + 1, n, s, h);
    return (fn - fp) / (T(2) * h);
}

template <typename T>
__global__ __launch_bounds__(256) void k_xc_div(GridDev g, const T* __restrict__ na, const T* __restrict__ nb,
                                                const T* __restrict__ ca, const T* __restrict__ cb,
                                                T* __restrict__ va, T* __restrict__ vb) {
    const int64_t n = (int64_t)g.nx * g.ny * g.nz, plane = (int64_t)g.nx * g.ny;
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
    const T tot = na[p] + nb[p];
    if (tot <= T(xc::kRhoFloor)) return;  // pass 1 already wrote 0
    const int i = (int)(p % g.nx), j = (int)((p / g.nx) % g.ny), k = (int)(p / plane);
    const int64_t sy = g.nx, sz = plane;
    const T hx = (T)g.dx, hy = (T)g.dy, hz = (T)g.dz;
    for (int s = 0; s < 2; ++s) {
        T* v = s == 0 ? va : vb;
        if (!v) continue;
        const T* f = s == 0 ? na : nb;
        const T* c = s == 0 ? ca : cb;
        const T cc = c[p];
        const T fx = cc * fd(f, p, i, g.nx, 1, hx);
        const T fy = cc * fd(f, p, j, g.ny, sy, hy);
        const T fz = cc * fd(f, p, k, g.nz, sz, hz);
        const T dvx = div_axis(c, f, p, i, g.nx, 1, hx, fx);
        const T dvy = div_axis(c, f, p, j, g.ny, sy, hy, fy);
        const T dvz = div_axis(c, f, p, k, g.nz, sz, hz, fz);
        v[p] = v[p] - (dvx + dvy + dvz);
    }
}

__global__ __launch_bounds__(256) void k_xc_points(int functional, const double* __restrict__ in, int64_t npts,
                                                   double* __restrict__ out) {
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= npts) return;
    const double* q = in + 4 * p;
    const xc::Out<double> r = xc::evaluate<double>(functional, q[0], q[1], q[2], q[3]);
    double* o = out + 5 * p;
    o[0] = r.eps; o[1] = r.va; o[2] = r.vb; o[3] = r.ca; o[4] = r.cb;
}

template <typename T>
int xc_t(dft_ctx* c, const dft_grid* gg, int functional, const T* na, const T* nb, T* va, T* vb, double* exc) {
    const int64_t n = grid_points(gg);
    GridDev g{(int)gg->nx, (int)gg->ny, (int)gg->nz, gg->dx, gg->dy, gg->dz};
    const int consistent = (functional & DFT_XC_FLUX_CONSISTENT) ? 1 : 0;
    functional &= ~DFT_XC_FLUX_CONSISTENT;
    const bool gga = functional == DFT_XC_PBE96;
    T* ca = nullptr;
    T* cb = nullptr;
    if (gga) {
        if (int rc = ensure_scratch(c, 2 * (size_t)n * sizeof(T))) return rc;
        ca = (T*)c->scratch;
        cb = ca + n;
    }
    const int blocks = (int)std::min<int64_t>((n + 255) / 256, 8 * (int64_t)c->num_cus);
    if (int rc = ensure_partials(c, (size_t)blocks + 1)) return rc;
    const double dv = gg->dx * gg->dy * gg->dz;
    k_xc_point<T><<<blocks, 256, 0, c->stream>>>(g, functional, consistent, na, nb, va, vb, ca, cb, dv, c->partials);
    DFT_CHECK_LAUNCH(c);
    if (gga) {
        k_xc_div<T><<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(g, na, nb, ca, cb, va, vb);
        DFT_CHECK_LAUNCH(c);
    }
    if (exc) return finish_sum(c, blocks, exc);
    return DFT_OK;
}

}  // namespace

extern "C" {

int dft_xc(dft_ctx* c, const dft_grid* g, int prec, int functional, const void* na, const void* nb, void* va, void* vb,
           double* exc_host) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, valid_grid(g) && na && nb && va, "dft_xc: bad arguments");
    const int base = functional & ~DFT_XC_FLUX_CONSISTENT;
    DFT_REQUIRE(c, base == DFT_XC_PZ81 || base == DFT_XC_PBE96, "dft_xc: unknown functional");
    if (prec == DFT_F64)
        return xc_t<double>(c, g, functional, (const double*)na, (const double*)nb, (double*)va, (double*)vb, exc_host);
    if (prec == DFT_F32)
        return xc_t<float>(c, g, functional, (const float*)na, (const float*)nb, (float*)va, (float*)vb, exc_host);
    return set_err(c, DFT_EINVAL, "dft_xc: bad precision");
}

int dft_xc_points(dft_ctx* c, int functional, const double* in, int64_t npts, double* out) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, in && out && npts >= 0, "dft_xc_points: bad arguments");
    DFT_REQUIRE(c, functional == DFT_XC_PZ81 || functional == DFT_XC_PBE96, "dft_xc_points: unknown functional");
    if (!npts) return DFT_OK;
    k_xc_points<<<(unsigned)((npts + 255) / 256), 256, 0, c->stream>>>(functional, in, npts, out);
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

}  // extern "C"
