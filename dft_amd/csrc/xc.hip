// xc.hip -- (c) exchange-correlation potential and energy on the grid.
//
// PZ81 (k_xc_lda): pointwise, one pass, both spins per point.
//
// PBE96 (k_xc_gga): one fused pass.  v_sigma = d f/d n_sigma - div F_sigma,
// with F_sigma = c'_sigma (finite-difference grad n_sigma) and c'_sigma the
// gradient coefficient c_sigma = d f / d|grad n_sigma| of the reference
// (ExchangeCorrelationMolecule.cpp:576-747), or c_sigma / |grad n_sigma| for
// the consistent flux.  A workgroup owns a TX x TY column of the grid over a
// chunk of z-planes and streams it plane by plane:
//   - LDS holds a ring of three n_alpha/n_beta planes with a 2-point x/y halo
//     (plane k .. k+2) and a ring of two c' planes over the tile plus a
//     1-point x/y halo (planes k, k+1);
//   - step b evaluates the functional at plane k+1 over the tile and its x/y
//     halo (|grad n| by the reference's central / one-sided differences,
//     ExchangeCorrelationMolecule.cpp:11-214), keeping d f/d n and the z-flux
//     of the thread's own points in registers;
//   - step c forms v at plane k from the fluxes at i-1, i, i+1 of each axis
//     exactly as calculateCartesianDivergence does (:216-572), one-sided at
//     faces, and writes it once.
// HBM traffic is the algorithmic n_alpha, n_beta in and v_alpha, v_beta out;
// the functional is evaluated (TX*TY + 2TX + 2TY)/(TX*TY) times per point
// in a plane (halo recomputation) plus two planes per z-chunk.
// In FP32 the functional is evaluated in FP64 where its FP32 form is unsafe:
// |grad n| / n > 1e4, or a non-finite FP32 result.  (The FP32 correlation is
// regrouped so that the tails' large b t^2 does not overflow it.)
#include "common.hpp"
#include "xc_functionals.hpp"

#include <algorithm>
#include <mutex>

using namespace dft;

// The allocation-free entry point's scratch: the workgroup energy partials
// of one XC call and their sum, in a fixed device array (one copy per GPU,
// held until the process ends).  Host code uses only the address that
// hipGetSymbolAddress returns for it on each device (dft_ctx::xc_scratch).
constexpr int64_t kXcScratchDoubles = 16384;
static __device__ __align__(256) double scratch[kXcScratchDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

struct GridDev {
    int nx, ny, nz;
    double dx, dy, dz;
};

// calculateCartesianDerivative (ExchangeCorrelationMolecule.cpp:11-170) from
// the values at i-1, i, i+1 along an axis of n points
template <typename T>
__device__ inline T fd3(T fm, T f0, T fp, int i, int n, T h) {
    if (i == 0) return (fp - f0) / h;
    if (i == n - 1) return (f0 - fm) / h;
    return (fp - fm) / (T(2) * h);
}

// calculateCartesianDivergence, one axis (:216-572): fluxes at i-1, i, i+1
template <typename T>
__device__ inline T div3(T Fm, T F0, T Fp, int i, int n, T h) {
    if (i == 0) return (Fp - F0) / h;
    if (i == n - 1) return (F0 - Fm) / h;
    return (Fp - Fm) / (T(2) * h);
}

template <typename T>
__device__ inline bool finite_out(const xc::Out<T>& r) {
    return isfinite(r.eps) && isfinite(r.va) && isfinite(r.vb) && isfinite(r.ca) && isfinite(r.cb);
}

// ---------------------------------------------------------------- PZ81
template <typename T>
__global__ __launch_bounds__(256) void k_xc_lda(int n, const T* __restrict__ na, const T* __restrict__ nb,
                                                T* __restrict__ va, T* __restrict__ vb, double dv,
                                                double* __restrict__ partials) {
    __shared__ double lds[4];
    double e = 0.0;
    for (int p = blockIdx.x * 256 + threadIdx.x; p < n; p += gridDim.x * 256) {
        const T a = na[p], b = nb[p];
        const T tot = a + b;
        T oa = T(0), ob = T(0);
        if (!(tot <= T(xc::kRhoFloor))) {
            const xc::Out<T> r = xc::pz81<T>(a, b);
            oa = r.va;
            ob = r.vb;
            e += (double)(tot * r.eps) * dv;
        }
        va[p] = oa;
        if (vb) vb[p] = ob;
    }
    e = block_sum(e, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = e;
}

// ---------------------------------------------------------------- PBE96
template <typename T>
struct GgaPoint {
    T dfa, dfb;  // d f / d n_sigma
    T cpa, cpb;  // flux coefficient c'_sigma
    T gza, gzb;  // finite-difference d n_sigma / dz
    double e;    // n eps dV
};

// The functional at one point, its |grad n| from the values at the six
// neighbours.
template <typename T>
__device__ inline GgaPoint<T> gga_point(T a, T b, T axm, T axp, T aym, T ayp, T azm, T azp, T bxm, T bxp, T bym, T byp,
                                        T bzm, T bzp, int x, int y, int z, const GridDev& g, int consistent, double dv) {
    GgaPoint<T> o{T(0), T(0), T(0), T(0), T(0), T(0), 0.0};
    const T hx = (T)g.dx, hy = (T)g.dy, hz = (T)g.dz;
    const T gax = fd3(axm, a, axp, x, g.nx, hx), gay = fd3(aym, a, ayp, y, g.ny, hy), gaz = fd3(azm, a, azp, z, g.nz, hz);
    const T gbx = fd3(bxm, b, bxp, x, g.nx, hx), gby = fd3(bym, b, byp, y, g.ny, hy), gbz = fd3(bzm, b, bzp, z, g.nz, hz);
    o.gza = gaz;
    o.gzb = gbz;
    const T tot = a + b;
    if (tot <= T(xc::kRhoFloor)) return o;
    const T ga = sqrt(gax * gax + gay * gay + gaz * gaz);
    const T gb = sqrt(gbx * gbx + gby * gby + gbz * gbz);
    T va, vb, ca, cb;
    if constexpr (sizeof(T) == 4) {
        bool done = false;
        if (!((double)ga + (double)gb > 1.0e4 * (double)tot)) {
            const xc::Out<T> r = xc::pbe96<T>(a, b, ga, gb);
            if (finite_out(r)) {
                va = r.va; vb = r.vb; ca = r.ca; cb = r.cb;
                o.e = (double)(tot * r.eps) * dv;
                done = true;
            }
        }
        if (!done) {
            const xc::Out<double> r = xc::pbe96<double>((double)a, (double)b, (double)ga, (double)gb);
            va = (T)r.va; vb = (T)r.vb; ca = (T)r.ca; cb = (T)r.cb;
            o.e = ((double)tot * r.eps) * dv;
        }
    } else {
        const xc::Out<T> r = xc::pbe96<T>(a, b, ga, gb);
        va = r.va; vb = r.vb; ca = r.ca; cb = r.cb;
        o.e = (double)(tot * r.eps) * dv;
    }
    if (consistent) {
        ca = ga > T(0) ? ca / ga : T(0);
        cb = gb > T(0) ? cb / gb : T(0);
    }
    o.dfa = va; o.dfb = vb; o.cpa = ca; o.cpb = cb;
    return o;
}

template <typename T, int TY>
struct GgaTile {
    static constexpr int TX = 32;
    static constexpr int RY = TY / 8;               // rows per thread (256 threads = 32 x 8)
    static constexpr int NX = TX + 4, NY = TY + 4;  // density planes: 2-point halo
    static constexpr int CX = TX + 2, CY = TY + 2;  // flux coefficients: 1-point halo
    static constexpr int HALO = 2 * TX + 2 * TY;
};

template <typename T, int TY>
__global__ __launch_bounds__(256) void k_xc_gga(GridDev g, int consistent, int tiles_x, int tiles_y, int tz,
                                                int total_tiles, const T* __restrict__ na, const T* __restrict__ nb,
                                                T* __restrict__ va, T* __restrict__ vb, double dv,
                                                double* __restrict__ partials) {
    using Tl = GgaTile<T, TY>;
    constexpr int TX = Tl::TX, RY = Tl::RY, NX = Tl::NX, NY = Tl::NY, CX = Tl::CX, CY = Tl::CY;
    __shared__ T sa[3][NY][NX], sb[3][NY][NX];
    __shared__ T ca[2][CY][CX], cb[2][CY][CX];
    __shared__ double red[4];

    const int tid = threadIdx.x;
    const int tile = (int)xcd_tile(blockIdx.x, total_tiles);
    const int txi = tile % tiles_x, tyi = (tile / tiles_x) % tiles_y, tzi = tile / (tiles_x * tiles_y);
    const int x0 = txi * TX, y0 = tyi * TY, z0 = tzi * tz;
    const int z1 = min(z0 + tz, g.nz);
    const unsigned plane = (unsigned)g.nx * (unsigned)g.ny;
    const int lxo = tid % TX, lyo = tid / TX;  // own column; rows lyo + 8 r

    auto slot3 = [](int p) { return ((p % 3) + 3) % 3; };
    auto load_plane = [&](int p) {
        const int s = slot3(p);
        for (int i = tid; i < NX * NY; i += 256) {
            const int ly = i / NX, lx = i - ly * NX;
            const int x = x0 - 2 + lx, y = y0 - 2 + ly;
            T a = T(0), b = T(0);
            if (p >= 0 && p < g.nz && x >= 0 && x < g.nx && y >= 0 && y < g.ny) {
                const unsigned q = (unsigned)p * plane + (unsigned)y * (unsigned)g.nx + (unsigned)x;
                a = na[q];
                b = nb[q];
            }
            sa[s][ly][lx] = a;
            sb[s][ly][lx] = b;
        }
    };
    // functional at point (cx, cy) of the flux-coefficient grid in plane p
    auto eval_at = [&](int cx, int cy, int p) -> GgaPoint<T> {
        const int x = x0 - 1 + cx, y = y0 - 1 + cy;
        if (p < 0 || p >= g.nz || x < 0 || x >= g.nx || y < 0 || y >= g.ny)
            return GgaPoint<T>{T(0), T(0), T(0), T(0), T(0), T(0), 0.0};
        const int s0 = slot3(p), sm = slot3(p - 1), sp = slot3(p + 1);
        const int lx = cx + 1, ly = cy + 1;
        return gga_point<T>(sa[s0][ly][lx], sb[s0][ly][lx], sa[s0][ly][lx - 1], sa[s0][ly][lx + 1], sa[s0][ly - 1][lx],
                            sa[s0][ly + 1][lx], sa[sm][ly][lx], sa[sp][ly][lx], sb[s0][ly][lx - 1], sb[s0][ly][lx + 1],
                            sb[s0][ly - 1][lx], sb[s0][ly + 1][lx], sb[sm][ly][lx], sb[sp][ly][lx], x, y, p, g,
                            consistent, dv);
    };

    // Per owned point, at the start of iteration k: d f/d n at plane k in
    // df[0], the z-flux at planes k-1 and k in fz[0], fz[1].  Step b for plane
    // k+1 fills df[1] and fz[2]; the queues then shift by one plane.
    T dfa[RY][2] = {}, dfb[RY][2] = {}, fza[RY][3] = {}, fzb[RY][3] = {};
    double e = 0.0;
    const T hx = (T)g.dx, hy = (T)g.dy, hz = (T)g.dz;

    load_plane(z0 - 2);
    load_plane(z0 - 1);
    // iterations z0-2 and z0-1 only fill the queues (planes z0-1 and z0)
    for (int k = z0 - 2; k < z1; ++k) {
        __syncthreads();  // plane k-1's readers are done with its ring slots
        load_plane(k + 2);
        __syncthreads();
        // ---- step b: the functional at plane p = k+1, tile and x/y halo
        {
            const int p = k + 1, cs = p & 1;
#pragma unroll
            for (int r = 0; r < RY; ++r) {
                const int cx = lxo + 1, cy = lyo + 8 * r + 1;
                const GgaPoint<T> o = eval_at(cx, cy, p);
                ca[cs][cy][cx] = o.cpa;
                cb[cs][cy][cx] = o.cpb;
                dfa[r][1] = o.dfa;
                dfb[r][1] = o.dfb;
                if (p >= z0 && p < z1) e += o.e;
                fza[r][2] = o.cpa * o.gza;
                fzb[r][2] = o.cpb * o.gzb;
            }
            for (int h = tid; h < Tl::HALO; h += 256) {
                int cx, cy;
                if (h < TY) { cx = 0; cy = 1 + h; }
                else if (h < 2 * TY) { cx = TX + 1; cy = 1 + h - TY; }
                else if (h < 2 * TY + TX) { cx = 1 + h - 2 * TY; cy = 0; }
                else { cx = 1 + h - 2 * TY - TX; cy = TY + 1; }
                const GgaPoint<T> o = eval_at(cx, cy, p);
                ca[cs][cy][cx] = o.cpa;
                cb[cs][cy][cx] = o.cpb;
            }
        }
        __syncthreads();
        // ---- step c: v at plane k
        const int s0 = slot3(k), cs = k & 1;
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            const int x = x0 + lxo, y = y0 + lyo + 8 * r;
            if (k >= z0 && x < g.nx && y < g.ny) {
                const int lx = lxo + 2, ly = lyo + 8 * r + 2, cx = lxo + 1, cy = lyo + 8 * r + 1;
                const unsigned q = (unsigned)k * plane + (unsigned)y * (unsigned)g.nx + (unsigned)x;
                const T tot = sa[s0][ly][lx] + sb[s0][ly][lx];
                T outa = T(0), outb = T(0);
                if (!(tot <= T(xc::kRhoFloor))) {
#pragma unroll
                    for (int sp = 0; sp < 2; ++sp) {
                        if (sp == 1 && !vb) break;
                        const T(*n)[NX] = sp == 0 ? sa[s0] : sb[s0];
                        const T(*c)[CX] = sp == 0 ? ca[cs] : cb[cs];
                        const T Fx0 = c[cy][cx] * fd3(n[ly][lx - 1], n[ly][lx], n[ly][lx + 1], x, g.nx, hx);
                        const T Fxm = x > 0 ? c[cy][cx - 1] * fd3(n[ly][lx - 2], n[ly][lx - 1], n[ly][lx], x - 1, g.nx, hx)
                                            : T(0);
                        const T Fxp = x < g.nx - 1
                                          ? c[cy][cx + 1] * fd3(n[ly][lx], n[ly][lx + 1], n[ly][lx + 2], x + 1, g.nx, hx)
                                          : T(0);
                        const T Fy0 = c[cy][cx] * fd3(n[ly - 1][lx], n[ly][lx], n[ly + 1][lx], y, g.ny, hy);
                        const T Fym = y > 0 ? c[cy - 1][cx] * fd3(n[ly - 2][lx], n[ly - 1][lx], n[ly][lx], y - 1, g.ny, hy)
                                            : T(0);
                        const T Fyp = y < g.ny - 1
                                          ? c[cy + 1][cx] * fd3(n[ly][lx], n[ly + 1][lx], n[ly + 2][lx], y + 1, g.ny, hy)
                                          : T(0);
                        const T* fz = sp == 0 ? fza[r] : fzb[r];
                        const T dvx = div3(Fxm, Fx0, Fxp, x, g.nx, hx);
                        const T dvy = div3(Fym, Fy0, Fyp, y, g.ny, hy);
                        const T dvz = div3(fz[0], fz[1], fz[2], k, g.nz, hz);
                        const T df = sp == 0 ? dfa[r][0] : dfb[r][0];
                        const T v = df - (dvx + dvy + dvz);
                        if (sp == 0) outa = v;
                        else outb = v;
                    }
                }
                va[q] = outa;
                if (vb) vb[q] = outb;
            }
            dfa[r][0] = dfa[r][1];
            dfb[r][0] = dfb[r][1];
            fza[r][0] = fza[r][1];
            fza[r][1] = fza[r][2];
            fzb[r][0] = fzb[r][1];
            fzb[r][1] = fzb[r][2];
        }
    }
    e = block_sum(e, red);
    if (tid == 0) partials[blockIdx.x] = e;
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

// Launch geometry of k_xc_gga: x/y tiles and the z-chunk length (enough
// workgroups to give every CU several).
void gga_shape(const dft_grid* gg, int ty_rows, int num_cus, int& tiles_x, int& tiles_y, int& tz, int& total) {
    tiles_x = (int)((gg->nx + 31) / 32);
    tiles_y = (int)((gg->ny + ty_rows - 1) / ty_rows);
    const int64_t txy = (int64_t)tiles_x * tiles_y;
    const int64_t want = 4 * (int64_t)num_cus;
    int64_t chunks = std::max<int64_t>(1, std::min<int64_t>(gg->nz, (want + txy - 1) / txy));
    tz = (int)((gg->nz + chunks - 1) / chunks);
    if (tz < 4) tz = (int)std::min<int64_t>(4, gg->nz);
    chunks = (gg->nz + tz - 1) / tz;
    total = (int)(txy * chunks);
}

constexpr int kGgaRowsF32 = 16, kGgaRowsF64 = 8;

// Workgroup energy partials one XC call writes.
int64_t xc_partials_needed(const dft_grid* gg, int prec, int functional, int num_cus) {
    functional &= ~DFT_XC_FLUX_CONSISTENT;
    if (functional == DFT_XC_PZ81) return std::min<int64_t>((grid_points(gg) + 255) / 256, 8 * (int64_t)num_cus);
    int tx, ty, tz, total;
    gga_shape(gg, prec == DFT_F32 ? kGgaRowsF32 : kGgaRowsF64, num_cus, tx, ty, tz, total);
    return total;
}

// Launch the XC kernels with the energy partials in `partials` (at least
// xc_partials_needed doubles); *nparts receives how many were written.
template <typename T>
int xc_launch(dft_ctx* c, const dft_grid* gg, int functional, const T* na, const T* nb, T* va, T* vb,
              double* partials, int64_t* nparts) {
    const int64_t n = grid_points(gg);
    GridDev g{(int)gg->nx, (int)gg->ny, (int)gg->nz, gg->dx, gg->dy, gg->dz};
    const int consistent = (functional & DFT_XC_FLUX_CONSISTENT) ? 1 : 0;
    functional &= ~DFT_XC_FLUX_CONSISTENT;
    const double dv = gg->dx * gg->dy * gg->dz;
    if (functional == DFT_XC_PZ81) {
        const int blocks = (int)std::min<int64_t>((n + 255) / 256, 8 * (int64_t)c->num_cus);
        k_xc_lda<T><<<blocks, 256, 0, c->stream>>>((int)n, na, nb, va, vb, dv, partials);
        DFT_CHECK_LAUNCH(c);
        *nparts = blocks;
        return DFT_OK;
    }
    constexpr int TY = sizeof(T) == 4 ? kGgaRowsF32 : kGgaRowsF64;
    int tx, ty, tz, total;
    gga_shape(gg, TY, c->num_cus, tx, ty, tz, total);
    k_xc_gga<T, TY><<<total, 256, 0, c->stream>>>(g, consistent, tx, ty, tz, total, na, nb, va, vb, dv, partials);
    DFT_CHECK_LAUNCH(c);
    *nparts = total;
    return DFT_OK;
}

int xc_args(dft_ctx* c, const dft_grid* g, int prec, int functional, const void* na, const void* nb, const void* va) {
    DFT_REQUIRE(c, valid_grid(g) && na && nb && va, "dft_xc: bad arguments");
    DFT_REQUIRE(c, grid_points(g) < ((int64_t)1 << 31), "dft_xc: grid beyond 2^31 points");
    const int base = functional & ~DFT_XC_FLUX_CONSISTENT;
    DFT_REQUIRE(c, base == DFT_XC_PZ81 || base == DFT_XC_PBE96, "dft_xc: unknown functional");
    DFT_REQUIRE(c, prec == DFT_F64 || prec == DFT_F32, "dft_xc: bad precision");
    return DFT_OK;
}

}  // namespace

namespace dft {
int xc_scratch_attach(dft_ctx* c) {
    if (c->device < 0 || c->device >= kMaxDevices) return set_err(c, DFT_EINVAL, "xc scratch: device index beyond 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[c->device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_err(c, DFT_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[c->device] = static_cast<double*>(p);
    }
    c->xc_scratch = g_scratch_ptr[c->device];
    return DFT_OK;
}
}  // namespace dft

extern "C" {

int dft_xc(dft_ctx* c, const dft_grid* g, int prec, int functional, const void* na, const void* nb, void* va, void* vb,
           double* exc_host) {
    if (!c) return DFT_EINVAL;
    if (int rc = xc_args(c, g, prec, functional, na, nb, va)) return rc;
    const int64_t need = xc_partials_needed(g, prec, functional, c->num_cus);
    if (int rc = ensure_partials(c, (size_t)need + 1)) return rc;
    int64_t parts = 0;
    const int rc = prec == DFT_F64 ? xc_launch<double>(c, g, functional, (const double*)na, (const double*)nb,
                                                        (double*)va, (double*)vb, c->partials, &parts)
                                    : xc_launch<float>(c, g, functional, (const float*)na, (const float*)nb,
                                                       (float*)va, (float*)vb, c->partials, &parts);
    if (rc) return rc;
    if (exc_host) return finish_sum(c, parts, exc_host);
    return DFT_OK;
}

int dft_xc_noalloc(dft_ctx* c, const dft_grid* g, int prec, int functional, const void* na, const void* nb, void* va,
                   void* vb, double* exc_host) {
    if (!c) return DFT_EINVAL;
    if (int rc = xc_args(c, g, prec, functional, na, nb, va)) return rc;
    DFT_REQUIRE(c, c->xc_scratch, "dft_xc_noalloc: the device scratch array was not attached");
    const int64_t need = xc_partials_needed(g, prec, functional, c->num_cus);
    if (need + 1 > kXcScratchDoubles)
        return set_err(c, DFT_EINVAL,
                       "dft_xc_noalloc: this grid needs " + std::to_string(need + 1) +
                           " energy partials, beyond the fixed scratch of " + std::to_string(kXcScratchDoubles));
    int64_t parts = 0;
    const int rc = prec == DFT_F64 ? xc_launch<double>(c, g, functional, (const double*)na, (const double*)nb,
                                                        (double*)va, (double*)vb, c->xc_scratch, &parts)
                                    : xc_launch<float>(c, g, functional, (const float*)na, (const float*)nb,
                                                       (float*)va, (float*)vb, c->xc_scratch, &parts);
    if (rc) return rc;
    if (exc_host) return finish_sum_at(c, c->xc_scratch, parts, exc_host);
    return DFT_OK;
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
