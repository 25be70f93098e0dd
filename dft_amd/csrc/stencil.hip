// stencil.hip -- (a) the 7-point Kohn-Sham operator on blocks of orbitals,
// the reference's Neumann-stencil kinetic energy, and the Davidson diagonal
// preconditioner.
//
// H psi layout: a workgroup owns a 64 (x) x 4*TY (y) column of the grid over
// a z-chunk and streams through it plane by plane.  Lane = x, so a wave is
// one unit-stride 64-point row segment (coalesced); each lane keeps TY
// consecutive y-rows and the z-1 / z / z+1 planes of NB orbitals in
// registers.  x-neighbours come from the adjacent lane by a DPP wave shift,
// the two values beyond the segment from one extra load per row; y-
// neighbours from the lane's own rows plus one halo row above and below;
// z-neighbours from the register queue.  V_eff is loaded once per plane and
// reused by the NB orbitals of the group.  Offsets are 32-bit from uniform
// column bases, which keeps the kernel small enough for several waves per SIMD.
#include "common.hpp"

#include <cmath>
#include <type_traits>

#include <algorithm>
#include <vector>

using namespace dft;

namespace {

// lane i <- lane i-1 (wave_shr:1); lane 0 <- fill
__device__ inline float shr1(float v, float fill) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(fill), __float_as_int(v), 0x138, 0xf, 0xf, false));
}
// lane i <- lane i+1 (wave_shl:1); lane 63 <- fill
__device__ inline float shl1(float v, float fill) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(fill), __float_as_int(v), 0x130, 0xf, 0xf, false));
}
__device__ inline double shr1(double v, double fill) {
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(fill), __double2loint(v), 0x138, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(fill), __double2hiint(v), 0x138, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}
__device__ inline double shl1(double v, double fill) {
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(fill), __double2loint(v), 0x130, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(fill), __double2hiint(v), 0x130, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}

template <typename T>
struct Coef {
    T cx, cy, cz, cc;  // 1/dx^2, 1/dy^2, 1/dz^2, -2(1/dx^2+1/dy^2+1/dz^2)
};

// SM = 0: H psi.  SM = 1, 2: instead of H psi, one step of the kinetic
// smoothing correction on psi = t with r in `R` (same layout as psi) and no
// V_eff read: J(t) = t + cf (r - (-1/2 Lap t + sigma t)), stored as
// out = a J(t) (SM = 1) or out = a J(t) + b out (SM = 2: out holds the
// previous iterate, read before it is overwritten: the Chebyshev recurrence).
// PF: the loads of plane k + 2 (centre) and k + 1 (halo, V_eff) are issued
// before plane k is computed, so a wave keeps a plane of loads in flight.
// Results are written with non-temporal stores: the output is not read again
// by this launch, and streaming it past the caches raised the sustained
// bandwidth by 5-15% on MI355X.
template <typename T, int TY, int NB, int SM = 0, bool PF = false>
__global__ __launch_bounds__(256) void k_hpsi(int nx, int ny, int nz, int zc, int nzc, Coef<T> co,
                                              const T* __restrict__ V, const T* __restrict__ psi, int64_t ld,
                                              T* __restrict__ out, int64_t ldo, const T* __restrict__ R = nullptr,
                                              int64_t ldr = 0, T sigma = T(0), T cf = T(0), T a = T(1),
                                              T b = T(0)) {
    constexpr bool SMOOTH = SM != 0;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t tile = xcd_tile(blockIdx.x + (int64_t)gridDim.x * (blockIdx.y + (int64_t)gridDim.y * blockIdx.z),
                                  (int64_t)gridDim.x * gridDim.y * gridDim.z);
    const int bx = (int)(tile % gridDim.x), by = (int)((tile / gridDim.x) % gridDim.y);
    const int bz = (int)(tile / ((int64_t)gridDim.x * gridDim.y));
    const int x0 = bx * 64, x = x0 + lane;
    const int ybase = by * (4 * TY) + w * TY;
    const int chunk = bz % nzc, grp = bz / nzc;
    const int k0 = chunk * zc, k1 = min(nz, k0 + zc);
    const int plane = nx * ny;  // the launcher guarantees nx*ny*nz < 2^31: 32-bit offsets
    const bool xin = x < nx;
    // x-halo: lane 0 reads x0-1 and lane 63 reads x0+64; the DPP shifts take
    // no fill value from the other lanes, which stay masked off
    int hx = x;
    bool hin = false;
    if (lane == 0) { hx = x0 - 1; hin = x0 > 0; }
    if (lane == 63) { hx = x0 + 64; hin = x0 + 64 < nx; }
    bool yin[TY];
    int roff[TY], hoff[TY];
#pragma unroll
    for (int r = 0; r < TY; ++r) {
        const int y = ybase + r;
        yin[r] = y < ny;
        roff[r] = y * nx + x;
        hoff[r] = y * nx + hx;
    }
    const bool ym_in = ybase > 0 && ybase - 1 < ny, yp_in = ybase + TY < ny;
    const int ymo = (ybase - 1) * nx + x, ypo = (ybase + TY) * nx + x;
    const T* P = psi + (int64_t)(grp * NB) * ld;
    T* Out = out + (int64_t)(grp * NB) * ldo;

    // plane k's x-halo values, y-halo rows and V_eff (sigma when smoothing)
    auto load_side = [&](int k, T (&h)[NB][TY], T (&lo)[NB], T (&hi)[NB], T (&vv)[TY]) {
        const int kp = k * plane;
#pragma unroll
        for (int r = 0; r < TY; ++r) vv[r] = (!SMOOTH && xin && yin[r]) ? V[kp + roff[r]] : sigma;
#pragma unroll
        for (int o = 0; o < NB; ++o) {
            const T* p = P + (int64_t)o * ld;
#pragma unroll
            for (int r = 0; r < TY; ++r) h[o][r] = (hin && yin[r]) ? p[kp + hoff[r]] : T(0);
            lo[o] = (xin && ym_in) ? p[kp + ymo] : T(0);
            hi[o] = (xin && yp_in) ? p[kp + ypo] : T(0);
        }
    };
    T prev[NB][TY], cur[NB][TY], nxt[NB][TY];
    T hv[NB][TY], ym[NB], yp[NB], v[TY];
#pragma unroll
    for (int o = 0; o < NB; ++o) {
        const T* p = P + (int64_t)o * ld;
#pragma unroll
        for (int r = 0; r < TY; ++r) {
            const bool ok = xin && yin[r];
            prev[o][r] = (ok && k0 > 0) ? p[(k0 - 1) * plane + roff[r]] : T(0);
            cur[o][r] = ok ? p[k0 * plane + roff[r]] : T(0);
            if constexpr (PF) nxt[o][r] = (ok && k0 + 1 < nz) ? p[(k0 + 1) * plane + roff[r]] : T(0);
        }
    }
    if constexpr (PF) load_side(k0, hv, ym, yp, v);
    for (int k = k0; k < k1; ++k) {
        const int kp = k * plane;
        T nn[NB][TY], hvn[NB][TY], ymn[NB], ypn[NB], vn[TY];
        if constexpr (PF) {
            const bool has_nn = k + 2 < nz && k + 1 < k1;
#pragma unroll
            for (int o = 0; o < NB; ++o) {
                const T* p = P + (int64_t)o * ld;
#pragma unroll
                for (int r = 0; r < TY; ++r) nn[o][r] = (xin && yin[r] && has_nn) ? p[kp + 2 * plane + roff[r]] : T(0);
            }
            if (k + 1 < k1) load_side(k + 1, hvn, ymn, ypn, vn);
        } else {
            const bool has_next = k + 1 < nz;
#pragma unroll
            for (int o = 0; o < NB; ++o) {
                const T* p = P + (int64_t)o * ld;
#pragma unroll
                for (int r = 0; r < TY; ++r) nxt[o][r] = (xin && yin[r] && has_next) ? p[kp + plane + roff[r]] : T(0);
            }
            load_side(k, hv, ym, yp, v);
        }
#pragma unroll
        for (int o = 0; o < NB; ++o) {
            T* q = Out + (int64_t)o * ldo;
#pragma unroll
            for (int r = 0; r < TY; ++r) {
                const T c = cur[o][r];
                const T xm = shr1(c, hv[o][r]);
                const T xp = shl1(c, hv[o][r]);
                const T ymv = r > 0 ? cur[o][r > 0 ? r - 1 : 0] : ym[o];
                const T ypv = r < TY - 1 ? cur[o][r < TY - 1 ? r + 1 : 0] : yp[o];
                // KohnShamHamiltonianMolecule.cpp:88-187 order: centre, x-, x+, y-, y+, z-, z+
                T lap = co.cc * c;
                lap += xm * co.cx;
                lap += xp * co.cx;
                lap += ymv * co.cy;
                lap += ypv * co.cy;
                lap += prev[o][r] * co.cz;
                lap += nxt[o][r] * co.cz;
                const T kin = T(-0.5) * lap;
                const T pot = v[r] * c;
                if (xin && yin[r]) {
                    T res;
                    if constexpr (SMOOTH) {
                        const T rv = R[(int64_t)(grp * NB + o) * ldr + kp + roff[r]];
                        const T jt = c + cf * (rv - (kin + pot));
                        if constexpr (SM == 2) res = a * jt + b * q[kp + roff[r]];
                        else res = a * jt;
                    } else {
                        res = kin + pot;
                    }
                    __builtin_nontemporal_store(res, q + kp + roff[r]);
                }
            }
        }
#pragma unroll
        for (int o = 0; o < NB; ++o)
#pragma unroll
            for (int r = 0; r < TY; ++r) {
                prev[o][r] = cur[o][r];
                cur[o][r] = nxt[o][r];
                if constexpr (PF) nxt[o][r] = nn[o][r];
            }
        if constexpr (PF) {
            if (k + 1 < k1) {
#pragma unroll
                for (int o = 0; o < NB; ++o) {
#pragma unroll
                    for (int r = 0; r < TY; ++r) hv[o][r] = hvn[o][r];
                    ym[o] = ymn[o];
                    yp[o] = ypn[o];
                }
#pragma unroll
                for (int r = 0; r < TY; ++r) v[r] = vn[r];
            }
        }
    }
}

template <typename T, int TY, int NB, int SM = 0, bool PF = false>
int launch_hpsi(dft_ctx* c, const dft_grid* g, const T* V, const T* psi, int64_t ld, int ngroups, T* out,
                int64_t ldo, const T* R = nullptr, int64_t ldr = 0, T sigma = T(0), T cf = T(0), T a = T(1),
                T b = T(0)) {
    const int nx = (int)g->nx, ny = (int)g->ny, nz = (int)g->nz;
    Coef<T> co;
    const double ix2 = 1.0 / (g->dx * g->dx), iy2 = 1.0 / (g->dy * g->dy), iz2 = 1.0 / (g->dz * g->dz);
    co.cx = (T)ix2;
    co.cy = (T)iy2;
    co.cz = (T)iz2;
    co.cc = (T)(-2.0 * (ix2 + iy2 + iz2));
    const int gx = (nx + 63) / 64, gy = (ny + 4 * TY - 1) / (4 * TY);
    const int64_t base = (int64_t)gx * gy * ngroups;
    const int64_t want = 32LL * c->num_cus;  // enough z-chunks for ~32 workgroups per CU
    int nzc = (int)std::max<int64_t>(1, std::min<int64_t>((want + base - 1) / base, std::max(1, nz / 8)));
    const int zc = (nz + nzc - 1) / nzc;
    nzc = (nz + zc - 1) / zc;
    DFT_REQUIRE(c, (int64_t)nzc * ngroups <= 65535 && gy <= 65535, "dft_hpsi: launch grid too large");
    dim3 grid(gx, gy, nzc * ngroups);
    k_hpsi<T, TY, NB, SM, PF><<<grid, 256, 0, c->stream>>>(nx, ny, nz, zc, nzc, co, V, psi, ld, out, ldo, R, ldr, sigma, cf,
                                                       a, b);
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

// Tile shape: TY y-rows per wave (a 64 x 4TY column per workgroup, so the
// y-halo is 2 rows per 4TY) and NB orbitals per group sharing each V_eff load.
// Chosen from a sweep of TY in {2, 4, 8}, NB in {2, 4, 8}, with and without
// the one-plane prefetch, at 160^3 x 32 and 256^3 x 96 on MI355X
// (non-temporal stores in all variants): FP32 blocks below 64 orbitals run
// TY = 4, NB = 2 (4.2 TB/s at 160^3 x 32); FP32 blocks of 64 and more TY = 8,
// NB = 4 without prefetch (4.67 TB/s at 256^3 x 96, against 4.38 for TY = 4
// with prefetch and 3.6-4.5 for the others); FP64 TY = 4, NB = 4 (4.0 and
// 4.8-5.1 TB/s).  The remainder of a block runs in groups of 2, then 1.
template <typename T, int NB, bool PF, int TY = 4>
int hpsi_groups(dft_ctx* c, const dft_grid* g, const T* V, const T* psi, int64_t ld, int nb, T* out, int64_t ldo) {
    const int ng = nb / NB;
    if (ng > 0)
        if (int rc = launch_hpsi<T, TY, NB, 0, PF>(c, g, V, psi, ld, ng, out, ldo)) return rc;
    int o = ng * NB;
    if (NB > 2 && nb - o >= 2) {
        const int n2 = (nb - o) / 2;
        if (int rc = launch_hpsi<T, TY, 2>(c, g, V, psi + (int64_t)o * ld, ld, n2, out + (int64_t)o * ldo, ldo)) return rc;
        o += 2 * n2;
    }
    for (; o < nb; ++o)
        if (int rc = launch_hpsi<T, TY, 1>(c, g, V, psi + (int64_t)o * ld, ld, 1, out + (int64_t)o * ldo, ldo)) return rc;
    return DFT_OK;
}

template <typename T>
int hpsi_t(dft_ctx* c, const dft_grid* g, const T* V, const T* psi, int64_t ld, int nb, T* out, int64_t ldo) {
    DFT_REQUIRE(c, g->nx * g->ny * g->nz < (1LL << 31), "dft_hpsi: grid exceeds 2^31 points");
    if constexpr (sizeof(T) == 4) {
        if (nb >= 64) return hpsi_groups<T, 4, false, 8>(c, g, V, psi, ld, nb, out, ldo);
        return hpsi_groups<T, 2, false>(c, g, V, psi, ld, nb, out, ldo);
    }
    return hpsi_groups<T, 4, false>(c, g, V, psi, ld, nb, out, ldo);
}

// Kinetic smoothing correction: t = r / D, then `sweeps` damped-Jacobi steps
// on (-1/2 Lap + sigma) t = r, ping-ponging between t and tmp (an even
// number of steps ends in t; an odd one is copied back).
template <typename T>
int kinetic_t(dft_ctx* c, const dft_grid* g, double sigma, double omega, int sweeps, const T* r, int64_t ldr, int nb,
              T* t, int64_t ldt, T* tmp) {
    const int64_t n = grid_points(g);
    const double D = 1.0 / (g->dx * g->dx) + 1.0 / (g->dy * g->dy) + 1.0 / (g->dz * g->dz) + sigma;
    if (int rc = dft_smooth_step(c, sizeof(T) == 8 ? DFT_F64 : DFT_F32, n, nb, 1.0 / D, r, ldr, nullptr, n, t, ldt, 1))
        return rc;
    T* src = t;
    T* dst = tmp;
    int64_t lds = ldt, ldd = n;
    for (int s = 0; s < sweeps; ++s) {
        constexpr int TY = 4, NB = 2;
        const int ng = nb / NB;
        if (ng > 0)
            if (int rc = launch_hpsi<T, TY, NB, 1>(c, g, nullptr, src, lds, ng, dst, ldd, r, ldr, (T)sigma, (T)(omega / D)))
                return rc;
        for (int o = ng * NB; o < nb; ++o)
            if (int rc = launch_hpsi<T, TY, 1, 1>(c, g, nullptr, src + (int64_t)o * lds, lds, 1, dst + (int64_t)o * ldd, ldd,
                                                  r + (int64_t)o * ldr, ldr, (T)sigma, (T)(omega / D)))
                return rc;
        std::swap(src, dst);
        std::swap(lds, ldd);
    }
    if (src != t)
        for (int o = 0; o < nb; ++o)
            DFT_HIP(c, hipMemcpyAsync(t + (int64_t)o * ldt, src + (int64_t)o * lds, sizeof(T) * n, hipMemcpyDeviceToDevice,
                                      c->stream));
    return DFT_OK;
}

// Chebyshev-accelerated kinetic smoothing: the same steps J(t) = t + (1/D)
// (r - A t), A = -1/2 Lap + sigma with zero ghosts, D = sum_a 1/h_a^2 + sigma
// (the centre of A's spectrum, so I - A/D has eigenvalues in [-rho, rho] with
// rho = sum_a cos(pi/(n_a+1))/h_a^2 / D), combined by the Chebyshev
// semi-iteration t_1 = r/D, t_2 = w_2 J(t_1), t_{k+1} = w_{k+1} (J(t_k) -
// t_{k-1}) + t_{k-1}, w_2 = 1/(1 - rho^2/2), w_{k+1} = 1/(1 - rho^2 w_k/4).
// The ping-pong keeps t_{k-1} in the destination buffer, so a step reads
// t_k, r and t_{k-1} and writes t_{k+1}.  The error of the solve of A t = r
// falls by ~rho/(1 + sqrt(1 - rho^2)) per step instead of ~rho.
template <typename T>
int kinetic_cheb_t(dft_ctx* c, const dft_grid* g, double sigma, int sweeps, const T* r, int64_t ldr, int nb, T* t,
                   int64_t ldt, T* tmp) {
    const int64_t n = grid_points(g);
    const double ix2 = 1.0 / (g->dx * g->dx), iy2 = 1.0 / (g->dy * g->dy), iz2 = 1.0 / (g->dz * g->dz);
    const double D = ix2 + iy2 + iz2 + sigma;
    const double kPiL = 3.141592653589793238462643383279502884;
    const double rho = (std::cos(kPiL / (g->nx + 1)) * ix2 + std::cos(kPiL / (g->ny + 1)) * iy2 +
                        std::cos(kPiL / (g->nz + 1)) * iz2) / D;
    if (int rc = dft_smooth_step(c, sizeof(T) == 8 ? DFT_F64 : DFT_F32, n, nb, 1.0 / D, r, ldr, nullptr, n, t, ldt, 1))
        return rc;
    T* src = t;
    T* dst = tmp;
    int64_t lds = ldt, ldd = n;
    double w = 1.0;
    for (int s = 0; s < sweeps; ++s) {
        w = s == 0 ? 1.0 / (1.0 - 0.5 * rho * rho) : 1.0 / (1.0 - 0.25 * rho * rho * w);
        const T a = (T)w, b = (T)(1.0 - w), cf = (T)(1.0 / D);
        constexpr int TY = 4, NB = 2;
        const int ng = nb / NB;
        // the first step has t_0 = 0: nothing to blend
        auto step = [&](auto nbc, const T* in, int64_t li, int groups, T* out, int64_t lo, const T* rr) -> int {
            constexpr int G = decltype(nbc)::value;
            if (s == 0) return launch_hpsi<T, TY, G, 1>(c, g, nullptr, in, li, groups, out, lo, rr, ldr, (T)sigma, cf, a);
            return launch_hpsi<T, TY, G, 2>(c, g, nullptr, in, li, groups, out, lo, rr, ldr, (T)sigma, cf, a, b);
        };
        if (ng > 0)
            if (int rc = step(std::integral_constant<int, NB>(), src, lds, ng, dst, ldd, r)) return rc;
        for (int o = ng * NB; o < nb; ++o)
            if (int rc = step(std::integral_constant<int, 1>(), src + (int64_t)o * lds, lds, 1, dst + (int64_t)o * ldd, ldd,
                              r + (int64_t)o * ldr))
                return rc;
        std::swap(src, dst);
        std::swap(lds, ldd);
    }
    if (src != t)
        for (int o = 0; o < nb; ++o)
            DFT_HIP(c, hipMemcpyAsync(t + (int64_t)o * ldt, src + (int64_t)o * lds, sizeof(T) * n, hipMemcpyDeviceToDevice,
                                      c->stream));
    return DFT_OK;
}

// Kinetic energy with the reference's face rule (TotalEnergyMolecule.cpp:80-186):
// only existing neighbours contribute (psi_nb - psi)/h^2.  One orbital per
// grid row (blockIdx.y), so every workgroup streams one column with
// coalesced centre and neighbour loads; per-workgroup FP64 partials.
template <typename T>
__global__ __launch_bounds__(256) void k_kinetic(int nx, int ny, int nz, double dx2, double dy2, double dz2,
                                                 const T* __restrict__ psi, int64_t ld,
                                                 const double* __restrict__ el, double* __restrict__ partials) {
    __shared__ double lds[4];
    const int o = blockIdx.y;
    const double w = el[o];
    double s = 0.0;
    if (w > 0.0) {
        // the caller guarantees fewer than 2^31 points: 32-bit index arithmetic
        const int n = nx * ny * nz, plane = nx * ny;
        const double ix2 = 1.0 / dx2, iy2 = 1.0 / dy2, iz2 = 1.0 / dz2;
        const T* q = psi + (int64_t)o * ld;
        const int stride = (int)gridDim.x * 256;
#pragma unroll 4
        for (int p = (int)blockIdx.x * 256 + (int)threadIdx.x; p < n; p += stride) {
            const int k = p / plane, r = p - k * plane, j = r / nx, i = r - j * nx;
            // a missing neighbour is read as the point itself, so its term (psi - psi) / h^2
            // is an exact zero: six unconditional loads the wave issues together
            const double c = (double)q[p];
            const double xm = (double)q[p - (i > 0 ? 1 : 0)], xp = (double)q[p + (i + 1 < nx ? 1 : 0)];
            const double ym = (double)q[p - (j > 0 ? nx : 0)], yp = (double)q[p + (j + 1 < ny ? nx : 0)];
            const double zm = (double)q[p - (k > 0 ? plane : 0)], zp = (double)q[p + (k + 1 < nz ? plane : 0)];
            double lap = 0.0;
            lap += (xm - c) * ix2;
            lap += (xp - c) * ix2;
            lap += (ym - c) * iy2;
            lap += (yp - c) * iy2;
            lap += (zm - c) * iz2;
            lap += (zp - c) * iz2;
            s += c * (-0.5 * lap);
        }
        s *= w;
    }
    s = block_sum(s, lds);
    if (threadIdx.x == 0) partials[(int64_t)o * gridDim.x + blockIdx.x] = s;
}

template <typename T>
__global__ __launch_bounds__(256) void k_precond_diag(int64_t n, double kin, const T* __restrict__ V,
                                                      const T* __restrict__ r, int64_t ldr,
                                                      const double* __restrict__ eps, T* __restrict__ t,
                                                      int64_t ldt) {
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int o = blockIdx.y;
    if (p >= n) return;
    // EigenvalueSolverMolecule.cpp:500-554
    const double d = eps[o] - (kin + (double)V[p]);
    const double rv = (double)r[(int64_t)o * ldr + p];
    t[(int64_t)o * ldt + p] = (T)((!isfinite(d) || fabs(d) < 1.0e-10) ? rv : rv / d);
}

}  // namespace

extern "C" {

int dft_hpsi(dft_ctx* c, const dft_grid* g, int prec, const void* veff, const void* psi, int64_t ld, int nb, void* out,
             int64_t ldo) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, valid_grid(g), "dft_hpsi: invalid grid");
    const int64_t n = grid_points(g);
    DFT_REQUIRE(c, veff && psi && out && nb >= 0 && ld >= n && ldo >= n, "dft_hpsi: bad arguments");
    DFT_REQUIRE(c, g->ny < (1 << 30) && g->nx < (1 << 30), "dft_hpsi: grid too large");
    if (nb == 0) return DFT_OK;
    if (prec == DFT_F64) return hpsi_t<double>(c, g, (const double*)veff, (const double*)psi, ld, nb, (double*)out, ldo);
    if (prec == DFT_F32) return hpsi_t<float>(c, g, (const float*)veff, (const float*)psi, ld, nb, (float*)out, ldo);
    return set_err(c, DFT_EINVAL, "dft_hpsi: bad precision");
}

int dft_kinetic_energy_occ(dft_ctx* c, const dft_grid* g, int prec, const void* psi, int64_t ld, int nb,
                           const double* occ_host, double* t_host) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, valid_grid(g) && psi && t_host && nb >= 0 && (nb == 0 || occ_host) && (prec == DFT_F64 || prec == DFT_F32),
                "dft_kinetic_energy: bad arguments");
    const int64_t n = grid_points(g);
    DFT_REQUIRE(c, ld >= n, "dft_kinetic_energy: ld < grid size");
    if (nb == 0) { *t_host = 0.0; return DFT_OK; }
    DFT_REQUIRE(c, nb <= 65535 && n < (1LL << 31), "dft_kinetic_energy: more than 65535 orbitals or 2^31 points");
    // ~8 workgroups per CU over all orbitals, at least one per orbital
    const int bx = (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8 * (int64_t)c->num_cus / nb));
    const size_t parts = (size_t)bx * nb;
    if (int rc = ensure_partials(c, parts + 1 + (size_t)nb)) return rc;
    double* el = c->partials + parts + 1;
    // the copy is synchronised below (finish_sum) before the host array can go away
    DFT_HIP(c, hipMemcpyAsync(el, occ_host, sizeof(double) * nb, hipMemcpyHostToDevice, c->stream));
    const double dx2 = g->dx * g->dx, dy2 = g->dy * g->dy, dz2 = g->dz * g->dz;
    const dim3 grid((unsigned)bx, (unsigned)nb);
    if (prec == DFT_F64)
        k_kinetic<double><<<grid, 256, 0, c->stream>>>((int)g->nx, (int)g->ny, (int)g->nz, dx2, dy2, dz2, (const double*)psi, ld, el, c->partials);
    else
        k_kinetic<float><<<grid, 256, 0, c->stream>>>((int)g->nx, (int)g->ny, (int)g->nz, dx2, dy2, dz2, (const float*)psi, ld, el, c->partials);
    DFT_CHECK_LAUNCH(c);
    double s = 0.0;
    if (int rc = finish_sum(c, (int64_t)parts, &s)) return rc;
    *t_host = s * (g->dx * g->dy * g->dz);
    return DFT_OK;
}

int dft_kinetic_energy(dft_ctx* c, const dft_grid* g, int prec, const void* psi, int64_t ld, int nb,
                       const int* electrons_host, double* t_host) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, nb >= 0 && (nb == 0 || electrons_host), "dft_kinetic_energy: bad arguments");
    std::vector<double> elh(nb);
    for (int o = 0; o < nb; ++o) elh[o] = (double)electrons_host[o];
    return dft_kinetic_energy_occ(c, g, prec, psi, ld, nb, elh.data(), t_host);
}

int dft_precond_kinetic(dft_ctx* c, const dft_grid* g, int prec, double sigma, double omega, int sweeps, const void* r,
                        int64_t ldr, int nb, void* t, int64_t ldt, void* tmp) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, valid_grid(g) && r && t && tmp && nb >= 0 && nb <= 65535 && sweeps >= 0 && sigma > 0.0 &&
                       ldr >= grid_points(g) && ldt >= grid_points(g),
                "dft_precond_kinetic: bad arguments");
    DFT_REQUIRE(c, g->nx * g->ny * g->nz < (1LL << 31), "dft_precond_kinetic: grid exceeds 2^31 points");
    if (!nb) return DFT_OK;
    if (prec == DFT_F64)
        return kinetic_t<double>(c, g, sigma, omega, sweeps, (const double*)r, ldr, nb, (double*)t, ldt, (double*)tmp);
    if (prec == DFT_F32)
        return kinetic_t<float>(c, g, sigma, omega, sweeps, (const float*)r, ldr, nb, (float*)t, ldt, (float*)tmp);
    return set_err(c, DFT_EINVAL, "dft_precond_kinetic: bad precision");
}

int dft_precond_kinetic_cheb(dft_ctx* c, const dft_grid* g, int prec, double sigma, int sweeps, const void* r,
                             int64_t ldr, int nb, void* t, int64_t ldt, void* tmp) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, valid_grid(g) && r && t && tmp && nb >= 0 && nb <= 65535 && sweeps >= 0 && sigma > 0.0 &&
                       ldr >= grid_points(g) && ldt >= grid_points(g),
                "dft_precond_kinetic_cheb: bad arguments");
    DFT_REQUIRE(c, g->nx * g->ny * g->nz < (1LL << 31), "dft_precond_kinetic_cheb: grid exceeds 2^31 points");
    if (!nb) return DFT_OK;
    if (prec == DFT_F64)
        return kinetic_cheb_t<double>(c, g, sigma, sweeps, (const double*)r, ldr, nb, (double*)t, ldt, (double*)tmp);
    if (prec == DFT_F32)
        return kinetic_cheb_t<float>(c, g, sigma, sweeps, (const float*)r, ldr, nb, (float*)t, ldt, (float*)tmp);
    return set_err(c, DFT_EINVAL, "dft_precond_kinetic_cheb: bad precision");
}

int dft_precond_diag(dft_ctx* c, const dft_grid* g, int prec, const void* veff, const void* r, int64_t ldr, int nb,
                     const double* eps, void* t, int64_t ldt) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, valid_grid(g) && veff && r && eps && t && nb >= 0 && nb < 65536, "dft_precond_diag: bad arguments");
    const int64_t n = grid_points(g);
    DFT_REQUIRE(c, ldr >= n && ldt >= n, "dft_precond_diag: leading dimension < grid size");
    if (nb == 0) return DFT_OK;
    const double kin = 1.0 / (g->dx * g->dx) + 1.0 / (g->dy * g->dy) + 1.0 / (g->dz * g->dz);
    dim3 grid((unsigned)((n + 255) / 256), nb);
    if (prec == DFT_F64)
        k_precond_diag<double><<<grid, 256, 0, c->stream>>>(n, kin, (const double*)veff, (const double*)r, ldr, eps, (double*)t, ldt);
    else
        k_precond_diag<float><<<grid, 256, 0, c->stream>>>(n, kin, (const float*)veff, (const float*)r, ldr, eps, (float*)t, ldt);
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

}  // extern "C"
