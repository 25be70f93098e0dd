// poisson.hip -- (b) Jacobi Poisson solver for V_H with the reference's
// boundary rule and stop rule.
//
// Two sweeps per launch (k_jacobi_tb, temporal blocking with overlapping
// tiles, no grid-wide wait): the same z-streaming layout as the H psi
// kernel (lane = x with DPP wave shifts for x-neighbours, R y-rows per lane,
// z-1/z/z+1 planes in registers, one register queue per sweep), updating
// interior points only; boundary points keep their values in both
// ping-pong buffers.  Each sweep reduces max|V_new - V_old| over the points
// a workgroup owns and atomically maxes it into its own slot, so a batch of
// sweeps runs without host round trips and the host finds the exact sweep
// at which the reference would have stopped.  The Chebyshev solver runs one
// sweep per launch (k_jacobi with its extrapolation epilogue).
#include "common.hpp"

#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

using namespace dft;

namespace {

constexpr double kPi = 3.141592653589793238462643383279502884;  // DFTConstants.h:7

__device__ inline float shr1(float v, float fill) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(fill), __float_as_int(v), 0x138, 0xf, 0xf, false));
}
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
struct JCoef {
    T cx, cy, cz, den, fourpi;
};

// One sweep over the tile (bx, by, bz): 64 x-points, 4*TY y-rows, zc planes.
// Returns this thread's max|V_new - V_old| over the interior points it updated.
// CHEB: Chebyshev semi-iteration on the Jacobi splitting; `out` holds the
// iterate before `cur_in` and receives omega*(J(cur) - prev) + prev.
template <typename T, int TY, bool CHEB = false>
__device__ inline T jacobi_tile(int nx, int ny, int nz, int zc, const JCoef<T>& co, const T* __restrict__ rho,
                                const T* __restrict__ cur_in, T* __restrict__ out, int bx, int by, int bz,
                                T omega = T(1)) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int x0 = bx * 64, x = x0 + lane;
    const int ybase = by * (4 * TY) + w * TY;
    const int k0 = bz * zc, k1 = min(nz, k0 + zc);
    const int plane = nx * ny;  // dft_poisson guarantees nx*ny*nz < 2^31: 32-bit offsets
    const bool xin = x < nx;
    const bool xint = x > 0 && x + 1 < nx;
    // x-halo: lane 0 reads x0-1, lane 63 reads x0+64; other lanes stay masked off
    int hx = x;
    bool hin = false;
    if (lane == 0) { hx = x0 - 1; hin = x0 > 0; }
    if (lane == 63) { hx = x0 + 64; hin = x0 + 64 < nx; }
    bool yin[TY], yint[TY];
    int roff[TY], hoff[TY];
#pragma unroll
    for (int r = 0; r < TY; ++r) {
        yin[r] = (ybase + r) < ny;
        yint[r] = (ybase + r) > 0 && (ybase + r) + 1 < ny;
        roff[r] = (ybase + r) * nx + x;
        hoff[r] = (ybase + r) * nx + hx;
    }
    const bool ym_in = ybase > 0 && ybase - 1 < ny, yp_in = ybase + TY < ny;
    const int ymo = (ybase - 1) * nx + x, ypo = (ybase + TY) * nx + x;
    T prev[TY], cur[TY], nxt[TY];
#pragma unroll
    for (int r = 0; r < TY; ++r) {
        const bool ok = xin && yin[r];
        prev[r] = (ok && k0 > 0) ? cur_in[(k0 - 1) * plane + roff[r]] : T(0);
        cur[r] = ok ? cur_in[k0 * plane + roff[r]] : T(0);
    }
    T md = T(0);
    for (int k = k0; k < k1; ++k) {
        const int pk = k * plane;
        const bool has_next = k + 1 < nz;
        const bool kint = k > 0 && k + 1 < nz;
#pragma unroll
        for (int r = 0; r < TY; ++r) nxt[r] = (xin && yin[r] && has_next) ? cur_in[pk + plane + roff[r]] : T(0);
        const T ym = (xin && ym_in) ? cur_in[pk + ymo] : T(0);
        const T yp = (xin && yp_in) ? cur_in[pk + ypo] : T(0);
#pragma unroll
        for (int r = 0; r < TY; ++r) {
            const T h = (hin && yin[r]) ? cur_in[pk + hoff[r]] : T(0);
            const T c = cur[r];
            const T xm = shr1(c, h);
            const T xp = shl1(c, h);
            const T ymv = r > 0 ? cur[r > 0 ? r - 1 : 0] : ym;
            const T ypv = r < TY - 1 ? cur[r < TY - 1 ? r + 1 : 0] : yp;
            if (xint && yint[r] && kint) {
                // HartreePotentialMolecule.cpp:295-322
                const T src = co.fourpi * rho[pk + roff[r]];
                T nv = (co.cx * (xm + xp) + co.cy * (ymv + ypv) + co.cz * (prev[r] + nxt[r]) + src) / co.den;
                if (CHEB) {
                    const T pv = out[pk + roff[r]];
                    nv = omega * (nv - pv) + pv;
                }
                const T d = nv > c ? nv - c : c - nv;
                md = md < d ? d : md;
                out[pk + roff[r]] = nv;
            }
        }
#pragma unroll
        for (int r = 0; r < TY; ++r) {
            prev[r] = cur[r];
            cur[r] = nxt[r];
        }
    }
    return md;
}

// workgroup max of a per-thread value into *slot (non-negative doubles)
__device__ inline void block_max_to(double m, double* red, unsigned long long* slot) {
    m = wave_max(m);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) red[w] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = red[0];
        for (int i = 1; i < 4; ++i) b = b < red[i] ? red[i] : b;
        if (b > 0.0) atomicMax(slot, dbits(b));
    }
    __syncthreads();
}

// one sweep per launch: one workgroup per tile
template <typename T, int TY, bool CHEB = false>
__global__ __launch_bounds__(256) void k_jacobi(int nx, int ny, int nz, int zc, JCoef<T> co, const T* __restrict__ rho,
                                                const T* __restrict__ cur_in, T* __restrict__ out,
                                                unsigned long long* __restrict__ slot, T omega) {
    __shared__ double red[4];
    const int64_t tile = xcd_tile(blockIdx.x + (int64_t)gridDim.x * (blockIdx.y + (int64_t)gridDim.y * blockIdx.z),
                                  (int64_t)gridDim.x * gridDim.y * gridDim.z);
    const int bx = (int)(tile % gridDim.x), by = (int)((tile / gridDim.x) % gridDim.y);
    const int bz = (int)(tile / ((int64_t)gridDim.x * gridDim.y));
    const T md = jacobi_tile<T, TY, CHEB>(nx, ny, nz, zc, co, rho, cur_in, out, bx, by, bz, omega);
    block_max_to((double)md, red, slot);
}

// ns <= KT sweeps per launch by temporal blocking, with no grid-wide wait.
// A workgroup loads an extended tile of 64 (x, one lane each) x 4R (y, R
// rows per thread) points and streams it along z; the tile overlaps its
// neighbours by KT points on each side in x and y, and its z-range by KT
// planes, so every sweep it needs is computed inside the workgroup (the
// overlap is computed twice, by both neighbours, with identical inputs and
// operations, hence identical values).  Sweep t (level t) is pipelined one
// plane behind sweep t - 1: at step s the thread loads plane s of the input
// (level 0) and computes plane s - t of level t from level t - 1's planes
// s - t - 1 .. s - t + 1, which it holds in a three-plane register queue per
// level.  x-neighbours come from the adjacent lanes (DPP); y-neighbours from
// the thread's own rows, or across waves from the row each wave publishes
// per level and plane in LDS (two plane slots per level: what a level reads
// at step s was written at step s - 1, so one barrier per step suffices).
// Values near the extended tile's edges are wrong after a few levels (their
// neighbours lie outside the tile); they never reach a level-ns value of
// the owned region [KT, 64 - KT) x [KT, 4R - KT) x [z0, z1), which alone is
// stored.  Each level's max|V_new - V_old| over the owned interior points
// goes to its own slot, so the per-sweep stop rule and replay work as with
// one sweep per launch; the arithmetic per point is jacobi_tile's, so the
// results are bitwise those of ns single sweeps.
template <typename T, int R, int KT>
__global__ __launch_bounds__(256) void k_jacobi_tb(int nx, int ny, int nz, int zc, int ns, JCoef<T> co,
                                                   const T* __restrict__ rho, const T* __restrict__ in,
                                                   T* __restrict__ out, unsigned long long* __restrict__ slot) {
    constexpr int EY = 4 * R, OX = 64 - 2 * KT, OY = EY - 2 * KT;
    __shared__ T edge[KT][2][4][2][64];  // [level][plane parity][wave][first / last row][lane]
    __shared__ double red[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t tile = xcd_tile(blockIdx.x + (int64_t)gridDim.x * (blockIdx.y + (int64_t)gridDim.y * blockIdx.z),
                                  (int64_t)gridDim.x * gridDim.y * gridDim.z);
    const int bx = (int)(tile % gridDim.x), by = (int)((tile / gridDim.x) % gridDim.y);
    const int bz = (int)(tile / ((int64_t)gridDim.x * gridDim.y));
    const int plane = nx * ny;  // dft_poisson guarantees nx*ny*nz < 2^31: 32-bit offsets
    const int gx = bx * OX - KT + lane;
    const bool xg = gx >= 0 && gx < nx, xi = gx > 0 && gx + 1 < nx;
    const bool xown = lane >= KT && lane < 64 - KT && xg;
    bool yg[R], yi[R], yown[R];
    int off[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int ly = w * R + r, gy = by * OY - KT + ly;
        yg[r] = gy >= 0 && gy < ny;
        yi[r] = gy > 0 && gy + 1 < ny;
        yown[r] = ly >= KT && ly < EY - KT && yg[r];
        off[r] = gy * nx + gx;
    }
    const int z0 = bz * zc, z1 = min(nz, z0 + zc);
    // loads for step s: plane s of the input, and rho at plane s - t for level t; issued one step
    // ahead, so they are in flight while the previous step computes
    auto load_in = [&](int s, T (&v)[R]) {
        const bool pin = s >= 0 && s < nz;
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = (pin && xg && yg[r]) ? in[s * plane + off[r]] : T(0);
    };
    auto load_rho = [&](int s, T (&v)[KT][R]) {
#pragma unroll
        for (int t = 1; t <= KT; ++t) {
            const int p = s - t;
            const bool pint = t <= ns && p > 0 && p + 1 < nz;
#pragma unroll
            for (int r = 0; r < R; ++r) v[t - 1][r] = (pint && xi && yi[r]) ? rho[p * plane + off[r]] : T(0);
        }
    };
    T q[KT][3][R];  // level t's planes p - 1, p, p + 1 (oldest first)
#pragma unroll
    for (int t = 0; t < KT; ++t)
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int r = 0; r < R; ++r) q[t][j][r] = T(0);
    T md[KT];
#pragma unroll
    for (int t = 0; t < KT; ++t) md[t] = T(0);
    T nin[R], nrho[KT][R];
    load_in(z0 - ns, nin);
    load_rho(z0 - ns, nrho);
    for (int s = z0 - ns; s < z1 + ns; ++s) {
        T rh[KT][R];
#pragma unroll
        for (int r = 0; r < R; ++r) {  // level 0: plane s of the input
            q[0][0][r] = q[0][1][r];
            q[0][1][r] = q[0][2][r];
            q[0][2][r] = nin[r];
        }
#pragma unroll
        for (int t = 0; t < KT; ++t)
#pragma unroll
            for (int r = 0; r < R; ++r) rh[t][r] = nrho[t][r];
        if (s + 1 < z1 + ns) {
            load_in(s + 1, nin);
            load_rho(s + 1, nrho);
        }
        edge[0][s & 1][w][0][lane] = q[0][2][0];
        edge[0][s & 1][w][1][lane] = q[0][2][R - 1];
#pragma unroll
        for (int t = 1; t <= KT; ++t) {
            if (t > ns) break;
            const int p = s - t;
            const bool pint = p > 0 && p + 1 < nz;
            const bool pown = p >= z0 && p < z1;
            // rows of level t - 1's plane p next to this thread's rows, published at step s - 1
            const T above = w > 0 ? edge[t - 1][p & 1][w - 1][1][lane] : q[t - 1][1][0];
            const T below = w < 3 ? edge[t - 1][p & 1][w + 1][0][lane] : q[t - 1][1][R - 1];
            T nv[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const T c = q[t - 1][1][r];
                const T xm = shr1(c, c);
                const T xp = shl1(c, c);
                const T ymv = r > 0 ? q[t - 1][1][r > 0 ? r - 1 : 0] : above;
                const T ypv = r < R - 1 ? q[t - 1][1][r < R - 1 ? r + 1 : 0] : below;
                nv[r] = c;
                if (xi && yi[r] && pint) {
                    // HartreePotentialMolecule.cpp:295-322
                    const T src = co.fourpi * rh[t - 1][r];
                    const T v = (co.cx * (xm + xp) + co.cy * (ymv + ypv) + co.cz * (q[t - 1][0][r] + q[t - 1][2][r]) + src) /
                                co.den;
                    nv[r] = v;
                    if (xown && yown[r] && pown) {
                        const T d = v > c ? v - c : c - v;
                        md[t - 1] = md[t - 1] < d ? d : md[t - 1];
                    }
                }
            }
            if (t == ns) {
#pragma unroll
                for (int r = 0; r < R; ++r)
                    if (xown && yown[r] && pown && xi && yi[r] && pint) out[p * plane + off[r]] = nv[r];
            } else if (t < KT) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    q[t][0][r] = q[t][1][r];
                    q[t][1][r] = q[t][2][r];
                    q[t][2][r] = nv[r];
                }
                edge[t][p & 1][w][0][lane] = nv[0];
                edge[t][p & 1][w][1][lane] = nv[R - 1];
            }
        }
        lds_barrier();
    }
#pragma unroll
    for (int t = 0; t < KT; ++t) {
        if (t >= ns) break;
        block_max_to((double)md[t], red, slot + t);
    }
}

// Boundary values (HartreePotentialMolecule.cpp:37-95, 171-202): 0 on a
// face, else Q/|r| measured from the origin.  Written to both buffers.
template <typename T>
__global__ __launch_bounds__(256) void k_boundary(int nx, int ny, int nz, double xmin, double xmax, double ymin,
                                                  double ymax, double zmin, double zmax, double dx, double dy,
                                                  double dz, double q, T* __restrict__ a,
                                                  T* __restrict__ b /* may be null */) {
    const int64_t n = (int64_t)nx * ny * nz;
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
    const int i = (int)(p % nx), j = (int)((p / nx) % ny), k = (int)(p / ((int64_t)nx * ny));
    if (!(i == 0 || i + 1 == nx || j == 0 || j + 1 == ny || k == 0 || k + 1 == nz)) return;
    const double x = xmin + (double)i * dx, y = ymin + (double)j * dy, z = zmin + (double)k * dz;
    const double d[6] = {fabs(x - xmin), fabs(x - xmax), fabs(y - ymin), fabs(y - ymax), fabs(z - zmin), fabs(z - zmax)};
    double m = INFINITY;
    for (int s = 0; s < 6; ++s) {
        const double t = d[s] * d[s];
        m = t < m ? t : m;
    }
    double val = 0.0;
    if (m > 0.0) {
        const double r = sqrt(x * x + y * y + z * z);
        if (r > 0.0) val = q / r;
    }
    a[p] = (T)val;
    if (b) b[p] = (T)val;
}

template <typename T>
__global__ __launch_bounds__(256) void k_charge(const T* __restrict__ rho, int64_t n, double* partials) {
    __shared__ double lds[4];
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) s += (double)rho[i];
    s = block_sum(s, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// Sweep blocking: KT = 2 sweeps per launch on 64 x 16 tiles (R = 4 rows per
// thread; owned 60 x 12), zc planes per workgroup.  Chosen from a sweep of
// R in {4, 8}, KT in {1, 2, 4} and the z-chunking at 48^3 .. 256^3 in FP64
// and FP32 on MI355X: KT = 2, R = 4 was fastest at every size, 1.2-2.1x the
// one-sweep-per-launch kernel (7.4 us against 15.3 us per sweep at 48^3,
// 13.4 against 17.3 at 96^3, 24 against 33 at 160^3, 77 against 109 at
// 256^3, FP64); KT = 4 loses to the wider overlap and the longer pipeline.
constexpr int kTbRows = 4, kTbSweeps = 2;

struct TbShape {
    int zc = 0;
    dim3 grid;
};

TbShape tb_shape(const dft_grid* g, int num_cus) {
    TbShape t;
    const int64_t nx = g->nx, ny = g->ny, nz = g->nz;
    const int OX = 64 - 2 * kTbSweeps, OY = 4 * kTbRows - 2 * kTbSweeps;
    const int64_t gx = (nx + OX - 1) / OX, gy = (ny + OY - 1) / OY;
    // z-chunks: about 4 workgroups per CU with chunks of at least 2 KT planes; small grids
    // (below 64^3) 8 per CU with chunks down to KT planes, where latency, not traffic, sets the rate
    const bool small = nx * ny * nz < 262144;
    const int64_t wpc = small ? 8 : 4, minz = (small ? 1 : 2) * kTbSweeps;
    int64_t nzc = std::max<int64_t>(1, (wpc * num_cus + gx * gy - 1) / (gx * gy));
    nzc = std::min<int64_t>(nzc, std::max<int64_t>(1, nz / minz));
    t.zc = (int)((nz + nzc - 1) / nzc);
    nzc = (nz + t.zc - 1) / t.zc;
    t.grid = dim3((unsigned)gx, (unsigned)gy, (unsigned)nzc);
    return t;
}

template <typename T>
struct Jacobi {
    dft_ctx* c;
    const dft_grid* g;
    JCoef<T> co;
    dim3 grid;
    int zc;
    const T* rho;
    TbShape tb;

    // ns <= kTbSweeps sweeps in one launch
    int block(const T* in, T* out, int ns, unsigned long long* slot) {
        k_jacobi_tb<T, kTbRows, kTbSweeps><<<tb.grid, 256, 0, c->stream>>>((int)g->nx, (int)g->ny, (int)g->nz, tb.zc, ns, co,
                                                                           rho, in, out, slot);
        DFT_CHECK_LAUNCH(c);
        return DFT_OK;
    }

    // nsweeps sweeps starting from buffer a, ping-ponging a -> b -> a per launch; *launches
    // receives the number of launches (the result is in b when it is odd)
    int sweeps(T* a, T* b, int nsweeps, unsigned long long* slot, int* launches) {
        int L = 0;
        for (int s = 0; s < nsweeps; ++L) {
            const int ns = std::min(kTbSweeps, nsweeps - s);
            if (int rc = block((L & 1) ? b : a, (L & 1) ? a : b, ns, slot + s)) return rc;
            s += ns;
        }
        *launches = L;
        return DFT_OK;
    }
};

// Boundary values (and zero interior when asked) in v and tmp for a solve
// with charge density rho; tmp receives a copy of v's interior.
template <typename T>
int poisson_prepare(dft_ctx* c, const dft_grid* g, const T* rho, T* v, T* tmp, int zero_start) {
    const int64_t n = grid_points(g);
    const size_t bytes = (size_t)n * sizeof(T);
    const int nblk = (int)std::min<int64_t>((n + 255) / 256, 4 * (int64_t)c->num_cus);
    if (int rc = ensure_partials(c, (size_t)nblk + 2)) return rc;
    // total charge Q = sum(rho) dV (HartreePotentialMolecule.cpp:17-35), on the device
    k_charge<T><<<nblk, 256, 0, c->stream>>>(rho, n, c->partials);
    DFT_CHECK_LAUNCH(c);
    double q = 0.0;
    if (int rc = finish_sum(c, nblk, &q)) return rc;
    const double qd = q * (g->dx * g->dy * g->dz);
    if (zero_start) DFT_HIP(c, hipMemsetAsync(v, 0, bytes, c->stream));
    DFT_HIP(c, hipMemcpyAsync(tmp, v, bytes, hipMemcpyDeviceToDevice, c->stream));
    k_boundary<T><<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>((int)g->nx, (int)g->ny, (int)g->nz, g->xmin, g->xmax,
                                                                        g->ymin, g->ymax, g->zmin, g->zmax, g->dx, g->dy,
                                                                        g->dz, qd, v, tmp);
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

template <typename T>
void jacobi_setup(dft_ctx* c, const dft_grid* g, const T* rho, Jacobi<T>& J) {
    J.c = c;
    J.g = g;
    const double ix2 = 1.0 / (g->dx * g->dx), iy2 = 1.0 / (g->dy * g->dy), iz2 = 1.0 / (g->dz * g->dz);
    J.co.cx = (T)ix2;
    J.co.cy = (T)iy2;
    J.co.cz = (T)iz2;
    J.co.den = (T)(2.0 * (ix2 + iy2 + iz2));
    J.co.fourpi = (T)(4.0 * kPi);
    J.rho = rho;
    const int nx = (int)g->nx, ny = (int)g->ny, nz = (int)g->nz;
    const int gx = (nx + 63) / 64, gy = (ny + 15) / 16;
    const int64_t base = (int64_t)gx * gy;
    int nzc = (int)std::max<int64_t>(1, std::min<int64_t>((8LL * c->num_cus + base - 1) / base, std::max(1, nz / 8)));
    J.zc = (nz + nzc - 1) / nzc;
    nzc = (nz + J.zc - 1) / J.zc;
    J.grid = dim3(gx, gy, nzc);
    J.tb = tb_shape(g, c->num_cus);
}

template <typename T>
int poisson_t(dft_ctx* c, const dft_grid* g, const T* rho, T* v, int64_t max_sweeps, double tol, int zero_start,
              int batch, int64_t* sweeps_done, double* last_maxdiff) {
    const int64_t n = grid_points(g);
    const size_t bytes = (size_t)n * sizeof(T);
    const bool exact = tol > 0.0;
    if (int rc = ensure_scratch(c, bytes * (exact ? 2 : 1))) return rc;
    T* tmp = (T*)c->scratch;
    T* snap = exact ? (T*)((char*)c->scratch + bytes) : nullptr;
    if (int rc = poisson_prepare<T>(c, g, rho, v, tmp, zero_start)) return rc;
    Jacobi<T> J;
    jacobi_setup<T>(c, g, rho, J);

    if (batch < 1) batch = 1;
    if (batch > (int)c->sweep_max_cap / 2) batch = (int)c->sweep_max_cap / 2;
    T* bufs[2] = {v, tmp};
    int curi = 0;  // buffer holding the latest iterate
    int64_t done = 0;
    double md = 0.0;
    std::vector<unsigned long long> hm(batch);
    while (done < max_sweeps) {
        const int nb = (int)std::min<int64_t>(batch, max_sweeps - done);
        DFT_HIP(c, hipMemsetAsync(c->sweep_max, 0, sizeof(unsigned long long) * nb, c->stream));
        if (exact) DFT_HIP(c, hipMemcpyAsync(snap, bufs[curi], bytes, hipMemcpyDeviceToDevice, c->stream));
        const int start = curi;
        int launches = 0;
        if (int rc = J.sweeps(bufs[curi], bufs[curi ^ 1], nb, c->sweep_max, &launches)) return rc;
        if (launches & 1) curi ^= 1;
        if (!exact && done + nb < max_sweeps) { done += nb; continue; }
        DFT_HIP(c, hipMemcpyAsync(hm.data(), c->sweep_max, sizeof(unsigned long long) * nb, hipMemcpyDeviceToHost, c->stream));
        DFT_HIP(c, hipStreamSynchronize(c->stream));
        int stop = -1;
        if (exact)
            for (int s = 0; s < nb; ++s) {
                double d;
                memcpy(&d, &hm[s], sizeof d);
                if (d < tol) { stop = s; break; }
            }
        if (stop < 0) {
            done += nb;
            memcpy(&md, &hm[nb - 1], sizeof md);
            continue;
        }
        // the reference stops after sweep `stop`; replay the batch up to it
        memcpy(&md, &hm[stop], sizeof md);
        if (stop < nb - 1) {
            curi = start;
            DFT_HIP(c, hipMemcpyAsync(bufs[curi], snap, bytes, hipMemcpyDeviceToDevice, c->stream));
            // replay into the upper slots (the batch's maxima are already read)
            if (int rc = J.sweeps(bufs[curi], bufs[curi ^ 1], stop + 1, c->sweep_max + c->sweep_max_cap / 2, &launches))
                return rc;
            if (launches & 1) curi ^= 1;
        }
        done += stop + 1;
        break;
    }
    if (curi != 0) DFT_HIP(c, hipMemcpyAsync(v, tmp, bytes, hipMemcpyDeviceToDevice, c->stream));
    DFT_HIP(c, hipStreamSynchronize(c->stream));
    if (sweeps_done) *sweeps_done = done;
    if (last_maxdiff) *last_maxdiff = md;
    return DFT_OK;
}


// Chebyshev semi-iteration on the Jacobi splitting (Golub & Varga 1961): with
// rho_J the spectral radius of the Jacobi iteration matrix, known in closed
// form for the Dirichlet box,
//   rho_J = sum_a cos(pi / (n_a - 1)) / h_a^2 / sum_a 1 / h_a^2,
// v_1 = J(v_0), v_{k+1} = w_{k+1} (J(v_k) - v_{k-1}) + v_{k-1},
// w_2 = 1 / (1 - rho_J^2 / 2), w_{k+1} = 1 / (1 - rho_J^2 w_k / 4).
// Same fixed point as the reference's Jacobi, in O(n) instead of O(n^2)
// sweeps.  Stops at the end of the first batch containing a sweep with
// max|v_{k+1} - v_k| < tol, or when a whole batch lowered the smallest
// max|dV| by less than 30% (the rounding floor of the precision: FP32
// cannot reach tolerances near 1e-8 on V ~ 10).
template <typename T>
int poisson_cheb_t(dft_ctx* c, const dft_grid* g, const T* rho, T* v, int64_t max_sweeps, double tol, int zero_start,
                   int batch, int64_t* sweeps_done, double* last_maxdiff) {
    const int64_t n = grid_points(g);
    const size_t bytes = (size_t)n * sizeof(T);
    if (int rc = ensure_scratch(c, bytes)) return rc;
    T* tmp = (T*)c->scratch;
    if (int rc = poisson_prepare<T>(c, g, rho, v, tmp, zero_start)) return rc;
    Jacobi<T> J;
    jacobi_setup<T>(c, g, rho, J);
    const double ix2 = 1.0 / (g->dx * g->dx), iy2 = 1.0 / (g->dy * g->dy), iz2 = 1.0 / (g->dz * g->dz);
    const double rj = (ix2 * std::cos(kPi / (double)(g->nx - 1)) + iy2 * std::cos(kPi / (double)(g->ny - 1)) +
                       iz2 * std::cos(kPi / (double)(g->nz - 1))) /
                      (ix2 + iy2 + iz2);
    if (batch < 1) batch = 1;
    if (batch > (int)c->sweep_max_cap) batch = (int)c->sweep_max_cap;
    T* bufs[2] = {v, tmp};
    int curi = 0;
    int64_t done = 0;
    double md = 0.0, w = 1.0, best = INFINITY;
    std::vector<unsigned long long> hm(batch);
    const int nx = (int)g->nx, ny = (int)g->ny, nz = (int)g->nz;
    while (done < max_sweeps) {
        const int nb = (int)std::min<int64_t>(batch, max_sweeps - done);
        DFT_HIP(c, hipMemsetAsync(c->sweep_max, 0, sizeof(unsigned long long) * nb, c->stream));
        for (int s = 0; s < nb; ++s) {
            const int64_t k = done + s;  // this sweep makes v_{k+1}
            if (k == 0) {
                k_jacobi<T, 4, false><<<J.grid, 256, 0, c->stream>>>(nx, ny, nz, J.zc, J.co, rho, bufs[curi], bufs[curi ^ 1],
                                                                     c->sweep_max + s, T(1));
            } else {
                w = (k == 1) ? 1.0 / (1.0 - 0.5 * rj * rj) : 1.0 / (1.0 - 0.25 * rj * rj * w);
                k_jacobi<T, 4, true><<<J.grid, 256, 0, c->stream>>>(nx, ny, nz, J.zc, J.co, rho, bufs[curi], bufs[curi ^ 1],
                                                                    c->sweep_max + s, (T)w);
            }
            DFT_CHECK_LAUNCH(c);
            curi ^= 1;
        }
        DFT_HIP(c, hipMemcpyAsync(hm.data(), c->sweep_max, sizeof(unsigned long long) * nb, hipMemcpyDeviceToHost, c->stream));
        DFT_HIP(c, hipStreamSynchronize(c->stream));
        done += nb;
        bool stop = false;
        double bmin = INFINITY;
        for (int s = 0; s < nb; ++s) {
            double d;
            memcpy(&d, &hm[s], sizeof d);
            md = d;
            bmin = std::min(bmin, d);
            if (d < tol) {
                stop = true;
                break;
            }
        }
        if (stop) break;
        if (bmin >= 0.7 * best) break;  // stagnated at the rounding floor
        best = std::min(best, bmin);
    }
    if (curi != 0) DFT_HIP(c, hipMemcpyAsync(v, tmp, bytes, hipMemcpyDeviceToDevice, c->stream));
    DFT_HIP(c, hipStreamSynchronize(c->stream));
    if (sweeps_done) *sweeps_done = done;
    if (last_maxdiff) *last_maxdiff = md;
    return DFT_OK;
}

// ---- direct solve of the Jacobi fixed point -------------------------------
// The interior equations of the Jacobi iteration at convergence are
// Lap_h V = -4 pi rho with the boundary values fixed (HartreePotentialMolecule.cpp:
// the update V = (sum_a (V_a+ + V_a-)/h_a^2 + 4 pi rho) / (2 sum_a 1/h_a^2)).
// With zero boundary values moved to the right-hand side, the interior
// operator is separable and each 1D factor is diagonalised by the
// orthogonal DST-I matrix S[p][i] = sqrt(2/(n+1)) sin(pi (p+1)(i+1)/(n+1)),
// eigenvalue -4 sin^2(pi (p+1) / (2 (n+1))) / h^2.  So
//   V = S_x S_y S_z [ (S_x S_y S_z b) / (mu_x + mu_y + mu_z) ],
// each 1D transform a dense FP64 GEMM along one axis (rocBLAS, MFMA):
// 2 n flops per point per transform, 12 n per solve (n <= a few hundred).
__global__ __launch_bounds__(256) void k_dst_matrix(int n, double* __restrict__ S) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)n * n) return;
    const int p = (int)(e % n), i = (int)(e / n);
    const int64_t t = ((int64_t)(p + 1) * (i + 1)) % (2LL * (n + 1));  // exact argument reduction
    S[e] = sqrt(2.0 / (n + 1)) * sinpi((double)t / (double)(n + 1));
}

// b (interior points, x fastest) = -4 pi rho - (boundary neighbours) / h^2
template <typename T>
__global__ __launch_bounds__(256) void k_dst_rhs(int nx, int ny, int nz, double ix2, double iy2, double iz2,
                                                 const T* __restrict__ rho, const T* __restrict__ v,
                                                 double* __restrict__ b) {
    const int mx = nx - 2, my = ny - 2, mz = nz - 2;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)mx * my * mz) return;
    const int i = (int)(e % mx) + 1, j = (int)((e / mx) % my) + 1, k = (int)(e / ((int64_t)mx * my)) + 1;
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    const int64_t p = k * sz + j * sy + i;
    double r = -4.0 * kPi * (double)rho[p];
    if (i == 1) r -= (double)v[p - 1] * ix2;
    if (i == nx - 2) r -= (double)v[p + 1] * ix2;
    if (j == 1) r -= (double)v[p - sy] * iy2;
    if (j == ny - 2) r -= (double)v[p + sy] * iy2;
    if (k == 1) r -= (double)v[p - sz] * iz2;
    if (k == nz - 2) r -= (double)v[p + sz] * iz2;
    b[e] = r;
}

// divide the transformed right-hand side by the operator's eigenvalues
__global__ __launch_bounds__(256) void k_dst_scale(int mx, int my, int mz, double ix2, double iy2, double iz2,
                                                   double* __restrict__ B) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)mx * my * mz) return;
    const int p = (int)(e % mx), q = (int)((e / mx) % my), r = (int)(e / ((int64_t)mx * my));
    const double sx = sinpi((double)(p + 1) / (2.0 * (mx + 1)));
    const double sy = sinpi((double)(q + 1) / (2.0 * (my + 1)));
    const double sz = sinpi((double)(r + 1) / (2.0 * (mz + 1)));
    const double mu = -4.0 * (sx * sx * ix2 + sy * sy * iy2 + sz * sz * iz2);
    B[e] /= mu;
}

template <typename T>
__global__ __launch_bounds__(256) void k_dst_store(int nx, int ny, int nz, const double* __restrict__ U,
                                                   T* __restrict__ v) {
    const int mx = nx - 2, my = ny - 2, mz = nz - 2;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)mx * my * mz) return;
    const int i = (int)(e % mx) + 1, j = (int)((e / mx) % my) + 1, k = (int)(e / ((int64_t)mx * my)) + 1;
    v[((int64_t)k * ny + j) * nx + i] = (T)U[e];
}

// out = the 1D transform of in along x, y, z in turn (S symmetric, so the
// same three products apply the inverse)
int dst3(dft_ctx* c, rocblas_handle h, int mx, int my, int mz, const double* Sx, const double* Sy, const double* Sz,
         double* a, double* b) {
    const double one = 1.0, zero = 0.0;
    rocblas_status st;
    st = rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, mx, my * mz, mx, &one, Sx, mx, a, mx, &zero,
                       b, mx);
    if (st != rocblas_status_success) return set_err(c, DFT_EHIP, "dft_poisson_direct: x transform failed");
    st = rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, mx, my, my, &one, b, mx,
                                       (rocblas_stride)mx * my, Sy, my, 0, &zero, a, mx, (rocblas_stride)mx * my, mz);
    if (st != rocblas_status_success) return set_err(c, DFT_EHIP, "dft_poisson_direct: y transform failed");
    st = rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, mx * my, mz, mz, &one, a, mx * my, Sz, mz,
                       &zero, b, mx * my);
    if (st != rocblas_status_success) return set_err(c, DFT_EHIP, "dft_poisson_direct: z transform failed");
    return DFT_OK;
}

template <typename T>
int poisson_direct_t(dft_ctx* c, const dft_grid* g, const T* rho, T* v) {
    const int64_t n = grid_points(g);
    const int nx = (int)g->nx, ny = (int)g->ny, nz = (int)g->nz;
    // total charge and the boundary values (HartreePotentialMolecule.cpp:17-35 and the boundary rule)
    const int nblk = (int)std::min<int64_t>((n + 255) / 256, 4 * (int64_t)c->num_cus);
    if (int rc = ensure_partials(c, (size_t)nblk + 2)) return rc;
    k_charge<T><<<nblk, 256, 0, c->stream>>>(rho, n, c->partials);
    DFT_CHECK_LAUNCH(c);
    double q = 0.0;
    if (int rc = finish_sum(c, nblk, &q)) return rc;
    const double qd = q * (g->dx * g->dy * g->dz);
    k_boundary<T><<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(nx, ny, nz, g->xmin, g->xmax, g->ymin, g->ymax,
                                                                        g->zmin, g->zmax, g->dx, g->dy, g->dz, qd, v,
                                                                        (T*)nullptr);
    DFT_CHECK_LAUNCH(c);
    const int mx = nx - 2, my = ny - 2, mz = nz - 2;
    if (mx < 1 || my < 1 || mz < 1) return DFT_OK;  // no interior
    const int64_t ni = (int64_t)mx * my * mz;
    const size_t need = sizeof(double) * ((size_t)2 * ni + (size_t)mx * mx + (size_t)my * my + (size_t)mz * mz);
    if (int rc = ensure_scratch(c, need)) return rc;
    double* w1 = (double*)c->scratch;
    double* w2 = w1 + ni;
    double* Sx = w2 + ni;
    double* Sy = Sx + (size_t)mx * mx;
    double* Sz = Sy + (size_t)my * my;
    k_dst_matrix<<<(unsigned)(((int64_t)mx * mx + 255) / 256), 256, 0, c->stream>>>(mx, Sx);
    k_dst_matrix<<<(unsigned)(((int64_t)my * my + 255) / 256), 256, 0, c->stream>>>(my, Sy);
    k_dst_matrix<<<(unsigned)(((int64_t)mz * mz + 255) / 256), 256, 0, c->stream>>>(mz, Sz);
    const double ix2 = 1.0 / (g->dx * g->dx), iy2 = 1.0 / (g->dy * g->dy), iz2 = 1.0 / (g->dz * g->dz);
    const unsigned gb = (unsigned)((ni + 255) / 256);
    k_dst_rhs<T><<<gb, 256, 0, c->stream>>>(nx, ny, nz, ix2, iy2, iz2, rho, v, w1);
    DFT_CHECK_LAUNCH(c);
    rocblas_handle h;
    if (int rc = dft::rocblas_for(c, reinterpret_cast<void**>(&h))) return rc;
    if (int rc = dst3(c, h, mx, my, mz, Sx, Sy, Sz, w1, w2)) return rc;  // transform -> w2
    k_dst_scale<<<gb, 256, 0, c->stream>>>(mx, my, mz, ix2, iy2, iz2, w2);
    DFT_CHECK_LAUNCH(c);
    if (int rc = dst3(c, h, mx, my, mz, Sx, Sy, Sz, w2, w1)) return rc;  // inverse -> w1
    k_dst_store<T><<<gb, 256, 0, c->stream>>>(nx, ny, nz, w1, v);
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

}  // namespace

extern "C" {

int dft_poisson(dft_ctx* c, const dft_grid* g, int prec, const void* rho, void* v, int64_t max_sweeps, double tol,
                int zero_start, int sweeps_per_launch, int64_t* sweeps_done, double* last_maxdiff) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, valid_grid(g) && rho && v && max_sweeps >= 0, "dft_poisson: bad arguments");
    DFT_REQUIRE(c, g->nx * g->ny * g->nz < (1LL << 31), "dft_poisson: grid exceeds 2^31 points");
    if (prec == DFT_F64)
        return poisson_t<double>(c, g, (const double*)rho, (double*)v, max_sweeps, tol, zero_start, sweeps_per_launch,
                                 sweeps_done, last_maxdiff);
    if (prec == DFT_F32)
        return poisson_t<float>(c, g, (const float*)rho, (float*)v, max_sweeps, tol, zero_start, sweeps_per_launch,
                                sweeps_done, last_maxdiff);
    return set_err(c, DFT_EINVAL, "dft_poisson: bad precision");
}

int dft_poisson_direct(dft_ctx* c, const dft_grid* g, int prec, const void* rho, void* v) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, g && rho && v && g->nx >= 1 && g->ny >= 1 && g->nz >= 1 && g->nx <= 4096 && g->ny <= 4096 &&
                       g->nz <= 4096,
                "dft_poisson_direct: bad arguments");
    if (prec == DFT_F64) return poisson_direct_t<double>(c, g, (const double*)rho, (double*)v);
    if (prec == DFT_F32) return poisson_direct_t<float>(c, g, (const float*)rho, (float*)v);
    return set_err(c, DFT_EINVAL, "dft_poisson_direct: bad precision");
}

int dft_poisson_chebyshev(dft_ctx* c, const dft_grid* g, int prec, const void* rho, void* v, int64_t max_sweeps,
                          double tol, int zero_start, int batch, int64_t* sweeps_done, double* last_maxdiff) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, valid_grid(g) && rho && v && max_sweeps >= 0 && tol >= 0.0, "dft_poisson_chebyshev: bad arguments");
    DFT_REQUIRE(c, g->nx >= 3 && g->ny >= 3 && g->nz >= 3, "dft_poisson_chebyshev: grid needs interior points");
    DFT_REQUIRE(c, g->nx * g->ny * g->nz < (1LL << 31), "dft_poisson_chebyshev: grid exceeds 2^31 points");
    if (prec == DFT_F64)
        return poisson_cheb_t<double>(c, g, (const double*)rho, (double*)v, max_sweeps, tol, zero_start, batch, sweeps_done,
                                      last_maxdiff);
    if (prec == DFT_F32)
        return poisson_cheb_t<float>(c, g, (const float*)rho, (float*)v, max_sweeps, tol, zero_start, batch, sweeps_done,
                                     last_maxdiff);
    return set_err(c, DFT_EINVAL, "dft_poisson_chebyshev: bad precision");
}

int dft_hartree_energy(dft_ctx* c, const dft_grid* g, int prec, const void* rho, const void* vh, double* e) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, valid_grid(g) && rho && vh && e, "dft_hartree_energy: bad arguments");
    double s = 0.0;
    if (int rc = dft_reduce(c, DFT_RED_DOT, prec, rho, vh, grid_points(g), &s)) return rc;
    *e = 0.5 * s * (g->dx * g->dy * g->dz);  // HartreePotentialMolecule.cpp:375-378
    return DFT_OK;
}

}  // extern "C"
