// poisson.hip -- (b) Jacobi Poisson solver for V_H with the reference's
// boundary rule and stop rule.
//
// One sweep = one launch of k_jacobi: the same z-streaming layout as the
// H psi kernel (lane = x with DPP wave shifts for x-neighbours, TY y-rows
// per lane, z-1/z/z+1 planes in registers), updating interior points only;
// boundary points keep their values in both ping-pong buffers.  Each sweep
// reduces max|V_new - V_old| over its workgroup and atomically maxes it into
// its own slot, so a batch of sweeps runs without host round trips and the
// host finds the exact sweep at which the reference would have stopped.
#include "common.hpp"

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

template <typename T, int TY>
__global__ __launch_bounds__(256) void k_jacobi(int nx, int ny, int nz, int zc, JCoef<T> co, const T* __restrict__ rho,
                                                const T* __restrict__ cur_in, T* __restrict__ out,
                                                unsigned long long* __restrict__ slot) {
    __shared__ double red[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int x0 = blockIdx.x * 64, x = x0 + lane;
    const int ybase = blockIdx.y * (4 * TY) + w * TY;
    const int k0 = blockIdx.z * zc, k1 = min(nz, k0 + zc);
    const int64_t plane = (int64_t)nx * ny;
    const bool xin = x < nx;
    const bool xint = x > 0 && x + 1 < nx;
    const bool has_xm = x0 > 0, has_xp = x0 + 64 < nx;
    bool yin[TY], yint[TY];
    int64_t rowoff[TY];
#pragma unroll
    for (int r = 0; r < TY; ++r) {
        yin[r] = (ybase + r) < ny;
        yint[r] = (ybase + r) > 0 && (ybase + r) + 1 < ny;
        rowoff[r] = (int64_t)(ybase + r) * nx;
    }
    const bool ym_in = ybase > 0 && ybase - 1 < ny, yp_in = ybase + TY < ny;
    const int64_t ymoff = (int64_t)(ybase - 1) * nx, ypoff = (int64_t)(ybase + TY) * nx;
    T prev[TY], cur[TY], nxt[TY];
#pragma unroll
    for (int r = 0; r < TY; ++r) {
        const bool ok = xin && yin[r];
        prev[r] = (ok && k0 > 0) ? cur_in[(int64_t)(k0 - 1) * plane + rowoff[r] + x] : T(0);
        cur[r] = ok ? cur_in[(int64_t)k0 * plane + rowoff[r] + x] : T(0);
    }
    T md = T(0);
    for (int k = k0; k < k1; ++k) {
        const int64_t pk = (int64_t)k * plane;
        const bool has_next = k + 1 < nz;
        const bool kint = k > 0 && k + 1 < nz;
        const T* p = cur_in + pk;
#pragma unroll
        for (int r = 0; r < TY; ++r) nxt[r] = (xin && yin[r] && has_next) ? p[plane + rowoff[r] + x] : T(0);
        const T ym = (xin && ym_in) ? p[ymoff + x] : T(0);
        const T yp = (xin && yp_in) ? p[ypoff + x] : T(0);
#pragma unroll
        for (int r = 0; r < TY; ++r) {
            T hm = T(0), hp = T(0);
            if (lane == 0 && has_xm && yin[r]) hm = p[rowoff[r] + x0 - 1];
            if (lane == 63 && has_xp && yin[r]) hp = p[rowoff[r] + x0 + 64];
            const T c = cur[r];
            const T xm = shr1(c, hm);
            const T xp = shl1(c, hp);
            const T ymv = r > 0 ? cur[r > 0 ? r - 1 : 0] : ym;
            const T ypv = r < TY - 1 ? cur[r < TY - 1 ? r + 1 : 0] : yp;
            if (xint && yint[r] && kint) {
                // HartreePotentialMolecule.cpp:295-322
                const T src = co.fourpi * rho[pk + rowoff[r] + x];
                const T nv = (co.cx * (xm + xp) + co.cy * (ymv + ypv) + co.cz * (prev[r] + nxt[r]) + src) / co.den;
                const T d = nv > c ? nv - c : c - nv;
                md = md < d ? d : md;
                out[pk + rowoff[r] + x] = nv;
            }
        }
#pragma unroll
        for (int r = 0; r < TY; ++r) {
            prev[r] = cur[r];
            cur[r] = nxt[r];
        }
    }
    double m = wave_max((double)md);
    if (lane == 0) red[w] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = red[0];
        for (int i = 1; i < 4; ++i) b = b < red[i] ? red[i] : b;
        if (b > 0.0) atomicMax(slot, dbits(b));
    }
}

// Boundary values (HartreePotentialMolecule.cpp:37-95, 171-202): 0 on a
// face, else Q/|r| measured from the origin.  Written to both buffers.
template <typename T>
__global__ __launch_bounds__(256) void k_boundary(int nx, int ny, int nz, double xmin, double xmax, double ymin,
                                                  double ymax, double zmin, double zmax, double dx, double dy,
                                                  double dz, double q, T* __restrict__ a,
                                                  T* __restrict__ b) {
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
    b[p] = (T)val;
}

template <typename T>
__global__ __launch_bounds__(256) void k_charge(const T* __restrict__ rho, int64_t n, double* partials) {
    __shared__ double lds[4];
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) s += (double)rho[i];
    s = block_sum(s, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

template <typename T>
struct Jacobi {
    dft_ctx* c;
    const dft_grid* g;
    JCoef<T> co;
    dim3 grid;
    int zc;
    const T* rho;

    int sweep(const T* in, T* out, unsigned long long* slot) {
        k_jacobi<T, 4><<<grid, 256, 0, c->stream>>>((int)g->nx, (int)g->ny, (int)g->nz, zc, co, rho, in, out, slot);
        DFT_CHECK_LAUNCH(c);
        return DFT_OK;
    }
};

template <typename T>
int poisson_t(dft_ctx* c, const dft_grid* g, const T* rho, T* v, int64_t max_sweeps, double tol, int zero_start,
              int batch, int64_t* sweeps_done, double* last_maxdiff) {
    const int64_t n = grid_points(g);
    const size_t bytes = (size_t)n * sizeof(T);
    const bool exact = tol > 0.0;
    if (int rc = ensure_scratch(c, bytes * (exact ? 2 : 1))) return rc;
    T* tmp = (T*)c->scratch;
    T* snap = exact ? (T*)((char*)c->scratch + bytes) : nullptr;
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

    Jacobi<T> J;
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

    if (batch < 1) batch = 1;
    if (batch > (int)c->sweep_max_cap) batch = (int)c->sweep_max_cap;
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
        for (int s = 0; s < nb; ++s) {
            if (int rc = J.sweep(bufs[curi], bufs[curi ^ 1], c->sweep_max + s)) return rc;
            curi ^= 1;
        }
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
            for (int s = 0; s <= stop; ++s) {
                if (int rc = J.sweep(bufs[curi], bufs[curi ^ 1], c->sweep_max)) return rc;
                curi ^= 1;
            }
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

}  // namespace

extern "C" {

int dft_poisson(dft_ctx* c, const dft_grid* g, int prec, const void* rho, void* v, int64_t max_sweeps, double tol,
                int zero_start, int sweeps_per_launch, int64_t* sweeps_done, double* last_maxdiff) {
    DFT_REQUIRE(c, valid_grid(g) && rho && v && max_sweeps >= 0, "dft_poisson: bad arguments");
    DFT_REQUIRE(c, (g->nz + 0) <= 65535LL * 8 && g->ny < (1 << 30), "dft_poisson: grid too large");
    if (prec == DFT_F64)
        return poisson_t<double>(c, g, (const double*)rho, (double*)v, max_sweeps, tol, zero_start, sweeps_per_launch,
                                 sweeps_done, last_maxdiff);
    if (prec == DFT_F32)
        return poisson_t<float>(c, g, (const float*)rho, (float*)v, max_sweeps, tol, zero_start, sweeps_per_launch,
                                sweeps_done, last_maxdiff);
    return set_err(c, DFT_EINVAL, "dft_poisson: bad precision");
}

int dft_hartree_energy(dft_ctx* c, const dft_grid* g, int prec, const void* rho, const void* vh, double* e) {
    DFT_REQUIRE(c, valid_grid(g) && rho && vh && e, "dft_hartree_energy: bad arguments");
    double s = 0.0;
    if (int rc = dft_reduce(c, DFT_RED_DOT, prec, rho, vh, grid_points(g), &s)) return rc;
    *e = 0.5 * s * (g->dx * g->dy * g->dz);  // HartreePotentialMolecule.cpp:375-378
    return DFT_OK;
}

}  // extern "C"
