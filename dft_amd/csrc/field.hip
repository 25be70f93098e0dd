// field.hip -- (d) density accumulation and per-column orbital-block helpers,
// plus the one-time nuclear potential.
#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

using namespace dft;

namespace {

// n = (acc ? n : 0) + sum_o occ_o psi_o^2, orbitals in order
// (ElectronDensityMolecule.cpp:62-73: density += e * psi * psi).
template <typename T>
__global__ __launch_bounds__(256) void k_density(int64_t n, const T* __restrict__ psi, int64_t ld, int nb,
                                                 const double* __restrict__ occ, T* __restrict__ out, int acc) {
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
    T s = acc ? out[p] : T(0);
    for (int o = 0; o < nb; ++o) {
        const T v = psi[(int64_t)o * ld + p];
        const T e = (T)occ[o];
        if (e > T(0)) s += e * v * v;
    }
    out[p] = s;
}

// NuclearPotential.cpp:13-31,36-76
template <typename T>
__global__ __launch_bounds__(256) void k_vnuc(int nx, int ny, int nz, double xmin, double ymin, double zmin, double dx,
                                              double dy, double dz, int natoms, const double* __restrict__ nuc,
                                              T* __restrict__ v, int* __restrict__ singular) {
    const int64_t n = (int64_t)nx * ny * nz;
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
    const int i = (int)(p % nx), j = (int)((p / nx) % ny), k = (int)(p / ((int64_t)nx * ny));
    const double x = xmin + (double)i * dx, y = ymin + (double)j * dy, z = zmin + (double)k * dz;
    double s = 0.0;
    for (int a = 0; a < natoms; ++a) {
        const double ddx = x - nuc[4 * a + 1], ddy = y - nuc[4 * a + 2], ddz = z - nuc[4 * a + 3];
        const double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
        if (d2 == 0.0) { atomicOr(singular, 1); continue; }
        s -= nuc[4 * a] / sqrt(d2);
    }
    v[p] = (T)s;
}

template <typename T>
__global__ __launch_bounds__(256) void k_scale_cols(int64_t n, const T* __restrict__ x, int64_t ldx,
                                                    const double* __restrict__ s, T* __restrict__ y, int64_t ldy) {
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int j = blockIdx.y;
    if (p < n) y[(int64_t)j * ldy + p] = (T)((double)x[(int64_t)j * ldx + p] * s[j]);
}

template <typename T>
__global__ __launch_bounds__(256) void k_residual(int64_t n, const T* __restrict__ hx, int64_t ldh,
                                                  const T* __restrict__ x, int64_t ldx,
                                                  const double* __restrict__ th, T* __restrict__ r, int64_t ldr) {
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int j = blockIdx.y;
    if (p < n) r[(int64_t)j * ldr + p] = hx[(int64_t)j * ldh + p] - (T)th[j] * x[(int64_t)j * ldx + p];
}

// column norms^2: grid (blocks, nb); partial sums then a per-column final sum
template <typename T>
__global__ __launch_bounds__(256) void k_colnorm_part(int64_t n, const T* __restrict__ x, int64_t ldx,
                                                      double* __restrict__ part, int nblk) {
    __shared__ double lds[4];
    const int j = blockIdx.y;
    const T* c = x + (int64_t)j * ldx;
    double s = 0.0;
    for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < n; p += (int64_t)nblk * 256) {
        const double v = (double)c[p];
        s += v * v;
    }
    s = block_sum(s, lds);
    if (threadIdx.x == 0) part[(int64_t)j * nblk + blockIdx.x] = s;
}

__global__ __launch_bounds__(64) void k_colnorm_final(const double* __restrict__ part, int nblk, double* __restrict__ out) {
    const int j = blockIdx.x;
    double s = 0.0;
    for (int i = threadIdx.x; i < nblk; i += 64) s += part[(int64_t)j * nblk + i];
    s = wave_sum(s);
    if (threadIdx.x == 0) out[j] = s;
}

// first: t = c r;  else t = t + c (r - kt)   (per column, FP64 arithmetic)
template <typename T>
__global__ __launch_bounds__(256) void k_smooth_step(int64_t n, double cf, const T* __restrict__ r, int64_t ldr,
                                                     const T* __restrict__ kt, int64_t ldk, T* __restrict__ t,
                                                     int64_t ldt, int first) {
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int j = blockIdx.y;
    if (p >= n) return;
    const double rv = (double)r[(int64_t)j * ldr + p];
    T* tp = t + (int64_t)j * ldt + p;
    *tp = first ? (T)(cf * rv) : (T)((double)*tp + cf * (rv - (double)kt[(int64_t)j * ldk + p]));
}

template <typename T>
__global__ __launch_bounds__(256) void k_fill(int64_t n, double v, T* __restrict__ x) {
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p < n) x[p] = (T)v;
}

}  // namespace

extern "C" {

int dft_smooth_step(dft_ctx* c, int prec, int64_t n, int nb, double cf, const void* r, int64_t ldr, const void* kt,
                    int64_t ldk, void* t, int64_t ldt, int first) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, r && t && (first || kt) && n > 0 && nb >= 0 && nb <= 65535 && ldr >= n && ldt >= n && (first || ldk >= n),
                "dft_smooth_step: bad arguments");
    if (!nb) return DFT_OK;
    dim3 g((unsigned)((n + 255) / 256), nb);
    if (prec == DFT_F64)
        k_smooth_step<double><<<g, 256, 0, c->stream>>>(n, cf, (const double*)r, ldr, (const double*)kt, ldk, (double*)t, ldt, first);
    else if (prec == DFT_F32)
        k_smooth_step<float><<<g, 256, 0, c->stream>>>(n, cf, (const float*)r, ldr, (const float*)kt, ldk, (float*)t, ldt, first);
    else
        return set_err(c, DFT_EINVAL, "dft_smooth_step: bad precision");
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

int dft_fill(dft_ctx* c, int prec, int64_t n, double value, void* x) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, x && n > 0, "dft_fill: bad arguments");
    const unsigned g = (unsigned)((n + 255) / 256);
    if (prec == DFT_F64) k_fill<double><<<g, 256, 0, c->stream>>>(n, value, (double*)x);
    else if (prec == DFT_F32) k_fill<float><<<g, 256, 0, c->stream>>>(n, value, (float*)x);
    else return set_err(c, DFT_EINVAL, "dft_fill: bad precision");
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

int dft_density(dft_ctx* c, int prec, const void* psi, int64_t ld, int nb, const double* occ_host, int64_t n, void* out,
                int accumulate) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, psi && out && n > 0 && ld >= n && nb >= 0 && (nb == 0 || occ_host), "dft_density: bad arguments");
    if (int rc = ensure_partials(c, (size_t)nb + 8)) return rc;
    double* occ = c->partials;
    if (nb) DFT_HIP(c, hipMemcpyAsync(occ, occ_host, sizeof(double) * nb, hipMemcpyHostToDevice, c->stream));
    // the host array may be reused by the caller as soon as we return
    DFT_HIP(c, hipStreamSynchronize(c->stream));
    const unsigned g = (unsigned)((n + 255) / 256);
    if (prec == DFT_F64) k_density<double><<<g, 256, 0, c->stream>>>(n, (const double*)psi, ld, nb, occ, (double*)out, accumulate);
    else k_density<float><<<g, 256, 0, c->stream>>>(n, (const float*)psi, ld, nb, occ, (float*)out, accumulate);
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

int dft_nuclear_potential(dft_ctx* c, const dft_grid* g, int prec, int natoms, const double* nuclei_host, void* v) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, valid_grid(g) && v && natoms >= 0 && (natoms == 0 || nuclei_host), "dft_nuclear_potential: bad arguments");
    const int64_t n = grid_points(g);
    if (int rc = ensure_partials(c, (size_t)4 * natoms + 16)) return rc;
    double* nuc = c->partials;
    int* flag = (int*)(c->partials + 4 * natoms + 8);
    if (natoms) DFT_HIP(c, hipMemcpyAsync(nuc, nuclei_host, sizeof(double) * 4 * natoms, hipMemcpyHostToDevice, c->stream));
    DFT_HIP(c, hipMemsetAsync(flag, 0, sizeof(int), c->stream));
    const unsigned gb = (unsigned)((n + 255) / 256);
    if (prec == DFT_F64)
        k_vnuc<double><<<gb, 256, 0, c->stream>>>((int)g->nx, (int)g->ny, (int)g->nz, g->xmin, g->ymin, g->zmin, g->dx, g->dy, g->dz, natoms, nuc, (double*)v, flag);
    else
        k_vnuc<float><<<gb, 256, 0, c->stream>>>((int)g->nx, (int)g->ny, (int)g->nz, g->xmin, g->ymin, g->zmin, g->dx, g->dy, g->dz, natoms, nuc, (float*)v, flag);
    DFT_CHECK_LAUNCH(c);
    int h = 0;
    DFT_HIP(c, hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    DFT_HIP(c, hipStreamSynchronize(c->stream));
    if (h) return set_err(c, DFT_ENUMERIC, "NuclearPotential is singular at a nuclear position.");
    return DFT_OK;
}

int dft_scale_columns(dft_ctx* c, int prec, int64_t n, int nb, const void* x, int64_t ldx, const double* s, void* y,
                      int64_t ldy) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, x && y && s && n > 0 && nb >= 0 && nb <= 65535 && ldx >= n && ldy >= n, "dft_scale_columns: bad arguments");
    if (!nb) return DFT_OK;
    dim3 g((unsigned)((n + 255) / 256), nb);
    if (prec == DFT_F64) k_scale_cols<double><<<g, 256, 0, c->stream>>>(n, (const double*)x, ldx, s, (double*)y, ldy);
    else k_scale_cols<float><<<g, 256, 0, c->stream>>>(n, (const float*)x, ldx, s, (float*)y, ldy);
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

int dft_residual(dft_ctx* c, int prec, int64_t n, int nb, const void* hx, int64_t ldh, const void* x, int64_t ldx,
                 const double* th, void* r, int64_t ldr) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, hx && x && th && r && n > 0 && nb >= 0 && nb <= 65535, "dft_residual: bad arguments");
    if (!nb) return DFT_OK;
    dim3 g((unsigned)((n + 255) / 256), nb);
    if (prec == DFT_F64) k_residual<double><<<g, 256, 0, c->stream>>>(n, (const double*)hx, ldh, (const double*)x, ldx, th, (double*)r, ldr);
    else k_residual<float><<<g, 256, 0, c->stream>>>(n, (const float*)hx, ldh, (const float*)x, ldx, th, (float*)r, ldr);
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

int dft_column_norms2(dft_ctx* c, int prec, int64_t n, int nb, const void* x, int64_t ldx, double* out) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, x && out && n > 0 && nb >= 0 && nb <= 65535 && ldx >= n, "dft_column_norms2: bad arguments");
    if (!nb) return DFT_OK;
    const int nblk = (int)std::min<int64_t>((n + 255) / 256, std::max<int64_t>(8, (4 * (int64_t)c->num_cus + nb - 1) / nb));
    if (int rc = ensure_partials(c, (size_t)nblk * nb + 8)) return rc;
    dim3 g(nblk, nb);
    if (prec == DFT_F64) k_colnorm_part<double><<<g, 256, 0, c->stream>>>(n, (const double*)x, ldx, c->partials, nblk);
    else k_colnorm_part<float><<<g, 256, 0, c->stream>>>(n, (const float*)x, ldx, c->partials, nblk);
    DFT_CHECK_LAUNCH(c);
    k_colnorm_final<<<nb, 64, 0, c->stream>>>(c->partials, nblk, out);
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

}  // extern "C"
