// harvest.hip -- column-batched vector kernels for the projected CG solves
// that produce the learned preconditioner's training targets, and the
// patch-energy / patch-extraction kernels that cut training patches from
// full-grid residuals.
#include "common.hpp"

#include <algorithm>

using namespace dft;

namespace {

// out_j = x_j . y_j : grid (blocks, ncols); partials then a fixed-order final sum
template <typename T>
__global__ __launch_bounds__(256) void k_coldot_part(int64_t n, const T* __restrict__ x, int64_t ldx,
                                                     const T* __restrict__ y, int64_t ldy, double* __restrict__ part,
                                                     int nblk) {
    __shared__ double lds[4];
    const int j = blockIdx.y;
    const T* a = x + (int64_t)j * ldx;
    const T* b = y + (int64_t)j * ldy;
    double s = 0.0;
    for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < n; p += (int64_t)nblk * 256) s += (double)a[p] * (double)b[p];
    s = block_sum(s, lds);
    if (threadIdx.x == 0) part[(int64_t)j * nblk + blockIdx.x] = s;
}

__global__ __launch_bounds__(64) void k_coldot_final(const double* __restrict__ part, int nblk, double* __restrict__ out) {
    const int j = blockIdx.x;
    double s = 0.0;
    for (int i = threadIdx.x; i < nblk; i += 64) s += part[(int64_t)j * nblk + i];
    s = wave_sum(s);
    if (threadIdx.x == 0) out[j] = s;
}

// y_j = a_j x_j + b_j y_j
template <typename T>
__global__ __launch_bounds__(256) void k_cols_axpby(int64_t n, const double* __restrict__ a, const T* __restrict__ x,
                                                    int64_t ldx, const double* __restrict__ b, T* __restrict__ y,
                                                    int64_t ldy) {
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int j = blockIdx.y;
    if (p >= n) return;
    const double v = a[j] * (double)x[(int64_t)j * ldx + p] + b[j] * (double)y[(int64_t)j * ldy + p];
    y[(int64_t)j * ldy + p] = (T)v;
}

// energy of column j in the aligned P^3 block (bx, by, bz): one workgroup per (block, column)
template <typename T>
__global__ __launch_bounds__(256) void k_patch_energy(int nx, int ny, int nz, int P, int nbx, int nby,
                                                      const T* __restrict__ R, int64_t ldr, double* __restrict__ out) {
    __shared__ double lds[4];
    const int blk = blockIdx.x, j = blockIdx.y;
    const int bx = blk % nbx, by = (blk / nbx) % nby, bz = blk / (nbx * nby);
    const T* r = R + (int64_t)j * ldr;
    double s = 0.0;
    const int64_t P3 = (int64_t)P * P * P;
    for (int64_t q = threadIdx.x; q < P3; q += 256) {
        const int x = bx * P + (int)(q % P), y = by * P + (int)((q / P) % P), z = bz * P + (int)(q / ((int64_t)P * P));
        const double v = (double)r[((int64_t)z * ny + y) * nx + x];
        s += v * v;
    }
    s = block_sum(s, lds);
    if (threadIdx.x == 0) out[(int64_t)j * gridDim.x + blk] = s;
}

// Training patches: desc[p] = {column, x0, y0, z0}; in[p] = {r/s, (V - theta) h2}, tgt[p] = c / cs
// (s = max|r|, cs = max|c| of the column: the Davidson step depends only on
// the direction of each correction column, so both are normalised per column)
template <typename T>
__global__ __launch_bounds__(256) void k_extract(int nx, int ny, int P, const int4* __restrict__ desc,
                                                 const T* __restrict__ V, const T* __restrict__ R, int64_t ldr,
                                                 const T* __restrict__ C, int64_t ldc, const double* __restrict__ theta,
                                                 const double* __restrict__ scale, const double* __restrict__ cscale,
                                                 double h2, float* __restrict__ in, float* __restrict__ tgt) {
    const int64_t P3 = (int64_t)P * P * P;
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int p = blockIdx.y;
    if (q >= P3) return;
    const int4 d = desc[p];
    const int x = d.y + (int)(q % P), y = d.z + (int)((q / P) % P), z = d.w + (int)(q / ((int64_t)P * P));
    const int64_t v = ((int64_t)z * ny + y) * nx + x;
    const double s = scale[d.x];
    const double inv = s > 0.0 ? 1.0 / s : 0.0;
    in[((int64_t)p * 2) * P3 + q] = (float)((double)R[(int64_t)d.x * ldr + v] * inv);
    in[((int64_t)p * 2 + 1) * P3 + q] = (float)(((double)V[v] - theta[d.x]) * h2);
    const double cs = cscale[d.x];
    tgt[(int64_t)p * P3 + q] = (float)((double)C[(int64_t)d.x * ldc + v] * (cs > 0.0 ? 1.0 / cs : 0.0));
}

}  // namespace

extern "C" {

int dft_column_dots(dft_ctx* c, int prec, int64_t n, int nb, const void* x, int64_t ldx, const void* y, int64_t ldy,
                    double* out) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, x && y && out && n > 0 && nb >= 0 && nb <= 65535 && ldx >= n && ldy >= n, "dft_column_dots: bad arguments");
    if (!nb) return DFT_OK;
    const int nblk = (int)std::min<int64_t>((n + 255) / 256, std::max<int64_t>(8, (4 * (int64_t)c->num_cus + nb - 1) / nb));
    if (int rc = ensure_partials(c, (size_t)nblk * nb + 8)) return rc;
    dim3 g(nblk, nb);
    if (prec == DFT_F64)
        k_coldot_part<double><<<g, 256, 0, c->stream>>>(n, (const double*)x, ldx, (const double*)y, ldy, c->partials, nblk);
    else if (prec == DFT_F32)
        k_coldot_part<float><<<g, 256, 0, c->stream>>>(n, (const float*)x, ldx, (const float*)y, ldy, c->partials, nblk);
    else
        return set_err(c, DFT_EINVAL, "dft_column_dots: bad precision");
    DFT_CHECK_LAUNCH(c);
    k_coldot_final<<<nb, 64, 0, c->stream>>>(c->partials, nblk, out);
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

int dft_cols_axpby(dft_ctx* c, int prec, int64_t n, int nb, const double* a, const void* x, int64_t ldx, const double* b,
                   void* y, int64_t ldy) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, a && x && b && y && n > 0 && nb >= 0 && nb <= 65535 && ldx >= n && ldy >= n, "dft_cols_axpby: bad arguments");
    if (!nb) return DFT_OK;
    dim3 g((unsigned)((n + 255) / 256), nb);
    if (prec == DFT_F64)
        k_cols_axpby<double><<<g, 256, 0, c->stream>>>(n, a, (const double*)x, ldx, b, (double*)y, ldy);
    else if (prec == DFT_F32)
        k_cols_axpby<float><<<g, 256, 0, c->stream>>>(n, a, (const float*)x, ldx, b, (float*)y, ldy);
    else
        return set_err(c, DFT_EINVAL, "dft_cols_axpby: bad precision");
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

int dft_patch_energy(dft_ctx* c, const dft_grid* g, int prec, int patch, const void* R, int64_t ldr, int ncols, double* out) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, valid_grid(g) && R && out && patch >= 1 && ncols >= 1 && ncols <= 65535 && ldr >= grid_points(g),
                "dft_patch_energy: bad arguments");
    const int nbx = (int)(g->nx / patch), nby = (int)(g->ny / patch), nbz = (int)(g->nz / patch);
    DFT_REQUIRE(c, nbx >= 1 && nby >= 1 && nbz >= 1, "dft_patch_energy: patch larger than the grid");
    const dim3 grid((unsigned)(nbx * nby * nbz), ncols);
    if (prec == DFT_F64)
        k_patch_energy<double><<<grid, 256, 0, c->stream>>>((int)g->nx, (int)g->ny, (int)g->nz, patch, nbx, nby,
                                                            (const double*)R, ldr, out);
    else if (prec == DFT_F32)
        k_patch_energy<float><<<grid, 256, 0, c->stream>>>((int)g->nx, (int)g->ny, (int)g->nz, patch, nbx, nby,
                                                           (const float*)R, ldr, out);
    else
        return set_err(c, DFT_EINVAL, "dft_patch_energy: bad precision");
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

int dft_extract_patches(dft_ctx* c, const dft_grid* g, int prec, int patch, int npatch, const int* desc, int ncols,
                        const void* veff, const void* R, int64_t ldr, const void* C, int64_t ldc, const double* theta_dev,
                        const double* scale_dev, const double* cscale_dev, float* in, float* tgt) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, valid_grid(g) && desc && veff && R && C && theta_dev && scale_dev && cscale_dev && in && tgt && patch >= 1 &&
                       npatch >= 0 && npatch <= 65535 && ldr >= grid_points(g) && ldc >= grid_points(g) && patch <= g->nx && patch <= g->ny && patch <= g->nz,
                "dft_extract_patches: bad arguments");
    if (!npatch) return DFT_OK;
    for (int p = 0; p < npatch; ++p) {
        const int* d = desc + 4 * p;
        DFT_REQUIRE(c, d[0] >= 0 && d[0] < ncols && d[1] >= 0 && d[2] >= 0 && d[3] >= 0 && d[1] + patch <= g->nx && d[2] + patch <= g->ny &&
                           d[3] + patch <= g->nz,
                    "dft_extract_patches: patch outside the grid");
    }
    // descriptors to the device (the caller may reuse its array on return)
    if (int rc = ensure_partials(c, (size_t)npatch * 2)) return rc;
    int* desc_dev = reinterpret_cast<int*>(c->partials);
    DFT_HIP(c, hipMemcpyAsync(desc_dev, desc, sizeof(int) * 4 * (size_t)npatch, hipMemcpyHostToDevice, c->stream));
    DFT_HIP(c, hipStreamSynchronize(c->stream));
    const int64_t P3 = (int64_t)patch * patch * patch;
    const double h2 = 1.0 / (1.0 / (g->dx * g->dx) + 1.0 / (g->dy * g->dy) + 1.0 / (g->dz * g->dz));
    const dim3 grid((unsigned)((P3 + 255) / 256), npatch);
    if (prec == DFT_F64)
        k_extract<double><<<grid, 256, 0, c->stream>>>((int)g->nx, (int)g->ny, patch, (const int4*)desc_dev,
                                                       (const double*)veff, (const double*)R, ldr, (const double*)C, ldc,
                                                       theta_dev, scale_dev, cscale_dev, h2, in, tgt);
    else if (prec == DFT_F32)
        k_extract<float><<<grid, 256, 0, c->stream>>>((int)g->nx, (int)g->ny, patch, (const int4*)desc_dev,
                                                      (const float*)veff, (const float*)R, ldr, (const float*)C, ldc,
                                                      theta_dev, scale_dev, cscale_dev, h2, in, tgt);
    else
        return set_err(c, DFT_EINVAL, "dft_extract_patches: bad precision");
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

}  // extern "C"
