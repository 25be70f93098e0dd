// kinetic_dst.hip -- the kinetic correction t = (-1/2 Lap_h + sigma)^{-1} r
// applied exactly (to BF16 rounding) by sine transforms.
//
// With zero ghosts beyond the grid (the H psi stencil's boundary rule) the
// operator -1/2 Lap_h + sigma on the nx x ny x nz grid is separable: along an
// axis of n points the orthogonal DST-I matrix S[p][i] = sqrt(2/(n+1))
// sin(pi (p+1)(i+1)/(n+1)) diagonalises -1/2 d^2/dx^2 with eigenvalues
// (1 - cos(pi (p+1)/(n+1))) / h^2.  So
//   t = S_x S_y S_z [ (S_x S_y S_z r) / (l_x + l_y + l_z + sigma) ],
// six 1D transforms, each a GEMM along one axis for all columns at once:
// BF16 operands on the matrix cores with FP32 accumulation (rocBLAS gemm_ex).
// A Davidson correction only needs a direction, so BF16's ~3 digits cost
// nothing there, and each element moves ~36 bytes in all instead of ~132 for
// eight Chebyshev-combined stencil steps.
#include "common.hpp"

#include <rocblas/rocblas.h>

#include <cmath>

using namespace dft;

namespace {

__global__ __launch_bounds__(256) void k_dst_matrix_bf16(int n, __bf16* __restrict__ S) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)n * n) return;
    const int p = (int)(e % n), i = (int)(e / n);
    const int64_t a = ((int64_t)(p + 1) * (i + 1)) % (2LL * (n + 1));  // exact argument reduction
    S[e] = (__bf16)(sqrt(2.0 / (n + 1)) * sinpi((double)a / (double)(n + 1)));
}

template <typename T>
__global__ __launch_bounds__(256) void k_cols_to_bf16(int64_t n, const T* __restrict__ r, int64_t ldr,
                                                      __bf16* __restrict__ a) {
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int o = blockIdx.y;
    if (p < n) a[(int64_t)o * n + p] = (__bf16)(float)r[(int64_t)o * ldr + p];
}

// B[i + nx (j + ny (k + nz o))] /= l_x(i) + l_y(j) + l_z(k) + sigma_o: one
// workgroup per (z plane k, column o), the x and y eigenvalues in LDS, so the
// pass is a plain stream of the block.  sigma_o = sigma, or with per-column
// shifts max(sigma, -scale theta_o).
__global__ __launch_bounds__(256) void k_kin_scale(int nx, int ny, int nz, float ix2, float iy2, float iz2,
                                                   float sigma, const double* __restrict__ theta, float scale,
                                                   __bf16* __restrict__ B) {
    __shared__ float lx[1024], ly[1024];
    const int k = blockIdx.x, o = blockIdx.y;
    for (int i = threadIdx.x; i < nx; i += 256) lx[i] = (1.0f - cospif((float)(i + 1) / (float)(nx + 1))) * ix2;
    for (int j = threadIdx.x; j < ny; j += 256) ly[j] = (1.0f - cospif((float)(j + 1) / (float)(ny + 1))) * iy2;
    const float so = theta ? fmaxf(sigma, -scale * (float)theta[o]) : sigma;
    const float lz = (1.0f - cospif((float)(k + 1) / (float)(nz + 1))) * iz2 + so;
    __syncthreads();
    __bf16* plane = B + ((int64_t)o * nz + k) * ((int64_t)nx * ny);
    const int np = nx * ny;
    for (int e = threadIdx.x; e < np; e += 256) {
        const int j = e / nx, i = e - j * nx;
        plane[e] = (__bf16)((float)plane[e] / (lx[i] + ly[j] + lz));
    }
}

__global__ __launch_bounds__(256) void k_f32_to_f64_cols(int64_t n, const float* __restrict__ s, double* __restrict__ t,
                                                         int64_t ldt) {
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int o = blockIdx.y;
    if (p < n) t[(int64_t)o * ldt + p] = (double)s[(int64_t)o * n + p];
}

// d = the three 1D transforms of a (x, then y, then z) for nb columns; the
// last product writes d_type into d with column stride ldd_col.
int dst3_bf16(dft_ctx* c, rocblas_handle h, int nx, int ny, int nz, int nb, const __bf16* Sx, const __bf16* Sy,
              const __bf16* Sz, __bf16* a, __bf16* b, void* d, rocblas_datatype d_type, int64_t ldd_col) {
    const float one = 1.0f, zero = 0.0f;
    const int64_t n = (int64_t)nx * ny * nz;
    const rocblas_datatype bf = rocblas_datatype_bf16_r, f32 = rocblas_datatype_f32_r;
    rocblas_status st;
    // x: b = S_x a, a viewed as nx x (ny nz nb)
    st = rocblas_gemm_ex(h, rocblas_operation_none, rocblas_operation_none, nx, (rocblas_int)(n / nx * nb), nx, &one, Sx,
                         bf, nx, a, bf, nx, &zero, b, bf, nx, b, bf, nx, f32, rocblas_gemm_algo_standard, 0, 0);
    if (st != rocblas_status_success) return set_err(c, DFT_EHIP, "kinetic transform (x) failed");
    // y: each (nx x ny) slab times S_y
    st = rocblas_gemm_strided_batched_ex(h, rocblas_operation_none, rocblas_operation_none, nx, ny, ny, &one, b, bf, nx,
                                         (rocblas_stride)nx * ny, Sy, bf, ny, 0, &zero, a, bf, nx,
                                         (rocblas_stride)nx * ny, a, bf, nx, (rocblas_stride)nx * ny, nz * nb, f32,
                                         rocblas_gemm_algo_standard, 0, 0);
    if (st != rocblas_status_success) return set_err(c, DFT_EHIP, "kinetic transform (y) failed");
    // z: each column's (nx ny x nz) matrix times S_z
    st = rocblas_gemm_strided_batched_ex(h, rocblas_operation_none, rocblas_operation_none, nx * ny, nz, nz, &one, a, bf,
                                         nx * ny, (rocblas_stride)n, Sz, bf, nz, 0, &zero, d, d_type, nx * ny,
                                         (rocblas_stride)ldd_col, d, d_type, nx * ny, (rocblas_stride)ldd_col, nb, f32,
                                         rocblas_gemm_algo_standard, 0, 0);
    if (st != rocblas_status_success) return set_err(c, DFT_EHIP, "kinetic transform (z) failed");
    return DFT_OK;
}

template <typename T>
int kinetic_exact_t(dft_ctx* c, const dft_grid* g, double sigma, const double* theta, double scale, const T* r,
                    int64_t ldr, int nb, T* t, int64_t ldt) {
    const int nx = (int)g->nx, ny = (int)g->ny, nz = (int)g->nz;
    const int64_t n = (int64_t)nx * ny * nz;
    const bool f64 = sizeof(T) == 8;
    auto al = [](size_t x) { return (x + 255) / 256 * 256; };
    const size_t sb = al(sizeof(__bf16) * (size_t)nb * n);
    const size_t sf = f64 ? al(sizeof(float) * (size_t)nb * n) : 0;
    if (int rc = ensure_scratch(c, 2 * sb + sf)) return rc;
    char* w = static_cast<char*>(c->scratch);
    __bf16* A = reinterpret_cast<__bf16*>(w);
    __bf16* B = reinterpret_cast<__bf16*>(w + sb);
    float* F = f64 ? reinterpret_cast<float*>(w + 2 * sb) : nullptr;
    // the three transform matrices, built once per grid shape and kept by the context
    const size_t sx = al(sizeof(__bf16) * (size_t)nx * nx), sy = al(sizeof(__bf16) * (size_t)ny * ny);
    const size_t need = sx + sy + al(sizeof(__bf16) * (size_t)nz * nz);
    if (c->dst_shape[0] != nx || c->dst_shape[1] != ny || c->dst_shape[2] != nz || !c->dst_mats) {
        if (c->dst_mats) {
            DFT_HIP(c, hipStreamSynchronize(c->stream));
            DFT_HIP(c, hipFree(c->dst_mats));
            c->dst_mats = nullptr;
        }
        DFT_HIP(c, hipMalloc(&c->dst_mats, need));
        char* m = static_cast<char*>(c->dst_mats);
        k_dst_matrix_bf16<<<(unsigned)(((int64_t)nx * nx + 255) / 256), 256, 0, c->stream>>>(nx, reinterpret_cast<__bf16*>(m));
        k_dst_matrix_bf16<<<(unsigned)(((int64_t)ny * ny + 255) / 256), 256, 0, c->stream>>>(
            ny, reinterpret_cast<__bf16*>(m + sx));
        k_dst_matrix_bf16<<<(unsigned)(((int64_t)nz * nz + 255) / 256), 256, 0, c->stream>>>(
            nz, reinterpret_cast<__bf16*>(m + sx + sy));
        DFT_CHECK_LAUNCH(c);
        c->dst_shape[0] = nx;
        c->dst_shape[1] = ny;
        c->dst_shape[2] = nz;
    }
    const __bf16* Sx = reinterpret_cast<const __bf16*>(c->dst_mats);
    const __bf16* Sy = reinterpret_cast<const __bf16*>(static_cast<char*>(c->dst_mats) + sx);
    const __bf16* Sz = reinterpret_cast<const __bf16*>(static_cast<char*>(c->dst_mats) + sx + sy);
    const dim3 gc((unsigned)((n + 255) / 256), (unsigned)nb);
    k_cols_to_bf16<T><<<gc, 256, 0, c->stream>>>(n, r, ldr, A);
    DFT_CHECK_LAUNCH(c);
    void* hv = nullptr;
    if (int rc = rocblas_for(c, &hv)) return rc;
    rocblas_handle h = static_cast<rocblas_handle>(hv);
    if (int rc = dst3_bf16(c, h, nx, ny, nz, nb, Sx, Sy, Sz, A, B, B, rocblas_datatype_bf16_r, n)) return rc;
    k_kin_scale<<<dim3((unsigned)nz, (unsigned)nb), 256, 0, c->stream>>>(
        nx, ny, nz, (float)(1.0 / (g->dx * g->dx)), (float)(1.0 / (g->dy * g->dy)), (float)(1.0 / (g->dz * g->dz)),
        (float)sigma, theta, (float)scale, B);
    DFT_CHECK_LAUNCH(c);
    if (!f64) return dst3_bf16(c, h, nx, ny, nz, nb, Sx, Sy, Sz, B, A, t, rocblas_datatype_f32_r, ldt);
    if (int rc = dst3_bf16(c, h, nx, ny, nz, nb, Sx, Sy, Sz, B, A, F, rocblas_datatype_f32_r, n)) return rc;
    k_f32_to_f64_cols<<<gc, 256, 0, c->stream>>>(n, F, reinterpret_cast<double*>(t), ldt);
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

}  // namespace

namespace {

int kinetic_exact(dft_ctx* c, const dft_grid* g, int prec, double sigma, const double* theta, double scale,
                         const void* r, int64_t ldr, int nb, void* t, int64_t ldt) {
    DFT_REQUIRE(c, valid_grid(g) && r && t && nb >= 0 && nb <= 65535 && sigma > 0.0 && ldr >= grid_points(g) &&
                       ldt >= grid_points(g),
                "dft_precond_kinetic_exact: bad arguments");
    const int64_t n = grid_points(g);
    DFT_REQUIRE(c, n * (int64_t)nb / g->nx < (1LL << 31) && g->nx * g->ny < (1LL << 31) && g->nx <= 1024 &&
                       g->ny <= 1024 && g->nz <= 65535,
                "dft_precond_kinetic_exact: grid or block too large for the transforms");
    if (!nb) return DFT_OK;
    if (prec == DFT_F64)
        return kinetic_exact_t<double>(c, g, sigma, theta, scale, (const double*)r, ldr, nb, (double*)t, ldt);
    if (prec == DFT_F32)
        return kinetic_exact_t<float>(c, g, sigma, theta, scale, (const float*)r, ldr, nb, (float*)t, ldt);
    return set_err(c, DFT_EINVAL, "dft_precond_kinetic_exact: bad precision");
}

}  // namespace

extern "C" {

int dft_precond_kinetic_exact(dft_ctx* c, const dft_grid* g, int prec, double sigma, const void* r, int64_t ldr, int nb,
                              void* t, int64_t ldt) {
    if (!c) return DFT_EINVAL;
    return kinetic_exact(c, g, prec, sigma, nullptr, 0.0, r, ldr, nb, t, ldt);
}

int dft_precond_kinetic_exact_shifted(dft_ctx* c, const dft_grid* g, int prec, double sigma_min, double scale,
                                      const double* theta, const void* r, int64_t ldr, int nb, void* t, int64_t ldt) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, theta && scale > 0.0, "dft_precond_kinetic_exact_shifted: bad arguments");
    return kinetic_exact(c, g, prec, sigma_min, theta, scale, r, ldr, nb, t, ldt);
}

}  // extern "C"
