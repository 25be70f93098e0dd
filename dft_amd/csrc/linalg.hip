// linalg.hip -- dense FP64 subspace algebra on the device for the block
// Davidson (Rayleigh-Ritz eigensolve and Cholesky orthonormalisation of
// m x m matrices, m <= a few hundred), through rocSOLVER on the context
// stream.
#include "common.hpp"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

using namespace dft;

namespace {
__global__ __launch_bounds__(256) void k_to_f32(int m, int k, const double* __restrict__ c, int64_t ldc, float* __restrict__ o) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)m * k) return;
    const int i = (int)(e % m), j = (int)(e / m);
    o[e] = (float)c[(int64_t)j * ldc + i];
}
// a (upper, i <= j) <- tmp(j, i) from the lower triangle of tmp; zeros below
__global__ __launch_bounds__(256) void k_lower_t(int m, const double* __restrict__ tmp, double* __restrict__ a, int64_t lda) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)m * m) return;
    const int i = (int)(e % m), j = (int)(e / m);
    a[(int64_t)j * lda + i] = i <= j ? tmp[(int64_t)i * m + j] : 0.0;
}
__global__ __launch_bounds__(256) void k_copy_sq(int m, const double* __restrict__ a, int64_t lda, double* __restrict__ t) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)m * m) return;
    const int i = (int)(e % m), j = (int)(e / m);
    t[(int64_t)j * m + i] = a[(int64_t)j * lda + i];
}
// ---- tall-skinny block update on the VALU --------------------------------
// Y = alpha * A C + beta * Y with A: N x m (N grid points), C: m x k small.
// Each thread owns two grid points and KMAX output columns in registers; the
// m x k coefficients are read with wave-uniform addresses (scalar loads), so
// every loaded A value feeds KMAX FMAs with no LDS traffic.  Columns beyond
// KMAX run in further passes over A.  Used for FP32 blocks with k <= 64,
// where rocBLAS' SGEMM tiles waste most of their width.
template <typename T, int KMAX>
__global__ __launch_bounds__(256) void k_update_valu(int64_t n, int m, int kc, T alpha, const T* __restrict__ A,
                                                     int64_t lda, const T* __restrict__ Cm, int ldc, T beta,
                                                     T* __restrict__ Y, int64_t ldy) {
    const int64_t p0 = (int64_t)blockIdx.x * 512 + threadIdx.x, p1 = p0 + 256;
    const bool ok0 = p0 < n, ok1 = p1 < n;
    T acc0[KMAX], acc1[KMAX];
#pragma unroll
    for (int j = 0; j < KMAX; ++j) {
        acc0[j] = T(0);
        acc1[j] = T(0);
    }
    for (int i = 0; i < m; ++i) {
        const T* a = A + (int64_t)i * lda;
        const T v0 = ok0 ? a[p0] : T(0);
        const T v1 = ok1 ? a[p1] : T(0);
        const T* crow = Cm + i;  // C[i][j] at crow[j * ldc]
#pragma unroll
        for (int j = 0; j < KMAX; ++j) {
            const T cj = j < kc ? crow[(int64_t)j * ldc] : T(0);
            acc0[j] = fma(v0, cj, acc0[j]);
            acc1[j] = fma(v1, cj, acc1[j]);
        }
    }
#pragma unroll
    for (int j = 0; j < KMAX; ++j) {
        if (j >= kc) break;
        T* y = Y + (int64_t)j * ldy;
        if (ok0) y[p0] = beta == T(0) ? alpha * acc0[j] : alpha * acc0[j] + beta * y[p0];
        if (ok1) y[p1] = beta == T(0) ? alpha * acc1[j] : alpha * acc1[j] + beta * y[p1];
    }
}

template <typename T>
int update_valu(dft_ctx* c, int64_t n, int m, int k, T alpha, const T* A, int64_t lda, const T* Cm, int ldc, T beta, T* Y,
                int64_t ldy) {
    const unsigned blocks = (unsigned)((n + 511) / 512);
    for (int j0 = 0; j0 < k; j0 += 64) {
        const int kc = std::min(64, k - j0);
        const T* C0 = Cm + (int64_t)j0 * ldc;
        T* Y0 = Y + (int64_t)j0 * ldy;
        if (kc <= 16)
            k_update_valu<T, 16><<<blocks, 256, 0, c->stream>>>(n, m, kc, alpha, A, lda, C0, ldc, beta, Y0, ldy);
        else if (kc <= 32)
            k_update_valu<T, 32><<<blocks, 256, 0, c->stream>>>(n, m, kc, alpha, A, lda, C0, ldc, beta, Y0, ldy);
        else
            k_update_valu<T, 64><<<blocks, 256, 0, c->stream>>>(n, m, kc, alpha, A, lda, C0, ldc, beta, Y0, ldy);
        DFT_CHECK_LAUNCH(c);
    }
    return DFT_OK;
}

}  // namespace

namespace dft {

void release_blas(dft_ctx* c) {
    if (c->blas) rocblas_destroy_handle((rocblas_handle)c->blas);
    if (c->dev_info) (void)hipFree(c->dev_info);
    c->blas = nullptr;
    c->dev_info = nullptr;
}

static int blas(dft_ctx* c, rocblas_handle* h) {
    if (!c->blas) {
        rocblas_handle hh;
        if (rocblas_create_handle(&hh) != rocblas_status_success) return set_err(c, DFT_EHIP, "rocblas_create_handle failed");
        c->blas = (void*)hh;
        DFT_HIP(c, hipMalloc((void**)&c->dev_info, 4 * sizeof(int)));
    }
    if (rocblas_set_stream((rocblas_handle)c->blas, c->stream) != rocblas_status_success)
        return set_err(c, DFT_EHIP, "rocblas_set_stream failed");
    *h = (rocblas_handle)c->blas;
    return DFT_OK;
}

}  // namespace dft

extern "C" {


// Block update Y = alpha A C + beta Y (a plain tall-skinny GEMM, N x m times
// m x k): rocBLAS' 64-bit-index GEMM on the context stream; FP32 blocks use
// an FP32 copy of the small coefficient matrix.
int dft_block_update(dft_ctx* c, int prec, int64_t n, int m, int k, double alpha, const void* a, int64_t lda,
                     const double* cm, int64_t ldc, double beta, void* y, int64_t ldy) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, a && cm && y && n > 0 && m > 0 && k >= 0 && lda >= n && ldy >= n && ldc >= m, "dft_block_update: bad arguments");
    if (k == 0) return DFT_OK;
    rocblas_handle h;
    if (int rc = blas(c, &h)) return rc;
    rocblas_status st;
    if (prec == DFT_F64) {
        st = rocblas_dgemm_64(h, rocblas_operation_none, rocblas_operation_none, n, k, m, &alpha, (const double*)a, lda,
                              cm, ldc, &beta, (double*)y, ldy);
    } else if (prec == DFT_F32) {
        if (int rc = ensure_scratch(c, sizeof(float) * (size_t)m * k)) return rc;
        float* c32 = (float*)c->scratch;
        k_to_f32<<<(unsigned)(((int64_t)m * k + 255) / 256), 256, 0, c->stream>>>(m, k, cm, ldc, c32);
        DFT_CHECK_LAUNCH(c);
        const float al = (float)alpha, be = (float)beta;
        if (k <= 64 && m <= 512)
            return update_valu<float>(c, n, m, k, al, (const float*)a, lda, c32, m, be, (float*)y, ldy);
        st = rocblas_sgemm_64(h, rocblas_operation_none, rocblas_operation_none, n, k, m, &al, (const float*)a, lda, c32,
                              m, &be, (float*)y, ldy);
    } else {
        return set_err(c, DFT_EINVAL, "dft_block_update: bad precision");
    }
    if (st != rocblas_status_success) return set_err(c, DFT_EHIP, "rocblas gemm failed");
    return DFT_OK;
}

int dft_syev(dft_ctx* c, int m, double* a, int64_t lda, double* w) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, a && w && m > 0 && lda >= m && m <= 4096, "dft_syev: bad arguments");
    rocblas_handle h;
    if (int rc = blas(c, &h)) return rc;
    if (int rc = ensure_scratch(c, sizeof(double) * (size_t)m)) return rc;
    rocblas_status st = rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_lower, m, a, (rocblas_int)lda, w,
                                         (double*)c->scratch, c->dev_info);
    if (st != rocblas_status_success) return set_err(c, DFT_EHIP, "rocsolver_dsyevd failed");
    int info = 0;
    DFT_HIP(c, hipMemcpyAsync(&info, c->dev_info, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    DFT_HIP(c, hipStreamSynchronize(c->stream));
    if (info != 0) return set_err(c, DFT_ENUMERIC, "dft_syev: eigensolver did not converge");
    return DFT_OK;
}

int dft_chol_inv(dft_ctx* c, int m, double* a, int64_t lda) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, a && m > 0 && lda >= m && m <= 4096, "dft_chol_inv: bad arguments");
    rocblas_handle h;
    if (int rc = blas(c, &h)) return rc;
    rocblas_status st = rocsolver_dpotrf(h, rocblas_fill_lower, m, a, (rocblas_int)lda, c->dev_info);
    if (st != rocblas_status_success) return set_err(c, DFT_EHIP, "rocsolver_dpotrf failed");
    int info = 0;
    DFT_HIP(c, hipMemcpyAsync(&info, c->dev_info, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    DFT_HIP(c, hipStreamSynchronize(c->stream));
    if (info != 0) return set_err(c, DFT_ENUMERIC, "dft_chol_inv: matrix is not positive definite");
    st = rocsolver_dtrtri(h, rocblas_fill_lower, rocblas_diagonal_non_unit, m, a, (rocblas_int)lda, c->dev_info);
    if (st != rocblas_status_success) return set_err(c, DFT_EHIP, "rocsolver_dtrtri failed");
    DFT_HIP(c, hipMemcpyAsync(&info, c->dev_info, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    DFT_HIP(c, hipStreamSynchronize(c->stream));
    if (info != 0) return set_err(c, DFT_ENUMERIC, "dft_chol_inv: singular triangular factor");
    // a <- (L^{-1})^T: upper triangular, zero below the diagonal
    if (int rc = ensure_scratch(c, sizeof(double) * (size_t)m * m)) return rc;
    double* t = (double*)c->scratch;
    const unsigned gb = (unsigned)(((int64_t)m * m + 255) / 256);
    k_copy_sq<<<gb, 256, 0, c->stream>>>(m, a, lda, t);
    k_lower_t<<<gb, 256, 0, c->stream>>>(m, t, a, lda);
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

}  // extern "C"
