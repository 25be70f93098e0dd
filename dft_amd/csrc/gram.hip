// gram.hip -- (e) the Davidson Gram / Rayleigh-Ritz products A^T B and the
// tall-skinny block update Y = A C, on the CDNA4 matrix cores.
//
// A^T B with A: N x ma, B: N x mb, N (grid points) >> ma, mb (<= a few
// hundred).  The long dimension N is split across workgroups; each
// workgroup owns a 64 x 64 output tile (four waves, each 32 x 32 = 2 x 2
// MFMA 16x16 tiles) over one K-range and streams its A and B columns
// straight from HBM into MFMA operands: an operand lane holds a short
// contiguous run of one column, so a wave touches a handful of 128-byte
// lines per step and the next K steps hit the same lines.  Partial tiles
// go to an FP64 buffer and a second kernel sums the K-splits in a fixed
// order (deterministic).
//   F64:         v_mfma_f64_16x16x4_f64
//   F32:         v_mfma_f32_16x16x4_f32
//   SPLIT_BF16:  x = hi + lo (two bf16), A^T B ~ hi.hi + hi.lo + lo.hi with
//                v_mfma_f32_16x16x32_bf16, FP32 accumulation per K-range.
// MFMA 16x16 operand layout: lane l feeds row/column (l % 16) and K slice
// (l / 16); the 16x16 result sits in column l % 16 of each lane, rows
// 4*(l/16) + r for the f32-accumulating instructions and (l/16) + 4*r for
// v_mfma_f64_16x16x4_f64 (r = result register 0..3).
#include "common.hpp"

#include <algorithm>

using namespace dft;

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

__device__ inline __bf16 to_bf16(float x) { return (__bf16)x; }  // round to nearest even

// result row of register r held by quarter-wave q (see the layout note above)
template <bool F64>
__device__ inline int res_row(int q, int r) { return F64 ? q + 4 * r : 4 * q + r; }

// ---------------------------------------------------------------- Gram
template <int MODE>
struct GramT;

template <>
struct GramT<DFT_GRAM_F64> {
    using In = double;
    using Acc = d4;
    static constexpr int KSTEP = 4;
};
template <>
struct GramT<DFT_GRAM_F32> {
    using In = float;
    using Acc = f4;
    static constexpr int KSTEP = 4;
};
template <>
struct GramT<DFT_GRAM_SPLIT_BF16> {
    using In = float;
    using Acc = f4;
    static constexpr int KSTEP = 32;
};

template <int MODE>
__global__ __launch_bounds__(256) void k_gram(int64_t n, int ma, int mb, int ntm, const typename GramT<MODE>::In* __restrict__ A,
                                              int64_t lda, const typename GramT<MODE>::In* __restrict__ B, int64_t ldb,
                                              int64_t kchunk, double* __restrict__ part) {
    using Tr = GramT<MODE>;
    using In = typename Tr::In;
    using Acc = typename Tr::Acc;
    constexpr int KS = Tr::KSTEP;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int tile = blockIdx.x, ks = blockIdx.y;
    const int i0 = (tile % ntm) * 64 + (w >> 1) * 32;
    const int j0 = (tile / ntm) * 64 + (w & 1) * 32;
    const int64_t kbeg = (int64_t)ks * kchunk, kend = min(n, kbeg + kchunk);
    const int r16 = lane & 15, kq = lane >> 4;
    Acc acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = Acc{0, 0, 0, 0};
    const int ia[2] = {i0 + r16, i0 + 16 + r16};
    const int jb[2] = {j0 + r16, j0 + 16 + r16};
    const bool va[2] = {ia[0] < ma, ia[1] < ma};
    const bool vb[2] = {jb[0] < mb, jb[1] < mb};
    const In* pa[2] = {A + (int64_t)(va[0] ? ia[0] : 0) * lda, A + (int64_t)(va[1] ? ia[1] : 0) * lda};
    const In* pb[2] = {B + (int64_t)(vb[0] ? jb[0] : 0) * ldb, B + (int64_t)(vb[1] ? jb[1] : 0) * ldb};

    if constexpr (MODE != DFT_GRAM_SPLIT_BF16) {
        for (int64_t k = kbeg; k < kend; k += KS) {
            const int64_t kk = k + kq;
            const bool kin = kk < kend;
            In xa[2], xb[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) xa[a] = (va[a] && kin) ? pa[a][kk] : In(0);
#pragma unroll
            for (int b = 0; b < 2; ++b) xb[b] = (vb[b] && kin) ? pb[b][kk] : In(0);
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    if constexpr (MODE == DFT_GRAM_F64)
                        acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[a], xb[b], acc[a][b], 0, 0, 0);
                    else
                        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa[a], xb[b], acc[a][b], 0, 0, 0);
                }
        }
    } else {
        for (int64_t k = kbeg; k < kend; k += KS) {
            const int64_t kk = k + 8 * kq;
            bf8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const float x = (va[a] && kk + e < kend) ? pa[a][kk + e] : 0.0f;
                    const __bf16 h = to_bf16(x);
                    ah[a][e] = h;
                    al[a][e] = to_bf16(x - (float)h);
                }
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const float x = (vb[b] && kk + e < kend) ? pb[b][kk + e] : 0.0f;
                    const __bf16 h = to_bf16(x);
                    bh[b][e] = h;
                    bl[b][e] = to_bf16(x - (float)h);
                }
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], bh[b], acc[a][b], 0, 0, 0);
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bl[b], acc[a][b], 0, 0, 0);
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bh[b], acc[a][b], 0, 0, 0);
                }
        }
    }
    // partial tile -> part[ks][j*ma + i]
    double* P = part + (int64_t)ks * ma * mb;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int j = j0 + 16 * b + r16;
            if (j >= mb) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + 16 * a + res_row<MODE == DFT_GRAM_F64>(kq, r);
                if (i < ma) P[(int64_t)j * ma + i] = (double)acc[a][b][r];
            }
        }
}

__global__ __launch_bounds__(256) void k_gram_reduce(const double* __restrict__ part, int nks, int ma, int mb,
                                                     double alpha, double beta, double* __restrict__ out, int64_t ldo) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t tot = (int64_t)ma * mb;
    if (e >= tot) return;
    double s = 0.0;
    for (int k = 0; k < nks; ++k) s += part[(int64_t)k * tot + e];
    const int i = (int)(e % ma), j = (int)(e / ma);
    double* o = out + (int64_t)j * ldo + i;
    *o = beta == 0.0 ? alpha * s : alpha * s + beta * *o;
}

// ------------------------------------------------------- block update Y = a A C + b Y
// A: N x m (column-major), C: m x k FP64, Y: N x k.  Workgroup = 4 waves x
// 16 rows = 64 rows of Y and WN*16 columns; the A operand (16 rows x 4 K)
// is shared by the WN MFMA tiles of the wave.
template <typename T, int WN>
__global__ __launch_bounds__(256) void k_update(int64_t n, int m, int k, double alpha, const T* __restrict__ A,
                                                int64_t lda, const double* __restrict__ C, int64_t ldc, double beta,
                                                T* __restrict__ Y, int64_t ldy) {
    using Acc = typename std::conditional<sizeof(T) == 8, d4, f4>::type;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t r0 = (int64_t)blockIdx.x * 64 + w * 16;
    const int j0 = blockIdx.y * (16 * WN);
    const int r16 = lane & 15, kq = lane >> 4;
    const int64_t row = r0 + r16;
    const bool rin = row < n;
    Acc acc[WN];
#pragma unroll
    for (int b = 0; b < WN; ++b) acc[b] = Acc{0, 0, 0, 0};
    int jc[WN];
    bool jv[WN];
#pragma unroll
    for (int b = 0; b < WN; ++b) {
        jc[b] = j0 + 16 * b + r16;
        jv[b] = jc[b] < k;
    }
    for (int kk = 0; kk < m; kk += 4) {
        const int q = kk + kq;
        const bool qin = q < m;
        const T a = (rin && qin) ? A[(int64_t)q * lda + row] : T(0);
#pragma unroll
        for (int b = 0; b < WN; ++b) {
            const T cv = (jv[b] && qin) ? (T)C[(int64_t)jc[b] * ldc + q] : T(0);
            if constexpr (sizeof(T) == 8)
                acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, cv, acc[b], 0, 0, 0);
            else
                acc[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, cv, acc[b], 0, 0, 0);
        }
    }
#pragma unroll
    for (int b = 0; b < WN; ++b) {
        const int j = j0 + 16 * b + r16;
        if (j >= k) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t i = r0 + res_row<sizeof(T) == 8>(kq, r);
            if (i >= n) continue;
            T* y = Y + (int64_t)j * ldy + i;
            const double v = alpha * (double)acc[b][r];
            *y = beta == 0.0 ? (T)v : (T)(v + beta * (double)*y);
        }
    }
}

template <typename T>
int update_t(dft_ctx* c, int64_t n, int m, int k, double alpha, const T* A, int64_t lda, const double* C, int64_t ldc,
             double beta, T* Y, int64_t ldy) {
    const unsigned gx = (unsigned)((n + 63) / 64);
    // the widest column block up to 128 columns; A is read once per block
    if (k <= 16) {
        k_update<T, 1><<<dim3(gx, (k + 15) / 16), 256, 0, c->stream>>>(n, m, k, alpha, A, lda, C, ldc, beta, Y, ldy);
    } else if (k <= 32) {
        k_update<T, 2><<<dim3(gx, (k + 31) / 32), 256, 0, c->stream>>>(n, m, k, alpha, A, lda, C, ldc, beta, Y, ldy);
    } else if (k <= 64) {
        k_update<T, 4><<<dim3(gx, (k + 63) / 64), 256, 0, c->stream>>>(n, m, k, alpha, A, lda, C, ldc, beta, Y, ldy);
    } else {
        k_update<T, 8><<<dim3(gx, (k + 127) / 128), 256, 0, c->stream>>>(n, m, k, alpha, A, lda, C, ldc, beta, Y, ldy);
    }
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

}  // namespace

extern "C" {

int dft_gram(dft_ctx* c, int mode, int64_t n, int ma, int mb, double alpha, const void* a, int64_t lda, const void* b,
             int64_t ldb, double beta, double* out, int64_t ldo) {
    DFT_REQUIRE(c, a && b && out && n > 0 && ma > 0 && mb > 0 && lda >= n && ldb >= n && ldo >= ma, "dft_gram: bad arguments");
    DFT_REQUIRE(c, mode >= DFT_GRAM_F64 && mode <= DFT_GRAM_SPLIT_BF16, "dft_gram: bad mode");
    const int ntm = (ma + 63) / 64, ntn = (mb + 63) / 64;
    const int ntiles = ntm * ntn;
    const int kstep = mode == DFT_GRAM_SPLIT_BF16 ? 32 : 4;
    // K-split: enough workgroups to fill the chip twice over, K-chunks a multiple of the MFMA K step
    int64_t nks = std::max<int64_t>(1, (8LL * c->num_cus + ntiles - 1) / ntiles);
    int64_t kchunk = (n + nks - 1) / nks;
    kchunk = std::max<int64_t>(kchunk, 512);
    kchunk = (kchunk + kstep - 1) / kstep * kstep;
    nks = (n + kchunk - 1) / kchunk;
    DFT_REQUIRE(c, nks <= 65535, "dft_gram: too many K splits");
    if (int rc = ensure_partials(c, (size_t)nks * ma * mb + 8)) return rc;
    dim3 grid(ntiles, (unsigned)nks);
    if (mode == DFT_GRAM_F64)
        k_gram<DFT_GRAM_F64><<<grid, 256, 0, c->stream>>>(n, ma, mb, ntm, (const double*)a, lda, (const double*)b, ldb, kchunk, c->partials);
    else if (mode == DFT_GRAM_F32)
        k_gram<DFT_GRAM_F32><<<grid, 256, 0, c->stream>>>(n, ma, mb, ntm, (const float*)a, lda, (const float*)b, ldb, kchunk, c->partials);
    else
        k_gram<DFT_GRAM_SPLIT_BF16><<<grid, 256, 0, c->stream>>>(n, ma, mb, ntm, (const float*)a, lda, (const float*)b, ldb, kchunk, c->partials);
    DFT_CHECK_LAUNCH(c);
    const int64_t tot = (int64_t)ma * mb;
    k_gram_reduce<<<(unsigned)((tot + 255) / 256), 256, 0, c->stream>>>(c->partials, (int)nks, ma, mb, alpha, beta, out, ldo);
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

int dft_block_update(dft_ctx* c, int prec, int64_t n, int m, int k, double alpha, const void* a, int64_t lda,
                     const double* cm, int64_t ldc, double beta, void* y, int64_t ldy) {
    DFT_REQUIRE(c, a && cm && y && n > 0 && m >= 0 && k >= 0 && lda >= n && ldy >= n && ldc >= m, "dft_block_update: bad arguments");
    DFT_REQUIRE(c, (n + 63) / 64 < (1LL << 31), "dft_block_update: n too large");
    if (k == 0) return DFT_OK;
    if (prec == DFT_F64) return update_t<double>(c, n, m, k, alpha, (const double*)a, lda, cm, ldc, beta, (double*)y, ldy);
    if (prec == DFT_F32) return update_t<float>(c, n, m, k, alpha, (const float*)a, lda, cm, ldc, beta, (float*)y, ldy);
    return set_err(c, DFT_EINVAL, "dft_block_update: bad precision");
}

}  // extern "C"
