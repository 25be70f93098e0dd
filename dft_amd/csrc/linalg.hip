// linalg.hip -- dense FP64 subspace algebra on the device for the block
// Davidson (Rayleigh-Ritz eigensolve and Cholesky orthonormalisation of
// m x m matrices, m <= a few hundred), through rocSOLVER on the context
// stream.
#include "common.hpp"

#include <atomic>

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
// ---- tall-skinny block update on the matrix cores -------------------------
// Y^T (k x N) = alpha C^T (k x m) [A | A2]^T (m x N) + beta Y^T, as 16 x 16
// MFMA tiles: operand A = C^T from LDS (lane: row j = l % 16, K slice
// i = l / 16), operand B = the grid columns straight from HBM, result rows =
// output columns j, result column = grid point.  Inner rows i < m1 come from
// A, rows m1 <= i < m from A2 (m1 = m: A alone), so two column blocks in
// separate buffers enter one pass.  A wave owns 16 VEC grid points and all
// k <= 96 output columns (6 accumulator tiles) and strides over the points;
// C is staged once per workgroup.
//
// VEC = 1: lane l' = l % 16 owns point p0 + l'.  VEC = 16 / sizeof(T): lane
// l' owns points p0 + VEC l' ... + VEC - 1, and MFMA c of a K step takes the
// points p0 + VEC l' + c as its columns; every lane then loads and stores its
// points as one 16-byte vector (a quarter-wave moves 256 contiguous bytes of
// a column) and one LDS read of C^T feeds VEC MFMAs.  The vector form needs
// n, lda, lda2, ldy multiples of VEC and 16-byte aligned A, A2, Y.
//
// NORM: part[j * gridDim.x + blockIdx.x] receives the workgroup's FP64 sum of
// squares of the stored column j of Y (the residual norms of the Davidson
// iteration come out of the same pass).
//
// Y may alias columns of A or A2 (an in-place basis update): a wave reads
// every input row at its points before it writes its outputs there, and no
// other wave touches those points.
template <typename T>
struct UpdMfma;
template <>
struct UpdMfma<double> {
    typedef double Acc __attribute__((ext_vector_type(4)));
    static __device__ inline Acc mfma(double a, double b, Acc c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    static __device__ inline int row(int q, int r) { return q + 4 * r; }
};
template <>
struct UpdMfma<float> {
    typedef float Acc __attribute__((ext_vector_type(4)));
    static __device__ inline Acc mfma(float a, float b, Acc c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    static __device__ inline int row(int q, int r) { return 4 * q + r; }
};

template <typename T, int VEC>
struct alignas(sizeof(T) * VEC) Pack {
    T v[VEC];
};

constexpr int kUpdLdsBytes = 48 * 1024;

template <typename T, int NJT, int VEC, bool NORM>
__global__ __launch_bounds__(256) void k_update_mfma(int64_t n, int m1, int m, int k, T alpha, const T* A,
                                                     int64_t lda, const T* A2, int64_t lda2,
                                                     const T* __restrict__ Cm, int ldc, T beta, T* Y,
                                                     int64_t ldy, double* __restrict__ part) {
    using U = UpdMfma<T>;
    using Acc = typename U::Acc;
    using P = Pack<T, VEC>;
    extern __shared__ __attribute__((aligned(16))) unsigned char upd_lds[];
    T* ct = reinterpret_cast<T*>(upd_lds);  // C^T padded to NJT*16 rows: ct[i * (NJT*16) + j]
    constexpr int KP = NJT * 16;
    const int mp = (m + 3) / 4 * 4;
    for (int q = threadIdx.x; q < mp * KP; q += 256) {
        const int i = q / KP, j = q % KP;
        ct[q] = (i < m && j < k) ? Cm[(int64_t)j * ldc + i] : T(0);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r16 = lane & 15, qd = lane >> 4;
    constexpr int TP = 16 * VEC;
    const int64_t ntile = (n + TP - 1) / TP;
    double s2[NJT][4];
#pragma unroll
    for (int jt = 0; jt < NJT; ++jt)
#pragma unroll
        for (int r = 0; r < 4; ++r) s2[jt][r] = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * 4 + w; t < ntile; t += (int64_t)gridDim.x * 4) {
        const int64_t p = t * TP + VEC * r16;
        const bool pin = p < n;  // n % VEC == 0: all VEC points or none
        Acc acc[VEC][NJT];
#pragma unroll
        for (int c = 0; c < VEC; ++c)
#pragma unroll
            for (int jt = 0; jt < NJT; ++jt) acc[c][jt] = Acc{0, 0, 0, 0};
        for (int i4 = 0; i4 < mp; i4 += 32) {
            P b[8];
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                const int i = i4 + 4 * s + qd;
                if (pin && i < m) {
                    const T* src = i < m1 ? A + (int64_t)i * lda : A2 + (int64_t)(i - m1) * lda2;
                    b[s] = *reinterpret_cast<const P*>(src + p);
                } else {
#pragma unroll
                    for (int c = 0; c < VEC; ++c) b[s].v[c] = T(0);
                }
            }
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                const int ib = i4 + 4 * s;
                if (ib >= mp) break;
#pragma unroll
                for (int jt = 0; jt < NJT; ++jt) {
                    const T a = ct[(ib + qd) * KP + jt * 16 + r16];
#pragma unroll
                    for (int c = 0; c < VEC; ++c) acc[c][jt] = U::mfma(a, b[s].v[c], acc[c][jt]);
                }
            }
        }
        if (!pin) continue;
#pragma unroll
        for (int jt = 0; jt < NJT; ++jt)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int j = jt * 16 + U::row(qd, r);
                if (j >= k) continue;
                P* y = reinterpret_cast<P*>(Y + (int64_t)j * ldy + p);
                P o;
#pragma unroll
                for (int c = 0; c < VEC; ++c) o.v[c] = alpha * acc[c][jt][r];
                if (beta != T(0)) {
                    const P yo = *y;
#pragma unroll
                    for (int c = 0; c < VEC; ++c) o.v[c] += beta * yo.v[c];
                }
                *y = o;
                if constexpr (NORM) {
#pragma unroll
                    for (int c = 0; c < VEC; ++c) s2[jt][r] += (double)o.v[c] * (double)o.v[c];
                }
            }
    }
    if constexpr (NORM) {
        __shared__ double red[4][KP];
#pragma unroll
        for (int jt = 0; jt < NJT; ++jt)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                double v = s2[jt][r];
#pragma unroll
                for (int o = 1; o < 16; o <<= 1) v += __shfl_xor(v, o, 16);
                if (r16 == 0) red[w][jt * 16 + U::row(qd, r)] = v;
            }
        __syncthreads();
        for (int j = threadIdx.x; j < k; j += 256)
            part[(int64_t)j * gridDim.x + blockIdx.x] = (red[0][j] + red[1][j]) + (red[2][j] + red[3][j]);
    }
}

// per-column sums of the NORM partials: out[j] = sum_b part[j * nblk + b]
__global__ __launch_bounds__(64) void k_sum_parts(const double* __restrict__ part, int nblk, double* __restrict__ out) {
    const int j = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += 64) s += part[(int64_t)j * nblk + b];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (threadIdx.x == 0) out[j] = s;
}

// true when the MFMA path applies (k <= 96 and C^T fits the LDS budget)
template <typename T>
bool update_mfma_fits(int m, int k) {
    const int njt = (k + 15) / 16;
    return k >= 1 && k <= 96 && (int64_t)((m + 3) / 4 * 4) * njt * 16 * (int64_t)sizeof(T) <= kUpdLdsBytes;
}

// largest inner-dimension chunk (a multiple of 4) whose C^T fits the LDS budget,
// balanced over ceil(m / max) chunks; 0 when chunking does not pay.  Chunks
// re-read and re-write Y, and past k = 48 the kernel's MFMA rate (~50 TF/s
// FP32, ~20 TF/s FP64 at 288 x 96) falls behind rocBLAS' tuned GEMM on these
// compute-bound shapes (measured on MI355X at N = 256^3: 19.7 / 54.9 ms chunked
// against 17.7 / 17.6 ms for the library, FP32 / FP64).
template <typename T>
int update_chunk(int m, int k) {
    if (k < 1 || k > 48) return 0;
    const int njt = (k + 15) / 16;
    const int mx = (int)(kUpdLdsBytes / (njt * 16 * (int64_t)sizeof(T))) / 4 * 4;
    if (mx < 4) return 0;
    const int nch = (m + mx - 1) / mx;
    return ((m + nch - 1) / nch + 3) / 4 * 4;
}

// Y = alpha [A | A2] C + beta Y (m1 inner rows from A); with norm2 != null,
// norm2[j] = ||Y[:, j]||^2 (device FP64, k values) from the same pass.
template <typename T>
int update_mfma(dft_ctx* c, int64_t n, int m1, int m, int k, T alpha, const T* A, int64_t lda, const T* A2,
                int64_t lda2, const T* Cm, int ldc, T beta, T* Y, int64_t ldy, double* norm2 = nullptr) {
    const int njt = (k + 15) / 16;
    const size_t lds = (size_t)((m + 3) / 4 * 4) * njt * 16 * sizeof(T);
    constexpr int VEC = 16 / sizeof(T);
    if (m1 >= m) A2 = A, lda2 = lda;
    const bool vec = n % VEC == 0 && lda % VEC == 0 && lda2 % VEC == 0 && ldy % VEC == 0 &&
                     reinterpret_cast<uintptr_t>(A) % 16 == 0 && reinterpret_cast<uintptr_t>(A2) % 16 == 0 &&
                     reinterpret_cast<uintptr_t>(Y) % 16 == 0;
    const int64_t ntile = vec ? (n + 16 * VEC - 1) / (16 * VEC) : (n + 15) / 16;
    const unsigned blocks = (unsigned)std::min<int64_t>((ntile + 3) / 4, 8LL * c->num_cus);
    double* part = nullptr;
    if (norm2) {
        if (int rc = ensure_partials(c, (size_t)blocks * k + 8)) return rc;
        part = c->partials;
    }
#define DFT_UPD_K(NJ, V, NRM)                                                                                         \
    k_update_mfma<T, NJ, V, NRM><<<blocks, 256, lds, c->stream>>>(n, m1, m, k, alpha, A, lda, A2, lda2, Cm, ldc, beta, \
                                                                  Y, ldy, part)
#define DFT_UPD(NJ)                                                                                                   \
    if (vec) {                                                                                                        \
        if (norm2) DFT_UPD_K(NJ, VEC, true);                                                                          \
        else DFT_UPD_K(NJ, VEC, false);                                                                               \
    } else {                                                                                                          \
        if (norm2) DFT_UPD_K(NJ, 1, true);                                                                            \
        else DFT_UPD_K(NJ, 1, false);                                                                                 \
    }
    switch (njt) {
        case 1: DFT_UPD(1); break;
        case 2: DFT_UPD(2); break;
        case 3: DFT_UPD(3); break;
        case 4: DFT_UPD(4); break;
        case 5: DFT_UPD(5); break;
        default: DFT_UPD(6); break;
    }
#undef DFT_UPD
#undef DFT_UPD_K
    DFT_CHECK_LAUNCH(c);
    if (norm2) {
        k_sum_parts<<<k, 64, 0, c->stream>>>(part, (int)blocks, norm2);
        DFT_CHECK_LAUNCH(c);
    }
    return DFT_OK;
}


// ---- FP32 block update on the bf16 matrix cores (three-piece split) -------
// Y = [A | A2] (alpha C) + beta Y for FP32 blocks: N grid points, m inner
// rows (m1 from A, the rest from A2), k <= 96 output columns.  Every FP32
// operand value x is split as x1 + x2 + x3 (24 significand bits: exact for
// normal values), and a product keeps the six largest terms, a3b1 + a1b3 +
// a2b2 + a2b1 + a1b2 + a1b1 (smallest first), each a
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation: FP32-level products at
// the bf16 MFMA rate, so the pass is bound by its HBM stream (N (m + k)
// floats) rather than by FP32 MFMA.  Both operands are split by
// rounding, x1 = bf16(x), x2 = bf16(x - x1), x3 = bf16(x - x1 - x2), so
// |x2| <= 2^-8 |x|, |x3| <= 2^-16 |x| and the dropped terms a2b3 + a3b2 +
// a3b3 stay below 2^-23 |ab|.
// Layout: MFMA operand A = (alpha C)^T, row j = output column, K = inner
// row; operand B = the grid columns, K = inner row, column = grid point.
// Lane l of a wave: points p0 + 4 (l % 16) .. + 3 (16 bytes per inner row),
// K slice 8 (l / 16) .. + 7 of each 32-row step; result register r of tile
// jt = output column jt 16 + 4 (l / 16) + r at those four points (one
// 16-byte store).  The split coefficients are prepared once per call in
// operand order ([32-row step][jt][piece][lane][8]) and staged through LDS
// KSC steps at a time for the 8 waves (512 points) of a workgroup.
// The grid values reach LDS by direct loads (global_load_lds_dwordx4), one
// 32-row step ahead, into a per-wave staging buffer: no registers hold them
// in flight, so two waves per SIMD fit the 256-register budget while the
// next step's loads overlap this step's products.  Lane l's 16 bytes of
// direct load e land at stage[w][64 e + l] (the instruction writes M0 +
// 16 x lane), which is where lane l reads them back.
// Y may alias columns of A or A2: a wave reads every inner row at its
// points before it writes there.  NORM: part[j * gridDim.x + b] = the
// workgroup's sum of squares of stored column j.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// bf16(lo) | bf16(hi) << 16, rounded to nearest even (one v_cvt_pk_bf16_f32)
__device__ __forceinline__ uint32_t pack_bf16x2(float lo, float hi) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){lo, hi}, bf16x2));
}

constexpr int kS3Waves = 8, kS3Steps = 4;  // waves per workgroup, 32-row steps per coefficient chunk

__global__ __launch_bounds__(256) void k_split_coef(int m, int k, int njt, int nks, double alpha,
                                                    const double* __restrict__ cm, int64_t ldc,
                                                    bf16x8* __restrict__ out) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;  // one (step, jt, lane) triple
    const int64_t total = (int64_t)nks * njt * 64;
    if (q >= total) return;
    const int lane = (int)(q % 64), jt = (int)((q / 64) % njt), ks = (int)(q / (64 * njt));
    const int j = jt * 16 + (lane & 15), ib = ks * 32 + 8 * (lane >> 4);
    bf16x8 p1, p2, p3;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int i = ib + e;
        const double x = (i < m && j < k) ? alpha * cm[(int64_t)j * ldc + i] : 0.0;
        const __bf16 h1 = (__bf16)(float)x;
        const double r1 = x - (double)(float)h1;
        const __bf16 h2 = (__bf16)(float)r1;
        const double r2 = r1 - (double)(float)h2;
        p1[e] = h1;
        p2[e] = h2;
        p3[e] = (__bf16)(float)r2;
    }
    bf16x8* o = out + (((int64_t)ks * njt + jt) * 3) * 64 + lane;
    o[0] = p1;
    o[64] = p2;
    o[128] = p3;
}

template <int NJT, bool NORM, int KSC = kS3Steps>
__global__ __launch_bounds__(512) void k_update_split3(int64_t n, int m1, int m, int k, const float* A, int64_t lda,
                                                        const float* A2, int64_t lda2,
                                                        const bf16x8* __restrict__ cs, float beta, float* Y,
                                                        int64_t ldy, double* __restrict__ part) {
    constexpr int W = kS3Waves;
    __shared__ bf16x8 lds[KSC * NJT * 3 * 64];
    __shared__ f32x4 stage[W][8 * 64];
    __shared__ double red[NORM ? W : 1][NORM ? NJT * 16 : 1];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r16 = lane & 15, qd = lane >> 4;
    const int nks = (m + 31) / 32;
    const int64_t ntile = (n + 63) / 64;
    if constexpr (NORM)
        for (int q = threadIdx.x; q < W * NJT * 16; q += 512) (&red[0][0])[q] = 0.0;
    for (int64_t t0 = (int64_t)blockIdx.x * W; t0 < ntile; t0 += (int64_t)gridDim.x * W) {
        const int64_t p = (t0 + w) * 64 + 4 * r16;
        const bool pin = p < n;
        const int64_t pc = pin ? p : 0;  // lanes past the grid load point 0 and store nothing
        // rows past m read row m - 1, whose coefficients are zero
        // per-lane byte offsets of this lane's first row of a step, from the
        // step's first row in A and in A2 (32-bit: the 8 rows of a lane group
        // and the point lie within 4 GiB of the step's row base)
        const uint32_t offA = (uint32_t)(((int64_t)8 * qd * lda + pc) * 4);
        const uint32_t offA2 = (uint32_t)(((int64_t)8 * qd * lda2 + pc) * 4);
        auto issue = [&](int ks) {
            const int r0 = ks * 32;
            if (r0 + 31 < m && (r0 + 31 < m1 || r0 >= m1)) {
                // the step's 32 rows lie in one of A, A2: a uniform row base
                // and stride, and one per-lane offset
                const bool inA = r0 < m1;
                const char* base = inA ? (const char*)(A + (int64_t)r0 * lda) : (const char*)(A2 + (int64_t)(r0 - m1) * lda2);
                const int64_t stride = (inA ? lda : lda2) * 4;
                const uint32_t off = inA ? offA : offA2;
#pragma unroll
                for (int e = 0; e < 8; ++e)
                    __builtin_amdgcn_global_load_lds((const void*)(base + e * stride + off),
                                                     (__attribute__((address_space(3))) void*)&stage[w][64 * e], 16, 0, 0);
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int i = min(r0 + 8 * qd + e, m - 1);
                    const float* s = i < m1 ? A + (int64_t)i * lda : A2 + (int64_t)(i - m1) * lda2;
                    __builtin_amdgcn_global_load_lds((const void*)(s + pc),
                                                     (__attribute__((address_space(3))) void*)&stage[w][64 * e], 16, 0, 0);
                }
            }
        };
        f32x4 acc[4][NJT];
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int jt = 0; jt < NJT; ++jt) acc[c][jt] = f32x4{0.f, 0.f, 0.f, 0.f};
        issue(0);
        for (int ks0 = 0; ks0 < nks; ks0 += KSC) {
            __syncthreads();  // the previous chunk's operands are consumed
            const int nsteps = min(KSC, nks - ks0);
            const bf16x8* src = cs + (int64_t)ks0 * NJT * 3 * 64;
            for (int q = threadIdx.x; q < nsteps * NJT * 3 * 64; q += 512) lds[q] = src[q];
            __syncthreads();
            for (int kk = 0; kk < nsteps; ++kk) {
                // this step's direct loads have landed; read them, and let the
                // reads finish before the next step's loads overwrite the buffer
                asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
                f32x4 x[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) x[e] = stage[w][64 * e + lane];
                asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
                if (ks0 + kk + 1 < nks) issue(ks0 + kk + 1);
                // pieces by rounding, two values per conversion: the packed
                // pair's halves are the rounded values' top bits
                u32x4 b1[4], b2[4], b3[4];
#pragma unroll
                for (int c = 0; c < 4; ++c)
#pragma unroll
                    for (int e = 0; e < 8; e += 2) {
                        const float v0 = x[e][c], v1 = x[e + 1][c];
                        const uint32_t h = pack_bf16x2(v0, v1);
                        const float r0 = v0 - __uint_as_float(h << 16), r1 = v1 - __uint_as_float(h & 0xffff0000u);
                        const uint32_t g = pack_bf16x2(r0, r1);
                        const float s0 = r0 - __uint_as_float(g << 16), s1 = r1 - __uint_as_float(g & 0xffff0000u);
                        b1[c][e / 2] = h;
                        b2[c][e / 2] = g;
                        b3[c][e / 2] = pack_bf16x2(s0, s1);
                    }
#pragma unroll
                for (int jt = 0; jt < NJT; ++jt) {
                    const bf16x8* blk = lds + ((kk * NJT + jt) * 3) * 64;
                    const bf16x8 a1 = blk[lane], a2 = blk[64 + lane], a3 = blk[128 + lane];
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        const bf16x8 p1 = __builtin_bit_cast(bf16x8, b1[c]), p2 = __builtin_bit_cast(bf16x8, b2[c]),
                                     p3 = __builtin_bit_cast(bf16x8, b3[c]);
                        f32x4 s = acc[c][jt];
                        s = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a3, p1, s, 0, 0, 0);
                        s = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, p3, s, 0, 0, 0);
                        s = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2, p2, s, 0, 0, 0);
                        s = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2, p1, s, 0, 0, 0);
                        s = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, p2, s, 0, 0, 0);
                        acc[c][jt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, p1, s, 0, 0, 0);
                    }
                }
            }
        }
#pragma unroll
        for (int jt = 0; jt < NJT; ++jt)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int j = jt * 16 + 4 * qd + r;
                f32x4 o = f32x4{acc[0][jt][r], acc[1][jt][r], acc[2][jt][r], acc[3][jt][r]};
                const bool st = pin && j < k;
                if (st) {
                    f32x4* y = reinterpret_cast<f32x4*>(Y + (int64_t)j * ldy + p);
                    if (beta != 0.f) o += beta * *y;
                    *y = o;
                }
                if constexpr (NORM) {
                    double v = st ? ((double)o[0] * o[0] + (double)o[1] * o[1]) + ((double)o[2] * o[2] + (double)o[3] * o[3])
                                  : 0.0;
#pragma unroll
                    for (int s = 1; s < 16; s <<= 1) v += __shfl_xor(v, s, 16);
                    if (r16 == 0) red[w][j] += v;
                }
            }
    }
    if constexpr (NORM) {
        __syncthreads();
        for (int j = threadIdx.x; j < k; j += 512) {
            double s = 0.0;
#pragma unroll
            for (int q = 0; q < W; ++q) s += red[q][j];
            part[(int64_t)j * gridDim.x + blockIdx.x] = s;
        }
    }
}

// true when the bf16 split update takes the shape (FP32, k <= 96, 16-byte
// aligned columns)
bool split3_fits(int64_t n, int k, int64_t lda, int64_t lda2, int64_t ldy, const void* a, const void* a2,
                 const void* y) {
    // the kernel's per-lane offsets from a step's first row are 32-bit
    const bool off32 = 4 * (24 * std::max(lda, lda2) + n) < (int64_t)1 << 32;
    return k >= 1 && k <= 96 && n % 4 == 0 && lda % 4 == 0 && lda2 % 4 == 0 && ldy % 4 == 0 && off32 &&
           reinterpret_cast<uintptr_t>(a) % 16 == 0 && reinterpret_cast<uintptr_t>(a2) % 16 == 0 &&
           reinterpret_cast<uintptr_t>(y) % 16 == 0;
}

// Y = alpha [A | A2] C + beta Y on the bf16 matrix cores (C: FP64, m x k,
// column-major, ldc); norm2 (device, k doubles) as update_mfma.
int update_split3(dft_ctx* c, int64_t n, int m1, int m, int k, double alpha, const float* A, int64_t lda,
                  const float* A2, int64_t lda2, const double* Cm, int64_t ldc, float beta, float* Y, int64_t ldy,
                  double* norm2 = nullptr) {
    if (m1 >= m) A2 = A, lda2 = lda;
    const int njt = (k + 15) / 16, nks = (m + 31) / 32;
    const size_t need = (size_t)nks * njt * 3 * 64 * sizeof(bf16x8);
    if (c->coef3_cap < need) {
        if (c->coef3) {
            DFT_HIP(c, hipStreamSynchronize(c->stream));
            DFT_HIP(c, hipFree(c->coef3));
            c->coef3 = nullptr;
            c->coef3_cap = 0;
        }
        DFT_HIP(c, hipMalloc(&c->coef3, need));
        c->coef3_cap = need;
    }
    bf16x8* cs = static_cast<bf16x8*>(c->coef3);
    const int64_t pq = (int64_t)nks * njt * 64;
    k_split_coef<<<(unsigned)((pq + 255) / 256), 256, 0, c->stream>>>(m, k, njt, nks, alpha, Cm, ldc, cs);
    DFT_CHECK_LAUNCH(c);
    const int64_t ntile = (n + 63) / 64;
    const unsigned blocks = (unsigned)std::min<int64_t>((ntile + kS3Waves - 1) / kS3Waves, 2LL * c->num_cus);
    double* part = nullptr;
    if (norm2) {
        if (int rc = ensure_partials(c, (size_t)blocks * k + 8)) return rc;
        part = c->partials;
    }
#define DFT_S3(NJ)                                                                                                  \
    if (norm2)                                                                                                      \
        k_update_split3<NJ, true><<<blocks, 512, 0, c->stream>>>(n, m1, m, k, A, lda, A2, lda2, cs, beta, Y, ldy, part); \
    else                                                                                                            \
        k_update_split3<NJ, false><<<blocks, 512, 0, c->stream>>>(n, m1, m, k, A, lda, A2, lda2, cs, beta, Y, ldy, part)
    switch (njt) {
        case 1: DFT_S3(1); break;
        case 2: DFT_S3(2); break;
        case 3: DFT_S3(3); break;
        case 4: DFT_S3(4); break;
        case 5: DFT_S3(5); break;
        default: DFT_S3(6); break;
    }
#undef DFT_S3
    DFT_CHECK_LAUNCH(c);
    if (norm2) {
        k_sum_parts<<<k, 64, 0, c->stream>>>(part, (int)blocks, norm2);
        DFT_CHECK_LAUNCH(c);
    }
    return DFT_OK;
}

// coefficients of the fused Davidson residual, 2m x k column-major:
// rows i < m: Q[i][j]; rows m + i: -theta[j] Q[i][j]
template <typename T>
__global__ __launch_bounds__(256) void k_ritz_coef(int m, int k, const double* __restrict__ q, int64_t ldq,
                                                   const double* __restrict__ theta, T* __restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)2 * m * k) return;
    const int i = (int)(e % (2 * m)), j = (int)(e / (2 * m));
    out[e] = i < m ? (T)q[(int64_t)j * ldq + i] : (T)(-theta[j] * q[(int64_t)j * ldq + (i - m)]);
}
// -theta[j] Q[i][j], m x k column-major (the two-pass form's second operand)
__global__ __launch_bounds__(256) void k_neg_q_theta(int m, int k, const double* __restrict__ q, int64_t ldq,
                                                     const double* __restrict__ theta, double* __restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)m * k) return;
    const int i = (int)(e % m), j = (int)(e / m);
    out[e] = -theta[j] * q[(int64_t)j * ldq + i];
}

// ---- small symmetric eigensolver (the Davidson's Rayleigh-Ritz step) -------
// The k lowest eigenpairs of a symmetric m x m matrix in one workgroup
// (rocSOLVER's syevd spends ~1.4 ms per call at m = 87 in dozens of
// latency-bound launches; this is one launch with every array in LDS):
//  1. Householder tridiagonalisation T = H^T A H (reflectors in the LAPACK
//     dsytd2 convention: v[0] = 1, H_j = I - tau_j v_j v_j^T);
//  2. the k lowest eigenvalues of T by Sturm-count multisection, 256 / k
//     counts per level and step;
//  3. eigenvectors of T by inverse iteration, one thread per level, four
//     iterations from a fixed pseudo-random start.  T - lambda I = L D L^T
//     without pivoting (the Sturm recurrence; pivots below eps ||T|| are
//     raised to that size), the inverse pivots kept for the later solves.
//     Levels closer than 1e-9 ||T|| (degenerate to working precision)
//     iterate together in one thread, each vector orthogonalised against the
//     earlier ones of its group every iteration (LAPACK dstein's scheme);
//  4. one modified Gram-Schmidt pass over the k vectors: independently
//     computed vectors of levels a gap g apart overlap by ~eps ||T|| / g,
//     and removing that overlap changes each residual by ~eps ||T|| only;
//  5. back-transform z = H_0 ... H_{m-3} x.
// LDS: A (m x (m+1)), the k vectors and the k inverse-pivot rows (k x m
// each) are dynamic; the fast path needs m (m+1) + 2 k m <= kEigLdsDoubles.
constexpr int kEigMax = 96;
constexpr int kEigLdsDoubles = 19000;  // + ~10 KB of static LDS <= 160 KB
constexpr int kEigThreads = 1024;

__device__ inline double eig_wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
// sum over an aligned group of G lanes (G a power of two <= 64)
__device__ inline double eig_group_sum(double v, int G) {
    for (int o = G >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__global__ __launch_bounds__(kEigThreads) void k_syev_small(int m, int k, const double* __restrict__ a, int64_t lda,
                                                            double* __restrict__ z, int64_t ldz,
                                                            double* __restrict__ w) {
    constexpr int NT = kEigThreads;
    extern __shared__ double eig_dyn[];
    const int LD = m + 1;
    double* A = eig_dyn;        // A[i * LD + j]
    double* X = A + m * LD;     // eigenvectors of T, k x m
    double* DI = X + k * m;     // inverse LDL^T pivots, one row per group
    __shared__ double vb[kEigMax], pb[kEigMax];
    __shared__ double d[kEigMax], e[kEigMax], e2[kEigMax], tau[kEigMax], lam[kEigMax];
    __shared__ double blo[kEigMax], bhi[kEigMax], xs[256];
    __shared__ int cnt[256];
    __shared__ double sc[4];
    __shared__ int nit_s;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int q = tid; q < m * m; q += NT) {
        const int i = q % m, j = q / m;  // A[i][j] from the lower triangle
        A[i * LD + j] = i >= j ? a[(int64_t)j * lda + i] : a[(int64_t)i * lda + j];
    }
    __syncthreads();
    // 1. tridiagonalisation
    for (int j = 0; j + 2 < m; ++j) {
        const int r = m - j - 1;
        if (wv == 0) {
            double s = 0.0;
            for (int i = 1 + lane; i < r; i += 64) {
                const double x = A[(j + 1 + i) * LD + j];
                s += x * x;
            }
            s = eig_wave_sum(s);
            const double alpha = A[(j + 1) * LD + j];
            double t = 0.0, beta = alpha, scale = 0.0;
            if (s > 0.0) {
                beta = -copysign(sqrt(alpha * alpha + s), alpha);
                t = (beta - alpha) / beta;
                scale = 1.0 / (alpha - beta);
            }
            for (int i = lane; i < r; i += 64) vb[i] = i == 0 ? 1.0 : A[(j + 1 + i) * LD + j] * scale;
            if (lane == 0) {
                tau[j] = t;
                d[j] = A[j * LD + j];
                e[j] = beta;
            }
        }
        __syncthreads();
        const double t = tau[j];
        if (t != 0.0) {
            {  // p = tau A22 v, eight lanes per row
                const int row = tid >> 3, part = tid & 7;
                double s0 = 0.0, s1 = 0.0;
                if (row < r) {
                    const double* rp = &A[(j + 1 + row) * LD + j + 1];
                    int l = part;
                    for (; l + 8 < r; l += 16) {
                        s0 += rp[l] * vb[l];
                        s1 += rp[l + 8] * vb[l + 8];
                    }
                    if (l < r) s0 += rp[l] * vb[l];
                }
                const double s = eig_group_sum(s0 + s1, 8);
                if (row < r && part == 0) pb[row] = t * s;
            }
            __syncthreads();
            if (wv == 0) {
                double s = 0.0;
                for (int i = lane; i < r; i += 64) s += pb[i] * vb[i];
                s = eig_wave_sum(s);
                const double K = 0.5 * t * s;
                for (int i = lane; i < r; i += 64) pb[i] -= K * vb[i];
            }
            __syncthreads();
            {  // A22 -= v w^T + w v^T: a thread owns columns tl, tl + 32, tl + 64 (v, w in registers)
                const int ti = tid >> 5, tl = tid & 31;
                const bool c1 = tl + 32 < r, c2 = tl + 64 < r;
                const double v0 = tl < r ? vb[tl] : 0.0, w0 = tl < r ? pb[tl] : 0.0;
                const double v1 = c1 ? vb[tl + 32] : 0.0, w1 = c1 ? pb[tl + 32] : 0.0;
                const double v2 = c2 ? vb[tl + 64] : 0.0, w2 = c2 ? pb[tl + 64] : 0.0;
                if (tl < r)
                    for (int i = ti; i < r; i += 32) {
                        const double vi = vb[i], wi = pb[i];
                        double* rp = &A[(j + 1 + i) * LD + j + 1 + tl];
                        rp[0] -= vi * w0 + wi * v0;
                        if (c1) rp[32] -= vi * w1 + wi * v1;
                        if (c2) rp[64] -= vi * w2 + wi * v2;
                    }
            }
        }
        for (int i = tid; i < r; i += NT) A[(j + 1 + i) * LD + j] = vb[i];  // reflector, v[0] = 1
        __syncthreads();
    }
    if (tid == 0) {
        for (int i = m >= 2 ? m - 2 : 0; i < m; ++i) d[i] = A[i * LD + i];
        if (m >= 2) e[m - 2] = A[(m - 1) * LD + m - 2];
        double lo = d[0], hi = d[0], emax2 = 0.0;
        for (int i = 0; i < m; ++i) {
            const double rad = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i + 1 < m ? fabs(e[i]) : 0.0);
            lo = fmin(lo, d[i] - rad);
            hi = fmax(hi, d[i] + rad);
            if (i + 1 < m) {
                e2[i] = e[i] * e[i];
                emax2 = fmax(emax2, e2[i]);
            }
        }
        const double eps = 1.1102230246251565e-16;
        // scale of T, floored so that tolerances and pivots derived from it stay normal numbers
        const double tn = fmax(fmax(fabs(lo), fabs(hi)), 1e-150);
        lo -= 4.0 * eps * tn;
        hi += 4.0 * eps * tn;
        sc[0] = lo;
        sc[1] = hi;
        sc[2] = tn;
        sc[3] = 2.2250738585072014e-308 * fmax(1.0, emax2);  // pivmin
        const int groups = 256 / k;
        const double ratio = (hi - lo) / (eps * tn);
        // bounded for any input (non-finite entries give garbage, never a long loop)
        const double steps = ratio > 1.0 ? ceil(log(ratio) / log((double)(groups + 1))) + 2.0 : 0.0;
        nit_s = steps >= 0.0 && steps <= 128.0 ? (int)steps : (steps > 128.0 ? 128 : 0);
    }
    __syncthreads();
    // 2. multisection for eigenvalue indices 0..k-1: 256 / k points per level and
    //    step (more counting threads only contend for the FP64 pipes)
    const double tn = sc[2], pivmin = sc[3];
    {
        const int groups = 256 / k;
        const int idx = tid % k, sgrp = tid / k;
        const bool act = sgrp < groups;
        if (tid < k) {
            blo[tid] = sc[0];
            bhi[tid] = sc[1];
        }
        __syncthreads();
        const int nit = nit_s;
        for (int it = 0; it < nit; ++it) {
            if (act) {
                const double L = blo[idx], U = bhi[idx];
                const double x = L + (U - L) * (double)(sgrp + 1) / (double)(groups + 1);
                double q = d[0] - x;
                if (fabs(q) <= pivmin) q = -pivmin;
                int c = q < 0.0;
                for (int i = 1; i < m; ++i) {
                    // e2 / q with the hardware reciprocal and one Newton step (relative
                    // error ~1e-16: the count stays that of a matrix within rounding of T)
                    double rq = __builtin_amdgcn_rcp(q);
                    rq = fma(rq, fma(-q, rq, 1.0), rq);
                    q = d[i] - x - e2[i - 1] * rq;
                    if (fabs(q) <= pivmin) q = -pivmin;
                    c += q < 0.0;
                }
                xs[tid] = x;
                cnt[tid] = c;
            }
            __syncthreads();
            if (tid < k) {
                double L = blo[tid], U = bhi[tid];
                for (int s = 0; s < groups; ++s) {
                    const int t2 = s * k + tid;
                    if (cnt[t2] <= tid) L = fmax(L, xs[t2]);
                    else U = fmin(U, xs[t2]);
                }
                blo[tid] = L;
                bhi[tid] = U;
            }
            __syncthreads();
        }
        if (tid < k) lam[tid] = 0.5 * (blo[tid] + bhi[tid]);
        __syncthreads();
    }
    // 3. inverse iteration, one thread per group of (numerically) degenerate levels
    const double eps = 1.1102230246251565e-16;
    const double gtol = 1e-9 * tn, pertol = 10.0 * eps * tn, tiny = eps * tn;
    if (tid < k && (tid == 0 || lam[tid] - lam[tid - 1] > gtol)) {
        double* di = DI + tid * m;
        int c1 = tid + 1;
        while (c1 < k && lam[c1] - lam[c1 - 1] <= gtol) ++c1;
        double prev = 0.0;
        for (int jj = tid; jj < c1; ++jj) {
            double sh = lam[jj];
            if (jj > tid && sh - prev < pertol) sh = prev + pertol;
            prev = sh;
            double* x = X + jj * m;
            uint64_t st = 0x9e3779b97f4a7c15ULL * (uint64_t)(jj + 1);
            for (int i = 0; i < m; ++i) {
                st ^= st >> 12;
                st ^= st << 25;
                st ^= st >> 27;
                x[i] = (double)((st * 0x2545F4914F6CDD1DULL) >> 11) * (2.0 / 9007199254740992.0) - 1.0;
            }
            double scale = 1.0;  // 1 / max|x| of the previous solve (|x| can reach ~1/tiny)
            for (int itr = 0; itr < 3; ++itr) {
                // forward: L y = scale x (the first pass also factors T - sh I)
                x[0] *= scale;
                if (itr == 0) {
                    double dc = d[0] - sh;
                    if (fabs(dc) < tiny) dc = dc < 0.0 ? -tiny : tiny;
                    di[0] = 1.0 / dc;
                    for (int i = 1; i < m; ++i) {
                        const double l = e[i - 1] * di[i - 1];
                        dc = (d[i] - sh) - l * e[i - 1];
                        if (fabs(dc) < tiny) dc = dc < 0.0 ? -tiny : tiny;
                        di[i] = 1.0 / dc;
                        x[i] = x[i] * scale - l * x[i - 1];
                    }
                } else {
#pragma unroll 4
                    for (int i = 1; i < m; ++i) x[i] = x[i] * scale - (e[i - 1] * di[i - 1]) * x[i - 1];
                }
                // backward: L^T x = D^{-1} y, tracking max|x|
                x[m - 1] *= di[m - 1];
                double xm = fabs(x[m - 1]);
#pragma unroll 4
                for (int i = m - 2; i >= 0; --i) {
                    x[i] = (x[i] - e[i] * x[i + 1]) * di[i];
                    xm = fmax(xm, fabs(x[i]));
                }
                scale = xm > 0.0 ? 1.0 / xm : 1.0;
                if (jj > tid) {  // earlier (normalised) vectors of the group
                    for (int i = 0; i < m; ++i) x[i] *= scale;
                    scale = 1.0;
                    for (int c = tid; c < jj; ++c) {
                        const double* y = X + c * m;
                        double s = 0.0;
                        for (int i = 0; i < m; ++i) s += y[i] * x[i];
                        for (int i = 0; i < m; ++i) x[i] -= s * y[i];
                    }
                }
            }
            double nn = 0.0;
            for (int i = 0; i < m; ++i) {
                x[i] *= scale;
                nn += x[i] * x[i];
            }
            const double inv = 1.0 / sqrt(fmax(nn, 1e-300));
            for (int i = 0; i < m; ++i) x[i] *= inv;
        }
    }
    __syncthreads();
    // 4. modified Gram-Schmidt over the k vectors in level order, G lanes per vector
    int G = 64;
    while (G > 1 && G * k > NT) G >>= 1;
    const int grp = tid / G, gl = tid % G;
    for (int c = 0; c < k; ++c) {
        double* xc = X + c * m;
        if (grp == c) {
            double s = 0.0;
            for (int i = gl; i < m; i += G) s += xc[i] * xc[i];
            s = eig_group_sum(s, G);
            const double inv = 1.0 / sqrt(fmax(s, 1e-300));
            for (int i = gl; i < m; i += G) xc[i] *= inv;
        }
        __syncthreads();
        if (grp > c && grp < k) {
            double* xj = X + grp * m;
            double s = 0.0;
            for (int i = gl; i < m; i += G) s += xc[i] * xj[i];
            s = eig_group_sum(s, G);
            for (int i = gl; i < m; i += G) xj[i] -= s * xc[i];
        }
        __syncthreads();
    }
    // 5. back-transform, G lanes per column (a column's lanes share one wave)
    if (grp < k) {
        double* x = X + grp * m;
        for (int j = m - 3; j >= 0; --j) {
            const double tj = tau[j];
            if (tj == 0.0) continue;
            const int r = m - j - 1;
            double s = 0.0;
            for (int i = gl; i < r; i += G) s += A[(j + 1 + i) * LD + j] * x[j + 1 + i];
            s = tj * eig_group_sum(s, G);
            for (int i = gl; i < r; i += G) x[j + 1 + i] -= s * A[(j + 1 + i) * LD + j];
        }
        // sign convention: the largest component positive
        double best = 0.0;
        int bi = 0;
        for (int i = gl; i < m; i += G)
            if (fabs(x[i]) > best) {
                best = fabs(x[i]);
                bi = i;
            }
        for (int o = G >> 1; o > 0; o >>= 1) {
            const double ob = __shfl_xor(best, o);
            const int oi = __shfl_xor(bi, o);
            if (ob > best || (ob == best && oi < bi)) {
                best = ob;
                bi = oi;
            }
        }
        const double sg = x[bi] < 0.0 ? -1.0 : 1.0;
        for (int i = gl; i < m; i += G) z[(int64_t)grp * ldz + i] = sg * x[i];
        if (gl == 0) w[grp] = lam[grp];
    }
}

// ---- tall-skinny block update on the VALU --------------------------------
// Y = alpha * A C + beta * Y with A: N x m (N grid points), C: m x k small.
// Each thread owns two grid points and KMAX output columns in registers; the
// m x k coefficients are read with wave-uniform addresses (scalar loads), so
// every loaded A value feeds KMAX FMAs with no LDS traffic.  Columns beyond
// KMAX run in further passes over A.  Used for FP32 blocks with k <= 64,
// where rocBLAS' SGEMM tiles waste most of their width.
template <typename T, int KMAX>
__global__ __launch_bounds__(256) void k_update_valu(int64_t n, int m, int kc, T alpha, const T* A,
                                                     int64_t lda, const T* __restrict__ Cm, int ldc, T beta,
                                                     T* Y, int64_t ldy) {
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

int rocblas_for(dft_ctx* c, void** handle) {
    rocblas_handle h;
    if (int rc = blas(c, &h)) return rc;
    *handle = (void*)h;
    return 0;
}

}  // namespace dft

extern "C" {


// Block update Y = alpha A C + beta Y (a tall-skinny GEMM, N x m times m x k).
// k <= 96: the MFMA streaming kernel in one pass when C^T fits the LDS budget;
// k <= 48 beyond it: the same kernel over chunks of the inner dimension
// accumulated into Y (Y apart from A);
// FP32 blocks with k <= 64 and m <= 512 that alias A: the VALU kernel;
// otherwise rocBLAS' 64-bit-index GEMM, through a staging copy when Y overlaps
// A.  FP32 blocks use an FP32 copy of the small coefficient matrix.
int dft_block_update(dft_ctx* c, int prec, int64_t n, int m, int k, double alpha, const void* a, int64_t lda,
                     const double* cm, int64_t ldc, double beta, void* y, int64_t ldy) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, a && cm && y && n > 0 && m > 0 && k >= 0 && lda >= n && ldy >= n && ldc >= m &&
                       (prec == DFT_F64 || prec == DFT_F32),
                "dft_block_update: bad arguments");
    if (k == 0) return DFT_OK;
    // FP32 beyond the FP32 MFMA kernel's LDS budget: the bf16 split kernel (in place, any m)
    if (prec == DFT_F32 && !update_mfma_fits<float>(m, k) && split3_fits(n, k, lda, lda, ldy, a, a, y))
        return update_split3(c, n, m, m, k, alpha, (const float*)a, lda, nullptr, 0, cm, ldc, (float)beta, (float*)y, ldy);
    rocblas_handle h;
    if (int rc = blas(c, &h)) return rc;
    const size_t es = prec == DFT_F64 ? sizeof(double) : sizeof(float);
    // y may overlap a (an in-place update of a basis block): the MFMA kernel and the
    // single-launch VALU kernel handle that; a library GEMM writes through a staging copy
    const char* ab = static_cast<const char*>(a);
    char* yb = static_cast<char*>(y);
    const bool alias = yb < ab + es * (size_t)m * lda && ab < yb + es * (size_t)k * ldy;
    const bool mfma = prec == DFT_F64 ? update_mfma_fits<double>(m, k) : update_mfma_fits<float>(m, k);
    // C beyond the LDS budget: the MFMA kernel over chunks of the inner dimension,
    // accumulating into Y (not in place: a later chunk would read overwritten columns)
    const int mc = mfma || alias ? 0 : (prec == DFT_F64 ? update_chunk<double>(m, k) : update_chunk<float>(m, k));
    const bool valu = prec == DFT_F32 && !mfma && !mc && k <= 64 && m <= 512;
    const bool staged = alias && !mfma && !valu;
    const size_t off = prec == DFT_F32 ? (sizeof(float) * (size_t)m * k + 255) / 256 * 256 : 0;
    const size_t need = off + (staged ? es * (size_t)n * k : 0);
    if (need)
        if (int rc = ensure_scratch(c, need)) return rc;
    void* out = staged ? static_cast<void*>(static_cast<char*>(c->scratch) + off) : y;
    const int64_t ldo = staged ? n : ldy;
    if (staged && beta != 0.0)
        DFT_HIP(c, hipMemcpy2DAsync(out, es * ldo, y, es * ldy, es * n, k, hipMemcpyDeviceToDevice, c->stream));
    rocblas_status st = rocblas_status_success;
    if (prec == DFT_F64) {
        if (mfma)
            return update_mfma<double>(c, n, m, m, k, alpha, (const double*)a, lda, nullptr, 0, cm, (int)ldc, beta,
                                       (double*)y, ldy);
        if (mc) {
            for (int r0 = 0; r0 < m; r0 += mc) {
                const int rows = std::min(mc, m - r0);
                if (int rc = update_mfma<double>(c, n, rows, rows, k, alpha, (const double*)a + (int64_t)r0 * lda, lda,
                                                 nullptr, 0, cm + r0, (int)ldc, r0 == 0 ? beta : 1.0, (double*)y, ldy))
                    return rc;
            }
            return DFT_OK;
        }
        st = rocblas_dgemm_64(h, rocblas_operation_none, rocblas_operation_none, n, k, m, &alpha, (const double*)a, lda,
                              cm, ldc, &beta, (double*)out, ldo);
    } else {
        float* c32 = (float*)c->scratch;
        k_to_f32<<<(unsigned)(((int64_t)m * k + 255) / 256), 256, 0, c->stream>>>(m, k, cm, ldc, c32);
        DFT_CHECK_LAUNCH(c);
        const float al = (float)alpha, be = (float)beta;
        if (mfma) return update_mfma<float>(c, n, m, m, k, al, (const float*)a, lda, nullptr, 0, c32, m, be, (float*)y, ldy);
        if (mc) {
            for (int r0 = 0; r0 < m; r0 += mc) {
                const int rows = std::min(mc, m - r0);
                if (int rc = update_mfma<float>(c, n, rows, rows, k, al, (const float*)a + (int64_t)r0 * lda, lda, nullptr,
                                                0, c32 + r0, m, r0 == 0 ? be : 1.0f, (float*)y, ldy))
                    return rc;
            }
            return DFT_OK;
        }
        if (valu) return update_valu<float>(c, n, m, k, al, (const float*)a, lda, c32, m, be, (float*)y, ldy);
        st = rocblas_sgemm_64(h, rocblas_operation_none, rocblas_operation_none, n, k, m, &al, (const float*)a, lda, c32,
                              m, &be, (float*)out, ldo);
    }
    if (st != rocblas_status_success) return set_err(c, DFT_EHIP, "rocblas gemm failed");
    if (staged)
        DFT_HIP(c, hipMemcpy2DAsync(y, es * ldy, out, es * ldo, es * n, k, hipMemcpyDeviceToDevice, c->stream));
    return DFT_OK;
}

int dft_block_update_in_place(int prec, int m, int k) {
    if (m <= 0 || k <= 0) return 0;
    if (prec == DFT_F64) return update_mfma_fits<double>(m, k) ? 1 : 0;
    // FP32: the FP32 MFMA kernel, the bf16 split kernel (k <= 96) or the VALU kernel
    if (prec == DFT_F32) return update_mfma_fits<float>(m, k) || k <= 96 ? 1 : 0;
    return 0;
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

// The k lowest eigenpairs of the symmetric matrix a (lower triangle read):
// eigenvalues ascending in w[0..k) (w has room for m values), eigenvectors in
// z's first k columns; a is not modified.  One workgroup when m <= 96 and
// m (m + 1) + 2 k m <= 19000 doubles (the LDS it needs); rocSOLVER dsyevd on
// a copy otherwise.
int dft_syev_lowest(dft_ctx* c, int m, int k, const double* a, int64_t lda, double* z, int64_t ldz, double* w) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, a && z && w && m > 0 && k > 0 && k <= m && lda >= m && ldz >= m && m <= 4096 && z != a,
                "dft_syev_lowest: bad arguments");
    const int64_t nd = (int64_t)m * (m + 1) + 2LL * k * m;
    if (m <= kEigMax && nd <= kEigLdsDoubles) {
        const size_t bytes = sizeof(double) * (size_t)nd;
        // the kernel's dynamic LDS limit, raised once per device
        static std::atomic<uint64_t> attr_done{0};
        const uint64_t bit = 1ULL << (c->device & 63);
        if (!(attr_done.load() & bit)) {
            DFT_HIP(c, hipFuncSetAttribute(reinterpret_cast<const void*>(k_syev_small),
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)(sizeof(double) * kEigLdsDoubles)));
            attr_done.fetch_or(bit);
        }
        k_syev_small<<<1, kEigThreads, bytes, c->stream>>>(m, k, a, lda, z, ldz, w);
        DFT_CHECK_LAUNCH(c);
        return DFT_OK;
    }
    if (ldz == lda) {
        DFT_HIP(c, hipMemcpyAsync(z, a, sizeof(double) * (size_t)lda * m, hipMemcpyDeviceToDevice, c->stream));
    } else {
        DFT_HIP(c, hipMemcpy2DAsync(z, sizeof(double) * ldz, a, sizeof(double) * lda, sizeof(double) * m, m,
                                    hipMemcpyDeviceToDevice, c->stream));
    }
    return dft_syev(c, m, z, ldz, w);
}

// Davidson residual block straight from the basis V (N x m) and W = H V:
// R[:, j] = W Q[:, j] - theta[j] V Q[:, j] and rn2[j] = ||R[:, j]||^2 for
// j < k.  One MFMA pass over [W | V] with the 2m x k coefficients
// [Q ; -Q diag(theta)] when they fit the LDS budget (norms from the same
// pass), else (k <= 48) over chunks of those 2m rows (norms from the last
// chunk); otherwise R = W Q, then R += V (-Q theta), then the norms.
int dft_ritz_residual(dft_ctx* c, int prec, int64_t n, int m, int k, const void* v, int64_t ldv, const void* w,
                      int64_t ldw, const double* q, int64_t ldq, const double* theta, void* r, int64_t ldr, double* rn2) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, v && w && q && theta && r && rn2 && n > 0 && m > 0 && k >= 0 && ldv >= n && ldw >= n && ldr >= n &&
                       ldq >= m && (prec == DFT_F64 || prec == DFT_F32),
                "dft_ritz_residual: bad arguments");
    if (k == 0) return DFT_OK;
    const unsigned gc = (unsigned)(((int64_t)2 * m * k + 255) / 256);
    if (prec == DFT_F64 && update_mfma_fits<double>(2 * m, k)) {
        if (int rc = ensure_scratch(c, sizeof(double) * (size_t)2 * m * k)) return rc;
        double* cf = (double*)c->scratch;
        k_ritz_coef<double><<<gc, 256, 0, c->stream>>>(m, k, q, ldq, theta, cf);
        DFT_CHECK_LAUNCH(c);
        return update_mfma<double>(c, n, m, 2 * m, k, 1.0, (const double*)w, ldw, (const double*)v, ldv, cf, 2 * m, 0.0,
                                   (double*)r, ldr, rn2);
    }
    if (prec == DFT_F32 && update_mfma_fits<float>(2 * m, k)) {
        if (int rc = ensure_scratch(c, sizeof(float) * (size_t)2 * m * k)) return rc;
        float* cf = (float*)c->scratch;
        k_ritz_coef<float><<<gc, 256, 0, c->stream>>>(m, k, q, ldq, theta, cf);
        DFT_CHECK_LAUNCH(c);
        return update_mfma<float>(c, n, m, 2 * m, k, 1.0f, (const float*)w, ldw, (const float*)v, ldv, cf, 2 * m, 0.0f,
                                  (float*)r, ldr, rn2);
    }
    // FP32 beyond the LDS budget: one bf16 split pass over [W | V] with its norms
    if (prec == DFT_F32 && split3_fits(n, k, ldw, ldv, ldr, w, v, r)) {
        if (int rc = ensure_scratch(c, sizeof(double) * (size_t)2 * m * k)) return rc;
        double* cf = (double*)c->scratch;
        k_ritz_coef<double><<<gc, 256, 0, c->stream>>>(m, k, q, ldq, theta, cf);
        DFT_CHECK_LAUNCH(c);
        return update_split3(c, n, m, 2 * m, k, 1.0, (const float*)w, ldw, (const float*)v, ldv, cf, 2 * m, 0.0f,
                             (float*)r, ldr, rn2);
    }
    // 2m x k beyond the LDS budget: chunks of the [W | V] rows, the norms from the last chunk's epilogue
    const int mc = prec == DFT_F64 ? update_chunk<double>(2 * m, k) : update_chunk<float>(2 * m, k);
    if (mc) {
        const size_t es = prec == DFT_F64 ? sizeof(double) : sizeof(float);
        if (int rc = ensure_scratch(c, es * (size_t)2 * m * k)) return rc;
        if (prec == DFT_F64) k_ritz_coef<double><<<gc, 256, 0, c->stream>>>(m, k, q, ldq, theta, (double*)c->scratch);
        else k_ritz_coef<float><<<gc, 256, 0, c->stream>>>(m, k, q, ldq, theta, (float*)c->scratch);
        DFT_CHECK_LAUNCH(c);
        for (int r0 = 0; r0 < 2 * m; r0 += mc) {
            const int rows = std::min(mc, 2 * m - r0), r1 = r0 + rows;
            double* nrm = r1 == 2 * m ? rn2 : nullptr;
            // rows r0..r1 of [W | V]: W's rows below m, V's from m
            const int m1 = r0 >= m ? 0 : std::min(rows, m - r0);
            int rc;
            if (prec == DFT_F64) {
                const double* a1 = r0 < m ? (const double*)w + (int64_t)r0 * ldw : (const double*)v + (int64_t)(r0 - m) * ldv;
                const double* a2 = r0 < m ? (const double*)v : nullptr;
                rc = m1 == 0 || m1 == rows
                         ? update_mfma<double>(c, n, rows, rows, k, 1.0, a1, m1 == 0 ? ldv : ldw, nullptr, 0,
                                               (const double*)c->scratch + r0, 2 * m, r0 == 0 ? 0.0 : 1.0, (double*)r, ldr,
                                               nrm)
                         : update_mfma<double>(c, n, m1, rows, k, 1.0, a1, ldw, a2, ldv, (const double*)c->scratch + r0,
                                               2 * m, r0 == 0 ? 0.0 : 1.0, (double*)r, ldr, nrm);
            } else {
                const float* a1 = r0 < m ? (const float*)w + (int64_t)r0 * ldw : (const float*)v + (int64_t)(r0 - m) * ldv;
                const float* a2 = r0 < m ? (const float*)v : nullptr;
                rc = m1 == 0 || m1 == rows
                         ? update_mfma<float>(c, n, rows, rows, k, 1.0f, a1, m1 == 0 ? ldv : ldw, nullptr, 0,
                                              (const float*)c->scratch + r0, 2 * m, r0 == 0 ? 0.0f : 1.0f, (float*)r, ldr,
                                              nrm)
                         : update_mfma<float>(c, n, m1, rows, k, 1.0f, a1, ldw, a2, ldv, (const float*)c->scratch + r0,
                                              2 * m, r0 == 0 ? 0.0f : 1.0f, (float*)r, ldr, nrm);
            }
            if (rc) return rc;
        }
        return DFT_OK;
    }
    // two passes; -Q theta sits past the FP32 coefficient copy dft_block_update keeps at the scratch start
    const size_t off = ((sizeof(float) * (size_t)m * k) + 255) / 256 * 256;
    if (int rc = ensure_scratch(c, off + sizeof(double) * (size_t)m * k)) return rc;
    double* nq = (double*)((char*)c->scratch + off);
    k_neg_q_theta<<<(unsigned)(((int64_t)m * k + 255) / 256), 256, 0, c->stream>>>(m, k, q, ldq, theta, nq);
    DFT_CHECK_LAUNCH(c);
    if (int rc = dft_block_update(c, prec, n, m, k, 1.0, w, ldw, q, ldq, 0.0, r, ldr)) return rc;
    if (int rc = dft_block_update(c, prec, n, m, k, 1.0, v, ldv, nq, m, 1.0, r, ldr)) return rc;
    return dft_column_norms2(c, prec, n, k, r, ldr, rn2);
}

}  // extern "C"
