// gram.hip -- (e) the Davidson Gram / Rayleigh-Ritz products A^T B on the
// CDNA4 matrix cores.
//
// A^T B with A: N x ma, B: N x mb, N (grid points) >> ma, mb (<= a few
// hundred): bandwidth-bound.  Two kernels:
//  * k_gram_stream (ma <= 96, mb <= 32, FP32 data: every Davidson product of
//    the mixed phase): no LDS.  A wave owns a contiguous range of points and
//    all output tiles; an operand lane loads 8 consecutive points of one
//    column (32 bytes) per step straight into MFMA operands.
//  * k_gram (any shape): 96 x 96 output tiles of 4 waves, 32-point K slabs
//    staged through LDS with the next slab prefetched in registers.
// The long dimension N is split across workgroups; partial tiles go to an
// FP64 buffer and two kernels sum them in a fixed order (deterministic).
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
typedef __bf16 bf4 __attribute__((ext_vector_type(4)));

__device__ inline __bf16 to_bf16(float x) { return (__bf16)x; }  // round to nearest even
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u2 __attribute__((ext_vector_type(2)));
// bf16(lo) | bf16(hi) << 16, rounded to nearest even (one v_cvt_pk_bf16_f32)
__device__ __forceinline__ uint32_t pack_bf16x2(float lo, float hi) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){lo, hi}, bf16x2));
}

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

// Workgroup = 4 waves in 2 x 2, each wave 3 x 3 MFMA 16x16 tiles: a 96 x 96
// output tile over one K-chunk.  Only the tile's valid columns (ca of A, cb
// of B) are read: each K-slab (KS rows) is loaded with 16-byte pieces into
// registers while the MFMAs of the previous slab run, then stored to LDS in
// operand order (for SPLIT_BF16 split into hi/lo bf16 on the way in).
// Waves and MFMA sub-tiles wholly outside the valid columns skip their work.
template <int MODE>
struct Slab;
template <>
struct Slab<DFT_GRAM_F64> {
    static constexpr int KS = 32, PAD = 33;  // doubles per column in LDS
};
template <>
struct Slab<DFT_GRAM_F32> {
    static constexpr int KS = 32, PAD = 33;  // floats per column
};
template <>
struct Slab<DFT_GRAM_SPLIT_BF16> {
    static constexpr int KS = 32, PAD = 40;  // bf16 per column (80-byte stride, 16-byte aligned)
};

constexpr int GT = 96;  // output tile edge

template <int MODE>
__global__ __launch_bounds__(256) void k_gram(int64_t n, int ma, int mb, int ntm, int ntiles, int nks,
                                              const typename GramT<MODE>::In* __restrict__ A, int64_t lda,
                                              const typename GramT<MODE>::In* __restrict__ B, int64_t ldb,
                                              int64_t kchunk, double* __restrict__ part) {
    using In = typename GramT<MODE>::In;
    using Acc = typename GramT<MODE>::Acc;
    constexpr int KS = Slab<MODE>::KS, PAD = Slab<MODE>::PAD;
    constexpr bool SPLIT = MODE == DFT_GRAM_SPLIT_BF16;
    using Ld = typename std::conditional<SPLIT, __bf16, In>::type;
    // LDS: [2 operands][GT columns][PAD]; SPLIT keeps hi and lo planes
    __shared__ __attribute__((aligned(16))) Ld lds[(SPLIT ? 2 : 1) * 2 * GT * PAD];
    Ld* ldsA = lds;
    Ld* ldsB = lds + GT * PAD;
    Ld* ldsAl = lds + 2 * GT * PAD;  // SPLIT lo planes
    Ld* ldsBl = lds + 3 * GT * PAD;

    // XCD-aware mapping: the ntiles tiles of one K-chunk run on one XCD
    const int b = blockIdx.x;
    const int xcd = b & 7, slot = b >> 3;
    const int tile = slot % ntiles;
    const int ks = (slot / ntiles) * 8 + xcd;
    const int i0 = (tile % ntm) * GT, j0 = (tile / ntm) * GT;
    const int ca = min(GT, ma - i0), cb = min(GT, mb - j0);  // valid columns of this tile
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wi = (w >> 1) * 48, wj = (w & 1) * 48;  // wave's 48 x 48 block inside the tile
    const int r16 = lane & 15, kq = lane >> 4;
    const int64_t kbeg = (int64_t)ks * kchunk, kend = min(n, kbeg + kchunk);
    const bool wave_live = wi < ca && wj < cb;

    Acc acc[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[a][c] = Acc{0, 0, 0, 0};

    // invalid columns stay zero for the whole loop
    for (int q = threadIdx.x; q < (SPLIT ? 4 : 2) * GT * PAD; q += 256) {
        const int plane = q / (GT * PAD), col = (q % (GT * PAD)) / PAD;
        const bool isA = (plane & 1) == 0;
        if (isA ? col >= ca : col >= cb) lds[q] = Ld(0);
    }

    // slab pieces: 16 bytes = EPV elements; PPC pieces per column; NP pieces per thread
    constexpr int EPV = 16 / sizeof(In);
    constexpr int PPC = KS / EPV;
    constexpr int NP = (2 * GT * PPC + 255) / 256;
    const int npieces = (ca + cb) * PPC;
    In reg[NP][EPV];

    auto load = [&](int64_t k0) {
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            const int q = threadIdx.x + 256 * p;
            if (q >= npieces) break;
            const int cc = q / PPC, lp = q % PPC;
            const bool isA = cc < ca;
            const In* src = isA ? A + (int64_t)(i0 + cc) * lda : B + (int64_t)(j0 + cc - ca) * ldb;
            const int64_t kk = k0 + lp * EPV;
            if (kk + EPV <= kend && ((((uintptr_t)(src + kk)) & 15) == 0)) {
                if constexpr (sizeof(In) == 8) {
                    const double2 v = *reinterpret_cast<const double2*>(src + kk);
                    reg[p][0] = v.x;
                    reg[p][1] = v.y;
                } else {
                    const float4 v = *reinterpret_cast<const float4*>(src + kk);
                    reg[p][0] = v.x;
                    reg[p][1] = v.y;
                    reg[p][2] = v.z;
                    reg[p][3] = v.w;
                }
            } else {
#pragma unroll
                for (int e = 0; e < EPV; ++e) reg[p][e] = kk + e < kend ? src[kk + e] : In(0);
            }
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            const int q = threadIdx.x + 256 * p;
            if (q >= npieces) break;
            const int cc = q / PPC, lp = q % PPC;
            const bool isA = cc < ca;
            const int lc = isA ? cc : cc - ca;
            Ld* dst = (isA ? ldsA : ldsB) + lc * PAD + lp * EPV;
            if constexpr (SPLIT) {
                Ld* dstl = (isA ? ldsAl : ldsBl) + lc * PAD + lp * EPV;
#pragma unroll
                for (int e = 0; e < EPV; ++e) {
                    const __bf16 h = (__bf16)reg[p][e];
                    dst[e] = h;
                    dstl[e] = (__bf16)(reg[p][e] - (float)h);
                }
            } else {
#pragma unroll
                for (int e = 0; e < EPV; ++e) dst[e] = reg[p][e];
            }
        }
    };

    if (kbeg < kend) load(kbeg);
    for (int64_t k0 = kbeg; k0 < kend; k0 += KS) {
        __syncthreads();  // the previous slab's MFMA reads are done
        store();
        __syncthreads();
        if (k0 + KS < kend) load(k0 + KS);  // in flight during this slab's MFMAs
        if (!wave_live) continue;
        // ---- MFMA over the slab
        if constexpr (SPLIT) {
            bf8 ah[3], al[3], bh[3], bl[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int col = wi + 16 * a + r16;
                ah[a] = *reinterpret_cast<const bf8*>(ldsA + col * PAD + 8 * kq);
                al[a] = *reinterpret_cast<const bf8*>(ldsAl + col * PAD + 8 * kq);
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int col = wj + 16 * c + r16;
                bh[c] = *reinterpret_cast<const bf8*>(ldsB + col * PAD + 8 * kq);
                bl[c] = *reinterpret_cast<const bf8*>(ldsBl + col * PAD + 8 * kq);
            }
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                if (wi + 16 * a >= ca) continue;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    if (wj + 16 * c >= cb) continue;
                    acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], bh[c], acc[a][c], 0, 0, 0);
                    acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bl[c], acc[a][c], 0, 0, 0);
                    acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bh[c], acc[a][c], 0, 0, 0);
                }
            }
        } else {
#pragma unroll
            for (int k4 = 0; k4 < KS; k4 += 4) {
                In xa[3], xb[3];
#pragma unroll
                for (int a = 0; a < 3; ++a) xa[a] = ldsA[(wi + 16 * a + r16) * PAD + k4 + kq];
#pragma unroll
                for (int c = 0; c < 3; ++c) xb[c] = ldsB[(wj + 16 * c + r16) * PAD + k4 + kq];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    if (wi + 16 * a >= ca) continue;
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        if (wj + 16 * c >= cb) continue;
                        if constexpr (MODE == DFT_GRAM_F64)
                            acc[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[a], xb[c], acc[a][c], 0, 0, 0);
                        else
                            acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa[a], xb[c], acc[a][c], 0, 0, 0);
                    }
                }
            }
        }
    }
    // partial tile -> part[ks][j*ma + i] (K-chunks past the end write zeros)
    double* P = part + (int64_t)ks * ma * mb;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int j = j0 + wj + 16 * c + r16;
            if (j >= mb) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + wi + 16 * a + res_row<MODE == DFT_GRAM_F64>(kq, r);
                if (i < ma) P[(int64_t)j * ma + i] = (double)acc[a][c][r];
            }
        }
}

// First stage of the K-split reduction: segment s of `step` partials -> part2[s]
// (enough workgroups to keep many loads in flight; fixed summation order).
// Streaming Gram for small outputs (SPLIT_BF16 or F32 on FP32 data): no LDS
// slabs.  Lane (c, q) = (l % 16, l / 16) loads points p + 8q .. p + 8q + 7 of
// output row (A column) 16 ta + c and of output column (B column) 16 tb + c:
// for SPLIT that is the K slice of one v_mfma_f32_16x16x32_bf16 (hi.hi +
// hi.lo + lo.hi); for F32 MFMA t of eight v_mfma_f32_16x16x4_f32 takes point
// p + 8q + t.  The 4 waves' tiles are summed in LDS and one FP64 partial per
// workgroup goes to part[blockIdx.x * ma * mb + j * ma + i].
template <int MODE, int TA, int TB>
__global__ __launch_bounds__(256) void k_gram_stream(int64_t n, int ma, int mb, const float* __restrict__ A,
                                                     int64_t lda, const float* __restrict__ B, int64_t ldb, int64_t kw,
                                                     double* __restrict__ part) {
    constexpr bool SPLIT = MODE == DFT_GRAM_SPLIT_BF16;
    __shared__ float red[4][TA * TB * 256];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r16 = lane & 15, q = lane >> 4;
    const int64_t k0 = ((int64_t)blockIdx.x * 4 + w) * kw, k1 = min(n, k0 + kw);
    const float* ap[TA];
    const float* bp[TB];
    bool av[TA], bv[TB];
#pragma unroll
    for (int t = 0; t < TA; ++t) {
        const int col = 16 * t + r16;
        av[t] = col < ma;
        ap[t] = A + (int64_t)(av[t] ? col : 0) * lda;
    }
#pragma unroll
    for (int t = 0; t < TB; ++t) {
        const int col = 16 * t + r16;
        bv[t] = col < mb;
        bp[t] = B + (int64_t)(bv[t] ? col : 0) * ldb;
    }
    f4 acc[TA][TB];
#pragma unroll
    for (int a = 0; a < TA; ++a)
#pragma unroll
        for (int b = 0; b < TB; ++b) acc[a][b] = f4{0, 0, 0, 0};
    for (int64_t p = k0; p < k1; p += 32) {
        const int64_t pp = p + 8 * q;
        const bool full = pp + 8 <= k1;
        float xa[TA][8], xb[TB][8];
        auto load8 = [&](const float* src, bool valid, float* x) {
            if (valid && full) {
                const float4 u = *reinterpret_cast<const float4*>(src + pp);
                const float4 v = *reinterpret_cast<const float4*>(src + pp + 4);
                x[0] = u.x; x[1] = u.y; x[2] = u.z; x[3] = u.w;
                x[4] = v.x; x[5] = v.y; x[6] = v.z; x[7] = v.w;
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) x[e] = (valid && pp + e < k1) ? src[pp + e] : 0.0f;
            }
        };
#pragma unroll
        for (int t = 0; t < TA; ++t) load8(ap[t], av[t], xa[t]);
#pragma unroll
        for (int t = 0; t < TB; ++t) load8(bp[t], bv[t], xb[t]);
        if constexpr (SPLIT) {
            bf8 bh[TB], bl[TB];
#pragma unroll
            for (int t = 0; t < TB; ++t)
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const __bf16 h = to_bf16(xb[t][e]);
                    bh[t][e] = h;
                    bl[t][e] = to_bf16(xb[t][e] - (float)h);
                }
#pragma unroll
            for (int a = 0; a < TA; ++a) {
                bf8 ah, al;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const __bf16 h = to_bf16(xa[a][e]);
                    ah[e] = h;
                    al[e] = to_bf16(xa[a][e] - (float)h);
                }
#pragma unroll
                for (int b = 0; b < TB; ++b) {
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[b], acc[a][b], 0, 0, 0);
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[b], acc[a][b], 0, 0, 0);
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[b], acc[a][b], 0, 0, 0);
                }
            }
        } else {
#pragma unroll
            for (int a = 0; a < TA; ++a)
#pragma unroll
                for (int b = 0; b < TB; ++b)
#pragma unroll
                    for (int t = 0; t < 8; ++t)
                        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa[a][t], xb[b][t], acc[a][b], 0, 0, 0);
        }
    }
    // sum the 4 waves, then one FP64 partial per output element
#pragma unroll
    for (int a = 0; a < TA; ++a)
#pragma unroll
        for (int b = 0; b < TB; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) red[w][(a * TB + b) * 256 + (4 * q + r) * 16 + r16] = acc[a][b][r];
    __syncthreads();
    double* P = part + (int64_t)blockIdx.x * ma * mb;
    for (int e = threadIdx.x; e < TA * TB * 256; e += 256) {
        const int tile = e >> 8, a = tile / TB, b = tile % TB;
        const int i = 16 * a + ((e & 255) >> 4), j = 16 * b + (e & 15);
        if (i < ma && j < mb)
            P[(int64_t)j * ma + i] = ((double)red[0][e] + (double)red[1][e]) + ((double)red[2][e] + (double)red[3][e]);
    }
}

// Wide Gram for FP32 data on the bf16 matrix cores (SPLIT_BF16: x = hi + lo,
// three products; SPLIT6_BF16: x = x1 + x2 + x3, six products, FP32-level),
// outputs up to 288 x 96.  One workgroup of 12 waves per CU computes the
// whole output over its share of the points (every G-th 32-point step), so
// each operand value is read from HBM once and split into bf16 pieces once.  Per step of 32 KS points the threads
// load the step's A and B columns in 16-byte chunks (8 KS threads per column:
// 128 KS contiguous bytes), split them and write the pieces to LDS in MFMA
// operand order slab[piece][slice][strip][lane] (strip = 16 columns, A's
// strips first; slice = 32 points).  Two LDS buffers: step s + 1 is split and
// stored while step s's products run, and the loads of step s + 3 are in
// flight meanwhile.  The waves form a 6 x 2 grid: wave (wa, wb) owns A
// strips wa SA .. + SA - 1 and B strips wb SB .. + SB - 1, A's pieces held in
// registers across its B strips.  Output element (i, j) of tile (a, b) sits
// in lane 16 (i % 16 / 4) + j % 16, register i % 4 (the f32 layout above).
constexpr int kWideWaves = 12, kWideWA = 6, kWideWB = 2;
constexpr int kWideMaxTA = 18, kWideMaxTB = 6;  // ma <= 288, mb <= 96
constexpr int kWideStrips = kWideMaxTA + kWideMaxTB;
constexpr int kWideKS = 1;                                                               // 32-point slices per step
constexpr int kWideChunks = kWideStrips * 16 * 8 * kWideKS / (64 * kWideWaves);         // 16-byte chunks per thread

template <int MODE, int SA, int SB>
__global__ __launch_bounds__(64 * kWideWaves) void k_gram_wide(int64_t n, int ma, int mb, const float* __restrict__ A,
                                                                int64_t lda, const float* __restrict__ B, int64_t ldb,
                                                                double* __restrict__ part) {
    constexpr int P = MODE == DFT_GRAM_SPLIT6_BF16 ? 3 : 2;  // bf16 pieces per value
    constexpr int KS = kWideKS, SP = 32 * KS, CPC = 8 * KS;  // slices, points and chunks per column per step
    constexpr int NT = 64 * kWideWaves;
    constexpr int PIECE = KS * kWideStrips * 64 * 8;  // bf16 per piece plane
    constexpr int BUF = P * PIECE;                    // bf16 per step buffer
    __shared__ __attribute__((aligned(16))) __bf16 sl[2 * BUF];  // two step buffers
    const int ta = (ma + 15) >> 4, tb = (mb + 15) >> 4;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r16 = lane & 15, q = lane >> 4;
    const int wa = w % kWideWA, wb = w / kWideWA;
    // steps: workgroup b takes the whole steps b, b + G, b + 2G, ... (G = gridDim.x), so
    // at any time the workgroups read neighbouring segments of each column; the
    // partial last step, if any, goes to the workgroup whose turn it is
    const int64_t G = gridDim.x, total_full = n / SP;
    const int64_t nfull = total_full > blockIdx.x ? (total_full - blockIdx.x + G - 1) / G : 0;
    const int64_t first = (int64_t)blockIdx.x * SP, stride = G * SP;
    const int64_t ptail = total_full * SP;
    const bool tail = ptail < n && total_full % G == blockIdx.x;

    // this thread's chunks: chunk c = threadIdx.x + NT u -> column cc = c / CPC of
    // the strip list, points 4 h .. + 3 of the step (h = c % CPC).  NT is a
    // multiple of 16 CPC, so h is the same for every u and cc = cc0 + (NT / CPC) u
    // moves by whole strips: the LDS offset is linear in u.  A chunk of a
    // padding column (or past the strip list) reads column 0 of A and stores
    // zeros; slots past the strip list are never read.
    static_assert(NT % (16 * CPC) == 0, "chunk map");
    constexpr int CSTEP = NT / CPC;  // columns between a thread's chunks
    const int h = threadIdx.x % CPC, cc0 = threadIdx.x / CPC;
    const int off = 4 * h;  // the chunks' first point within the step
    // operand lane ci + 16 (hh / 2) of slice ks holds points 8 (hh / 2) .. + 7; a chunk is its half hh % 2
    const int dst0 = (((h >> 3) * kWideStrips + (cc0 >> 4)) * 64 + (cc0 & 15) + 16 * ((h & 7) >> 1)) * 8 + 4 * (h & 1);
    constexpr int DSTEP = CSTEP / 16 * 64 * 8;  // bf16 between a thread's chunks in a piece plane
    const float* src[kWideChunks];
    bool colv[kWideChunks];
#pragma unroll
    for (int u = 0; u < kWideChunks; ++u) {
        const int cc = cc0 + CSTEP * u;
        const int s = cc >> 4, ci = cc & 15;
        const bool isA = s < ta;
        const int col = 16 * (isA ? s : s - ta) + ci;
        colv[u] = s < ta + tb && col < (isA ? ma : mb);
        src[u] = (colv[u] ? (isA ? A + (int64_t)col * lda : B + (int64_t)col * ldb) : A) + off;
    }
    float raw0[kWideChunks][4], raw1[kWideChunks][4];
    auto load_full = [&](float (&raw)[kWideChunks][4], int64_t p) {  // every point of the step is below k1
#pragma unroll
        for (int u = 0; u < kWideChunks; ++u) {
            // padding columns are zeroed when the step is stored, so no
            // instruction touches the loaded registers before then
            const float4 v = *reinterpret_cast<const float4*>(src[u] + p);
            raw[u][0] = v.x;
            raw[u][1] = v.y;
            raw[u][2] = v.z;
            raw[u][3] = v.w;
        }
    };
    auto load_tail = [&](float (&raw)[kWideChunks][4], int64_t p) {
#pragma unroll
        for (int u = 0; u < kWideChunks; ++u)
#pragma unroll
            for (int e = 0; e < 4; ++e) raw[u][e] = (colv[u] && p + off + e < n) ? src[u][p + e] : 0.0f;
    };

    f4 acc[SA][SB];
#pragma unroll
    for (int a = 0; a < SA; ++a)
#pragma unroll
        for (int b = 0; b < SB; ++b) acc[a][b] = f4{0, 0, 0, 0};
    // operand strips of this wave, clamped into the strip list (tiles past the
    // output are computed on valid data and never stored)
    int sa_[SA], sb_[SB];
#pragma unroll
    for (int a = 0; a < SA; ++a) sa_[a] = min(wa * SA + a, ta - 1);
#pragma unroll
    for (int b = 0; b < SB; ++b) sb_[b] = ta + min(wb * SB + b, tb - 1);
    constexpr int PV = PIECE / 8;  // bf8 per piece plane

    auto store = [&](const float (&raw)[kWideChunks][4], int buf) {  // split a step's values into LDS buffer buf
#pragma unroll
        for (int u = 0; u < kWideChunks; ++u) {
            // two values per conversion; the packed pair's halves are the
            // rounded values' top bits
            u2 x1, x2, x3;
#pragma unroll
            for (int e = 0; e < 4; e += 2) {
                const float v0 = colv[u] ? raw[u][e] : 0.0f, v1 = colv[u] ? raw[u][e + 1] : 0.0f;
                const uint32_t h = pack_bf16x2(v0, v1);
                const float r0 = v0 - __uint_as_float(h << 16), r1 = v1 - __uint_as_float(h & 0xffff0000u);
                const uint32_t g = pack_bf16x2(r0, r1);
                x1[e / 2] = h;
                x2[e / 2] = g;
                if constexpr (P == 3)
                    x3[e / 2] = pack_bf16x2(r0 - __uint_as_float(g << 16), r1 - __uint_as_float(g & 0xffff0000u));
            }
            __bf16* d = sl + buf * BUF + dst0 + DSTEP * u;
            *reinterpret_cast<u2*>(d) = x1;
            *reinterpret_cast<u2*>(d + PIECE) = x2;
            if constexpr (P == 3) *reinterpret_cast<u2*>(d + 2 * PIECE) = x3;
        }
    };
    auto products = [&](int buf) {  // this wave's tiles over the step in LDS buffer buf
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const bf8* sk = reinterpret_cast<const bf8*>(sl + buf * BUF) + ks * kWideStrips * 64 + lane;
            bf8 av[SA][P];
#pragma unroll
            for (int a = 0; a < SA; ++a)
#pragma unroll
                for (int pc = 0; pc < P; ++pc) av[a][pc] = sk[pc * PV + sa_[a] * 64];
#pragma unroll
            for (int b = 0; b < SB; ++b) {
                bf8 bv[P];
#pragma unroll
                for (int pc = 0; pc < P; ++pc) bv[pc] = sk[pc * PV + sb_[b] * 64];
                // the SA tiles' chains interleaved, smallest terms first:
                // x1 y1 + (x1 y2 + x2 y1) [+ (x1 y3 + x2 y2 + x3 y1)]
                if constexpr (P == 3) {
#pragma unroll
                    for (int a = 0; a < SA; ++a) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a][2], bv[0], acc[a][b], 0, 0, 0);
#pragma unroll
                    for (int a = 0; a < SA; ++a) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a][0], bv[2], acc[a][b], 0, 0, 0);
#pragma unroll
                    for (int a = 0; a < SA; ++a) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a][1], bv[1], acc[a][b], 0, 0, 0);
                }
#pragma unroll
                for (int a = 0; a < SA; ++a) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a][1], bv[0], acc[a][b], 0, 0, 0);
#pragma unroll
                for (int a = 0; a < SA; ++a) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a][0], bv[1], acc[a][b], 0, 0, 0);
#pragma unroll
                for (int a = 0; a < SA; ++a) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a][0], bv[0], acc[a][b], 0, 0, 0);
            }
        }
    };

    // Whole steps with two LDS buffers and one barrier per step: at step s the
    // workgroup stores step s + 1 (raw set (s + 1) % 2) into the other buffer,
    // reloads that raw set with step s + 3 and runs step s's products; then the
    // partial step, if any.
    int64_t st = 0;
    if (nfull > 0) {
        load_full(raw0, first);
        if (nfull > 1) load_full(raw1, first + stride);
        store(raw0, 0);
        if (nfull > 2) load_full(raw0, first + 2 * stride);
        __syncthreads();
    }
    for (; st + 4 < nfull; st += 2) {  // no conditions: a straight-line body for the wait-count pass
        const int64_t p = first + st * stride;
        store(raw1, 1);
        load_full(raw1, p + 3 * stride);
        products(0);
        __syncthreads();
        store(raw0, 0);
        load_full(raw0, p + 4 * stride);
        products(1);
        __syncthreads();
    }
    for (; st < nfull; st += 2) {
        const int64_t p = first + st * stride;
        if (st + 1 < nfull) store(raw1, 1);
        if (st + 3 < nfull) load_full(raw1, p + 3 * stride);
        products(0);
        __syncthreads();
        if (st + 1 < nfull) {
            if (st + 2 < nfull) store(raw0, 0);
            if (st + 4 < nfull) load_full(raw0, p + 4 * stride);
            products(1);
            __syncthreads();
        }
    }
    if (tail) {
        load_tail(raw0, ptail);
        store(raw0, 0);
        __syncthreads();
        products(0);
    }
    double* P0 = part + (int64_t)blockIdx.x * ma * mb;
#pragma unroll
    for (int a = 0; a < SA; ++a)
#pragma unroll
        for (int b = 0; b < SB; ++b) {
            const int j = 16 * (wb * SB + b) + r16;
            if (wa * SA + a >= ta || wb * SB + b >= tb || j >= mb) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = 16 * (wa * SA + a) + 4 * q + r;
                if (i < ma) P0[(int64_t)j * ma + i] = (double)acc[a][b][r];
            }
        }
}

__global__ __launch_bounds__(256) void k_gram_reduce1(const double* __restrict__ part, int nks, int64_t tot, int step,
                                                      double* __restrict__ part2) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int seg = blockIdx.y;
    if (e >= tot) return;
    const int k0 = seg * step, k1 = min(nks, k0 + step);
    double s = 0.0;
    for (int k = k0; k < k1; ++k) s += part[(int64_t)k * tot + e];
    part2[(int64_t)seg * tot + e] = s;
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

template <int MODE, int TA, int TB>
void launch_stream(dft_ctx* c, unsigned nblk, int64_t n, int ma, int mb, const float* a, int64_t lda, const float* b,
                   int64_t ldb, int64_t kw) {
    k_gram_stream<MODE, TA, TB><<<nblk, 256, 0, c->stream>>>(n, ma, mb, a, lda, b, ldb, kw, c->partials);
}

template <int MODE>
void launch_stream_ta(dft_ctx* c, unsigned nblk, int64_t n, int ma, int mb, const float* a, int64_t lda, const float* b,
                      int64_t ldb, int64_t kw) {
    const int ta = (ma + 15) / 16, tb = (mb + 15) / 16;
#define DFT_GS(TA)                                                                                     \
    if (tb == 1) launch_stream<MODE, TA, 1>(c, nblk, n, ma, mb, a, lda, b, ldb, kw);                   \
    else launch_stream<MODE, TA, 2>(c, nblk, n, ma, mb, a, lda, b, ldb, kw)
    switch (ta) {
        case 1: DFT_GS(1); break;
        case 2: DFT_GS(2); break;
        case 3: DFT_GS(3); break;
        case 4: DFT_GS(4); break;
        case 5: DFT_GS(5); break;
        default: DFT_GS(6); break;
    }
#undef DFT_GS
}

int gram_stream(dft_ctx* c, int mode, int64_t n, int ma, int mb, double alpha, const float* a, int64_t lda,
                const float* b, int64_t ldb, double beta, double* out, int64_t ldo) {
    // two workgroups per CU, each wave a contiguous run of 32-point steps
    const int64_t waves = 8LL * c->num_cus;
    int64_t kw = (n + waves - 1) / waves;
    kw = std::max<int64_t>(32, (kw + 31) / 32 * 32);
    const int64_t nblk = (n + 4 * kw - 1) / (4 * kw);
    const int64_t tot = (int64_t)ma * mb;
    const int nseg = (int)std::min<int64_t>(nblk, 32), step = (int)((nblk + nseg - 1) / nseg);
    if (int rc = ensure_partials(c, (size_t)(nblk + nseg) * tot + 8)) return rc;
    if (mode == DFT_GRAM_SPLIT_BF16)
        launch_stream_ta<DFT_GRAM_SPLIT_BF16>(c, (unsigned)nblk, n, ma, mb, a, lda, b, ldb, kw);
    else
        launch_stream_ta<DFT_GRAM_F32>(c, (unsigned)nblk, n, ma, mb, a, lda, b, ldb, kw);
    DFT_CHECK_LAUNCH(c);
    double* part2 = c->partials + nblk * tot;
    k_gram_reduce1<<<dim3((unsigned)((tot + 255) / 256), nseg), 256, 0, c->stream>>>(c->partials, (int)nblk, tot, step,
                                                                                    part2);
    DFT_CHECK_LAUNCH(c);
    k_gram_reduce<<<(unsigned)((tot + 255) / 256), 256, 0, c->stream>>>(part2, nseg, ma, mb, alpha, beta, out, ldo);
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

// true when the wide kernel takes the shape: FP32 data in 16-byte aligned columns, ma <= 288, mb <= 96
bool wide_fits(int ma, int mb, const void* a, int64_t lda, const void* b, int64_t ldb) {
    return ma <= 16 * kWideMaxTA && mb <= 16 * kWideMaxTB && lda % 4 == 0 && ldb % 4 == 0 &&
           reinterpret_cast<uintptr_t>(a) % 16 == 0 && reinterpret_cast<uintptr_t>(b) % 16 == 0;
}

template <int MODE, int SA, int SB>
void launch_wide(dft_ctx* c, unsigned nblk, int64_t n, int ma, int mb, const float* a, int64_t lda, const float* b,
                 int64_t ldb) {
    k_gram_wide<MODE, SA, SB><<<nblk, 64 * kWideWaves, 0, c->stream>>>(n, ma, mb, a, lda, b, ldb, c->partials);
}

template <int MODE, int SA>
void launch_wide_sb(dft_ctx* c, int sb, unsigned nblk, int64_t n, int ma, int mb, const float* a, int64_t lda,
                    const float* b, int64_t ldb) {
    if (sb == 1) launch_wide<MODE, SA, 1>(c, nblk, n, ma, mb, a, lda, b, ldb);
    else if (sb == 2) launch_wide<MODE, SA, 2>(c, nblk, n, ma, mb, a, lda, b, ldb);
    else launch_wide<MODE, SA, 3>(c, nblk, n, ma, mb, a, lda, b, ldb);
}

int gram_wide(dft_ctx* c, int mode, int64_t n, int ma, int mb, double alpha, const float* a, int64_t lda,
              const float* b, int64_t ldb, double beta, double* out, int64_t ldo) {
    const int ta = (ma + 15) / 16, tb = (mb + 15) / 16;
    const int sa = (ta + kWideWA - 1) / kWideWA, sb = (tb + kWideWB - 1) / kWideWB;
    // one workgroup per CU (at most one per 32-point step)
    const int64_t nblk = std::max<int64_t>(1, std::min<int64_t>(c->num_cus, (n + 31) / 32));
    const int64_t tot = (int64_t)ma * mb;
    const int nseg = (int)std::min<int64_t>(nblk, 32), step = (int)((nblk + nseg - 1) / nseg);
    if (int rc = ensure_partials(c, (size_t)(nblk + nseg) * tot + 8)) return rc;
#define DFT_GW(M)                                                                          \
    if (sa == 1) launch_wide_sb<M, 1>(c, sb, (unsigned)nblk, n, ma, mb, a, lda, b, ldb);        \
    else if (sa == 2) launch_wide_sb<M, 2>(c, sb, (unsigned)nblk, n, ma, mb, a, lda, b, ldb);   \
    else launch_wide_sb<M, 3>(c, sb, (unsigned)nblk, n, ma, mb, a, lda, b, ldb)
    if (mode == DFT_GRAM_SPLIT6_BF16) {
        DFT_GW(DFT_GRAM_SPLIT6_BF16);
    } else {
        DFT_GW(DFT_GRAM_SPLIT_BF16);
    }
#undef DFT_GW
    DFT_CHECK_LAUNCH(c);
    double* part2 = c->partials + nblk * tot;
    k_gram_reduce1<<<dim3((unsigned)((tot + 255) / 256), nseg), 256, 0, c->stream>>>(c->partials, (int)nblk, tot, step,
                                                                                    part2);
    DFT_CHECK_LAUNCH(c);
    k_gram_reduce<<<(unsigned)((tot + 255) / 256), 256, 0, c->stream>>>(part2, nseg, ma, mb, alpha, beta, out, ldo);
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

}  // namespace

extern "C" {

int dft_gram(dft_ctx* c, int mode, int64_t n, int ma, int mb, double alpha, const void* a, int64_t lda, const void* b,
             int64_t ldb, double beta, double* out, int64_t ldo) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, a && b && out && n > 0 && ma > 0 && mb > 0 && lda >= n && ldb >= n && ldo >= ma, "dft_gram: bad arguments");
    DFT_REQUIRE(c, mode >= DFT_GRAM_F64 && mode <= DFT_GRAM_SPLIT6_BF16, "dft_gram: bad mode");
    // SPLIT6 has FP32-product accuracy: where the FP32 MFMA is not the bound (small outputs) or the
    // wide kernel does not take the shape, the FP32 kernels run it
    const bool wide = (mode == DFT_GRAM_SPLIT_BF16 || mode == DFT_GRAM_SPLIT6_BF16) && wide_fits(ma, mb, a, lda, b, ldb);
    if (mode != DFT_GRAM_F64 && ma <= 96 && mb <= 32 && lda % 4 == 0 && ldb % 4 == 0 &&
        reinterpret_cast<uintptr_t>(a) % 16 == 0 && reinterpret_cast<uintptr_t>(b) % 16 == 0)
        return gram_stream(c, mode == DFT_GRAM_SPLIT6_BF16 ? DFT_GRAM_F32 : mode, n, ma, mb, alpha, (const float*)a,
                           lda, (const float*)b, ldb, beta, out, ldo);
    if (wide) return gram_wide(c, mode, n, ma, mb, alpha, (const float*)a, lda, (const float*)b, ldb, beta, out, ldo);
    if (mode == DFT_GRAM_SPLIT6_BF16) mode = DFT_GRAM_F32;
    const int ntm = (ma + GT - 1) / GT, ntn = (mb + GT - 1) / GT;
    const int ntiles = ntm * ntn;
    const int ks = mode == DFT_GRAM_SPLIT_BF16 ? 32 : 16;
    // K-split: about four workgroups per CU in total, a multiple of 8 chunks
    // (one group of chunks per XCD), chunks a multiple of the slab depth
    int64_t nks = std::max<int64_t>(1, (4LL * c->num_cus + ntiles - 1) / ntiles);
    int64_t kchunk = std::max<int64_t>((n + nks - 1) / nks, 4 * ks);
    kchunk = (kchunk + ks - 1) / ks * ks;
    nks = (n + kchunk - 1) / kchunk;
    nks = (nks + 7) / 8 * 8;
    DFT_REQUIRE(c, nks * ntiles < (1LL << 31), "dft_gram: too many workgroups");
    const int64_t tot = (int64_t)ma * mb;
    const int nseg = (int)std::min<int64_t>(nks, 32), step = (int)((nks + nseg - 1) / nseg);
    if (int rc = ensure_partials(c, (size_t)(nks + nseg) * tot + 8)) return rc;
    const unsigned nblk = (unsigned)(nks * ntiles);
    if (mode == DFT_GRAM_F64)
        k_gram<DFT_GRAM_F64><<<nblk, 256, 0, c->stream>>>(n, ma, mb, ntm, ntiles, (int)nks, (const double*)a, lda, (const double*)b, ldb, kchunk, c->partials);
    else if (mode == DFT_GRAM_F32)
        k_gram<DFT_GRAM_F32><<<nblk, 256, 0, c->stream>>>(n, ma, mb, ntm, ntiles, (int)nks, (const float*)a, lda, (const float*)b, ldb, kchunk, c->partials);
    else
        k_gram<DFT_GRAM_SPLIT_BF16><<<nblk, 256, 0, c->stream>>>(n, ma, mb, ntm, ntiles, (int)nks, (const float*)a, lda, (const float*)b, ldb, kchunk, c->partials);
    DFT_CHECK_LAUNCH(c);
    double* part2 = c->partials + nks * tot;
    k_gram_reduce1<<<dim3((unsigned)((tot + 255) / 256), nseg), 256, 0, c->stream>>>(c->partials, (int)nks, tot, step, part2);
    DFT_CHECK_LAUNCH(c);
    k_gram_reduce<<<(unsigned)((tot + 255) / 256), 256, 0, c->stream>>>(part2, nseg, ma, mb, alpha, beta, out, ldo);
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

}  // extern "C"
