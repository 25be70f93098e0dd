// cnn.hip -- the learned Davidson preconditioner: a 3D CNN of four 3x3x3
// convolutions (channels 2 -> C -> C -> C -> 1, SiLU after the first three,
// zero padding), BF16 parameters, FP32 activations and accumulation.
//
// Parameter layout (one array, FP32 master copy and BF16 working copy):
//   for layer l = 0..3: W_l[cout][cin][27] (tap = kz*9 + ky*3 + kx, the
//   output at v reads the input at v + (kx-1, ky-1, kz-1)), then b_l[cout].
// Activation layout: [item][channel][z][y][x], x fastest (the grid layout).
//
// Kernels:
//   k_conv16_mfma -- the 16 -> 16 layers (forward and data gradient) as an
//                implicit GEMM on the bf16 matrix cores (below).
//                The 16 -> 1 output layer runs on the same kernel with one
//                non-zero row of A.
//   k_conv    -- direct convolution (the 2 -> 16 layer and the 1 -> 16 data
//                gradient).  A workgroup owns a 16x4x4 output tile
//                (one voxel per thread) and stages the (18x6x6) x cin input
//                halo tile in LDS; the FP32 copy of the BF16 weights is read
//                with wave-uniform addresses (scalar loads), so each LDS read
//                of an input value feeds cout FMAs.  Modes: forward (bias,
//                optional SiLU applied to the input on load) and the data
//                gradient (flipped, transposed kernel, times SiLU'(z) of the
//                layer below in the epilogue).
//   k_wgrad   -- weight/bias gradient partials per workgroup over a tile
//                loop (LDS-staged input and output-gradient tiles), reduced
//                in a fixed order by k_reduce_cols (deterministic).
//   k_mse, k_adam, k_to_bf16 and the preconditioner's input/output maps.
#include "common.hpp"

#include <algorithm>
#include <cmath>

using namespace dft;

namespace {

constexpr int TX = 16, TY = 4, TZ = 4;  // output tile (256 voxels)
constexpr int HX = TX + 2, HY = TY + 2, HZ = TZ + 2;
constexpr int HV = HX * HY * HZ;  // 648 halo-tile voxels

struct NetLayout {
    int cin[4], cout[4];
    int64_t w[4], b[4], total;
};

inline NetLayout net_layout(int C) {
    NetLayout L;
    const int ci[4] = {2, C, C, C}, co[4] = {C, C, C, 1};
    int64_t off = 0;
    for (int l = 0; l < 4; ++l) {
        L.cin[l] = ci[l];
        L.cout[l] = co[l];
        L.w[l] = off;
        off += (int64_t)co[l] * ci[l] * 27;
        L.b[l] = off;
        off += co[l];
    }
    L.total = off;
    return L;
}

inline int64_t align64(int64_t n) { return (n + 63) / 64 * 64; }

__device__ inline float bf16_to_f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
// round to nearest even (finite inputs)
__device__ inline uint16_t f_to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ inline float silu(float z) { return z / (1.0f + expf(-z)); }
__device__ inline float dsilu(float z) {
    const float s = 1.0f / (1.0f + expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

// Stage the input halo tile of one item into LDS (zero outside the grid).
template <int NI, bool IN_SILU>
__device__ inline void load_halo(float* xs, const float* __restrict__ inb, int nx, int ny, int nz, int64_t N, int x0,
                                 int y0, int z0) {
    for (int q = threadIdx.x; q < NI * HV; q += 256) {
        const int i = q / HV, r = q - i * HV;
        const int lx = r % HX, ly = (r / HX) % HY, lz = r / (HX * HY);
        const int gx = x0 - 1 + lx, gy = y0 - 1 + ly, gz = z0 - 1 + lz;
        float v = 0.0f;
        if (gx >= 0 && gx < nx && gy >= 0 && gy < ny && gz >= 0 && gz < nz) {
            v = inb[(int64_t)i * N + ((int64_t)gz * ny + gy) * nx + gx];
            if (IN_SILU) v = silu(v);
        }
        xs[q] = v;
    }
}

// out[item][o][v] = (bias[o]) + sum_{i,tap} K(o,i,tap) * f(in[item][i][v + off(tap)])
//   forward:   K(o,i,tap) = W[o][i][tap]            (W of shape [NO][NI][27])
//   transpose: K(o,i,tap) = W[i][o][26 - tap]       (W of shape [NI][NO][27])
// EPI_DSILU multiplies the result by SiLU'(zepi[item][o][v]).
template <int NI, int NO, bool IN_SILU, bool TRANS, bool BIAS, bool EPI_DSILU, bool OUT_SILU = false>
__global__ __launch_bounds__(256) void k_conv(int nx, int ny, int nz, int tiles_z, const float* __restrict__ w,
                                              const float* __restrict__ bias, const float* __restrict__ in,
                                              float* __restrict__ out, const float* __restrict__ zepi) {
    __shared__ float xs[NI * HV];
    const int bz = blockIdx.z % tiles_z, item = blockIdx.z / tiles_z;
    const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TY, z0 = bz * TZ;
    const int64_t N = (int64_t)nx * ny * nz;
    load_halo<NI, IN_SILU>(xs, in + (int64_t)item * NI * N, nx, ny, nz, N, x0, y0, z0);
    __syncthreads();
    const int t = threadIdx.x;
    const int tx = t % TX, ty = (t / TX) % TY, tz = t / (TX * TY);
    float acc[NO];
#pragma unroll
    for (int o = 0; o < NO; ++o) acc[o] = BIAS ? bias[o] : 0.0f;
    for (int i = 0; i < NI; ++i) {
#pragma unroll
        for (int kz = 0; kz < 3; ++kz)
#pragma unroll
            for (int ky = 0; ky < 3; ++ky)
#pragma unroll
                for (int kx = 0; kx < 3; ++kx) {
                    const int tap = kz * 9 + ky * 3 + kx;
                    const float xv = xs[i * HV + ((tz + kz) * HY + (ty + ky)) * HX + tx + kx];
#pragma unroll
                    for (int o = 0; o < NO; ++o) {
                        const float k = TRANS ? w[((int64_t)i * NO + o) * 27 + (26 - tap)] : w[((int64_t)o * NI + i) * 27 + tap];
                        acc[o] = fmaf(k, xv, acc[o]);
                    }
                }
    }
    const int gx = x0 + tx, gy = y0 + ty, gz = z0 + tz;
    if (gx < nx && gy < ny && gz < nz) {
        const int64_t v = ((int64_t)gz * ny + gy) * nx + gx;
#pragma unroll
        for (int o = 0; o < NO; ++o) {
            const int64_t p = ((int64_t)item * NO + o) * N + v;
            float r = acc[o];
            if (EPI_DSILU) r *= dsilu(zepi[p]);
            if (OUT_SILU) r = silu(r);
            out[p] = r;
        }
    }
}

// The 16 -> 16 convolutions (layers 1 and 2, 92% of the network's flops) as an
// implicit GEMM on the bf16 matrix cores: out[o][v] = sum_K A[o][K] B[K][v] with
// K = tap * 16 + i (27 taps x 16 channels = 432, padded to 14 steps of 32).
//   A (v_mfma_f32_16x16x32_bf16 operand A, 16 x 32): the BF16 parameters as
//     they are (exact); lane l holds output channel l % 16 and K = 32 s + 8 (l / 16)
//     .. + 7, i.e. tap 2 s + (l / 16) / 2 and channels 8 ((l / 16) % 2) .. + 7,
//     all 14 steps in registers for the whole tile.
//   B (32 x 16): the input halo tile in LDS, channel-last, each FP32 value split
//     into bf16 hi + lo (two MFMAs per step, lo first: ~2^-17 relative, the FP32
//     kernel's 3e-5 bar with room); lane l reads the 8 channels of its K slice
//     at voxel (x0 + l % 16 + kx, ...) with one 16-byte load per piece.
//   D (16 x 16): lane l, register r = output channel 4 (l / 16) + r at voxel
//     x0 + l % 16; wave w takes the output rows (ty, tz) 4w .. 4w + 3 of the
//     16 x 4 x 4 tile.
// Modes as k_conv: forward (bias, SiLU of the input on load) and the data
// gradient (flipped, transposed kernel; times SiLU'(z) in the epilogue).
typedef __bf16 bfv8 __attribute__((ext_vector_type(8)));
typedef float fv4 __attribute__((ext_vector_type(4)));
constexpr int kConvKSteps = 14;

// A operands of one 16 -> 16 layer in operand order, wp[s * 64 + lane] (built once per
// layer and call by k_pack_w16 from the BF16 parameters; TRANS: the flipped, transposed
// kernel of the data gradient).  The padding taps (27) are zero.
template <bool TRANS, int NOUT = 16>
__global__ __launch_bounds__(256) void k_pack_w16(const uint16_t* __restrict__ wbf, bfv8* __restrict__ wp) {
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= kConvKSteps * 64) return;
    const int s = q / 64, lane = q % 64, r16 = lane & 15, kq = lane >> 4;
    const int tap = 2 * s + (kq >> 1), i0 = 8 * (kq & 1);
    bfv8 v;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int i = i0 + e;
        uint16_t bits = 0;
        if (tap < 27 && r16 < NOUT) bits = TRANS ? wbf[(i * 16 + r16) * 27 + 26 - tap] : wbf[(r16 * 16 + i) * 27 + tap];
        v[e] = __builtin_bit_cast(__bf16, bits);
    }
    wp[q] = v;
}

template <bool IN_SILU, bool TRANS, bool BIAS, bool EPI_DSILU, bool OUT_SILU, int NOUT = 16>
__global__ __launch_bounds__(256) void k_conv16_mfma(int nx, int ny, int nz, int tiles_z,
                                                     const bfv8* __restrict__ wp, const float* __restrict__ bias,
                                                     const float* __restrict__ in, float* __restrict__ out,
                                                     const float* __restrict__ zepi) {
    __shared__ __attribute__((aligned(16))) __bf16 xh[HV * 16];
    __shared__ __attribute__((aligned(16))) __bf16 xl[HV * 16];
    const int bz = blockIdx.z % tiles_z, item = blockIdx.z / tiles_z;
    const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TY, z0 = bz * TZ;
    const int64_t N = (int64_t)nx * ny * nz;
    const float* inb = in + (int64_t)item * 16 * N;
    // a thread stages whole halo voxels: 16 channel loads (each coalesced over the
    // threads' consecutive voxels), then the voxel's 32 bytes of hi and of lo pieces
    for (int r = threadIdx.x; r < HV; r += 256) {
        const int lx = r % HX, ly = (r / HX) % HY, lz = r / (HX * HY);
        const int gx = x0 - 1 + lx, gy = y0 - 1 + ly, gz = z0 - 1 + lz;
        const bool inside = gx >= 0 && gx < nx && gy >= 0 && gy < ny && gz >= 0 && gz < nz;
        const int64_t v = inside ? ((int64_t)gz * ny + gy) * nx + gx : 0;
        bfv8 h[2], l[2];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            float x = inside ? inb[(int64_t)i * N + v] : 0.0f;
            if (IN_SILU) x = silu(x);
            const __bf16 hb = (__bf16)x;
            h[i >> 3][i & 7] = hb;
            l[i >> 3][i & 7] = (__bf16)(x - (float)hb);
        }
        bfv8* dh = reinterpret_cast<bfv8*>(xh + r * 16);
        bfv8* dl = reinterpret_cast<bfv8*>(xl + r * 16);
        dh[0] = h[0];
        dh[1] = h[1];
        dl[0] = l[0];
        dl[1] = l[1];
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r16 = lane & 15, kq = lane >> 4;
    bfv8 wa[kConvKSteps];
#pragma unroll
    for (int s = 0; s < kConvKSteps; ++s) wa[s] = wp[s * 64 + lane];
    __syncthreads();
    fv4 acc[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[j] = fv4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < kConvKSteps; ++s) {
        const int tap0 = 2 * s + (kq >> 1);
        const int tap = tap0 < 27 ? tap0 : 0;  // padding: any in-tile voxel (zero weights)
        const int kz = tap / 9, ky = (tap / 3) % 3, kx = tap % 3, i0 = 8 * (kq & 1);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int row = 4 * w + j, ty = row % TY, tz = row / TY;
            const int vox = ((tz + kz) * HY + (ty + ky)) * HX + r16 + kx;
            const bfv8 bh = *reinterpret_cast<const bfv8*>(xh + vox * 16 + i0);
            const bfv8 bl = *reinterpret_cast<const bfv8*>(xl + vox * 16 + i0);
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa[s], bl, acc[j], 0, 0, 0);
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa[s], bh, acc[j], 0, 0, 0);
        }
    }
    const int gx = x0 + r16;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int row = 4 * w + j, gy = y0 + row % TY, gz = z0 + row / TY;
        if (gx >= nx || gy >= ny || gz >= nz) continue;
        const int64_t v = ((int64_t)gz * ny + gy) * nx + gx;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int o = 4 * kq + r;
            if (o >= NOUT) continue;
            const int64_t p = ((int64_t)item * NOUT + o) * N + v;
            float val = acc[j][r];
            if (BIAS) val += bias[o];
            if (EPI_DSILU) val *= dsilu(zepi[p]);
            if (OUT_SILU) val = silu(val);
            out[p] = val;
        }
    }
}

// Partial gradients of one layer over a strided set of output tiles:
//   dW[o][i][tap] = sum_v dz[o][v] * f(x[i][v + off(tap)]),  db[o] = sum_v dz[o][v]
// Thread t owns the (o, i) pair t % (NO*NI) with its 27 taps in registers and
// the tile rows (vz, vy) = t / (NO*NI), +G, ...; a row's 3x3 input rows of 18
// values are read once from LDS and reused by the 16 x positions and 3 kx
// taps.  Each (workgroup, row group) writes one partial row of NITEMS values:
// the weights in W layout, then the NO biases.
template <int NI, int NO, bool IN_SILU>
__global__ __launch_bounds__(256) void k_wgrad(int nx, int ny, int nz, int tiles_x, int tiles_y, int tiles_z, int64_t ntiles,
                                               const float* __restrict__ in, const float* __restrict__ dz,
                                               float* __restrict__ partials) {
    constexpr int NP = NO * NI, G = 256 / NP, NW = NP * 27, NITEMS = NW + NO;
    static_assert(NP <= 256, "one (o, i) pair per thread");
    __shared__ float xs[NI * HV];
    __shared__ float ds[NO * 256];
    const int64_t N = (int64_t)nx * ny * nz;
    const int t = threadIdx.x;
    const int pair = t % NP, grp = t / NP;
    const int o = pair / NI, i = pair % NI;
    const bool active = grp < G;
    float acc[27];
#pragma unroll
    for (int k = 0; k < 27; ++k) acc[k] = 0.0f;
    float bacc = 0.0f;
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int bx = (int)(tile % tiles_x), by = (int)((tile / tiles_x) % tiles_y);
        const int bz = (int)((tile / ((int64_t)tiles_x * tiles_y)) % tiles_z);
        const int64_t item = tile / ((int64_t)tiles_x * tiles_y * tiles_z);
        const int x0 = bx * TX, y0 = by * TY, z0 = bz * TZ;
        __syncthreads();  // the previous tile's reads are done
        load_halo<NI, IN_SILU>(xs, in + item * NI * N, nx, ny, nz, N, x0, y0, z0);
        {
            const int tx = t % TX, ty = (t / TX) % TY, tz = t / (TX * TY);
            const int gx = x0 + tx, gy = y0 + ty, gz = z0 + tz;
            const bool ok = gx < nx && gy < ny && gz < nz;
            const int64_t v = ((int64_t)gz * ny + gy) * nx + gx;
#pragma unroll
            for (int q = 0; q < NO; ++q) ds[q * 256 + t] = ok ? dz[(item * NO + q) * N + v] : 0.0f;
        }
        __syncthreads();
        if (active) {
            for (int row = grp; row < TZ * TY; row += G) {
                const int vz = row / TY, vy = row % TY;
                float d[TX];
#pragma unroll
                for (int vx = 0; vx < TX; ++vx) d[vx] = ds[o * 256 + row * TX + vx];
                if (i == 0) {
#pragma unroll
                    for (int vx = 0; vx < TX; ++vx) bacc += d[vx];
                }
#pragma unroll
                for (int kz = 0; kz < 3; ++kz)
#pragma unroll
                    for (int ky = 0; ky < 3; ++ky) {
                        const float* xr = xs + i * HV + ((vz + kz) * HY + (vy + ky)) * HX;
                        float xrow[HX];
#pragma unroll
                        for (int q = 0; q < HX; ++q) xrow[q] = xr[q];
#pragma unroll
                        for (int kx = 0; kx < 3; ++kx) {
                            float s = 0.0f;
#pragma unroll
                            for (int vx = 0; vx < TX; ++vx) s = fmaf(d[vx], xrow[vx + kx], s);
                            acc[kz * 9 + ky * 3 + kx] += s;
                        }
                    }
            }
        }
    }
    if (active) {
        float* prow = partials + ((int64_t)blockIdx.x * G + grp) * NITEMS;
#pragma unroll
        for (int k = 0; k < 27; ++k) prow[pair * 27 + k] = acc[k];
        if (i == 0) prow[NW + o] = bacc;
    }
}

// out[q] = sum_b partials[b * nitems + q] (fixed order, FP64 accumulation)
__global__ __launch_bounds__(256) void k_reduce_cols(int nblk, int nitems, const float* __restrict__ partials,
                                                     float* __restrict__ out) {
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= nitems) return;
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += (double)partials[(int64_t)b * nitems + q];
    out[q] = (float)s;
}

// d = 2 (out - target) / M; partial sums of (out - target)^2
__global__ __launch_bounds__(256) void k_mse(int64_t M, const float* __restrict__ out, const float* __restrict__ tgt,
                                             float* __restrict__ d, double* __restrict__ partials) {
    __shared__ double lds[4];
    const float scale = (float)(2.0 / (double)M);
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < M; i += (int64_t)gridDim.x * 256) {
        const float e = out[i] - tgt[i];
        d[i] = scale * e;
        s += (double)e * (double)e;
    }
    s = block_sum(s, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_bf16_to_f(int64_t n, const uint16_t* __restrict__ h, float* __restrict__ f) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) f[i] = bf16_to_f(h[i]);
}

__global__ __launch_bounds__(256) void k_to_bf16(int64_t n, const float* __restrict__ w, uint16_t* __restrict__ h) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) h[i] = f_to_bf16(w[i]);
}

__global__ __launch_bounds__(256) void k_adam(int64_t n, float* __restrict__ w, uint16_t* __restrict__ h,
                                              const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                                              float lr, float b1, float b2, float eps, float c1, float c2) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / c1, vh = vi / c2;
    const float wi = w[i] - lr * mh / (sqrtf(vh) + eps);
    w[i] = wi;
    h[i] = f_to_bf16(wi);
}

// in[b][0][v] = r_b[v] / s_b, in[b][1][v] = (V[v] - theta_b) * h2   (s_b = 0 -> 0)
template <typename T>
__global__ __launch_bounds__(256) void k_cnn_prep(int64_t N, int ncols, const T* __restrict__ V, const T* __restrict__ R,
                                                  int64_t ldr, const double* __restrict__ theta,
                                                  const double* __restrict__ scale, double h2, float* __restrict__ in) {
    const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int b = blockIdx.y;
    if (v >= N || b >= ncols) return;
    const double s = scale[b];
    const double inv = s > 0.0 ? 1.0 / s : 0.0;
    in[((int64_t)b * 2) * N + v] = (float)((double)R[(int64_t)b * ldr + v] * inv);
    in[((int64_t)b * 2 + 1) * N + v] = (float)(((double)V[v] - theta[b]) * h2);
}

// T_b[v] = out[b][v] * s_b * h2
template <typename T>
__global__ __launch_bounds__(256) void k_cnn_finish(int64_t N, int ncols, const float* __restrict__ out,
                                                    const double* __restrict__ scale, double h2, T* __restrict__ Tm,
                                                    int64_t ldt) {
    const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int b = blockIdx.y;
    if (v >= N || b >= ncols) return;
    Tm[(int64_t)b * ldt + v] = (T)((double)out[(int64_t)b * N + v] * scale[b] * h2);
}

// The network's inputs on a sub-box of the grid: box = {x0, y0, z0, bx, by,
// bz}; in[b][0][v] = r_b[g(v)] / s_b, in[b][1][v] = (V[g(v)] - theta_b) h2,
// v over the box (x fastest), g(v) the grid index.
template <typename T>
__global__ __launch_bounds__(256) void k_cnn_prep_box(int nx, int ny, int x0, int y0, int z0, int bx, int by, int bz,
                                                      int ncols, const T* __restrict__ V, const T* __restrict__ R,
                                                      int64_t ldr, const double* __restrict__ theta,
                                                      const double* __restrict__ scale, double h2,
                                                      float* __restrict__ in) {
    const int64_t nb = (int64_t)bx * by * bz;
    const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int b = blockIdx.y;
    if (v >= nb || b >= ncols) return;
    const int lx = (int)(v % bx), ly = (int)((v / bx) % by), lz = (int)(v / ((int64_t)bx * by));
    const int64_t g = ((int64_t)(z0 + lz) * ny + (y0 + ly)) * nx + (x0 + lx);
    const double s = scale[b];
    const double inv = s > 0.0 ? 1.0 / s : 0.0;
    in[((int64_t)b * 2) * nb + v] = (float)((double)R[(int64_t)b * ldr + g] * inv);
    in[((int64_t)b * 2 + 1) * nb + v] = (float)(((double)V[g] - theta[b]) * h2);
}

// Linear ramp over `taper` points at a box face that lies inside the grid
// (1 at and beyond taper + 1 points from the face), 1 at a face on the grid
// boundary, where the network's zero padding is the orbitals' own boundary.
__device__ inline double box_ramp(int l, int b, int lo_inside, int hi_inside, int taper) {
    double w = 1.0;
    if (taper > 0) {
        if (lo_inside) w = fmin(w, (double)(l + 1) / (double)(taper + 1));
        if (hi_inside) w = fmin(w, (double)(b - l) / (double)(taper + 1));
    }
    return w;
}

// T_b[g(v)] += out[b][v] s_b h2 w(v) over the box (w: the face ramps).
template <typename T>
__global__ __launch_bounds__(256) void k_cnn_add_box(int nx, int ny, int nz, int x0, int y0, int z0, int bx, int by,
                                                     int bz, int taper, int ncols, const float* __restrict__ out,
                                                     const double* __restrict__ scale, double h2, T* __restrict__ Tm,
                                                     int64_t ldt) {
    const int64_t nb = (int64_t)bx * by * bz;
    const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int b = blockIdx.y;
    if (v >= nb || b >= ncols) return;
    const int lx = (int)(v % bx), ly = (int)((v / bx) % by), lz = (int)(v / ((int64_t)bx * by));
    const int64_t g = ((int64_t)(z0 + lz) * ny + (y0 + ly)) * nx + (x0 + lx);
    const double w = box_ramp(lx, bx, x0 > 0, x0 + bx < nx, taper) * box_ramp(ly, by, y0 > 0, y0 + by < ny, taper) *
                     box_ramp(lz, bz, z0 > 0, z0 + bz < nz, taper);
    T* t = Tm + (int64_t)b * ldt + g;
    *t = (T)((double)*t + (double)out[(int64_t)b * nb + v] * scale[b] * h2 * w);
}

// max |X_b| per column: partials[b * nblk + blk]
template <typename T>
__global__ __launch_bounds__(256) void k_colmax(int64_t n, const T* __restrict__ X, int64_t ldx, int nblk,
                                                double* __restrict__ partials) {
    __shared__ double lds[4];
    const int b = blockIdx.y;
    const T* x = X + (int64_t)b * ldx;
    double m = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) m = fmax(m, fabs((double)x[i]));
    m = wave_max(m);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) lds[w] = m;
    __syncthreads();
    if (threadIdx.x == 0) partials[(int64_t)b * nblk + blockIdx.x] = fmax(fmax(lds[0], lds[1]), fmax(lds[2], lds[3]));
}
__global__ __launch_bounds__(64) void k_colmax_final(int ncols, int nblk, const double* __restrict__ partials,
                                                     double* __restrict__ out) {
    const int b = blockIdx.x;
    double m = 0.0;
    for (int i = threadIdx.x; i < nblk; i += 64) m = fmax(m, partials[(int64_t)b * nblk + i]);
    m = wave_max(m);
    if (threadIdx.x == 0 && b < ncols) out[b] = m;
}

// ---- host helpers ---------------------------------------------------------

inline dim3 conv_grid(int64_t nx, int64_t ny, int64_t nz, int nbatch, int* tiles_z) {
    *tiles_z = (int)((nz + TZ - 1) / TZ);
    return dim3((unsigned)((nx + TX - 1) / TX), (unsigned)((ny + TY - 1) / TY), (unsigned)(*tiles_z * nbatch));
}

int check_dims(dft_ctx* c, int C, int64_t nx, int64_t ny, int64_t nz, int nbatch) {
    DFT_REQUIRE(c, C == 16, "dft_cnn: only 16 channels are built");
    DFT_REQUIRE(c, nx >= 1 && ny >= 1 && nz >= 1 && nbatch >= 1, "dft_cnn: bad dimensions");
    DFT_REQUIRE(c, (ny + TY - 1) / TY <= 65535 && ((nz + TZ - 1) / TZ) * (int64_t)nbatch <= 65535 &&
                       (nx + TX - 1) / TX <= (1LL << 31) - 1,
                "dft_cnn: launch grid too large");
    return DFT_OK;
}

// float copy of the BF16 parameters into dst
int weights_to_float(dft_ctx* c, int64_t count, const uint16_t* w_bf16, float* dst) {
    k_bf16_to_f<<<(unsigned)((count + 255) / 256), 256, 0, c->stream>>>(count, w_bf16, dst);
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

template <int NI, int NO, bool IN_SILU, bool TRANS, bool BIAS, bool EPI, bool OUT_SILU = false>
int conv(dft_ctx* c, int64_t nx, int64_t ny, int64_t nz, int nbatch, const float* w, const float* b, const float* in,
         float* out, const float* zepi) {
    int tz = 0;
    const dim3 grid = conv_grid(nx, ny, nz, nbatch, &tz);
    k_conv<NI, NO, IN_SILU, TRANS, BIAS, EPI, OUT_SILU>
        <<<grid, 256, 0, c->stream>>>((int)nx, (int)ny, (int)nz, tz, w, b, in, out, zepi);
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

// floats of one layer's packed A operands (14 steps x 64 lanes x 16 bytes)
constexpr int64_t kPackFloats = kConvKSteps * 64 * 4;

// wpack: kPackFloats floats of device memory for the layer's packed weights
template <bool IN_SILU, bool TRANS, bool BIAS, bool EPI, bool OUT_SILU = false, int NOUT = 16>
int conv16_mfma(dft_ctx* c, int64_t nx, int64_t ny, int64_t nz, int nbatch, const uint16_t* wbf, const float* b,
                const float* in, float* out, const float* zepi, float* wpack) {
    bfv8* wp = reinterpret_cast<bfv8*>(wpack);
    k_pack_w16<TRANS, NOUT><<<(kConvKSteps * 64 + 255) / 256, 256, 0, c->stream>>>(wbf, wp);
    DFT_CHECK_LAUNCH(c);
    int tz = 0;
    const dim3 grid = conv_grid(nx, ny, nz, nbatch, &tz);
    k_conv16_mfma<IN_SILU, TRANS, BIAS, EPI, OUT_SILU, NOUT>
        <<<grid, 256, 0, c->stream>>>((int)nx, (int)ny, (int)nz, tz, wp, b, in, out, zepi);
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

constexpr int kWgradBlocks = 512;

template <int NI, int NO, bool IN_SILU>
int wgrad(dft_ctx* c, int64_t nx, int64_t ny, int64_t nz, int nbatch, const float* in, const float* dz, float* partials,
          float* grad_w) {
    constexpr int NITEMS = NO * NI * 27 + NO, G = 256 / (NO * NI);
    const int txs = (int)((nx + TX - 1) / TX), tys = (int)((ny + TY - 1) / TY), tzs = (int)((nz + TZ - 1) / TZ);
    const int64_t ntiles = (int64_t)txs * tys * tzs * nbatch;
    const int nblk = (int)std::min<int64_t>(ntiles, kWgradBlocks);
    k_wgrad<NI, NO, IN_SILU><<<nblk, 256, 0, c->stream>>>((int)nx, (int)ny, (int)nz, txs, tys, tzs, ntiles, in, dz, partials);
    DFT_CHECK_LAUNCH(c);
    k_reduce_cols<<<(NITEMS + 255) / 256, 256, 0, c->stream>>>(nblk * G, NITEMS, partials, grad_w);
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

}  // namespace

extern "C" {

int dft_cnn_param_count(int channels, int64_t* count) {
    if (!count || channels < 1) return DFT_EINVAL;
    *count = net_layout(channels).total;
    return DFT_OK;
}

int dft_cnn_act_count(int channels, int64_t nx, int64_t ny, int64_t nz, int nbatch, int training, int64_t* count) {
    if (!count || channels < 1 || nx < 1 || ny < 1 || nz < 1 || nbatch < 1) return DFT_EINVAL;
    const int64_t layer = (int64_t)channels * nx * ny * nz * nbatch;
    // FP32 weights, the layers, then the packed A operands of layers 1, 2 and 3
    *count = align64(net_layout(channels).total) + align64((training ? 3 : 2) * layer) + 3 * kPackFloats;
    return DFT_OK;
}

int dft_cnn_scratch_count(int channels, int64_t nx, int64_t ny, int64_t nz, int nbatch, int64_t* count) {
    if (!count || channels < 1 || nx < 1 || ny < 1 || nz < 1 || nbatch < 1) return DFT_EINVAL;
    const int64_t layer = (int64_t)channels * nx * ny * nz * nbatch;
    // partial rows per layer: (256 / (cout*cin)) row groups x (cout*cin*27 + cout) items
    const NetLayout L = net_layout(channels);
    int64_t rows = 0;
    for (int l = 0; l < 4; ++l) {
        const int64_t np = (int64_t)L.cout[l] * L.cin[l];
        rows = std::max<int64_t>(rows, std::max<int64_t>(1, 256 / np) * (np * 27 + L.cout[l]));
    }
    *count = align64(L.total) + 2 * align64(layer) + align64((int64_t)kWgradBlocks * rows) + 2 * kPackFloats;
    return DFT_OK;
}

int dft_cnn_forward(dft_ctx* c, int channels, int64_t nx, int64_t ny, int64_t nz, int nbatch, const uint16_t* w_bf16,
                    const float* in, float* out, float* act, int training) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, w_bf16 && in && out && act, "dft_cnn_forward: null argument");
    if (int rc = check_dims(c, channels, nx, ny, nz, nbatch)) return rc;
    const NetLayout L = net_layout(16);
    float* wf = act;
    const int64_t layer = 16 * nx * ny * nz * (int64_t)nbatch;
    float* a0 = act + align64(L.total);
    float* a1 = a0 + layer;
    float* a2 = training ? a1 + layer : a0;  // inference: ping-pong
    float* wpk = a0 + align64((training ? 3 : 2) * layer);  // packed A operands, layers 1, 2 and 3
    if (int rc = weights_to_float(c, L.total, w_bf16, wf)) return rc;
    if (training) {  // the pre-activations z are kept for the backward pass; SiLU is applied on load
        if (int rc = conv<2, 16, false, false, true, false>(c, nx, ny, nz, nbatch, wf + L.w[0], wf + L.b[0], in, a0, nullptr)) return rc;
        if (int rc = conv16_mfma<true, false, true, false>(c, nx, ny, nz, nbatch, w_bf16 + L.w[1], wf + L.b[1], a0, a1, nullptr, wpk)) return rc;
        if (int rc = conv16_mfma<true, false, true, false>(c, nx, ny, nz, nbatch, w_bf16 + L.w[2], wf + L.b[2], a1, a2, nullptr, wpk + kPackFloats)) return rc;
        return conv16_mfma<true, false, true, false, false, 1>(c, nx, ny, nz, nbatch, w_bf16 + L.w[3], wf + L.b[3], a2, out,
                                                              nullptr, wpk + 2 * kPackFloats);
    }
    // inference: each layer stores SiLU(z) once, instead of its consumer applying it at every halo read
    if (int rc = conv<2, 16, false, false, true, false, true>(c, nx, ny, nz, nbatch, wf + L.w[0], wf + L.b[0], in, a0, nullptr)) return rc;
    if (int rc = conv16_mfma<false, false, true, false, true>(c, nx, ny, nz, nbatch, w_bf16 + L.w[1], wf + L.b[1], a0, a1, nullptr, wpk)) return rc;
    if (int rc = conv16_mfma<false, false, true, false, true>(c, nx, ny, nz, nbatch, w_bf16 + L.w[2], wf + L.b[2], a1, a2, nullptr, wpk + kPackFloats)) return rc;
    return conv16_mfma<false, false, true, false, false, 1>(c, nx, ny, nz, nbatch, w_bf16 + L.w[3], wf + L.b[3], a2, out,
                                                             nullptr, wpk + 2 * kPackFloats);
}

int dft_cnn_backward(dft_ctx* c, int channels, int64_t nx, int64_t ny, int64_t nz, int nbatch, const uint16_t* w_bf16,
                     const float* in, const float* act, const float* out, const float* target, float* grad,
                     float* scratch, double* loss_host) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, w_bf16 && in && act && out && target && grad && scratch, "dft_cnn_backward: null argument");
    if (int rc = check_dims(c, channels, nx, ny, nz, nbatch)) return rc;
    const NetLayout L = net_layout(16);
    const int64_t N = nx * ny * nz, layer = 16 * N * (int64_t)nbatch, M = N * nbatch;
    const float* z0 = act + align64(L.total);
    const float* z1 = z0 + layer;
    const float* z2 = z1 + layer;
    float* wf = scratch;
    float* A = scratch + align64(L.total);
    float* B = A + align64(layer);
    float* part = B + align64(layer);
    const NetLayout L0 = net_layout(16);
    int64_t rows = 0;
    for (int l = 0; l < 4; ++l) {
        const int64_t np = (int64_t)L0.cout[l] * L0.cin[l];
        rows = std::max<int64_t>(rows, std::max<int64_t>(1, 256 / np) * (np * 27 + L0.cout[l]));
    }
    float* wpk = part + align64((int64_t)kWgradBlocks * rows);  // packed A operands of the data gradients
    if (int rc = weights_to_float(c, L.total, w_bf16, wf)) return rc;
    // dL/dout and the loss
    const int nblk = (int)std::min<int64_t>((M + 255) / 256, 4 * (int64_t)c->num_cus);
    if (int rc = ensure_partials(c, (size_t)nblk)) return rc;
    k_mse<<<nblk, 256, 0, c->stream>>>(M, out, target, A, c->partials);
    DFT_CHECK_LAUNCH(c);
    // layer 3: input SiLU(z2), output gradient A (1 channel)
    if (int rc = wgrad<16, 1, true>(c, nx, ny, nz, nbatch, z2, A, part, grad + L.w[3])) return rc;
    if (int rc = conv<1, 16, false, true, false, true>(c, nx, ny, nz, nbatch, wf + L.w[3], nullptr, A, B, z2)) return rc;
    // layer 2: dz2 in B
    if (int rc = wgrad<16, 16, true>(c, nx, ny, nz, nbatch, z1, B, part, grad + L.w[2])) return rc;
    if (int rc = conv16_mfma<false, true, false, true>(c, nx, ny, nz, nbatch, w_bf16 + L.w[2], nullptr, B, A, z1, wpk)) return rc;
    // layer 1: dz1 in A
    if (int rc = wgrad<16, 16, true>(c, nx, ny, nz, nbatch, z0, A, part, grad + L.w[1])) return rc;
    if (int rc = conv16_mfma<false, true, false, true>(c, nx, ny, nz, nbatch, w_bf16 + L.w[1], nullptr, A, B, z0, wpk + kPackFloats)) return rc;
    // layer 0: dz0 in B, raw input
    if (int rc = wgrad<2, 16, false>(c, nx, ny, nz, nbatch, in, B, part, grad + L.w[0])) return rc;
    if (loss_host) {
        double s = 0.0;
        if (int rc = finish_sum(c, nblk, &s)) return rc;
        *loss_host = s / (double)M;
    }
    return DFT_OK;
}

int dft_cnn_adam(dft_ctx* c, int64_t count, float* w, uint16_t* w_bf16, const float* grad, float* m, float* v, int step,
                 float lr, float beta1, float beta2, float eps) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, count > 0 && w && w_bf16 && grad && m && v && step >= 1, "dft_cnn_adam: bad arguments");
    const float c1 = (float)(1.0 - std::pow((double)beta1, step));
    const float c2 = (float)(1.0 - std::pow((double)beta2, step));
    k_adam<<<(unsigned)((count + 255) / 256), 256, 0, c->stream>>>(count, w, w_bf16, grad, m, v, lr, beta1, beta2, eps, c1, c2);
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

int dft_cnn_to_bf16(dft_ctx* c, int64_t count, const float* w, uint16_t* w_bf16) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, count > 0 && w && w_bf16, "dft_cnn_to_bf16: bad arguments");
    k_to_bf16<<<(unsigned)((count + 255) / 256), 256, 0, c->stream>>>(count, w, w_bf16);
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

int dft_column_maxabs(dft_ctx* c, int prec, int64_t n, int ncols, const void* X, int64_t ldx, double* out_dev) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, n > 0 && ncols > 0 && ncols <= 65535 && X && out_dev && ldx >= n, "dft_column_maxabs: bad arguments");
    const int nblk = (int)std::min<int64_t>((n + 255) / 256, 64);
    if (int rc = ensure_partials(c, (size_t)nblk * ncols)) return rc;
    const dim3 grid(nblk, ncols);
    if (prec == DFT_F64)
        k_colmax<double><<<grid, 256, 0, c->stream>>>(n, (const double*)X, ldx, nblk, c->partials);
    else if (prec == DFT_F32)
        k_colmax<float><<<grid, 256, 0, c->stream>>>(n, (const float*)X, ldx, nblk, c->partials);
    else
        return set_err(c, DFT_EINVAL, "dft_column_maxabs: bad precision");
    DFT_CHECK_LAUNCH(c);
    k_colmax_final<<<ncols, 64, 0, c->stream>>>(ncols, nblk, c->partials, out_dev);
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

int dft_cnn_prepare(dft_ctx* c, const dft_grid* g, int prec, const void* veff, const void* R, int64_t ldr, int ncols,
                    const double* theta_dev, const double* scale_dev, float* in) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, valid_grid(g) && veff && R && theta_dev && scale_dev && in && ncols > 0 && ncols <= 65535,
                "dft_cnn_prepare: bad arguments");
    const int64_t N = grid_points(g);
    const double h2 = 1.0 / (1.0 / (g->dx * g->dx) + 1.0 / (g->dy * g->dy) + 1.0 / (g->dz * g->dz));
    const dim3 grid((unsigned)((N + 255) / 256), ncols);
    if (prec == DFT_F64)
        k_cnn_prep<double><<<grid, 256, 0, c->stream>>>(N, ncols, (const double*)veff, (const double*)R, ldr, theta_dev,
                                                        scale_dev, h2, in);
    else if (prec == DFT_F32)
        k_cnn_prep<float><<<grid, 256, 0, c->stream>>>(N, ncols, (const float*)veff, (const float*)R, ldr, theta_dev,
                                                       scale_dev, h2, in);
    else
        return set_err(c, DFT_EINVAL, "dft_cnn_prepare: bad precision");
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

static int box_ok(const dft_grid* g, const int64_t* box) {
    return box && box[0] >= 0 && box[1] >= 0 && box[2] >= 0 && box[3] >= 1 && box[4] >= 1 && box[5] >= 1 &&
           box[0] + box[3] <= g->nx && box[1] + box[4] <= g->ny && box[2] + box[5] <= g->nz;
}

int dft_cnn_prepare_box(dft_ctx* c, const dft_grid* g, int prec, const void* veff, const void* R, int64_t ldr,
                        int ncols, const double* theta_dev, const double* scale_dev, const int64_t* box, float* in) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, valid_grid(g) && veff && R && theta_dev && scale_dev && in && ncols > 0 && ncols <= 65535 &&
                       ldr >= grid_points(g) && box_ok(g, box),
                "dft_cnn_prepare_box: bad arguments");
    const int64_t nb = box[3] * box[4] * box[5];
    const double h2 = 1.0 / (1.0 / (g->dx * g->dx) + 1.0 / (g->dy * g->dy) + 1.0 / (g->dz * g->dz));
    const dim3 grid((unsigned)((nb + 255) / 256), ncols);
#define DFT_PREP_BOX(E)                                                                                              \
    k_cnn_prep_box<E><<<grid, 256, 0, c->stream>>>((int)g->nx, (int)g->ny, (int)box[0], (int)box[1], (int)box[2],  \
                                                   (int)box[3], (int)box[4], (int)box[5], ncols, (const E*)veff,   \
                                                   (const E*)R, ldr, theta_dev, scale_dev, h2, in)
    if (prec == DFT_F64) DFT_PREP_BOX(double);
    else if (prec == DFT_F32) DFT_PREP_BOX(float);
    else return set_err(c, DFT_EINVAL, "dft_cnn_prepare_box: bad precision");
#undef DFT_PREP_BOX
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

int dft_cnn_add_box(dft_ctx* c, const dft_grid* g, int prec, int ncols, const float* out, const double* scale_dev,
                    const int64_t* box, int taper, void* T, int64_t ldt) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, valid_grid(g) && out && scale_dev && T && ncols > 0 && ncols <= 65535 && ldt >= grid_points(g) &&
                       taper >= 0 && box_ok(g, box),
                "dft_cnn_add_box: bad arguments");
    const int64_t nb = box[3] * box[4] * box[5];
    const double h2 = 1.0 / (1.0 / (g->dx * g->dx) + 1.0 / (g->dy * g->dy) + 1.0 / (g->dz * g->dz));
    const dim3 grid((unsigned)((nb + 255) / 256), ncols);
#define DFT_ADD_BOX(E)                                                                                               \
    k_cnn_add_box<E><<<grid, 256, 0, c->stream>>>((int)g->nx, (int)g->ny, (int)g->nz, (int)box[0], (int)box[1],    \
                                                  (int)box[2], (int)box[3], (int)box[4], (int)box[5], taper, ncols, \
                                                  out, scale_dev, h2, (E*)T, ldt)
    if (prec == DFT_F64) DFT_ADD_BOX(double);
    else if (prec == DFT_F32) DFT_ADD_BOX(float);
    else return set_err(c, DFT_EINVAL, "dft_cnn_add_box: bad precision");
#undef DFT_ADD_BOX
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

int dft_cnn_finish(dft_ctx* c, const dft_grid* g, int prec, int ncols, const float* out, const double* scale_dev, void* T,
                   int64_t ldt) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, valid_grid(g) && out && scale_dev && T && ncols > 0 && ncols <= 65535, "dft_cnn_finish: bad arguments");
    const int64_t N = grid_points(g);
    const double h2 = 1.0 / (1.0 / (g->dx * g->dx) + 1.0 / (g->dy * g->dy) + 1.0 / (g->dz * g->dz));
    const dim3 grid((unsigned)((N + 255) / 256), ncols);
    if (prec == DFT_F64)
        k_cnn_finish<double><<<grid, 256, 0, c->stream>>>(N, ncols, out, scale_dev, h2, (double*)T, ldt);
    else if (prec == DFT_F32)
        k_cnn_finish<float><<<grid, 256, 0, c->stream>>>(N, ncols, out, scale_dev, h2, (float*)T, ldt);
    else
        return set_err(c, DFT_EINVAL, "dft_cnn_finish: bad precision");
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

}  // extern "C"
