// common.hpp -- context, error plumbing and device helpers shared by the
// gfx950 kernel library (libdft_hip.so).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "dft_hip.h"

struct dft_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // FP64 partial sums for the two-stage reductions (device) and a pinned
    // host slot the final value is copied to.
    double* partials = nullptr;
    size_t partials_cap = 0;  // doubles
    double* host_scalar = nullptr;  // pinned, 64 doubles
    // Poisson: per-sweep max|dV| (bit patterns of non-negative doubles)
    unsigned long long* sweep_max = nullptr;
    size_t sweep_max_cap = 0;
    // general device scratch (grown on demand)
    void* scratch = nullptr;
    size_t scratch_cap = 0;
    int num_cus = 256;
    // BF16 sine-transform matrices of the exact kinetic correction, for the
    // grid shape dst_shape (built on first use)
    void* dst_mats = nullptr;
    int dst_shape[3] = {0, 0, 0};
    // rocBLAS handle for the dense FP64 subspace solvers (created on first use)
    void* blas = nullptr;
    // device address, on this context's device, of the fixed energy-partials
    // array of the allocation-free XC entry point (xc.hip), set at creation
    double* xc_scratch = nullptr;
    // three-piece bf16 split of a coefficient block for the bf16 MFMA update
    // (linalg.hip), grown on demand
    void* coef3 = nullptr;
    size_t coef3_cap = 0;
    int* dev_info = nullptr;
};

namespace dft {

inline int set_err(dft_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

#define DFT_HIP(ctx, call)                                                                    \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return dft::set_err((ctx), DFT_EHIP,                                              \
                                std::string(#call) + ": " + hipGetErrorString(e_));          \
    } while (0)

#define DFT_CHECK_LAUNCH(ctx)                                                                 \
    do {                                                                                      \
        hipError_t e_ = hipGetLastError();                                                    \
        if (e_ != hipSuccess)                                                                 \
            return dft::set_err((ctx), DFT_EHIP, std::string("launch: ") + hipGetErrorString(e_)); \
    } while (0)

#define DFT_REQUIRE(ctx, cond, msg)                                                           \
    do {                                                                                      \
        if (!(cond)) return dft::set_err((ctx), DFT_EINVAL, (msg));                           \
    } while (0)

// Grow ctx->partials to at least n doubles.
int ensure_partials(dft_ctx* c, size_t n);
// Destroy the rocBLAS handle (linalg.hip).
void release_blas(dft_ctx* c);
// The context's rocBLAS handle (a rocblas_handle), created on first use and
// bound to the context stream (linalg.hip).
int rocblas_for(dft_ctx* c, void** handle);
// Grow ctx->scratch to at least bytes.
int ensure_scratch(dft_ctx* c, size_t bytes);
// Sum of n FP64 device values already written to ctx->partials -> host.
int finish_sum(dft_ctx* c, int64_t n, double* out_host);
// The same for n values at `partials` (device); writes partials[n].
int finish_sum_at(dft_ctx* c, double* partials, int64_t n, double* out_host);
// Look up the allocation-free XC entry point's device array on c's device
// (the device must be current); sets c->xc_scratch (xc.hip).
int xc_scratch_attach(dft_ctx* c);

inline size_t prec_bytes(int prec) { return prec == DFT_F64 ? 8 : 4; }
inline int64_t grid_points(const dft_grid* g) { return g->nx * g->ny * g->nz; }
inline int valid_grid(const dft_grid* g) { return g && g->nx >= 2 && g->ny >= 2 && g->nz >= 2 && g->dx > 0 && g->dy > 0 && g->dz > 0; }

// ---- device helpers -------------------------------------------------------
template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
template <typename T>
__device__ inline T wave_max(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        T w = __shfl_xor(v, o, 64);
        v = v < w ? w : v;
    }
    return v;
}

// Block-wide sum into lane 0 of wave 0 (blockDim.x multiple of 64, <= 1024).
template <typename T>
__device__ inline T block_sum(T v, T* lds32) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) lds32[w] = v;
    __syncthreads();
    T s = 0;
    if (threadIdx.x == 0) {
        const int nw = blockDim.x >> 6;
        for (int i = 0; i < nw; ++i) s += lds32[i];
    }
    return s;
}

// Workgroup barrier for LDS traffic only: waits for this wave's LDS
// operations, not for its global loads, so loads issued for the next step
// stay in flight across it.
__device__ inline void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// XCD-aware tile order.  Workgroups are dispatched round-robin over the 8
// XCDs (linear workgroup id b runs on XCD b % 8); this maps b to a tile index
// so that each XCD owns one contiguous range of tiles, whose neighbouring
// tiles then share that XCD's L2 (halo lines are fetched from HBM once).
__device__ inline int64_t xcd_tile(int64_t b, int64_t total) {
    const int64_t per = total / 8, rem = total % 8;
    const int64_t xcd = b % 8, slot = b / 8;
    return xcd < rem ? xcd * (per + 1) + slot : rem * (per + 1) + (xcd - rem) * per + slot;
}

// Ordered bit pattern of a non-negative double for atomicMax.
__device__ inline unsigned long long dbits(double x) { return (unsigned long long)__double_as_longlong(x); }

}  // namespace dft
