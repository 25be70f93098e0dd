// context.hip -- context lifetime, device memory, events, elementwise
// helpers and the deterministic two-stage FP64 reductions.
#include "common.hpp"

#include <cmath>
#include <new>

using namespace dft;

namespace dft {

int ensure_partials(dft_ctx* c, size_t n) {
    if (c->partials_cap >= n) return 0;
    if (c->partials) {
        DFT_HIP(c, hipStreamSynchronize(c->stream));
        DFT_HIP(c, hipFree(c->partials));
        c->partials = nullptr;
    }
    size_t cap = n < 4096 ? 4096 : n;
    DFT_HIP(c, hipMalloc(&c->partials, cap * sizeof(double)));
    c->partials_cap = cap;
    return 0;
}

int ensure_scratch(dft_ctx* c, size_t bytes) {
    if (c->scratch_cap >= bytes) return 0;
    if (c->scratch) {
        DFT_HIP(c, hipStreamSynchronize(c->stream));
        DFT_HIP(c, hipFree(c->scratch));
        c->scratch = nullptr;
    }
    DFT_HIP(c, hipMalloc(&c->scratch, bytes));
    c->scratch_cap = bytes;
    return 0;
}

// One block sums n partials in a fixed order (deterministic), result in out[0].
__global__ __launch_bounds__(256) void k_final_sum(const double* __restrict__ p, int64_t n, double* out, int op_max) {
    __shared__ double lds[4];
    double s = op_max ? 0.0 : 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) {
        const double v = p[i];
        s = op_max ? (s < v ? v : s) : s + v;
    }
    if (op_max) {
        s = wave_max(s);
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        if (lane == 0) lds[w] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            double m = lds[0];
            for (int i = 1; i < 4; ++i) m = m < lds[i] ? lds[i] : m;
            out[0] = m;
        }
    } else {
        s = block_sum(s, lds);
        if (threadIdx.x == 0) out[0] = s;
    }
}

static int finish(dft_ctx* c, double* partials, int64_t n, double* out_host, int op_max) {
    k_final_sum<<<1, 256, 0, c->stream>>>(partials, n, partials + n, op_max);
    DFT_CHECK_LAUNCH(c);
    DFT_HIP(c, hipMemcpyAsync(c->host_scalar, partials + n, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    DFT_HIP(c, hipStreamSynchronize(c->stream));
    *out_host = c->host_scalar[0];
    return 0;
}

int finish_sum(dft_ctx* c, int64_t n, double* out_host) { return finish(c, c->partials, n, out_host, 0); }
int finish_sum_at(dft_ctx* c, double* partials, int64_t n, double* out_host) {
    return finish(c, partials, n, out_host, 0);
}

}  // namespace dft

template <typename T, int OP>
__global__ __launch_bounds__(256) void k_reduce(const T* __restrict__ a, const T* __restrict__ b, int64_t n,
                                                double* __restrict__ partials) {
    __shared__ double lds[4];
    double s = 0.0;
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
        const double x = (double)a[i];
        if (OP == DFT_RED_SUM) s += x;
        else if (OP == DFT_RED_DOT) s += x * (double)b[i];
        else if (OP == DFT_RED_MAXABS) { const double y = fabs(x); s = s < y ? y : s; }
        else { const double d = x - (double)b[i]; s += d * d; }
    }
    if (OP == DFT_RED_MAXABS) {
        s = wave_max(s);
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        if (lane == 0) lds[w] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            double m = lds[0];
            for (int i = 1; i < 4; ++i) m = m < lds[i] ? lds[i] : m;
            partials[blockIdx.x] = m;
        }
    } else {
        s = block_sum(s, lds);
        if (threadIdx.x == 0) partials[blockIdx.x] = s;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_axpby(int64_t n, double alpha, const T* __restrict__ x, double beta,
                                               T* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = (T)(alpha * (double)x[i] + beta * (double)y[i]);
}

template <typename T>
__global__ __launch_bounds__(256) void k_clamp_min(int64_t n, T lo, T* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n && !(y[i] >= lo)) y[i] = lo;  // NaN becomes lo too
}

template <typename T>
__global__ __launch_bounds__(256) void k_add3(int64_t n, const T* __restrict__ a, const T* __restrict__ b,
                                              const T* __restrict__ c, T* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    // (a + b) + c: MolecularSCFIteration.cpp:106-114 order
    T s = a[i] + b[i];
    if (c) s = s + c[i];
    out[i] = s;
}

template <typename S, typename D>
__global__ __launch_bounds__(256) void k_convert(const S* __restrict__ x, D* __restrict__ y, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = (D)x[i];
}

static inline unsigned grid1d(int64_t n) { return (unsigned)((n + 255) / 256); }

extern "C" {

int dft_grid_init(dft_grid* g, int64_t nx, int64_t ny, int64_t nz, double xmin, double xmax, double ymin, double ymax,
                  double zmin, double zmax) {
    if (!g || nx < 2 || ny < 2 || nz < 2) return DFT_EINVAL;
    if (xmax <= xmin || ymax <= ymin || zmax <= zmin) return DFT_EINVAL;
    g->nx = nx; g->ny = ny; g->nz = nz;
    g->xmin = xmin; g->xmax = xmax; g->ymin = ymin; g->ymax = ymax; g->zmin = zmin; g->zmax = zmax;
    g->dx = (xmax - xmin) / (double)(nx - 1);
    g->dy = (ymax - ymin) / (double)(ny - 1);
    g->dz = (zmax - zmin) / (double)(nz - 1);
    return DFT_OK;
}

int dft_ctx_create(int device, dft_ctx** out) {
    if (!out) return DFT_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) return DFT_EHIP;
    if (hipSetDevice(device) != hipSuccess) return DFT_EHIP;
    dft_ctx* c = new (std::nothrow) dft_ctx();
    if (!c) return DFT_EHIP;
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cus = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipHostMalloc((void**)&c->host_scalar, 64 * sizeof(double), hipHostMallocDefault) != hipSuccess ||
        hipMalloc((void**)&c->sweep_max, 4096 * sizeof(unsigned long long)) != hipSuccess) {
        delete c;
        return DFT_EHIP;
    }
    c->sweep_max_cap = 4096;
    if (ensure_partials(c, 8192) != 0 || dft::xc_scratch_attach(c) != 0) {
        (void)hipFree(c->partials);
        (void)hipFree(c->sweep_max);
        (void)hipHostFree(c->host_scalar);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return DFT_EHIP;
    }
    *out = c;
    return DFT_OK;
}

int dft_ctx_destroy(dft_ctx* c) {
    if (!c) return DFT_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->partials) (void)hipFree(c->partials);
    if (c->scratch) (void)hipFree(c->scratch);
    if (c->coef3) (void)hipFree(c->coef3);
    if (c->sweep_max) (void)hipFree(c->sweep_max);
    if (c->dst_mats) (void)hipFree(c->dst_mats);
    if (c->host_scalar) (void)hipHostFree(c->host_scalar);
    dft::release_blas(c);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return DFT_OK;
}

const char* dft_last_error(const dft_ctx* c) { return c ? c->err.c_str() : "null context"; }

int dft_sync(dft_ctx* c) {
    if (!c) return DFT_EINVAL;
    DFT_HIP(c, hipStreamSynchronize(c->stream));
    return DFT_OK;
}

void* dft_stream(dft_ctx* c) { return c ? (void*)c->stream : nullptr; }

int dft_alloc(dft_ctx* c, size_t bytes, void** p) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, p != nullptr, "dft_alloc: null out pointer");
    *p = nullptr;
    if (bytes == 0) return DFT_OK;
    DFT_HIP(c, hipMalloc(p, bytes));
    return DFT_OK;
}

int dft_free(dft_ctx* c, void* p) {
    if (!c) return DFT_EINVAL;
    if (!p) return DFT_OK;
    DFT_HIP(c, hipStreamSynchronize(c->stream));
    DFT_HIP(c, hipFree(p));
    return DFT_OK;
}

int dft_upload(dft_ctx* c, void* dst, const void* src, size_t bytes) {
    if (!c) return DFT_EINVAL;
    if (!bytes) return DFT_OK;
    DFT_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    DFT_HIP(c, hipStreamSynchronize(c->stream));
    return DFT_OK;
}

int dft_download(dft_ctx* c, void* dst, const void* src, size_t bytes) {
    if (!c) return DFT_EINVAL;
    if (!bytes) return DFT_OK;
    DFT_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    DFT_HIP(c, hipStreamSynchronize(c->stream));
    return DFT_OK;
}

int dft_copy(dft_ctx* c, void* dst, const void* src, size_t bytes) {
    if (!c) return DFT_EINVAL;
    if (!bytes) return DFT_OK;
    DFT_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
    return DFT_OK;
}

int dft_zero(dft_ctx* c, void* p, size_t bytes) {
    if (!c) return DFT_EINVAL;
    if (!bytes) return DFT_OK;
    DFT_HIP(c, hipMemsetAsync(p, 0, bytes, c->stream));
    return DFT_OK;
}

int dft_convert(dft_ctx* c, const void* x, int ps, void* y, int pd, int64_t n) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, n >= 0 && (ps == DFT_F64 || ps == DFT_F32) && (pd == DFT_F64 || pd == DFT_F32), "dft_convert: bad args");
    if (!n) return DFT_OK;
    if (ps == DFT_F64 && pd == DFT_F32) k_convert<double, float><<<grid1d(n), 256, 0, c->stream>>>((const double*)x, (float*)y, n);
    else if (ps == DFT_F32 && pd == DFT_F64) k_convert<float, double><<<grid1d(n), 256, 0, c->stream>>>((const float*)x, (double*)y, n);
    else return dft_copy(c, y, x, (size_t)n * prec_bytes(ps));
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

int dft_event_create(dft_ctx* c, void** ev) {
    if (!c) return DFT_EINVAL;
    hipEvent_t e;
    DFT_HIP(c, hipEventCreate(&e));
    *ev = (void*)e;
    return DFT_OK;
}
int dft_event_record(dft_ctx* c, void* ev) {
    if (!c) return DFT_EINVAL;
    DFT_HIP(c, hipEventRecord((hipEvent_t)ev, c->stream));
    return DFT_OK;
}
int dft_event_elapsed_ms(dft_ctx* c, void* a, void* b, float* ms) {
    if (!c) return DFT_EINVAL;
    DFT_HIP(c, hipEventSynchronize((hipEvent_t)b));
    DFT_HIP(c, hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
    return DFT_OK;
}
int dft_event_destroy(dft_ctx* c, void* ev) {
    if (!c) return DFT_EINVAL;
    DFT_HIP(c, hipEventDestroy((hipEvent_t)ev));
    return DFT_OK;
}

int dft_reduce(dft_ctx* c, int op, int prec, const void* a, const void* b, int64_t n, double* out) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, out && a && n >= 0 && op >= 0 && op <= 3, "dft_reduce: bad args");
    DFT_REQUIRE(c, (op != DFT_RED_DOT && op != DFT_RED_DIFF2) || b, "dft_reduce: second operand required");
    if (n == 0) { *out = 0.0; return DFT_OK; }
    int blocks = (int)std::min<int64_t>((n + 255) / 256, 4 * (int64_t)c->num_cus);
    if (int rc = ensure_partials(c, (size_t)blocks + 1)) return rc;
#define LAUNCH(T, OP) k_reduce<T, OP><<<blocks, 256, 0, c->stream>>>((const T*)a, (const T*)b, n, c->partials)
    if (prec == DFT_F64) {
        switch (op) {
            case DFT_RED_SUM: LAUNCH(double, DFT_RED_SUM); break;
            case DFT_RED_DOT: LAUNCH(double, DFT_RED_DOT); break;
            case DFT_RED_MAXABS: LAUNCH(double, DFT_RED_MAXABS); break;
            default: LAUNCH(double, DFT_RED_DIFF2); break;
        }
    } else {
        switch (op) {
            case DFT_RED_SUM: LAUNCH(float, DFT_RED_SUM); break;
            case DFT_RED_DOT: LAUNCH(float, DFT_RED_DOT); break;
            case DFT_RED_MAXABS: LAUNCH(float, DFT_RED_MAXABS); break;
            default: LAUNCH(float, DFT_RED_DIFF2); break;
        }
    }
#undef LAUNCH
    DFT_CHECK_LAUNCH(c);
    return finish(c, c->partials, blocks, out, op == DFT_RED_MAXABS);
}

int dft_axpby(dft_ctx* c, int prec, int64_t n, double alpha, const void* x, double beta, void* y) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, n >= 0 && x && y, "dft_axpby: bad args");
    if (!n) return DFT_OK;
    if (prec == DFT_F64) k_axpby<double><<<grid1d(n), 256, 0, c->stream>>>(n, alpha, (const double*)x, beta, (double*)y);
    else k_axpby<float><<<grid1d(n), 256, 0, c->stream>>>(n, alpha, (const float*)x, beta, (float*)y);
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

int dft_clamp_min(dft_ctx* c, int prec, int64_t n, double lo, void* y) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, n >= 0 && y && (prec == DFT_F64 || prec == DFT_F32), "dft_clamp_min: bad args");
    if (!n) return DFT_OK;
    if (prec == DFT_F64) k_clamp_min<double><<<grid1d(n), 256, 0, c->stream>>>(n, lo, (double*)y);
    else k_clamp_min<float><<<grid1d(n), 256, 0, c->stream>>>(n, (float)lo, (float*)y);
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

int dft_add3(dft_ctx* c, int prec, int64_t n, const void* a, const void* b, const void* d, void* out) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, n >= 0 && a && b && out, "dft_add3: bad args");
    if (!n) return DFT_OK;
    if (prec == DFT_F64) k_add3<double><<<grid1d(n), 256, 0, c->stream>>>(n, (const double*)a, (const double*)b, (const double*)d, (double*)out);
    else k_add3<float><<<grid1d(n), 256, 0, c->stream>>>(n, (const float*)a, (const float*)b, (const float*)d, (float*)out);
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

}  // extern "C"
