// setup.hip -- one-time fields: Davidson seed vectors, the hydrogenic
// initial density, and small FP64 matrix helpers for the subspace algebra.
#include "common.hpp"

#include <cmath>

using namespace dft;

namespace {

// buildInitialVector (EigenvalueSolverMolecule.cpp:177-283): Gaussian of
// width 0.8 + 0.18*idx about the box centre times one of nine polynomials.
template <typename T>
__global__ __launch_bounds__(256) void k_seed(int nx, int ny, int nz, double xmin, double xmax, double ymin, double ymax,
                                              double zmin, double zmax, double dx, double dy, double dz, int first,
                                              T* __restrict__ out, int64_t ld) {
    const int64_t n = (int64_t)nx * ny * nz;
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
    const int idx = first + (int)blockIdx.y;
    const int i = (int)(p % nx), j = (int)((p / nx) % ny), k = (int)(p / ((int64_t)nx * ny));
    const double cx = 0.5 * (xmin + xmax), cy = 0.5 * (ymin + ymax), cz = 0.5 * (zmin + zmax);
    const double sigma = 0.8 + 0.18 * (double)idx;
    const double factor = 1.0 / (2.0 * sigma * sigma);
    const double x = (xmin + (double)i * dx) - cx, y = (ymin + (double)j * dy) - cy, z = (zmin + (double)k * dz) - cz;
    const double r2 = x * x + y * y + z * z;
    const double ga = exp(-r2 * factor);
    double poly = 1.0;
    switch (idx % 9) {
        case 1: poly = x; break;
        case 2: poly = y; break;
        case 3: poly = z; break;
        case 4: poly = x * y; break;
        case 5: poly = x * z; break;
        case 6: poly = y * z; break;
        case 7: poly = x * x - y * y; break;
        case 8: poly = 2.0 * z * z - x * x - y * y; break;
        default: break;
    }
    out[(int64_t)blockIdx.y * ld + p] = (T)(poly * ga);
}

// Sum over atoms of Z * (Z^3/pi) exp(-2 Z r): hydrogenic 1s densities
// weighted by nuclear charge (unnormalised).
template <typename T>
__global__ __launch_bounds__(256) void k_hydrogenic(int nx, int ny, int nz, double xmin, double ymin, double zmin,
                                                    double dx, double dy, double dz, int natoms,
                                                    const double* __restrict__ nuc, T* __restrict__ out) {
    const int64_t n = (int64_t)nx * ny * nz;
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
    const int i = (int)(p % nx), j = (int)((p / nx) % ny), k = (int)(p / ((int64_t)nx * ny));
    const double x = xmin + (double)i * dx, y = ymin + (double)j * dy, z = zmin + (double)k * dz;
    double s = 0.0;
    for (int a = 0; a < natoms; ++a) {
        const double Z = nuc[4 * a];
        const double ddx = x - nuc[4 * a + 1], ddy = y - nuc[4 * a + 2], ddz = z - nuc[4 * a + 3];
        const double r = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
        s += Z * (Z * Z * Z / 3.141592653589793238462643383279502884) * exp(-2.0 * Z * r);
    }
    out[p] = (T)s;
}

// dst(j, i) = src(i, j) for an r x c block (column-major, FP64)
// Counter-based uniform values in [-1, 1): splitmix64 of (seed, column, point).
template <typename T>
__global__ __launch_bounds__(256) void k_random(int64_t n, uint64_t seed, T* __restrict__ out, int64_t ld) {
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int j = blockIdx.y;
    if (p >= n) return;
    uint64_t z = seed + 0x9E3779B97F4A7C15ULL * ((uint64_t)j * 0x100000000ULL + (uint64_t)p + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    out[(int64_t)j * ld + p] = (T)((double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0);
}

__global__ __launch_bounds__(256) void k_transpose(int r, int c, const double* __restrict__ src, int64_t lds,
                                                   double* __restrict__ dst, int64_t ldd) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)r * c) return;
    const int i = (int)(e % r), j = (int)(e / r);
    dst[(int64_t)i * ldd + j] = src[(int64_t)j * lds + i];
}

}  // namespace

extern "C" {

int dft_seed_vectors(dft_ctx* c, const dft_grid* g, int prec, int first_index, int count, void* out, int64_t ld) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, valid_grid(g) && out && count >= 0 && count <= 65535 && first_index >= 0, "dft_seed_vectors: bad arguments");
    const int64_t n = grid_points(g);
    DFT_REQUIRE(c, ld >= n, "dft_seed_vectors: ld < grid size");
    if (!count) return DFT_OK;
    dim3 grid((unsigned)((n + 255) / 256), count);
    if (prec == DFT_F64)
        k_seed<double><<<grid, 256, 0, c->stream>>>((int)g->nx, (int)g->ny, (int)g->nz, g->xmin, g->xmax, g->ymin, g->ymax, g->zmin, g->zmax, g->dx, g->dy, g->dz, first_index, (double*)out, ld);
    else
        k_seed<float><<<grid, 256, 0, c->stream>>>((int)g->nx, (int)g->ny, (int)g->nz, g->xmin, g->xmax, g->ymin, g->ymax, g->zmin, g->zmax, g->dx, g->dy, g->dz, first_index, (float*)out, ld);
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

int dft_hydrogenic_density(dft_ctx* c, const dft_grid* g, int prec, int natoms, const double* nuclei_host, double electrons,
                           void* out) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, valid_grid(g) && out && natoms > 0 && nuclei_host, "dft_hydrogenic_density: bad arguments");
    const int64_t n = grid_points(g);
    if (int rc = ensure_partials(c, (size_t)4 * natoms + 8)) return rc;
    DFT_HIP(c, hipMemcpyAsync(c->partials, nuclei_host, sizeof(double) * 4 * natoms, hipMemcpyHostToDevice, c->stream));
    DFT_HIP(c, hipStreamSynchronize(c->stream));
    const unsigned gb = (unsigned)((n + 255) / 256);
    if (prec == DFT_F64)
        k_hydrogenic<double><<<gb, 256, 0, c->stream>>>((int)g->nx, (int)g->ny, (int)g->nz, g->xmin, g->ymin, g->zmin, g->dx, g->dy, g->dz, natoms, c->partials, (double*)out);
    else
        k_hydrogenic<float><<<gb, 256, 0, c->stream>>>((int)g->nx, (int)g->ny, (int)g->nz, g->xmin, g->ymin, g->zmin, g->dx, g->dy, g->dz, natoms, c->partials, (float*)out);
    DFT_CHECK_LAUNCH(c);
    double s = 0.0;
    if (int rc = dft_reduce(c, DFT_RED_SUM, prec, out, nullptr, n, &s)) return rc;
    s *= g->dx * g->dy * g->dz;
    const double f = s > 0.0 ? electrons / s : 0.0;
    return dft_axpby(c, prec, n, 0.0, out, f, out);
}

int dft_transpose(dft_ctx* c, int rows, int cols, const double* src, int64_t lds, double* dst, int64_t ldd) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, src && dst && rows >= 0 && cols >= 0 && lds >= rows && ldd >= cols, "dft_transpose: bad arguments");
    const int64_t tot = (int64_t)rows * cols;
    if (!tot) return DFT_OK;
    k_transpose<<<(unsigned)((tot + 255) / 256), 256, 0, c->stream>>>(rows, cols, src, lds, dst, ldd);
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

int dft_random_vectors(dft_ctx* c, int prec, int64_t n, int count, uint64_t seed, void* out, int64_t ld) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, out && n > 0 && count >= 0 && count <= 65535 && ld >= n, "dft_random_vectors: bad arguments");
    if (!count) return DFT_OK;
    const dim3 grid((unsigned)((n + 255) / 256), count);
    if (prec == DFT_F64) k_random<double><<<grid, 256, 0, c->stream>>>(n, seed, (double*)out, ld);
    else if (prec == DFT_F32) k_random<float><<<grid, 256, 0, c->stream>>>(n, seed, (float*)out, ld);
    else return set_err(c, DFT_EINVAL, "dft_random_vectors: bad precision");
    DFT_CHECK_LAUNCH(c);
    return DFT_OK;
}

}  // extern "C"
