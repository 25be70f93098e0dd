// radial.hip -- the reference's radial (atomic) Kohn-Sham stack on the GPU,
// batched: one 256-thread workgroup per atom runs that atom's whole SCF, so
// a table of atoms (Z = 1..30 in two functionals) is one launch.
//
// Per SCF iteration (SelfConsistentField.cpp, AtomicDFT.cpp):
//  * effective potentials: the radial Hartree potential by workgroup prefix
//    and suffix scans of the trapezoid terms (HartreePotential.cpp:8-86), XC
//    with radial differences through the same FP64 device functionals as the
//    3D path (ExchangeCorrelation.cpp:11-279), V = V_xc - Z/r + V_H;
//  * every occupied orbital: waves 0 and 1 take the orbitals in turn.  The
//    eigenvalue is the reference's Sturm-count bisection
//    (EigenvalueSolver.cpp:37-215) run as a multisection: the 63 midpoints of
//    the next six bisection levels are counted at once, one per lane, and
//    the lanes then walk the bisection path, so the result is the
//    sequential bisection's to the bit.  The vector is the reference's
//    inverse iteration (:217-312): the Thomas solve runs on one lane over
//    LDS rows, the norms, dot products and differences on the wave;
//  * output densities, energies on them (TotalEnergy.cpp), dRho, |dE| and the
//    KS residual as fixed-order workgroup reductions; the reference's stop
//    rule and adaptive linear mixing.
// Every loop has the reference's bound (250 SCF iterations, 100 bracket
// doublings, 300 bisections, 1000 inverse iterations); a numerical failure
// sets the job's status and ends its SCF.
#include "common.hpp"
#include "xc_functionals.hpp"

#include <cmath>
#include <vector>

namespace {

constexpr int kRThreads = 256;       // one atom per workgroup
constexpr int kROrbWaves = 2;        // waves that solve orbitals
constexpr int kRMaxPoints = 2048;    // LDS rows per orbital wave: 3 x 16 KB
constexpr int kRMaxOrb = 14;         // 7 subshells x 2 spins
constexpr double kPi = 3.141592653589793238462643383279502884;  // DFTConstants.h:7
constexpr double kEps = 1.0e-14;                                  // DFTConstants.h:8
constexpr double kRhoFloor = 1.0e-20;                             // DFTConstants.h:9

struct RadialJob {
    int Z, functional, norb, points;
    double dr;
    // per orbital: n, l, spin (0 alpha, 1 beta), electrons, eigenvalue index within its l and spin
    int on[kRMaxOrb], ol[kRMaxOrb], ospin[kRMaxOrb], oe[kRMaxOrb], oidx[kRMaxOrb];
};

// output row per job: {kinetic, external, hartree, xc, total, dRho, |dE|, KS residual,
// iterations, converged, status, mixing, eig[14]}
constexpr int kROut = 12 + kRMaxOrb;
// work arrays per job, each `points` doubles
enum { W_NA, W_NB, W_OA, W_OB, W_VA, W_VB, W_VH, W_RHO, W_DDA, W_DDB, W_CFA, W_CFB, W_T1, W_T2, W_IV0, W_IV1,
       W_U, W_COUNT = W_U + kRMaxOrb };

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}
__device__ inline double wave_min(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
    return v;
}

// workgroup sum in a fixed order (waves' sums, then wave 0 .. 3)
__device__ double block_sum(double v, double* red) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    v = wave_sum(v);
    if (lane == 0) red[w] = v;
    __syncthreads();
    const double s = (red[0] + red[1]) + (red[2] + red[3]);
    __syncthreads();
    return s;
}

__device__ inline double rpt(int i, double dr) { return (double)(i + 1) * dr; }  // RadialGrid.cpp:17-19

// work array k of the job whose arrays start at base (n = J.points in scope)
#define WA(k) (base + (int64_t)(k) * n)

// inclusive scan of t[0..n) (t[0] the first term) into q, forward or backward, by
// contiguous chunks per thread; red holds kRThreads doubles
__device__ void block_scan(const double* t, double* q, int n, bool backward, double* red) {
    const int tid = threadIdx.x;
    const int chunk = (n + kRThreads - 1) / kRThreads;
    const int b = tid * chunk, e = min(n, b + chunk);
    double s = 0.0;
    for (int k = b; k < e; ++k) s += t[backward ? n - 1 - k : k];
    red[tid] = s;
    __syncthreads();
    // exclusive prefix of the chunk totals (Hillis-Steele over kRThreads values)
    for (int o = 1; o < kRThreads; o <<= 1) {
        const double x = tid >= o ? red[tid - o] : 0.0;
        __syncthreads();
        red[tid] += x;
        __syncthreads();
    }
    double acc = tid > 0 ? red[tid - 1] : 0.0;
    for (int k = b; k < e; ++k) {
        const int i = backward ? n - 1 - k : k;
        acc += t[i];
        q[i] = acc;
    }
    __syncthreads();
}

// HartreePotential.cpp:8-86 on the workgroup: V_H = Q(r)/r + outer integral
__device__ void hartree(const double* rho, double* vh, double* t1, double* t2, int n, double dr, double* red) {
    for (int i = threadIdx.x; i < n; i += kRThreads) {
        // enclosed-charge term i (i >= 1) and outer term i (i <= n - 2), the reference's expressions
        double a = 0.0, b = 0.0;
        if (i >= 1) {
            const double r0 = rpt(i - 1, dr), r1 = rpt(i, dr);
            const double f0 = 4.0 * kPi * r0 * r0 * rho[i - 1];
            const double f1 = 4.0 * kPi * r1 * r1 * rho[i];
            a = 0.5 * (f0 + f1) * dr;
        }
        if (i <= n - 2) {
            const double r0 = rpt(i, dr), r1 = rpt(i + 1, dr);
            const double f0 = 4.0 * kPi * r0 * rho[i];
            const double f1 = 4.0 * kPi * r1 * rho[i + 1];
            b = 0.5 * (f0 + f1) * dr;
        }
        t1[i] = a;
        t2[i] = b;
    }
    __syncthreads();
    block_scan(t1, t1, n, false, red);  // Q[i] = sum_{k <= i} a_k
    block_scan(t2, t2, n, true, red);   // O[i] = sum_{k >= i} b_k
    for (int i = threadIdx.x; i < n; i += kRThreads) vh[i] = t1[i] / rpt(i, dr) + t2[i];
    __syncthreads();
}

// ExchangeCorrelation.cpp:11-59
__device__ inline double rgrad(const double* f, int i, int n, double dr) {
    if (i == 0) return (f[1] - f[0]) / (rpt(1, dr) - rpt(0, dr));
    if (i == n - 1) return (f[n - 1] - f[n - 2]) / (rpt(n - 1, dr) - rpt(n - 2, dr));
    return (f[i + 1] - f[i - 1]) / (rpt(i + 1, dr) - rpt(i - 1, dr));
}
// ExchangeCorrelation.cpp:61-163: (1/r^2) d(r^2 c)/dr
__device__ inline double rdiv(const double* c, int i, int n, double dr) {
    int i0 = i - 1, i1 = i + 1;
    if (i == 0) i0 = 0, i1 = 1;
    if (i == n - 1) i0 = n - 2, i1 = n - 1;
    const double ra = rpt(i0, dr), rb = rpt(i1, dr), ri = rpt(i, dr);
    return ((rb * rb * c[i1] - ra * ra * c[i0]) / (rb - ra)) / (ri * ri);
}

__device__ inline dft::xc::Out<double> xc_eval(int functional, double a, double b, double ga, double gb) {
    if (functional == 1) return dft::xc::pz81<double>(a, b);
    return dft::xc::pbe96<double>(a, b, ga, gb);
}

// V_sigma = V_xc,sigma - Z/r + V_H (SelfConsistentField.cpp buildEffectivePotentials)
__device__ void effective_potentials(const RadialJob& J, double* base, double* red) {
    const int n = J.points;
    const double dr = J.dr;
    double *na = WA(W_NA), *nb = WA(W_NB);
    for (int i = threadIdx.x; i < n; i += kRThreads) {
        WA(W_RHO)[i] = na[i] + nb[i];
        const dft::xc::Out<double> r = xc_eval(J.functional, na[i], nb[i], rgrad(na, i, n, dr), rgrad(nb, i, n, dr));
        WA(W_DDA)[i] = r.va;
        WA(W_DDB)[i] = r.vb;
        WA(W_CFA)[i] = r.ca;
        WA(W_CFB)[i] = r.cb;
    }
    __syncthreads();
    hartree(WA(W_RHO), WA(W_VH), WA(W_T1), WA(W_T2), n, dr, red);
    for (int i = threadIdx.x; i < n; i += kRThreads) {
        const bool empty = na[i] + nb[i] <= kRhoFloor;
        const double xa = empty ? 0.0 : WA(W_DDA)[i] - rdiv(WA(W_CFA), i, n, dr);
        const double xb = empty ? 0.0 : WA(W_DDB)[i] - rdiv(WA(W_CFB), i, n, dr);
        const double vn = -(double)J.Z / rpt(i, dr);
        WA(W_VA)[i] = xa + (vn + WA(W_VH)[i]);
        WA(W_VB)[i] = xb + (vn + WA(W_VH)[i]);
    }
    __syncthreads();
}

// EigenvalueSolver.cpp:37-90: negative pivots of the LDL^T of H - e (one lane)
__device__ int sturm_count(const double* diag, double off2, int n, double e) {
    int count = 0;
    double prev = diag[0] - e;
    if (prev < 0.0) ++count;
    if (fabs(prev) < kEps) prev = copysign(kEps, prev == 0.0 ? 1.0 : prev);
    for (int i = 1; i < n; ++i) {
        const double pivot = diag[i] - e - off2 / prev;
        if (pivot < 0.0) ++count;
        prev = fabs(pivot) < kEps ? copysign(kEps, pivot == 0.0 ? 1.0 : pivot) : pivot;
    }
    return count;
}

// One orbital on one wave (EigenvalueSolver.cpp:314-403 solveOrbital).  lds: 3 rows of n
// doubles (H's diagonal, then H - shift; the Thomas pivots; the solution).  iv: the iterate
// (global, n).  Returns 0, or -2 on a numerical failure.
__device__ int solve_orbital(const double* V, int l, int index, int n, double dr, double* lds, double* iv, double* u,
                             double* eig) {
    const int lane = threadIdx.x & 63;
    double* sh = lds;
    double* dp = lds + n;
    double* x = lds + 2 * n;
    const double kd = 1.0 / (dr * dr);        // KohnShamHamiltonian.cpp:62-66
    const double ko = -0.5 * (1.0 / (dr * dr));
    const double cc = 0.5 * (double)(l * (l + 1));
    const double off2 = ko * ko;
    double mn = INFINITY, mx = -INFINITY;
    for (int i = lane; i < n; i += 64) {
        const double ri = rpt(i, dr);
        const double d = kd + cc / (ri * ri) + V[i];
        sh[i] = d;
        mn = fmin(mn, d);
        mx = fmax(mx, d);
    }
    mn = wave_min(mn);
    mx = wave_max(mx);
    __builtin_amdgcn_wave_barrier();
    // bracket [min - 10, max + 10], doubled until it holds eigenvalue `index` (lanes 0 and 1 count)
    double lo = mn - 10.0, hi = mx + 10.0;
    int c = sturm_count(sh, off2, n, lane == 0 ? lo : hi);
    int lc = __shfl(c, 0), hc = __shfl(c, 1);
    double width = hi - lo;
    for (int xpand = 0; xpand < 100; ++xpand) {
        if (lc <= index && index < hc) break;
        width *= 2.0;
        if (lc > index) lo -= width;
        if (hc <= index) hi += width;
        c = sturm_count(sh, off2, n, lane == 0 ? lo : hi);
        if (lc > index) lc = __shfl(c, 0);
        if (hc <= index) hc = __shfl(c, 1);
    }
    if (lc > index || index >= hc) return -2;
    // bisection, six levels per round: lane j < 63 counts at the midpoint of node j of the
    // next six levels' tree (breadth-first; children of j are 2j + 1 (left) and 2j + 2 (right))
    int it = 0;
    bool done = false;
    while (!done) {
        double nlo = lo, nhi = hi;
        if (lane < 63) {
            const int level = 31 - __clz(lane + 1), pos = lane + 1 - (1 << level);
            for (int bit = level - 1; bit >= 0; --bit) {
                const double m = 0.5 * (nlo + nhi);
                if ((pos >> bit) & 1) nlo = m;
                else nhi = m;
            }
        }
        const int cnt = lane < 63 ? sturm_count(sh, off2, n, 0.5 * (nlo + nhi)) : 0;
        int node = 0;
        for (int level = 0; level < 6 && !done; ++level) {
            const double mid = 0.5 * (lo + hi);
            if (__shfl(cnt, node) <= index) {
                lo = mid;
                node = 2 * node + 2;
            } else {
                hi = mid;
                node = 2 * node + 1;
            }
            ++it;
            const double tol = 1.0e-13 * fmax(1.0, fmax(fabs(lo), fabs(hi)));
            if (hi - lo <= tol || it >= 300) done = true;
        }
    }
    const double e = 0.5 * (lo + hi);
    // inverse iteration from sin(i + 1) at e + 1e-10 max(1, |e|)
    double s = 0.0;
    for (int i = lane; i < n; i += 64) {
        const double v = sin((double)(i + 1));
        iv[i] = v;
        s += v * v;
    }
    const double nrm0 = sqrt(wave_sum(s));
    if (nrm0 < kEps) return -2;
    for (int i = lane; i < n; i += 64) iv[i] /= nrm0;
    const double shift = e + 1.0e-10 * fmax(1.0, fabs(e));
    for (int i = lane; i < n; i += 64) sh[i] -= shift;
    __builtin_amdgcn_wave_barrier();
    for (int iter = 0; iter < 1000; ++iter) {
        for (int i = lane; i < n; i += 64) {
            dp[i] = sh[i];
            x[i] = iv[i];
        }
        __builtin_amdgcn_wave_barrier();
        int bad = 0;
        if (lane == 0) {  // NumericalMethods.cpp:88-138, the Thomas algorithm
            for (int i = 1; i < n && !bad; ++i) {
                if (fabs(dp[i - 1]) < kEps) {
                    bad = 1;
                    break;
                }
                const double f = ko / dp[i - 1];
                dp[i] -= f * ko;
                x[i] -= f * x[i - 1];
            }
            if (!bad && fabs(dp[n - 1]) < kEps) bad = 1;
            if (!bad) x[n - 1] /= dp[n - 1];
            for (int i = n - 1; i-- > 0 && !bad;) {
                if (fabs(dp[i]) < kEps) {
                    bad = 1;
                    break;
                }
                x[i] = (x[i] - ko * x[i + 1]) / dp[i];
            }
        }
        if (__shfl(bad, 0)) return -2;
        __builtin_amdgcn_wave_barrier();
        double q = 0.0;
        for (int i = lane; i < n; i += 64) q += x[i] * x[i];
        const double nn = sqrt(wave_sum(q));
        if (nn < kEps) return -2;
        double d = 0.0;
        for (int i = lane; i < n; i += 64) d += iv[i] * (x[i] / nn);
        const double sg = wave_sum(d) < 0.0 ? -1.0 : 1.0;
        double diff = 0.0;
        for (int i = lane; i < n; i += 64) {
            const double v = sg * (x[i] / nn);
            diff = fmax(diff, fabs(iv[i] - v));
            iv[i] = v;
        }
        diff = wave_max(diff);
        __builtin_amdgcn_wave_barrier();
        if (diff < 1.0e-12) break;
    }
    // NumericalMethods.cpp:40-60: u / sqrt(dr u.u)
    double q = 0.0;
    for (int i = lane; i < n; i += 64) q += iv[i] * iv[i];
    const double nrm = sqrt(dr * wave_sum(q));
    for (int i = lane; i < n; i += 64) u[i] = iv[i] / nrm;
    *eig = e;
    return 0;
}

// ElectronDensity.cpp:8-55
__device__ void spin_densities(const RadialJob& J, double* base, double* oa, double* ob) {
    const int n = J.points;
    for (int i = threadIdx.x; i < n; i += kRThreads) {
        const double ri = rpt(i, J.dr);
        double a = 0.0, b = 0.0;
        for (int k = 0; k < J.norb; ++k) {
            const double uk = WA(W_U + k)[i];
            const double t = (double)J.oe[k] * uk * uk / (4.0 * kPi * ri * ri);
            if (J.ospin[k] == 0) a += t;
            else b += t;
        }
        oa[i] = a;
        ob[i] = b;
    }
    __syncthreads();
}

// relative 4 pi r^2-weighted L2 difference
__device__ double density_difference(const double* old, const double* nw, int n, double dr, double* red) {
    double dn = 0.0, nn = 0.0;
    for (int i = threadIdx.x; i < n; i += kRThreads) {
        const double ri = rpt(i, dr);
        const double w = 4.0 * kPi * ri * ri;
        const double d = nw[i] - old[i];
        dn += d * d * w;
        nn += nw[i] * nw[i] * w;
    }
    const double a = sqrt(block_sum(dn, red) * dr), b = sqrt(block_sum(nn, red) * dr);
    return b <= kEps ? a : a / b;
}

__global__ __launch_bounds__(kRThreads) void k_radial_scf(const RadialJob* __restrict__ jobs, double* __restrict__ work,
                                                         int64_t stride, double* __restrict__ out) {
    extern __shared__ double lds[];  // [kROrbWaves][3][points], reductions (kRThreads), eigenvalues (kRMaxOrb)
    const RadialJob& J = jobs[blockIdx.x];  // read in place (indexed by orbital at run time)
    const int n = J.points, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const double dr = J.dr;
    double* red = lds + kROrbWaves * 3 * n;  // kRThreads doubles
    __shared__ int status;
    double* const base = work + (int64_t)blockIdx.x * stride;
    double* o = out + (int64_t)blockIdx.x * kROut;
    if (threadIdx.x == 0) status = 0;
    // starting orbitals u = r^(l+1) exp(-max(0.15, Z/n^2) r), normalised (SelfConsistentField.cpp)
    for (int k = 0; k < J.norb; ++k) {
        const double a = fmax(0.15, (double)J.Z / ((double)J.on[k] * (double)J.on[k]));
        double s = 0.0;
        for (int i = threadIdx.x; i < n; i += kRThreads) {
            const double ri = rpt(i, dr);
            const double v = pow(ri, (double)(J.ol[k] + 1)) * exp(-a * ri);
            WA(W_U + k)[i] = v;
            s += v * v;
        }
        const double nrm = sqrt(dr * block_sum(s, red));
        for (int i = threadIdx.x; i < n; i += kRThreads) WA(W_U + k)[i] /= nrm;
        __syncthreads();
    }
    spin_densities(J, base, WA(W_NA), WA(W_NB));
    double prev_e = INFINITY, prev_d = INFINITY, mixing = 0.25;  // DFTConstants.h:17
    int iters = 0, converged = 0;
    double res[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int it = 1; it <= 250; ++it) {  // DFTConstants.h:18
        iters = it;
        effective_potentials(J, base, red);
        if (w < kROrbWaves) {
            double* rows = lds + w * 3 * n;
            double* iv = WA(W_IV0 + w);
            for (int k = w; k < J.norb; k += kROrbWaves) {
                double e = 0.0;
                const int rc = solve_orbital(WA(J.ospin[k] == 0 ? W_VA : W_VB), J.ol[k], J.oidx[k], n, dr, rows, iv,
                                             WA(W_U + k), &e);
                if (rc != 0) {
                    if (lane == 0) status = rc;
                    break;
                }
                if (lane == 0) red[kRThreads - kRMaxOrb + k] = e;  // eigenvalues through LDS
            }
        }
        __syncthreads();
        if (status != 0) break;
        double* eig = lds + kROrbWaves * 3 * n + kRThreads;  // the eigenvalues, kept past the reductions
        if (threadIdx.x < J.norb) eig[threadIdx.x] = red[kRThreads - kRMaxOrb + threadIdx.x];
        __syncthreads();
        spin_densities(J, base, WA(W_OA), WA(W_OB));
        const double *oa = WA(W_OA), *ob = WA(W_OB);
        for (int i = threadIdx.x; i < n; i += kRThreads) WA(W_RHO)[i] = oa[i] + ob[i];
        __syncthreads();
        hartree(WA(W_RHO), WA(W_VH), WA(W_T1), WA(W_T2), n, dr, red);
        // energies on the output densities (TotalEnergy.cpp)
        const double kd = 1.0 / (dr * dr), ko = -0.5 / (dr * dr);
        double ekin = 0.0;
        for (int k = 0; k < J.norb; ++k) {
            const double ccl = 0.5 * (double)(J.ol[k] * (J.ol[k] + 1));
            const double* u = WA(W_U + k);
            double s = 0.0;
            for (int i = threadIdx.x; i < n; i += kRThreads) {
                const double ri = rpt(i, dr);
                double a = (kd + ccl / (ri * ri)) * u[i];
                if (i > 0) a += ko * u[i - 1];
                if (i + 1 < n) a += ko * u[i + 1];
                s += u[i] * a;
            }
            ekin += (double)J.oe[k] * block_sum(s, red) * dr;
        }
        double sx = 0.0, sh = 0.0, sxc = 0.0;
        for (int i = threadIdx.x; i < n; i += kRThreads) {
            const double ri = rpt(i, dr);
            const double rho = WA(W_RHO)[i];
            const double half = (i == 0 || i == n - 1) ? 0.5 : 1.0;
            sx += half * (-(double)J.Z * 4.0 * kPi * ri * rho);
            sh += half * (4.0 * kPi * ri * ri * rho * WA(W_VH)[i]);
            double eps = 0.0;
            if (rho > kRhoFloor) eps = xc_eval(J.functional, oa[i], ob[i], rgrad(oa, i, n, dr), rgrad(ob, i, n, dr)).eps;
            sxc += half * (4.0 * kPi * ri * ri * rho * eps);
        }
        const double eext = block_sum(sx, red) * dr;
        const double eh = 0.5 * block_sum(sh, red) * dr;
        const double exc = block_sum(sxc, red) * dr;
        const double etot = ekin + eext + eh + exc;
        // dRho = max(total, alpha, beta); |dE|; KS residual with this iteration's potentials
        for (int i = threadIdx.x; i < n; i += kRThreads) WA(W_T1)[i] = WA(W_NA)[i] + WA(W_NB)[i];
        __syncthreads();
        const double dd = density_difference(WA(W_T1), WA(W_RHO), n, dr, red);
        const double dsa = density_difference(WA(W_NA), oa, n, dr, red);
        const double dsb = density_difference(WA(W_NB), ob, n, dr, red);
        const double dmax = fmax(dd, fmax(dsa, dsb));
        const double de = isfinite(prev_e) ? fabs(etot - prev_e) : INFINITY;
        double ks = 0.0;
        for (int k = 0; k < J.norb; ++k) {
            const double ccl = 0.5 * (double)(J.ol[k] * (J.ol[k] + 1));
            const double* u = WA(W_U + k);
            const double* V = WA(J.ospin[k] == 0 ? W_VA : W_VB);
            double s = 0.0;
            for (int i = threadIdx.x; i < n; i += kRThreads) {
                const double ri = rpt(i, dr);
                double v = (kd + ccl / (ri * ri) + V[i]) * u[i];
                if (i > 0) v += (-0.5 * (1.0 / (dr * dr))) * u[i - 1];
                if (i + 1 < n) v += (-0.5 * (1.0 / (dr * dr))) * u[i + 1];
                const double r = v - eig[k] * u[i];
                s += r * r;
            }
            ks = fmax(ks, sqrt(block_sum(s, red) * dr));
        }
        res[0] = ekin, res[1] = eext, res[2] = eh, res[3] = exc, res[4] = etot;
        res[5] = dmax, res[6] = de, res[7] = ks;
        if (dmax < 1.0e-9 && de < 1.0e-11 && ks < 1.0e-7) {  // DFTConstants.h:20-22
            converged = 1;
            break;
        }
        if (isfinite(prev_d)) {
            if (dmax > prev_d * 1.05) mixing = fmax(0.10, mixing * 0.50);
            else if (dmax < prev_d) mixing = fmin(0.50, mixing * 1.10);
        }
        for (int i = threadIdx.x; i < n; i += kRThreads) {
            WA(W_NA)[i] = (1.0 - mixing) * WA(W_NA)[i] + mixing * oa[i];
            WA(W_NB)[i] = (1.0 - mixing) * WA(W_NB)[i] + mixing * ob[i];
        }
        __syncthreads();
        prev_d = dmax;
        prev_e = etot;
    }
    if (threadIdx.x == 0) {
        for (int q = 0; q < 8; ++q) o[q] = res[q];
        o[8] = iters;
        o[9] = converged;
        o[10] = status;
        o[11] = mixing;
        const double* eig = lds + kROrbWaves * 3 * n + kRThreads;
        for (int k = 0; k < kRMaxOrb; ++k) o[12 + k] = (k < J.norb && iters > 0) ? eig[k] : 0.0;
    }
}

// ElectronicConfiguration.cpp:16-145 (1s 2s 2p 3s 3p 4s 3d; Cr and Cu 4s1)
int configuration(int Z, int* sn, int* sl, int* se) {
    int k = 0;
    auto sub = [&](int n, int l, int e) { sn[k] = n, sl[k] = l, se[k] = e, ++k; };
    sub(1, 0, std::min(Z, 2));
    if (Z >= 3) sub(2, 0, std::min(Z - 2, 2));
    if (Z >= 5) sub(2, 1, std::min(Z - 4, 6));
    if (Z >= 11) sub(3, 0, std::min(Z - 10, 2));
    if (Z >= 13) sub(3, 1, std::min(Z - 12, 6));
    if (Z >= 19) sub(4, 0, (Z == 24 || Z == 29) ? 1 : (Z >= 20 ? 2 : 1));
    if (Z >= 21) sub(3, 2, Z == 24 ? 5 : (Z >= 29 ? 10 : Z - 20));
    return k;
}

}  // namespace

extern "C" {

int dft_radial_config(int Z, int* n, int* l, int* alpha, int* beta) {
    if (Z < 1 || Z > 30 || !n || !l || !alpha || !beta) return DFT_EINVAL;
    int sn[7], sl[7], se[7];
    const int k = configuration(Z, sn, sl, se);
    for (int i = 0; i < k; ++i) {
        n[i] = sn[i];
        l[i] = sl[i];
        alpha[i] = std::min(se[i], 2 * sl[i] + 1);
        beta[i] = se[i] - alpha[i];
    }
    return k;
}

int dft_radial_scf(dft_ctx* c, int njobs, const int* Z, const int* functional, int64_t points, double rmax,
                   double* results, double* densities) {
    if (!c) return DFT_EINVAL;
    DFT_REQUIRE(c, njobs > 0 && Z && functional && results, "dft_radial_scf: bad arguments");
    DFT_REQUIRE(c, points >= 2 && points <= kRMaxPoints, "dft_radial_scf: points must be in 2..2048");
    DFT_REQUIRE(c, rmax > 0.0 && std::isfinite(rmax), "dft_radial_scf: rmax must be positive");
    std::vector<RadialJob> jobs(njobs);
    for (int j = 0; j < njobs; ++j) {
        DFT_REQUIRE(c, Z[j] >= 1 && Z[j] <= 30, "dft_radial_scf: Z must be in 1..30");
        DFT_REQUIRE(c, functional[j] == DFT_XC_PZ81 || functional[j] == DFT_XC_PBE96,
                    "dft_radial_scf: unknown functional");
        RadialJob& J = jobs[j];
        J.Z = Z[j];
        J.functional = functional[j];
        J.points = (int)points;
        J.dr = rmax / (double)(points + 1);  // RadialGrid.cpp:7
        int an[7], al[7], aa[7], ab[7];
        const int ns = dft_radial_config(Z[j], an, al, aa, ab);
        J.norb = 0;
        for (int spin = 0; spin < 2; ++spin) {
            int seen[4] = {0, 0, 0, 0};
            for (int s = 0; s < ns; ++s) {
                const int e = spin == 0 ? aa[s] : ab[s];
                if (e <= 0) continue;
                const int k = J.norb++;
                J.on[k] = an[s], J.ol[k] = al[s], J.ospin[k] = spin, J.oe[k] = e, J.oidx[k] = seen[al[s]]++;
            }
        }
    }
    const int64_t stride = (int64_t)W_COUNT * points;
    RadialJob* djobs = nullptr;
    double *work = nullptr, *dout = nullptr;
    auto release = [&]() {
        if (djobs) (void)hipFree(djobs);
        if (work) (void)hipFree(work);
        if (dout) (void)hipFree(dout);
    };
    if (hipMalloc((void**)&djobs, sizeof(RadialJob) * njobs) != hipSuccess ||
        hipMalloc((void**)&work, sizeof(double) * stride * njobs) != hipSuccess ||
        hipMalloc((void**)&dout, sizeof(double) * kROut * njobs) != hipSuccess) {
        release();
        return dft::set_err(c, DFT_EHIP, "dft_radial_scf: out of device memory");
    }
    if (hipMemcpyAsync(djobs, jobs.data(), sizeof(RadialJob) * njobs, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipFuncSetAttribute(reinterpret_cast<const void*>(k_radial_scf), hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)(sizeof(double) * ((size_t)kROrbWaves * 3 * points + kRThreads + kRMaxOrb))) !=
            hipSuccess) {
        release();
        return dft::set_err(c, DFT_EHIP, "dft_radial_scf: set-up failed");
    }
    const size_t lds = sizeof(double) * ((size_t)kROrbWaves * 3 * points + kRThreads + kRMaxOrb);
    k_radial_scf<<<njobs, kRThreads, lds, c->stream>>>(djobs, work, stride, dout);
    int rc = DFT_OK;
    if (hipGetLastError() != hipSuccess) rc = dft::set_err(c, DFT_EHIP, "dft_radial_scf: launch failed");
    std::vector<double> host((size_t)kROut * njobs);
    if (rc == DFT_OK &&
        hipMemcpyAsync(host.data(), dout, sizeof(double) * host.size(), hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        rc = dft::set_err(c, DFT_EHIP, "dft_radial_scf: copy failed");
    if (rc == DFT_OK && densities)
        for (int j = 0; j < njobs && rc == DFT_OK; ++j)
            for (int s = 0; s < 2; ++s)
                if (hipMemcpyAsync(densities + ((int64_t)2 * j + s) * points, work + (int64_t)j * stride + (W_OA + s) * points,
                                   sizeof(double) * points, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
                    rc = dft::set_err(c, DFT_EHIP, "dft_radial_scf: copy failed");
    if (hipStreamSynchronize(c->stream) != hipSuccess && rc == DFT_OK)
        rc = dft::set_err(c, DFT_EHIP, "dft_radial_scf: kernel failed");
    release();
    if (rc != DFT_OK) return rc;
    int failed = 0;
    for (int j = 0; j < njobs; ++j) {
        const double* h = host.data() + (size_t)kROut * j;
        double* r = results + (size_t)kROut * j;
        for (int q = 0; q < kROut; ++q) r[q] = h[q];
        if (h[10] != 0.0) failed = 1;
    }
    if (failed) return dft::set_err(c, DFT_ENUMERIC, "dft_radial_scf: a tridiagonal solve met a singular pivot");
    return DFT_OK;
}

}  // extern "C"
