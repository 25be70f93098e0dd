/*
 * radial_oracle.c -- FP64 CPU restatement of the reference's radial (atomic)
 * Kohn-Sham stack: radial grid, aufbau configurations, the tridiagonal
 * Hamiltonian, Sturm-count bisection and inverse iteration, the radial
 * Hartree potential, XC with radial gradients, energies and the adaptive
 * linear-mixing SCF.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Nothing in the product library
 * links, loads or calls this file; tests/ use it to judge the HIP radial
 * solver.  Operation order follows the reference expression by expression.
 * Reference citations are src/dft/<file>:<line>.
 */
#include "ks_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static const double PI = 3.141592653589793238462643383279502884; /* DFTConstants.h:7 */
static const double EPS = 1.0e-14;                                  /* DFTConstants.h:8 */
static const double RHO_FLOOR = 1.0e-20;                            /* DFTConstants.h:9 */

/* ---------------------------------------------------------------- grid */
/* RadialGrid.cpp:5-20: r_i = (i + 1) rmax / (points + 1) */
static double radial_dr(long points, double rmax) { return rmax / (double)(points + 1); }

/* ---------------------------------------------------- configurations */
/* ElectronicConfiguration.cpp:16-145: subshells filled in the reference's
 * order (1s 2s 2p 3s 3p 4s 3d; Cr and Cu 4s1), each split into alpha =
 * min(e, 2l + 1) and beta = the rest. */
int og_radial_config(int Z, int* n, int* l, int* alpha, int* beta) {
    if (Z < 1 || Z > 30) return -1;
    int k = 0;
    int sub_n[7], sub_l[7], sub_e[7];
#define SUB(N, L, E) (sub_n[k] = (N), sub_l[k] = (L), sub_e[k] = (E), ++k)
    SUB(1, 0, Z < 2 ? Z : 2);
    if (Z >= 3) SUB(2, 0, Z - 2 < 2 ? Z - 2 : 2);
    if (Z >= 5) SUB(2, 1, Z - 4 < 6 ? Z - 4 : 6);
    if (Z >= 11) SUB(3, 0, Z - 10 < 2 ? Z - 10 : 2);
    if (Z >= 13) SUB(3, 1, Z - 12 < 6 ? Z - 12 : 6);
    if (Z >= 19) {
        int s = Z >= 20 ? 2 : 1;
        if (Z == 24 || Z == 29) s = 1;
        SUB(4, 0, s);
    }
    if (Z >= 21) {
        int d = Z - 20;
        if (Z == 24) d = 5;
        else if (Z >= 29) d = 10;
        SUB(3, 2, d);
    }
#undef SUB
    for (int i = 0; i < k; ++i) {
        const int cap = 2 * sub_l[i] + 1;
        n[i] = sub_n[i];
        l[i] = sub_l[i];
        alpha[i] = sub_e[i] < cap ? sub_e[i] : cap;
        beta[i] = sub_e[i] - alpha[i];
    }
    return k;
}

/* ------------------------------------------------------- tridiagonal */
/* KohnShamHamiltonian.cpp:7-99: diag = 1/dr^2 + l(l+1)/(2 r^2) + V, off = -1/(2 dr^2) */
static void build_h(long n, double dr, const double* v, int l, double* diag, double* off) {
    const double kd = 1.0 / (dr * dr);
    const double ko = -0.5 * (1.0 / (dr * dr));
    const double cc = 0.5 * (double)(l * (l + 1));
    for (long i = 0; i < n; ++i) {
        const double ri = (double)(i + 1) * dr;
        diag[i] = kd + cc / (ri * ri) + v[i];
    }
    for (long i = 0; i + 1 < n; ++i) off[i] = ko;
}

/* EigenvalueSolver.cpp:37-90: negative pivots of the LDL^T of (H - e), tiny pivots replaced by +-EPS */
static long sturm_count(long n, const double* diag, const double* off, double e) {
    long count = 0;
    double prev = diag[0] - e;
    if (prev < 0.0) ++count;
    if (fabs(prev) < EPS) prev = copysign(EPS, prev == 0.0 ? 1.0 : prev);
    for (long i = 1; i < n; ++i) {
        double p = prev;
        if (fabs(p) < EPS) p = copysign(EPS, p == 0.0 ? 1.0 : p);
        const double pivot = diag[i] - e - off[i - 1] * off[i - 1] / p;
        if (pivot < 0.0) ++count;
        prev = fabs(pivot) < EPS ? copysign(EPS, pivot == 0.0 ? 1.0 : pivot) : pivot;
    }
    return count;
}

/* EigenvalueSolver.cpp:92-215: bracket expansion (100 doublings), then 300 bisections to 1e-13 relative */
static int find_eigenvalue(long n, const double* diag, const double* off, long index, double lo, double hi,
                           double* out) {
    long lc = sturm_count(n, diag, off, lo), hc = sturm_count(n, diag, off, hi);
    double width = hi - lo;
    for (int x = 0; x < 100; ++x) {
        if (lc <= index && index < hc) break;
        width *= 2.0;
        if (lc > index) {
            lo -= width;
            lc = sturm_count(n, diag, off, lo);
        }
        if (hc <= index) {
            hi += width;
            hc = sturm_count(n, diag, off, hi);
        }
    }
    if (lc > index || index >= hc) return -2;
    for (int it = 0; it < 300; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (sturm_count(n, diag, off, mid) <= index) lo = mid;
        else hi = mid;
        const double a = fabs(lo), b = fabs(hi);
        const double tol = 1.0e-13 * fmax(1.0, fmax(a, b));
        if (hi - lo <= tol) break;
    }
    *out = 0.5 * (lo + hi);
    return 0;
}

/* NumericalMethods.cpp:88-138: Thomas algorithm (symmetric off-diagonals) */
static int solve_tridiagonal(long n, const double* diag_in, const double* off, const double* rhs, double* x,
                             double* work) {
    double* d = work;
    memcpy(d, diag_in, sizeof(double) * n);
    memcpy(x, rhs, sizeof(double) * n);
    for (long i = 1; i < n; ++i) {
        if (fabs(d[i - 1]) < EPS) return -2;
        const double f = off[i - 1] / d[i - 1];
        d[i] -= f * off[i - 1];
        x[i] -= f * x[i - 1];
    }
    if (fabs(d[n - 1]) < EPS) return -2;
    x[n - 1] /= d[n - 1];
    for (long i = n - 1; i-- > 0;) {
        if (fabs(d[i]) < EPS) return -2;
        x[i] = (x[i] - off[i] * x[i + 1]) / d[i];
    }
    return 0;
}

static double dot(long n, const double* a, const double* b) {
    double s = 0.0;
    for (long i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* EigenvalueSolver.cpp:217-312: inverse iteration from sin(i + 1) at e + 1e-10 max(1, |e|) */
static int solve_eigenvector(long n, const double* diag, const double* off, double e, double* v) {
    double* sh = malloc(sizeof(double) * n);
    double* nx = malloc(sizeof(double) * n);
    double* wk = malloc(sizeof(double) * n);
    int rc = 0;
    for (long i = 0; i < n; ++i) v[i] = sin((double)(i + 1));
    double nrm = sqrt(dot(n, v, v));
    for (long i = 0; i < n; ++i) v[i] /= nrm;
    const double shift = e + 1.0e-10 * fmax(1.0, fabs(e));
    for (long i = 0; i < n; ++i) sh[i] = diag[i] - shift;
    for (int it = 0; it < 1000; ++it) {
        if ((rc = solve_tridiagonal(n, sh, off, v, nx, wk)) != 0) break;
        const double nn = sqrt(dot(n, nx, nx));
        if (nn < EPS) {
            rc = -2;
            break;
        }
        for (long i = 0; i < n; ++i) nx[i] /= nn;
        if (dot(n, v, nx) < 0.0)
            for (long i = 0; i < n; ++i) nx[i] = -nx[i];
        double diff = 0.0;
        for (long i = 0; i < n; ++i) diff = fmax(diff, fabs(v[i] - nx[i]));
        memcpy(v, nx, sizeof(double) * n);
        if (diff < 1.0e-12) break;
    }
    free(sh);
    free(nx);
    free(wk);
    return rc;
}

/* NumericalMethods.cpp:40-60: u / sqrt(dr u.u) */
static void normalize(long n, double* u, double dr) {
    const double nrm = sqrt(dr * dot(n, u, u));
    for (long i = 0; i < n; ++i) u[i] /= nrm;
}

/* EigenvalueSolver.cpp:314-403 solveOrbital: eigenvalue number `index` of H(l), bracket
 * [min diag - 10, max diag + 10], its inverse-iteration vector normalised with dr */
int og_radial_orbital(long points, double rmax, const double* v, int l, long index, double* eig, double* u) {
    const double dr = radial_dr(points, rmax);
    double* diag = malloc(sizeof(double) * points);
    double* off = malloc(sizeof(double) * points);
    build_h(points, dr, v, l, diag, off);
    double mn = diag[0], mx = diag[0];
    for (long i = 1; i < points; ++i) {
        mn = fmin(mn, diag[i]);
        mx = fmax(mx, diag[i]);
    }
    int rc = find_eigenvalue(points, diag, off, index, mn - 10.0, mx + 10.0, eig);
    if (rc == 0) rc = solve_eigenvector(points, diag, off, *eig, u);
    if (rc == 0) normalize(points, u, dr);
    free(diag);
    free(off);
    return rc;
}

/* ------------------------------------------------------------ Hartree */
/* HartreePotential.cpp:8-86: trapezoid enclosed charge outward, 4 pi r rho inward; V = Q(r)/r + outer */
int og_radial_hartree(long points, double rmax, const double* rho, double* vh) {
    const double dr = radial_dr(points, rmax);
    double* q = calloc(points, sizeof(double));
    double* o = calloc(points, sizeof(double));
    for (long i = 1; i < points; ++i) {
        const double r0 = (double)i * dr, r1 = (double)(i + 1) * dr;
        const double f0 = 4.0 * PI * r0 * r0 * rho[i - 1];
        const double f1 = 4.0 * PI * r1 * r1 * rho[i];
        q[i] = q[i - 1] + 0.5 * (f0 + f1) * dr;
    }
    for (long i = points - 1; i-- > 0;) {
        const double r0 = (double)(i + 1) * dr, r1 = (double)(i + 2) * dr;
        const double f0 = 4.0 * PI * r0 * rho[i];
        const double f1 = 4.0 * PI * r1 * rho[i + 1];
        o[i] = o[i + 1] + 0.5 * (f0 + f1) * dr;
    }
    for (long i = 0; i < points; ++i) vh[i] = q[i] / ((double)(i + 1) * dr) + o[i];
    free(q);
    free(o);
    return 0;
}

/* HartreePotential.cpp:88-136: 1/2 trapezoid of 4 pi r^2 rho V_H */
static double hartree_energy(long n, double dr, const double* rho, const double* vh) {
    double e = 0.0;
    for (long i = 0; i < n; ++i) {
        const double ri = (double)(i + 1) * dr;
        const double f = 4.0 * PI * ri * ri * rho[i] * vh[i];
        e += (i == 0 || i == n - 1) ? 0.5 * f : f;
    }
    return 0.5 * e * dr;
}

/* ------------------------------------------------------------------ XC */
/* ExchangeCorrelation.cpp:11-59: central differences inside, one-sided at the ends */
static double radial_grad(long n, double dr, const double* f, long i) {
#define R(j) ((double)((j) + 1) * dr)
    if (i == 0) return (f[1] - f[0]) / (R(1) - R(0));
    if (i == n - 1) return (f[n - 1] - f[n - 2]) / (R(n - 1) - R(n - 2));
    return (f[i + 1] - f[i - 1]) / (R(i + 1) - R(i - 1));
#undef R
}

/* ExchangeCorrelation.cpp:61-163: (1/r^2) d(r^2 c)/dr with the same differences */
static double radial_div(long n, double dr, const double* c, long i) {
#define R(j) ((double)((j) + 1) * dr)
    if (i == 0) return ((R(1) * R(1) * c[1] - R(0) * R(0) * c[0]) / (R(1) - R(0))) / (R(0) * R(0));
    if (i == n - 1)
        return ((R(n - 1) * R(n - 1) * c[n - 1] - R(n - 2) * R(n - 2) * c[n - 2]) / (R(n - 1) - R(n - 2))) /
               (R(n - 1) * R(n - 1));
    return ((R(i + 1) * R(i + 1) * c[i + 1] - R(i - 1) * R(i - 1) * c[i - 1]) / (R(i + 1) - R(i - 1))) / (R(i) * R(i));
#undef R
}

/* ExchangeCorrelation.cpp:167-279: v_sigma = df/dn_sigma - div(c_sigma) (the radial flux is c itself) */
int og_radial_xc_potential(long points, double rmax, int functional, const double* na, const double* nb, int spin,
                           double* v) {
    if (spin != 0 && spin != 1) return -1;
    const double dr = radial_dr(points, rmax);
    double* dd = calloc(points, sizeof(double));
    double* cf = calloc(points, sizeof(double));
    for (long i = 0; i < points; ++i) {
        double in[4] = {na[i], nb[i], radial_grad(points, dr, na, i), radial_grad(points, dr, nb, i)}, out[5];
        if (og_xc_point(functional, in, out) != 0) {
            free(dd);
            free(cf);
            return -1;
        }
        dd[i] = spin == 0 ? out[1] : out[2];
        cf[i] = spin == 0 ? out[3] : out[4];
    }
    for (long i = 0; i < points; ++i) {
        if (na[i] + nb[i] <= RHO_FLOOR) {
            v[i] = 0.0;
            continue;
        }
        v[i] = dd[i] - radial_div(points, dr, cf, i);
    }
    free(dd);
    free(cf);
    return 0;
}

/* ExchangeCorrelation.cpp:281-370: trapezoid of 4 pi r^2 n eps_xc */
double og_radial_xc_energy(long points, double rmax, int functional, const double* na, const double* nb) {
    const double dr = radial_dr(points, rmax);
    double e = 0.0;
    for (long i = 0; i < points; ++i) {
        const double n = na[i] + nb[i];
        double eps = 0.0;
        if (n > RHO_FLOOR) {
            double in[4] = {na[i], nb[i], radial_grad(points, dr, na, i), radial_grad(points, dr, nb, i)}, out[5];
            og_xc_point(functional, in, out);
            eps = out[0];
        }
        const double ri = (double)(i + 1) * dr;
        const double f = 4.0 * PI * ri * ri * n * eps;
        e += (i == 0 || i == points - 1) ? 0.5 * f : f;
    }
    return e * dr;
}

/* ------------------------------------------------------------ the SCF */
typedef struct {
    int n, l, spin, electrons;
    double eig;
    double* u;
} rorb;

/* ElectronDensity.cpp:8-55: sum of electrons u^2 / (4 pi r^2) over the spin's occupied orbitals */
static void spin_density(long n, double dr, const rorb* o, int norb, int spin, double* out) {
    for (long i = 0; i < n; ++i) out[i] = 0.0;
    for (int k = 0; k < norb; ++k) {
        if (o[k].spin != spin || o[k].electrons <= 0) continue;
        for (long i = 0; i < n; ++i) {
            const double ri = (double)(i + 1) * dr;
            out[i] += (double)o[k].electrons * o[k].u[i] * o[k].u[i] / (4.0 * PI * ri * ri);
        }
    }
}

/* SelfConsistentField.cpp: relative 4 pi r^2-weighted L2 difference */
static double density_difference(long n, double dr, const double* old, const double* nw) {
    double dn = 0.0, nn = 0.0;
    for (long i = 0; i < n; ++i) {
        const double ri = (double)(i + 1) * dr;
        const double w = 4.0 * PI * ri * ri;
        const double d = nw[i] - old[i];
        dn += d * d * w;
        nn += nw[i] * nw[i] * w;
    }
    const double a = sqrt(dn * dr), b = sqrt(nn * dr);
    return b <= EPS ? a : a / b;
}

/* TotalEnergy.cpp:9-62: sum_k electrons dr u.(T + l(l+1)/2r^2) u with the three-point kinetic stencil */
static double kinetic_energy(long n, double dr, const rorb* o, int norb) {
    const double kd = 1.0 / (dr * dr), ko = -0.5 / (dr * dr);
    double e = 0.0;
    for (int k = 0; k < norb; ++k) {
        if (o[k].electrons <= 0) continue;
        const double cc = 0.5 * (double)(o[k].l * (o[k].l + 1));
        double s = 0.0;
        for (long i = 0; i < n; ++i) {
            const double ri = (double)(i + 1) * dr;
            double a = (kd + cc / (ri * ri)) * o[k].u[i];
            if (i > 0) a += ko * o[k].u[i - 1];
            if (i + 1 < n) a += ko * o[k].u[i + 1];
            s += o[k].u[i] * a;
        }
        e += (double)o[k].electrons * s * dr;
    }
    return e;
}

/* TotalEnergy.cpp:64-110: trapezoid of -Z 4 pi r rho */
static double external_energy(long n, double dr, const double* rho, int Z) {
    double e = 0.0;
    for (long i = 0; i < n; ++i) {
        const double ri = (double)(i + 1) * dr;
        const double f = -(double)Z * 4.0 * PI * ri * rho[i];
        e += (i == 0 || i == n - 1) ? 0.5 * f : f;
    }
    return e * dr;
}

/* SelfConsistentField.cpp: max over orbitals of sqrt(dr sum (H u - e u)^2) */
static double ks_residual(long n, double dr, const double* va, const double* vb, const rorb* o, int norb) {
    double* diag = malloc(sizeof(double) * n);
    double* off = malloc(sizeof(double) * n);
    double mx = 0.0;
    for (int k = 0; k < norb; ++k) {
        build_h(n, dr, o[k].spin == 0 ? va : vb, o[k].l, diag, off);
        double s = 0.0;
        for (long i = 0; i < n; ++i) {
            double v = diag[i] * o[k].u[i];
            if (i > 0) v += off[i - 1] * o[k].u[i - 1];
            if (i + 1 < n) v += off[i] * o[k].u[i + 1];
            const double r = v - o[k].eig * o[k].u[i];
            s += r * r;
        }
        mx = fmax(mx, sqrt(s * dr));
    }
    free(diag);
    free(off);
    return mx;
}

/* SelfConsistentField.cpp: V_sigma = V_xc,sigma - Z/r + V_H[n_alpha + n_beta] */
static int effective_potentials(long n, double rmax, int Z, int functional, const double* na, const double* nb,
                                double* va, double* vb) {
    const double dr = radial_dr(n, rmax);
    double* rho = malloc(sizeof(double) * n);
    double* vh = malloc(sizeof(double) * n);
    for (long i = 0; i < n; ++i) rho[i] = na[i] + nb[i];
    og_radial_hartree(n, rmax, rho, vh);
    int rc = og_radial_xc_potential(n, rmax, functional, na, nb, 0, va);
    if (rc == 0) rc = og_radial_xc_potential(n, rmax, functional, na, nb, 1, vb);
    for (long i = 0; i < n; ++i) {
        const double vn = -(double)Z / ((double)(i + 1) * dr);
        va[i] += vn + vh[i];
        vb[i] += vn + vh[i];
    }
    free(rho);
    free(vh);
    return rc;
}

/* AtomicDFT.cpp:6-41 + SelfConsistentField.cpp solveSelfConsistentField: hydrogenic-like start
 * (u = r^(l+1) exp(-max(0.15, Z/n^2) r)), then per iteration the effective potentials, every
 * occupied orbital (alpha states, then beta; the k-th state of a given l is eigenvalue k of H(l)),
 * output densities, energies on the output densities, dRho = max(total, alpha, beta), |dE|, the
 * KS residual; stop when dRho < 1e-9, |dE| < 1e-11 and KS < 1e-7; otherwise the mixing factor
 * (starting 0.25) halves (floor 0.1) if dRho grew by more than 5 %, grows by 10 % (cap 0.5) if it
 * fell, and both spin densities mix linearly with it.  res = {kinetic, external, hartree, xc,
 * total, dRho, |dE|, KS residual}; eig: alpha orbitals then beta; returns the iteration count,
 * negated when not converged within 250, or < -1000 on an error. */
int og_radial_scf(int Z, int functional, long points, double rmax, double* res, double* eig, int* norb_out,
                  double* na_out, double* nb_out) {
    int cn[7], cl[7], ca[7], cb[7];
    const int ns = og_radial_config(Z, cn, cl, ca, cb);
    if (ns < 0 || points < 2 || rmax <= 0.0) return -1001;
    const long n = points;
    const double dr = radial_dr(points, rmax);
    rorb orb[14];
    int norb = 0;
    for (int spin = 0; spin < 2; ++spin)
        for (int s = 0; s < ns; ++s) {
            const int e = spin == 0 ? ca[s] : cb[s];
            if (e <= 0) continue;
            orb[norb].n = cn[s];
            orb[norb].l = cl[s];
            orb[norb].spin = spin;
            orb[norb].electrons = e;
            orb[norb].eig = 0.0;
            orb[norb].u = malloc(sizeof(double) * n);
            ++norb;
        }
    double *na = malloc(sizeof(double) * n), *nb = malloc(sizeof(double) * n);
    double *va = malloc(sizeof(double) * n), *vb = malloc(sizeof(double) * n);
    double *oa = malloc(sizeof(double) * n), *ob = malloc(sizeof(double) * n);
    double *rho_old = malloc(sizeof(double) * n), *rho_out = malloc(sizeof(double) * n);
    double* vh = malloc(sizeof(double) * n);
    /* the starting orbitals */
    for (int k = 0; k < norb; ++k) {
        const double a = fmax(0.15, (double)Z / ((double)orb[k].n * (double)orb[k].n));
        for (long i = 0; i < n; ++i) {
            const double ri = (double)(i + 1) * dr;
            orb[k].u[i] = pow(ri, orb[k].l + 1) * exp(-a * ri);
        }
        normalize(n, orb[k].u, dr);
    }
    spin_density(n, dr, orb, norb, 0, na);
    spin_density(n, dr, orb, norb, 1, nb);
    double prev_e = INFINITY, prev_d = INFINITY, mixing = 0.25;
    int rc = 0, iters = 0, converged = 0;
    for (int it = 1; it <= 250; ++it) {
        iters = it;
        if ((rc = effective_potentials(n, rmax, Z, functional, na, nb, va, vb)) != 0) break;
        /* orbitals: per spin, the k-th state of each l is that l's k-th eigenvalue */
        for (int spin = 0; spin < 2 && rc == 0; ++spin) {
            int seen[4] = {0, 0, 0, 0};
            for (int k = 0; k < norb; ++k) {
                if (orb[k].spin != spin) continue;
                rc = og_radial_orbital(n, rmax, spin == 0 ? va : vb, orb[k].l, seen[orb[k].l]++, &orb[k].eig, orb[k].u);
                if (rc != 0) break;
            }
        }
        if (rc != 0) break;
        spin_density(n, dr, orb, norb, 0, oa);
        spin_density(n, dr, orb, norb, 1, ob);
        for (long i = 0; i < n; ++i) {
            rho_old[i] = na[i] + nb[i];
            rho_out[i] = oa[i] + ob[i];
        }
        og_radial_hartree(n, rmax, rho_out, vh);
        const double ek = kinetic_energy(n, dr, orb, norb);
        const double ee = external_energy(n, dr, rho_out, Z);
        const double eh = hartree_energy(n, dr, rho_out, vh);
        const double ex = og_radial_xc_energy(n, rmax, functional, oa, ob);
        const double et = ek + ee + eh + ex;
        const double dd = density_difference(n, dr, rho_old, rho_out);
        const double dsa = density_difference(n, dr, na, oa), dsb = density_difference(n, dr, nb, ob);
        const double dmax = fmax(dd, fmax(dsa, dsb));
        const double de = isfinite(prev_e) ? fabs(et - prev_e) : INFINITY;
        const double ks = ks_residual(n, dr, va, vb, orb, norb);
        res[0] = ek;
        res[1] = ee;
        res[2] = eh;
        res[3] = ex;
        res[4] = et;
        res[5] = dmax;
        res[6] = de;
        res[7] = ks;
        if (na_out) memcpy(na_out, oa, sizeof(double) * n);
        if (nb_out) memcpy(nb_out, ob, sizeof(double) * n);
        if (dmax < 1.0e-9 && de < 1.0e-11 && ks < 1.0e-7) {
            converged = 1;
            break;
        }
        if (isfinite(prev_d)) {
            if (dmax > prev_d * 1.05) mixing = fmax(0.10, mixing * 0.50);
            else if (dmax < prev_d) mixing = fmin(0.50, mixing * 1.10);
        }
        for (long i = 0; i < n; ++i) {
            na[i] = (1.0 - mixing) * na[i] + mixing * oa[i];
            nb[i] = (1.0 - mixing) * nb[i] + mixing * ob[i];
        }
        prev_d = dmax;
        prev_e = et;
    }
    for (int k = 0; k < norb; ++k) {
        if (eig) eig[k] = orb[k].eig;
        free(orb[k].u);
    }
    if (norb_out) *norb_out = norb;
    free(na);
    free(nb);
    free(va);
    free(vb);
    free(oa);
    free(ob);
    free(rho_old);
    free(rho_out);
    free(vh);
    if (rc != 0) return -1000 + rc;
    return converged ? iters : -iters;
}
