// xc_functionals.hpp -- PZ81 and PBE96 as device functions templated on the
// arithmetic type.  With T = double the expressions keep the reference's
// operation order (PZ81.cpp:41-501, PBE96.cpp:99-966) so the FP64 path
// agrees with the CPU reference to a few ulps of the math library.  With
// T = float, log(1 + 1/q1) in the PW92 interpolation is evaluated as
// log1p(1/q1), and PBE's H and its derivatives are regrouped (pbe_c), which
// keeps FP32 finite down to the density floor; callers fall back to
// T = double for extreme relative gradients (|grad n| / n > 1e4) and on a
// non-finite FP32 result.
#pragma once

#include <hip/hip_runtime.h>

namespace dft {
namespace xc {

constexpr double kPi = 3.141592653589793238462643383279502884;  // DFTConstants.h:7
constexpr double kRhoFloor = 1.0e-20;                             // DFTConstants.h:9
constexpr double kExchangeCoefficient = 0.7385587663820223;       // DFTConstants.h:25

template <typename T>
struct Out {
    T eps, va, vb, ca, cb;
};

template <typename T>
__host__ __device__ inline T tmax(T a, T b) { return (a < b) ? b : a; }
template <typename T>
__host__ __device__ inline T tclamp(T v, T lo, T hi) { return (v < lo) ? lo : ((hi < v) ? hi : v); }

// Transcendental functions.  FP64 keeps the math library's (the reference's
// results to a few ulps).  FP32 uses the hardware's base-2 logarithm and
// exponential (v_log_f32, v_exp_f32, about 1 ulp each): log x = log2(x) ln 2,
// exp x = 2^(x log2 e), cbrt x = 2^(log2(x) / 3), x^y = 2^(y log2 x) for
// x >= 0 (every power here has a non-negative base); a few ulps of FP32,
// far inside the FP32 kernels' tolerances, at a fraction of the cost of the
// correctly rounded library versions.  log1p and expm1, which guard
// cancellations at low density, stay the library's.
template <typename T>
__device__ inline T xlog(T x) {
    if constexpr (sizeof(T) == 4) return __builtin_amdgcn_logf(x) * 0.693147180559945309f;
    else return log(x);
}
template <typename T>
__device__ inline T xexp(T x) {
    if constexpr (sizeof(T) == 4) return __builtin_amdgcn_exp2f(x * 1.44269504088896341f);
    else return exp(x);
}
template <typename T>
__device__ inline T xcbrt(T x) {
    if constexpr (sizeof(T) == 4) return __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(x) * (1.0f / 3.0f));
    else return cbrt(x);
}
template <typename T>
__device__ inline T xpow(T x, T y) {
    if constexpr (sizeof(T) == 4) return __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(x) * y);
    else return pow(x, y);
}

// ---------------------------------------------------------------- PZ81
template <typename T>
struct PZPar { T A, B, C, D, gamma, beta1, beta2; };

template <typename T>
__host__ __device__ inline PZPar<T> pz_unpol() { return {T(0.0311), T(-0.0480), T(0.0020), T(-0.0116), T(-0.1423), T(1.0529), T(0.3334)}; }  // PZ81.cpp:21-29
template <typename T>
__host__ __device__ inline PZPar<T> pz_pol() { return {T(0.01555), T(-0.0269), T(0.0007), T(-0.0048), T(-0.0843), T(1.3981), T(0.2611)}; }  // PZ81.cpp:31-39

template <typename T>
__device__ inline T pz_ec(T rs, const PZPar<T>& p) {  // PZ81.cpp:52-81
    if (rs < T(1)) return p.A * xlog(rs) + p.B + p.C * rs * xlog(rs) + p.D * rs;
    const T sr = sqrt(rs);
    const T den = T(1) + p.beta1 * sr + p.beta2 * rs;
    return p.gamma / den;
}
template <typename T>
__device__ inline T pz_dec(T rs, const PZPar<T>& p) {  // PZ81.cpp:83-120
    if (rs < T(1)) return p.A / rs + p.C * (xlog(rs) + T(1)) + p.D;
    const T sr = sqrt(rs);
    const T den = T(1) + p.beta1 * sr + p.beta2 * rs;
    const T dden = p.beta1 / (T(2) * sr) + p.beta2;
    return -p.gamma * dden / (den * den);
}

template <typename T>
__device__ inline Out<T> pz81(T a_in, T b_in) {  // PZ81.cpp:437-501 with 293-433
    Out<T> o{T(0), T(0), T(0), T(0), T(0)};
    const T a = tmax(a_in, T(0)), b = tmax(b_in, T(0));
    const T n = a + b;
    if (n <= T(kRhoFloor)) return o;
    const T zeta = tclamp((a - b) / n, T(-1), T(1));
    const T rs = xcbrt(T(3) / (T(4.0 * kPi) * n));
    const T f43 = T(4.0 / 3.0), f13 = T(1.0 / 3.0);
    const T p43 = xpow(T(1) + zeta, f43), m43 = xpow(T(1) - zeta, f43);
    const T p13 = xpow(T(1) + zeta, f13), m13 = xpow(T(1) - zeta, f13);
    const T fden = xpow(T(2), f43) - T(2);
    const T f = (p43 + m43 - T(2)) / fden;
    const T df = f43 * (p13 - m13) / fden;
    const T sx = (p43 + m43) / T(2);
    const T dsx = T(2.0 / 3.0) * (p13 - m13);
    const T ex = -T(kExchangeCoefficient) * xcbrt(n);
    const T exs = ex * sx;
    const PZPar<T> U = pz_unpol<T>(), P = pz_pol<T>();
    const T e0 = pz_ec(rs, U), e1 = pz_ec(rs, P);
    const T d0 = pz_dec(rs, U), d1 = pz_dec(rs, P);
    const T ec = e0 + f * (e1 - e0);
    o.eps = exs + ec;
    const T dexrs = -exs / rs;
    const T dcrs = d0 + f * (d1 - d0);
    const T exc = exs + ec;
    const T dxcrs = dexrs + dcrs;
    const T dxcz = ex * dsx + df * (e1 - e0);
    const T radial = exc - rs * dxcrs / T(3);
    o.va = radial + (T(1) - zeta) * dxcz;
    o.vb = radial - (T(1) + zeta) * dxcz;
    return o;
}

// ---------------------------------------------------------------- PBE96
// PBE96.cpp:10-54
constexpr double kKappa = 0.804, kMu = 0.2195149727645171;
constexpr double kGamma = 0.031090690869654895, kBeta = 0.06672455060314922;
constexpr double kDelta = kBeta / kGamma;
constexpr double kSpinGamma = 0.51984209978974632953;
constexpr double kSpinFzz = 8.0 / (9.0 * kSpinGamma);
constexpr double kEta = 1.0e-12;
constexpr double kExchangeA = -0.73855876638202240588;
constexpr double kKf = 3.09366772628014;

template <typename T>
__device__ inline T sign_of(T v) { return v > T(0) ? T(1) : (v < T(0) ? T(-1) : T(0)); }

template <typename T>
__device__ inline T log_1_plus_inv(T q1) {
    if constexpr (sizeof(T) == 4) return log1p(T(1) / q1);
    else return xlog(T(1) + T(1) / q1);
}

template <typename T>
__device__ inline T gcor2(T a, T a1, T b1, T b2, T b3, T b4, T sr, T& drs) {  // PBE96.cpp:99-179
    const T sr2 = sr * sr;
    const T q0 = T(-2) * a * (T(1) + a1 * sr2);
    const T q1 = T(2) * a * sr * (b1 + sr * (b2 + sr * (b3 + b4 * sr)));
    const T q2 = log_1_plus_inv(q1);
    const T e = q0 * q2;
    const T q3 = a * (b1 / sr + T(2) * b2 + sr * (T(3) * b3 + T(4) * b4 * sr));
    drs = T(-2) * a * a1 * q2 - q0 * q3 / (q1 * (T(1) + q1));
    return e;
}

template <typename T>
__device__ inline void pw92(T rs, T zeta, T& e, T& ers, T& ez) {  // PBE96.cpp:181-345
    const T sr = sqrt(rs);
    T eud, epd, ald;
    const T eu = gcor2(T(0.0310907), T(0.21370), T(7.5957), T(3.5876), T(1.6382), T(0.49294), sr, eud);
    const T ep = gcor2(T(0.01554535), T(0.20548), T(14.1189), T(6.1977), T(3.3662), T(0.62517), sr, epd);
    const T al = gcor2(T(0.0168869), T(0.11125), T(10.357), T(3.6231), T(0.88026), T(0.49671), sr, ald);
    const T z4 = zeta * zeta * zeta * zeta;
    const T f43 = T(4.0 / 3.0);
    const T f = (xpow(T(1) + zeta, f43) + xpow(T(1) - zeta, f43) - T(2)) / T(kSpinGamma);
    const T fzz = T(kSpinFzz);
    e = eu * (T(1) - f * z4) + ep * f * z4 - al * f * (T(1) - z4) / fzz;
    ers = eud * (T(1) - f * z4) + epd * f * z4 - ald * f * (T(1) - z4) / fzz;
    const T df = f43 * (xpow(T(1) + zeta, T(1.0 / 3.0)) - xpow(T(1) - zeta, T(1.0 / 3.0))) / T(kSpinGamma);
    ez = T(4) * zeta * zeta * zeta * f * (ep - eu + al / fzz) + df * (z4 * ep - z4 * eu - (T(1) - z4) * al / fzz);
}

template <typename T>
__device__ inline void pbe_x(T rho, T gm, T& ed, T& dd, T& gd) {  // PBE96.cpp:347-444
    ed = dd = gd = T(0);
    if (rho <= T(0)) return;
    const T rt = xcbrt(rho);
    const T exu = T(kExchangeA) * rt;
    const T exud = rho * exu;
    const T kf = T(kKf) * rt;
    const T den = T(2) * kf * rho;
    T s = T(0);
    if (den > T(0)) s = gm / den;
    const T s2 = s * s;
    const T mok = T(kMu / kKappa);
    const T p0 = T(1) + mok * s2;
    const T enh = T(1.0 + kKappa) - T(kKappa) / p0;
    const T fs = T(2.0 * kKappa) * mok / (p0 * p0);
    dd = exu * T(4.0 / 3.0) * (enh - s2 * fs);
    if (gm > T(0)) gd = T(kExchangeA) * T(0.161620459673995) * s * fs;
    ed = exud * enh;
}

template <typename T>
__device__ inline void pbe_c(T na, T nb, T ga, T gb, T out[4]) {  // PBE96.cpp:446-810
    out[0] = out[1] = out[2] = out[3] = T(0);
    const T n = na + nb;
    if (n <= T(0)) return;
    const T zeta = tclamp((na - nb) / n, T(-1), T(1));
    const T gm = fabs(ga + gb);
    const T kf = xcbrt(T(3.0 * kPi * kPi) * n);
    const T rs = xcbrt(T(3) / (T(4.0 * kPi) * n));
    const T phi = T(0.5) * (xpow(T(1) + zeta, T(2.0 / 3.0)) + xpow(T(1) - zeta, T(2.0 / 3.0)));
    const T ks = sqrt(T(4) * kf / T(kPi));
    const T tkp = T(2) * ks * phi;
    T t = T(0);
    if (tkp > T(0)) t = gm / (tkp * n);
    T ec, ecrs, ecz;
    pw92(rs, zeta, ec, ecrs, ecz);
    const T phi3 = phi * phi * phi;
    const T phi4 = phi3 * phi;
    const T ex = -ec / (T(kGamma) * phi3);
    T b;
    if (ex < T(700)) b = T(kDelta) / expm1(ex);
    else b = T(kDelta) * xexp(-ex);
    const T b2 = b * b, t2 = t * t;
    T h, hb, ht;
    if constexpr (sizeof(T) == 4) {
        // The same H, dH/db and dH/dt regrouped so that nothing grows beyond q5 ~ (b t^2)^2: in a
        // density's tail b t^2 reaches ~1e13 and the reference's q8 = q5^2 + ... and b t^6 (2 + b t^2)
        // would overflow FP32 (FP64 keeps the reference's form below).  With y = b t^2:
        //   q4 / q5 = 1 / (1 + y (y / (1 + y))),
        //   hb = -beta phi^3 [y (2 + y) / q5] [t^4 / (q5 + delta q4 t^2)],
        //   ht = 2 beta phi^3 [q9 / q5] / (q5 + delta q4 t^2).
        const T y = b * t2;
        const T q4 = T(1) + y;
        const T q5 = T(1) + y + y * y;
        h = phi3 * T(kBeta / kDelta) * xlog(T(1) + T(kDelta) * t2 / (T(1) + y * (y / q4)));
        const T q8r = q5 + T(kDelta) * q4 * t2;  // q8 / q5
        hb = -T(kBeta) * phi3 * (y * (T(2) + y) / q5) * (t2 * t2 / q8r);
        ht = T(2.0 * kBeta) * phi3 * ((T(1) + T(2) * y) / q5) / q8r;
    } else {
        const T t4 = t2 * t2, t6 = t4 * t2;
        const T q4 = T(1) + b * t2;
        const T q5 = T(1) + b * t2 + b2 * t4;
        h = phi3 * T(kBeta / kDelta) * xlog(T(1) + T(kDelta) * q4 * t2 / q5);
        const T q8 = q5 * q5 + T(kDelta) * q4 * q5 * t2;
        const T q9 = T(1) + T(2) * b * t2;
        hb = -T(kBeta) * phi3 * b * t6 * (T(2) + b * t2) / q8;
        ht = T(2.0 * kBeta) * phi3 * q9 / q8;
    }
    const T rsthrd = rs / T(3);
    const T fac = T(kDelta) / b + T(1);
    const T bec = b2 * fac / (T(kBeta) * phi3);
    const T hrs = -rsthrd * hb * bec * ecrs;
    const T comm = h + hrs - T(7) * t2 * ht / T(6);
    const T gz = (xpow(T(1) + zeta + T(kEta), T(-1.0 / 3.0)) - xpow(T(1) - zeta + T(kEta), T(-1.0 / 3.0))) / T(3);
    const T bg = T(-3) * b2 * ec * fac / (T(kBeta) * phi4);
    const T hz = T(3) * gz * h / phi + hb * (bg * gz + bec * ecz);
    const T pref = hz - gz * t2 * ht / phi;
    out[0] = ec + h;
    out[1] = (ec - rs * ecrs / T(3) - zeta * ecz) + (ecz) + comm + pref * (T(1) - zeta);
    out[2] = (ec - rs * ecrs / T(3) - zeta * ecz) - (ecz) + comm + pref * (T(-1) - zeta);
    if (gm > T(0)) out[3] = t * ht / tkp;
}

template <typename T>
__device__ inline Out<T> pbe96(T a_in, T b_in, T ga, T gb) {  // PBE96.cpp:814-966
    Out<T> o{T(0), T(0), T(0), T(0), T(0)};
    const T a = tmax(a_in, T(0)), b = tmax(b_in, T(0));
    const T n = a + b;
    if (n <= T(kRhoFloor)) return o;
    const T gam = fabs(ga), gbm = fabs(gb);
    T xa0, xa1, xa2, xb0, xb1, xb2, c[4];
    pbe_x(T(2) * a, T(2) * gam, xa0, xa1, xa2);
    pbe_x(T(2) * b, T(2) * gbm, xb0, xb1, xb2);
    pbe_c(a, b, ga, gb, c);
    const T exd = T(0.5) * (xa0 + xb0);
    o.eps = exd / n + c[0];
    o.va = xa1 + c[1];
    o.vb = xb1 + c[2];
    const T cs = sign_of(ga + gb);
    o.ca = xa2 * sign_of(ga) + c[3] * cs;
    o.cb = xb2 * sign_of(gb) + c[3] * cs;
    return o;
}

template <typename T>
__device__ inline Out<T> evaluate(int functional, T na, T nb, T ga, T gb) {
    if (functional == 1) return pz81(na, nb);
    return pbe96(na, nb, ga, gb);
}

}  // namespace xc
}  // namespace dft
