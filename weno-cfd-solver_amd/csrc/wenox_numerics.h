// wenox_numerics.h -- point-wise numerics of the 2D Euler WENO5 / HLLC
// right-hand side, shared by every kernel (device) and usable on the host.
//
// Each function restates one piece of the reference Rust crate weno_ext; the
// file:line it follows is given above it.  The translation unit that includes
// this header picks one of two arithmetic variants:
//
//   WX_FAST == 0  ("exact"): the reference's operation order, true IEEE
//                 divisions, built with -ffp-contract=off.  Bitwise equal to
//                 the reference semantics wherever division and sqrt are
//                 correctly rounded (they are on gfx950).
//   WX_FAST == 1  ("fast"):  algebraically equal rewrites that remove most
//                 divisions (one division per WENO5 evaluation instead of
//                 seven, reciprocals shared across a face) and FMA
//                 contraction.  Results differ from "exact" by rounding only.
//
// A third build, WX_TWIN == 1 (with WX_FAST == 0), follows the reference's
// pure-NumPy twin WENO_HLLC_2d.py instead of the Rust crate where the two
// differ: the exact sqrt(3)/12 in the Gauss-point pass (:5), HLLC's
// right-star coefficient rho_R (S* - u_R) / (S_R - S*) and wave speeds
// without the max(0, .) under the root (:453,476,528,547) with its masks
// assigned in order (later regions overwrite earlier ones, :493-506), and
// HLLE's middle state over the raw S_R - S_L with the unfloored face states,
// S_R <= 0 taking F_R (:601-621,647-667).
//
// T is double (FP64 path) or float (the bulk arithmetic of the FP32 path).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#ifndef WX_FAST
#error "WX_FAST must be defined to 0 or 1"
#endif
#ifndef WX_TWIN
#define WX_TWIN 0
#endif
#if WX_TWIN && WX_FAST
#error "the twin semantics are built with the exact operation order only"
#endif

#define WX_HD __host__ __device__ __forceinline__

namespace wx {

enum Dir : int { DIR_X = 0, DIR_Y = 1 };

// utils.rs:3-6
template <typename T> struct K {
    static constexpr T eps = T(1e-12);        // DEFAULT_EPS = RHO_MIN = P_MIN
#if WX_TWIN
    static constexpr T sqrt3_12 = T(0.14433756729740643);   // np.sqrt(3)/12 (WENO_HLLC_2d.py:5)
#else
    static constexpr T sqrt3_12 = T(0.1443375673);
#endif
};

template <typename T> WX_HD T sq(T x) { return x * x; }
template <typename T> WX_HD T vmax(T a, T b) { return fmax(a, b); }   // Rust f64::max
template <typename T> WX_HD T vmin(T a, T b) { return fmin(a, b); }   // Rust f64::min

// Reciprocal used only by the fast variant.  Divisors passed here are
// strictly positive and finite by construction (squares offset by 1e-12,
// floored densities, sums of positive terms).  v_rcp_f64 is good to about
// 24 bits on gfx950 (max relative error 4.6e-8 measured over 2^20 values
// spread across 1e-13..1e13); one Newton step brings it to 2.2e-15, within
// one ulp of the correctly rounded quotient.
WX_HD double rcp(double x)
{
#if WX_FAST && defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    return fma(r, e, r);
#else
    return 1.0 / x;
#endif
}
// Reciprocal for divisors that may be zero or infinite (the guarded Riemann
// denominators): v_rcp_f64 returns +-inf for +-0 and 0 for +-inf exactly as
// 1/x does; the Newton step is applied only to finite nonzero x.
WX_HD double rcp_z(double x)
{
#if WX_FAST && defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    double rr = fma(r, e, r);
    return (x == 0.0 || fabs(x) == __builtin_inf()) ? r : rr;
#else
    return 1.0 / x;
#endif
}
WX_HD float rcp(float x)
{
#if WX_FAST && defined(__HIP_DEVICE_COMPILE__)
    float r = __builtin_amdgcn_rcpf(x);
    return fmaf(r, fmaf(-x, r, 1.0f), r);
#else
    return 1.0f / x;
#endif
}
WX_HD float rcp_z(float x) { return 1.0f / x; }

// Square root and reciprocal square root of a positive finite x for the fast
// variant: v_rsq_f64 followed by the coupled Goldschmidt iteration for
// s ~ sqrt(x), h ~ 1/(2 sqrt(x)) (Markstein, "IA-64 and Elementary
// Functions", 2000, ch. 9), a residual correction of s (s + (x - s^2) h) and
// a second step on h.  Measured on gfx950 over 2^20 values spread across
// 1e-26..1e26: s equal to the correctly rounded root in every case, 1/s
// within 2.2e-15 relative.  The sequence costs one transcendental and nine
// VALU operations against about seventeen for a correctly rounded sqrt plus
// a reciprocal.
WX_HD void sqrt_rsqrt(double x, double& s, double& g)
{
#if WX_FAST && defined(__HIP_DEVICE_COMPILE__)
    double y = __builtin_amdgcn_rsq(x);
    double ss = x * y, h = 0.5 * y;
    double r = fma(-ss, h, 0.5);
    ss = fma(ss, r, ss);
    h = fma(h, r, h);
    double d = fma(-ss, ss, x);
    ss = fma(d, h, ss);
    r = fma(-ss, h, 0.5);
    h = fma(h, r, h);
    s = ss;
    g = h + h;
#else
    s = sqrt(x);
    g = 1.0 / s;
#endif
}
WX_HD void sqrt_rsqrt(float x, float& s, float& g)
{
    s = sqrtf(x);
    g = 1.0f / s;
}
// Square root of a finite x >= 0 (zero allowed) for the fast variant: the
// same iteration without the second step on h.  Arguments below 1e-300
// (zero, subnormals) return 0, NaN stays NaN, +inf gives NaN.
WX_HD double sqrt_z(double x)
{
#if WX_FAST && defined(__HIP_DEVICE_COMPILE__)
    double y = __builtin_amdgcn_rsq(x);
    double ss = x * y, h = 0.5 * y;
    double r = fma(-ss, h, 0.5);
    ss = fma(ss, r, ss);
    h = fma(h, r, h);
    double d = fma(-ss, ss, x);
    ss = fma(d, h, ss);
    return x > 1e-300 ? ss : (x == x ? 0.0 : x);
#else
    return sqrt(x);
#endif
}
WX_HD float sqrt_z(float x) { return sqrtf(x); }

// Roe-average ingredients from two positive densities a (left), b (right):
// k = sqrt(b / a), sqrt(a b), 1 / sqrt(a b), 1 / a and 1 / b, from one
// sqrt_rsqrt of a b (fast variant; a, b >= 1e-12 keep a b in range).
template <typename T> struct RoeW { T k, rt, irt, ia, ib; };
template <typename T> WX_HD RoeW<T> roe_weights(T a, T b)
{
    RoeW<T> w;
    T ab = a * b;
    sqrt_rsqrt(ab, w.rt, w.irt);
    T g2 = w.irt * w.irt;
    w.k = b * w.irt;
    w.ia = b * g2;
    w.ib = a * g2;
    return w;
}

template <typename T> struct V4 { T v[4]; };

// ---------------------------------------------------------------------
// utils.rs:74-108  Euler flux normal to x (dir 0) or y (dir 1).
// Density is floored, pressure is not.
// ---------------------------------------------------------------------
template <typename T> WX_HD void flux_uv(int dir, const T q[4], T rho, T u, T v, T gam, T f[4])
{
    T p = (gam - T(1)) * (q[3] - T(0.5) * rho * (u * u + v * v));
    if (dir == DIR_X) {
        f[0] = q[1];
        f[1] = q[1] * u + p;
        f[2] = q[2] * u;
        f[3] = u * (q[3] + p);
    } else {
        f[0] = q[2];
        f[1] = q[1] * v;
        f[2] = q[2] * v + p;
        f[3] = v * (q[3] + p);
    }
}
template <typename T> WX_HD void flux_pt(int dir, const T q[4], T gam, T f[4])
{
    T rho = vmax(q[0], K<T>::eps);
#if WX_FAST
    T ir = rcp(rho);
    flux_uv(dir, q, rho, q[1] * ir, q[2] * ir, gam, f);
#else
    flux_uv(dir, q, rho, q[1] / rho, q[2] / rho, gam, f);
#endif
}
// Two fluxes; the fast variant shares one reciprocal, 1/a = b / (a b) and
// 1/b = a / (a b) (floored densities keep a b in range).
template <typename T> WX_HD void flux_pt2(int dir, const T qa[4], const T qb[4], T gam, T fa[4], T fb[4])
{
#if WX_FAST
    T ra = vmax(qa[0], K<T>::eps), rb = vmax(qb[0], K<T>::eps);
    T r = rcp(ra * rb);
    T ia = rb * r, ib = ra * r;
    flux_uv(dir, qa, ra, qa[1] * ia, qa[2] * ia, gam, fa);
    flux_uv(dir, qb, rb, qb[1] * ib, qb[2] * ib, gam, fb);
#else
    flux_pt(dir, qa, gam, fa);
    flux_pt(dir, qb, gam, fb);
#endif
}

// utils.rs:9-35  conservative -> (p, a, rho_safe, u, v, h); floors on rho and p.
template <typename T> struct Prim { T p, a, rho, u, v, h; };
template <typename T> WX_HD Prim<T> con2primi_pt(const T q[4], T gam)
{
    Prim<T> r;
    r.rho = vmax(q[0], K<T>::eps);
    r.u = q[1] / r.rho;
    r.v = q[2] / r.rho;
    r.p = vmax((gam - T(1)) * (q[3] - T(0.5) * r.rho * (r.u * r.u + r.v * r.v)), K<T>::eps);
    r.a = sqrt(gam * r.p / r.rho);
    r.h = (q[3] + r.p) / r.rho;
    return r;
}

// Floored pressure only (the sensor's input), same arithmetic as con2primi_pt.
template <typename T> WX_HD T pressure_floored(const T q[4], T gam)
{
    T rho = vmax(q[0], K<T>::eps);
#if WX_FAST
    T ir = rcp(rho);
    T u = q[1] * ir, v = q[2] * ir;
#else
    T u = q[1] / rho, v = q[2] / rho;
#endif
    return vmax((gam - T(1)) * (q[3] - T(0.5) * rho * (u * u + v * v)), K<T>::eps);
}

// ---------------------------------------------------------------------
// weno.rs:13-63  Roe average for the characteristic basis (same for x, y).
// ---------------------------------------------------------------------
// The fast variant writes the weights as sqrt(rho_l) / (sqrt(rho_l) +
// sqrt(rho_r)) = 1 / (1 + k), k = sqrt(rho_r / rho_l), and takes rho~ =
// sqrt(rho_l rho_r) and 1 / rho~ from roe_weights; it also returns 1 / c.
template <typename T> struct Roe { T rho, u, v, h, c, ic, irho; };
template <typename T> WX_HD Roe<T> roe_avg(const T ul[4], const T ur[4], T gam)
{
    const T eps = K<T>::eps;
    T rho_l = vmax(ul[0], eps);
    T rho_r = vmax(ur[0], eps);
#if WX_FAST
    RoeW<T> rw = roe_weights(rho_l, rho_r);
    T il = rw.ia, ir = rw.ib, id = rcp(T(1) + rw.k);
    T u_l = ul[1] * il, v_l = ul[2] * il;
    T u_r = ur[1] * ir, v_r = ur[2] * ir;
    T p_l = vmax((gam - T(1)) * (ul[3] - T(0.5) * rho_l * (u_l * u_l + v_l * v_l)), eps);
    T p_r = vmax((gam - T(1)) * (ur[3] - T(0.5) * rho_r * (u_r * u_r + v_r * v_r)), eps);
    T h_l = (ul[3] + p_l) * il;
    T h_r = (ur[3] + p_r) * ir;
    Roe<T> r;
    r.u = (u_l + rw.k * u_r) * id;
    r.v = (v_l + rw.k * v_r) * id;
    r.h = (h_l + rw.k * h_r) * id;
    T kin = T(0.5) * (r.u * r.u + r.v * r.v);
    sqrt_rsqrt(vmax((gam - T(1)) * (r.h - kin), eps * vmax(il, ir)), r.c, r.ic);
    r.rho = rw.rt;
    r.irho = rw.irt;
    return r;
#else
    T sl = sqrt(rho_l), sr = sqrt(rho_r);
    T den = sl + sr;
    T u_l = ul[1] / rho_l, v_l = ul[2] / rho_l;
    T u_r = ur[1] / rho_r, v_r = ur[2] / rho_r;
    T p_l = vmax((gam - T(1)) * (ul[3] - T(0.5) * rho_l * (u_l * u_l + v_l * v_l)), eps);
    T p_r = vmax((gam - T(1)) * (ur[3] - T(0.5) * rho_r * (u_r * u_r + v_r * v_r)), eps);
    T h_l = (ul[3] + p_l) / rho_l;
    T h_r = (ur[3] + p_r) / rho_r;
    Roe<T> r;
    r.u = (sl * u_l + sr * u_r) / den;
    r.v = (sl * v_l + sr * v_r) / den;
    r.h = (sl * h_l + sr * h_r) / den;
    T kin = T(0.5) * (r.u * r.u + r.v * r.v);
    r.c = sqrt(vmax((gam - T(1)) * (r.h - kin), eps / vmin(rho_l, rho_r)));
    r.rho = sl * sr;
    r.ic = T(0);     // not used by the exact variant
    r.irho = T(0);
    return r;
#endif
}

// weno.rs:76-189  left (rows: u-c, entropy, shear, u+c) and right
// eigenvector matrices of the x- or y-flux Jacobian at the Roe state.
template <typename T> struct Mat4 { T m[4][4]; };

template <typename T> WX_HD void eig_left(int dir, const Roe<T>& s, T gam, T L[4][4])
{
    const T rho = s.rho, u = s.u, v = s.v, c = s.c;
    T g1 = gam - T(1);
    T kin = T(0.5) * (u * u + v * v);
    T un = dir == DIR_X ? u : v;
    T rc = rho * c;
    T inv_rc = T(1) / rc;
    T inv_c2 = T(1) / (c * c);
    L[0][0] = g1 * kin / (T(2) * c * c) + un / (T(2) * c);
    L[3][0] = g1 * kin / (T(2) * c * c) - un / (T(2) * c);
    L[1][0] = T(1) - g1 * kin * inv_c2;
    L[1][1] = g1 * u * inv_c2;
    L[1][2] = g1 * v * inv_c2;
    L[1][3] = -g1 * inv_c2;
    L[0][3] = g1 * inv_c2 / T(2);
    L[3][3] = g1 * inv_c2 / T(2);
    if (dir == DIR_X) {
        L[0][1] = -g1 * u * inv_c2 / T(2) - T(0.5) * inv_rc;
        L[0][2] = -g1 * v * inv_c2 / T(2);
        L[3][1] = -g1 * u * inv_c2 / T(2) + T(0.5) * inv_rc;
        L[3][2] = -g1 * v * inv_c2 / T(2);
        L[2][0] = -v / rho; L[2][1] = T(0); L[2][2] = T(1) / rho; L[2][3] = T(0);
    } else {
        L[0][1] = -g1 * u * inv_c2 / T(2);
        L[0][2] = -g1 * v * inv_c2 / T(2) - T(0.5) * inv_rc;
        L[3][1] = -g1 * u * inv_c2 / T(2);
        L[3][2] = -g1 * v * inv_c2 / T(2) + T(0.5) * inv_rc;
        L[2][0] = -u / rho; L[2][1] = T(1) / rho; L[2][2] = T(0); L[2][3] = T(0);
    }
}

template <typename T> WX_HD void eig_right(int dir, const Roe<T>& s, T R[4][4])
{
    const T rho = s.rho, u = s.u, v = s.v, h = s.h, c = s.c;
    T kin = T(0.5) * (u * u + v * v);
    R[0][0] = T(1); R[0][1] = T(1); R[0][2] = T(0); R[0][3] = T(1);
    if (dir == DIR_X) {
        R[1][0] = u - c; R[1][1] = u; R[1][2] = T(0); R[1][3] = u + c;
        R[2][0] = v; R[2][1] = v; R[2][2] = rho; R[2][3] = v;
        R[3][0] = h - u * c; R[3][1] = kin; R[3][2] = rho * v; R[3][3] = h + u * c;
    } else {
        R[1][0] = u; R[1][1] = u; R[1][2] = rho; R[1][3] = u;
        R[2][0] = v - c; R[2][1] = v; R[2][2] = T(0); R[2][3] = v + c;
        R[3][0] = h - v * c; R[3][1] = kin; R[3][2] = rho * u; R[3][3] = h + v * c;
    }
}

// weno.rs:197-205  accumulation from 0.0 in column order
template <typename T> WX_HD void mat4_vec4(const T M[4][4], const T x[4], T y[4])
{
#pragma unroll
    for (int r = 0; r < 4; ++r) {
#if WX_FAST
        y[r] = M[r][0] * x[0] + M[r][1] * x[1] + M[r][2] * x[2] + M[r][3] * x[3];
#else
        T acc = T(0);
#pragma unroll
        for (int c = 0; c < 4; ++c) acc += M[r][c] * x[c];
        y[r] = acc;
#endif
    }
}

// Fast-variant products with the reference's eigenvector matrices, from
// their entries' structure instead of 4x4 products (weno.rs:76-189):
//   rows 0 and 3 of L are A +- B with A = a c0 - hu c1 - hv c2 + h3 c3 and
//   B = b c0 - hr c_n (c_n the normal momentum); row 1 equals
//   e0 - 2 (a, -hu, -hv, h3), so L[1] c = c0 - 2 A; row 2 is
//   (c_t - w_t c0) / rho (c_t, w_t transverse momentum and velocity).
//   R c has y0 = c0 + c1 + c3, the normal-momentum row w_n y0 + c (c3 - c0),
//   the transverse one w_t y0 + rho c2 and the energy row
//   h y0 + (kin - h) c1 + w_n c (c3 - c0) + rho w_t c2.
// The exact variant keeps the matrices and the four-term accumulation.
template <typename T> struct LeftRows { T a, b, hu, hv, h3, hr, irho, wt; };
template <typename T> WX_HD LeftRows<T> left_rows(int dir, const Roe<T>& s, T gam)
{
    const T g1 = gam - T(1), u = s.u, v = s.v, ic = s.ic;
    const T h3 = T(0.5) * g1 * (ic * ic);
    LeftRows<T> L;
    L.a = T(0.5) * (u * u + v * v) * h3;
    L.b = (dir == DIR_X ? u : v) * (T(0.5) * ic);
    L.hu = h3 * u;
    L.hv = h3 * v;
    L.h3 = h3;
    L.hr = T(0.5) * (s.irho * ic);
    L.irho = s.irho;
    L.wt = dir == DIR_X ? v : u;
    return L;
}
template <typename T> WX_HD void left_project(int dir, const LeftRows<T>& L, const T x[4], T y[4])
{
    const T cn = dir == DIR_X ? x[1] : x[2];
    const T ct = dir == DIR_X ? x[2] : x[1];
    T A = L.a * x[0] - L.hu * x[1] - L.hv * x[2] + L.h3 * x[3];
    T B = L.b * x[0] - L.hr * cn;
    y[0] = A + B;
    y[1] = x[0] - T(2) * A;
    y[2] = L.irho * (ct - L.wt * x[0]);
    y[3] = A - B;
}
template <typename T> struct RightCols { T wn, wt, h, c, rho, kmh, wnc, rwt; };
template <typename T> WX_HD RightCols<T> right_cols(int dir, const Roe<T>& s)
{
    RightCols<T> R;
    R.wn = dir == DIR_X ? s.u : s.v;
    R.wt = dir == DIR_X ? s.v : s.u;
    R.h = s.h;
    R.c = s.c;
    R.rho = s.rho;
    R.kmh = T(0.5) * (s.u * s.u + s.v * s.v) - s.h;
    R.wnc = R.wn * s.c;
    R.rwt = s.rho * R.wt;
    return R;
}
template <typename T> WX_HD void right_project(int dir, const RightCols<T>& R, const T x[4], T y[4])
{
    T y0 = x[0] + x[1] + x[3];
    T d30 = x[3] - x[0];
    T yn = R.wn * y0 + R.c * d30;
    T yt = R.wt * y0 + R.rho * x[2];
    y[0] = y0;
    y[1] = dir == DIR_X ? yn : yt;
    y[2] = dir == DIR_X ? yt : yn;
    y[3] = R.h * y0 + R.kmh * x[1] + R.wnc * d30 + R.rwt * x[2];
}

// ---------------------------------------------------------------------
// WENO5-JS (weno.rs:214-257).  weno5_left: value at the right face of the
// centre cell u2 from u0..u4; weno5_right: value at the left face of the
// centre cell u3 from u1..u5.  eps = 1e-12, alpha_k = d_k / (eps+beta_k)^2.
// The fast variant evaluates sum(alpha_k p_k) / (6 sum(alpha_k)) with all
// three alphas scaled by prod_k (eps+beta_k)^2, i.e. one division, and works
// with 4 beta_k and 4 eps (beta4), which leaves the normalised weights
// unchanged and saves one multiplication per indicator.
// ---------------------------------------------------------------------
template <typename T> WX_HD T beta4(T d2, T s) { return T(13.0 / 3.0) * d2 * d2 + s * s; }

template <typename T> WX_HD T weno5_core(T b0, T b1, T b2, T d0, T d1, T d2, T p0, T p1, T p2)
{
#if WX_FAST
    const T eps = T(4) * K<T>::eps;   // the betas come from beta4
    T t0 = sq(eps + b0), t1 = sq(eps + b1), t2 = sq(eps + b2);
    if constexpr (sizeof(T) == 4) {
        // FP32: products of two (eps+beta)^2 underflow below 1e-38; use one
        // reciprocal per weight instead
        T a0 = d0 * rcp(t0), a1 = d1 * rcp(t1), a2 = d2 * rcp(t2);
        return (a0 * p0 + a1 * p1 + a2 * p2) * rcp(T(6) * (a0 + a1 + a2));
    } else {
        T a0 = d0 * (t1 * t2), a1 = d1 * (t0 * t2), a2 = d2 * (t0 * t1);
        return (a0 * p0 + a1 * p1 + a2 * p2) * rcp(T(6) * (a0 + a1 + a2));
    }
#else
    const T eps = K<T>::eps;
    T a0 = d0 / sq(eps + b0);
    T a1 = d1 / sq(eps + b1);
    T a2 = d2 / sq(eps + b2);
    T s = a0 + a1 + a2;
    T w0 = a0 / s, w1 = a1 / s, w2 = a2 / s;
    return (w0 * p0 + w1 * p1 + w2 * p2) / T(6);
#endif
}

template <typename T> WX_HD T weno5_left(T u0, T u1, T u2, T u3, T u4)
{
#if WX_FAST
    T b0 = beta4(u2 - T(2) * u3 + u4, T(3) * u2 - T(4) * u3 + u4);
    T b1 = beta4(u1 - T(2) * u2 + u3, u1 - u3);
    T b2 = beta4(u0 - T(2) * u1 + u2, u0 - T(4) * u1 + T(3) * u2);
#else
    const T c13 = T(13.0 / 12.0);
    T b0 = c13 * sq(u2 - T(2) * u3 + u4) + T(0.25) * sq(T(3) * u2 - T(4) * u3 + u4);
    T b1 = c13 * sq(u1 - T(2) * u2 + u3) + T(0.25) * sq(u1 - u3);
    T b2 = c13 * sq(u0 - T(2) * u1 + u2) + T(0.25) * sq(u0 - T(4) * u1 + T(3) * u2);
#endif
    T p0 = -u4 + T(5) * u3 + T(2) * u2;
    T p1 = -u1 + T(5) * u2 + T(2) * u3;
    T p2 = T(2) * u0 - T(7) * u1 + T(11) * u2;
    return weno5_core(b0, b1, b2, T(0.3), T(0.6), T(0.1), p0, p1, p2);
}

template <typename T> WX_HD T weno5_right(T u1, T u2, T u3, T u4, T u5)
{
#if WX_FAST
    T b0 = beta4(u3 - T(2) * u4 + u5, T(3) * u3 - T(4) * u4 + u5);
    T b1 = beta4(u2 - T(2) * u3 + u4, u2 - u4);
    T b2 = beta4(u1 - T(2) * u2 + u3, u1 - T(4) * u2 + T(3) * u3);
#else
    const T c13 = T(13.0 / 12.0);
    T b0 = c13 * sq(u3 - T(2) * u4 + u5) + T(0.25) * sq(T(3) * u3 - T(4) * u4 + u5);
    T b1 = c13 * sq(u2 - T(2) * u3 + u4) + T(0.25) * sq(u2 - u4);
    T b2 = c13 * sq(u1 - T(2) * u2 + u3) + T(0.25) * sq(u1 - T(4) * u2 + T(3) * u3);
#endif
    T p0 = T(2) * u5 - T(7) * u4 + T(11) * u3;
    T p1 = -u4 + T(5) * u3 + T(2) * u2;
    T p2 = -u1 + T(5) * u2 + T(2) * u3;
    return weno5_core(b0, b1, b2, T(0.1), T(0.6), T(0.3), p0, p1, p2);
}

// Both WENO5 values of one face: left from u0..u4 (weno5_left), right from
// u1..u5 (weno5_right).  The fast FP64 variant normalises the two with one
// reciprocal: left = nL / dL = nL dR / (dL dR), right = nR dL / (dL dR).  dL
// and dR are 6 sums of d_k times products of two (eps+beta)^2 factors (about
// beta^4 each), so dL dR stays in range while the indicators stay below
// about 1e37 (differences of the stencil values below about 1e18); at the
// other end it is at least about 1e-90.
template <typename T> WX_HD void weno5_pair(T u0, T u1, T u2, T u3, T u4, T u5, T& left, T& right)
{
#if WX_FAST
    if constexpr (sizeof(T) == 8) {
        const T eps = T(4) * K<T>::eps;
        auto nd = [&](T b0, T b1, T b2, T d0, T d1, T d2, T p0, T p1, T p2, T& num, T& den) {
            T t0 = sq(eps + b0), t1 = sq(eps + b1), t2 = sq(eps + b2);
            T a0 = d0 * (t1 * t2), a1 = d1 * (t0 * t2), a2 = d2 * (t0 * t1);
            num = a0 * p0 + a1 * p1 + a2 * p2;
            den = T(6) * (a0 + a1 + a2);
        };
        T nl, dl, nr, dr;
        nd(beta4(u2 - T(2) * u3 + u4, T(3) * u2 - T(4) * u3 + u4), beta4(u1 - T(2) * u2 + u3, u1 - u3),
           beta4(u0 - T(2) * u1 + u2, u0 - T(4) * u1 + T(3) * u2), T(0.3), T(0.6), T(0.1),
           -u4 + T(5) * u3 + T(2) * u2, -u1 + T(5) * u2 + T(2) * u3, T(2) * u0 - T(7) * u1 + T(11) * u2, nl, dl);
        nd(beta4(u3 - T(2) * u4 + u5, T(3) * u3 - T(4) * u4 + u5), beta4(u2 - T(2) * u3 + u4, u2 - u4),
           beta4(u1 - T(2) * u2 + u3, u1 - T(4) * u2 + T(3) * u3), T(0.1), T(0.6), T(0.3),
           T(2) * u5 - T(7) * u4 + T(11) * u3, -u4 + T(5) * u3 + T(2) * u2, -u1 + T(5) * u2 + T(2) * u3, nr, dr);
        const T r = rcp(dl * dr);
        left = nl * (dr * r);
        right = nr * (dl * r);
        return;
    }
#endif
    left = weno5_left(u0, u1, u2, u3, u4);
    right = weno5_right(u1, u2, u3, u4, u5);
}

// Fast FP64 Gauss-point pass (weno5_gq below) from the centre value q2 and
// the four first differences of its stencil, dmm = q1 - q0, dm = q2 - q1,
// dp = q3 - q2, dpp = q4 - q3.  The second differences and one-sided slopes
// of the smoothness indicators follow from them:
//   q2 - 2 q3 + q4 = dpp - dp          3 q2 - 4 q3 + q4 = (dpp - dp) - 2 dp
//   q1 - 2 q2 + q3 = dp - dm           q3 - q1 = dm + dp
//   q0 - 2 q1 + q2 = dm - dmm          3 q2 - 4 q1 + q0 = (dm - dmm) + 2 dm
// so a lane of the fused sweep forms one difference per value and takes the
// other three from its neighbours (its dp is lane l+1's dm, lane l+2's dmm
// and lane l-1's dpp); every engine passes the same differences and rounds
// alike.  The middle slope is q3 - q1 here, so its terms enter with the
// opposite sign.
WX_HD void weno5_gq_d(double q2, double dmm, double dm, double dp, double dpp, double& ga, double& gb)
{
    const double S = K<double>::sqrt3_12;
    const double dA = 0.1928406937, dB = 0.6111111111, dC = 0.1960481952;
    const double eps = 4.0 * K<double>::eps;   // with the betas from beta4
    const double c0 = dpp - dp, c1 = dp - dm, c2 = dm - dmm;
    const double s0 = fma(-2.0, dp, c0), s1 = dm + dp, s2 = fma(2.0, dm, c2);
    const double b0 = beta4(c0, s0), b1 = beta4(c1, s1), b2 = beta4(c2, s2);
    const double t0 = sq(eps + b0), t1 = sq(eps + b1), t2 = sq(eps + b2);
    const double e0 = t1 * t2, e1 = t0 * t2, e2 = t0 * t1;   // weights up to a common factor
    const double SA = S * dA, SB = S * dB, SC = S * dC;
    const double u0 = e0 * s0, u1 = e1 * s1, u2 = e2 * s2;
    const double na = SA * u0 - SB * u1 - SC * u2;
    const double nb = SC * u0 - SB * u1 - SA * u2;
    const double da = dA * e0 + dB * e1 + dC * e2;
    const double db = dC * e0 + dB * e1 + dA * e2;
    const double r = rcp(da * db);   // one reciprocal for both normalisations
    ga = q2 + na * (db * r);
    gb = q2 - nb * (da * r);
}

// weno.rs:262-303  Gauss-point values (ga at -h/(2 sqrt 3), gb at +h/(2 sqrt 3))
// from five face values, sharing the smoothness indicators.
template <typename T> WX_HD void weno5_gq(T q0, T q1, T q2, T q3, T q4, T& ga, T& gb)
{
    const T S = K<T>::sqrt3_12;
    const T dA = T(0.1928406937), dB = T(0.6111111111), dC = T(0.1960481952);
#if WX_FAST
    if constexpr (sizeof(T) == 8) {
        weno5_gq_d(q2, q1 - q0, q2 - q1, q3 - q2, q4 - q3, ga, gb);
        return;
    }
    const T eps = T(4) * K<T>::eps;   // with the betas from beta4
    T s0 = T(3) * q2 - T(4) * q3 + q4;
    T s1 = q1 - q3;
    T s2 = T(3) * q2 - T(4) * q1 + q0;
    T b0 = beta4(q2 - T(2) * q3 + q4, s0);
    T b1 = beta4(q1 - T(2) * q2 + q3, s1);
    T b2 = beta4(q0 - T(2) * q1 + q2, s2);
    T t0 = sq(eps + b0), t1 = sq(eps + b1), t2 = sq(eps + b2);
    T e0, e1, e2;   // weights up to a common factor (see weno5_core)
    if constexpr (sizeof(T) == 4) {
        e0 = rcp(t0); e1 = rcp(t1); e2 = rcp(t2);
    } else {
        e0 = t1 * t2; e1 = t0 * t2; e2 = t0 * t1;
    }
    // The candidates are q2 + S s0, q2 + S s1, q2 - S s2 (ga) and their
    // mirror images (gb); the weights sum to one, so
    //   ga = q2 + S (dA e0 s0 + dB e1 s1 - dC e2 s2) / (dA e0 + dB e1 + dC e2)
    //   gb = q2 - S (dC e0 s0 + dB e1 s1 - dA e2 s2) / (dC e0 + dB e1 + dA e2).
    const T SA = S * dA, SB = S * dB, SC = S * dC;
    T u0 = e0 * s0, u1 = e1 * s1, u2 = e2 * s2;
    T na = SA * u0 + SB * u1 - SC * u2;
    T nb = SC * u0 + SB * u1 - SA * u2;
    T da = dA * e0 + dB * e1 + dC * e2;
    T db = dC * e0 + dB * e1 + dA * e2;
    if constexpr (sizeof(T) == 8) {
        // one reciprocal for both normalisations (da db stays in range as in weno5_pair)
        const T r = rcp(da * db);
        ga = q2 + na * (db * r);
        gb = q2 - nb * (da * r);
    } else {
        ga = q2 + na * rcp(da);
        gb = q2 - nb * rcp(db);
    }
#else
    const T c13 = T(13.0 / 12.0);
    const T eps = K<T>::eps;
    T b0 = c13 * sq(q2 - T(2) * q3 + q4) + T(0.25) * sq(T(3) * q2 - T(4) * q3 + q4);
    T b1 = c13 * sq(q1 - T(2) * q2 + q3) + T(0.25) * sq(q1 - q3);
    T b2 = c13 * sq(q0 - T(2) * q1 + q2) + T(0.25) * sq(q0 - T(4) * q1 + T(3) * q2);
    {
        T a0 = dA / sq(eps + b0), a1 = dB / sq(eps + b1), a2 = dC / sq(eps + b2);
        T s = a0 + a1 + a2;
        T w0 = a0 / s, w1 = a1 / s, w2 = a2 / s;
        T p0 = q2 + (T(3) * q2 - T(4) * q3 + q4) * S;
        T p1 = q2 - (-q1 + q3) * S;
        T p2 = q2 - (T(3) * q2 - T(4) * q1 + q0) * S;
        ga = w0 * p0 + w1 * p1 + w2 * p2;
    }
    {
        T a0 = dC / sq(eps + b0), a1 = dB / sq(eps + b1), a2 = dA / sq(eps + b2);
        T s = a0 + a1 + a2;
        T w0 = a0 / s, w1 = a1 / s, w2 = a2 / s;
        T p0 = q2 - (T(3) * q2 - T(4) * q3 + q4) * S;
        T p1 = q2 - (q1 - q3) * S;
        T p2 = q2 + (T(3) * q2 - T(4) * q1 + q0) * S;
        gb = w0 * p0 + w1 * p1 + w2 * p2;
    }
#endif
}

// weno.rs:312-352  positivity limiter: density pass then pressure pass,
// scaling the face state toward the cell average ub.
template <typename T> WX_HD T pressure_pp(const T q[4], T gam)
{
    if (q[0] <= T(0)) return T(0);
#if WX_FAST
    return (gam - T(1)) * (q[3] - T(0.5) * (q[1] * q[1] + q[2] * q[2]) * rcp(q[0]));
#else
    return (gam - T(1)) * (q[3] - T(0.5) * (q[1] * q[1] + q[2] * q[2]) / q[0]);
#endif
}

template <typename T> WX_HD void pp_limit(T q[4], const T ub[4], T gam)
{
    const T eps = K<T>::eps;
    if (q[0] < eps) {
        T rb = vmax(ub[0], eps);
        T th = vmax(vmin((rb - eps) / (rb - q[0]), T(1)), T(0));
#pragma unroll
        for (int k = 0; k < 4; ++k) q[k] = th * (q[k] - ub[k]) + ub[k];
    }
    T pq = pressure_pp(q, gam);
    if (pq < eps) {
        T pb = vmax(pressure_pp(ub, gam), eps);
        T th = vmax(vmin((pb - eps) / (pb - pq), T(1)), T(0));
#pragma unroll
        for (int k = 0; k < 4; ++k) q[k] = th * (q[k] - ub[k]) + ub[k];
    }
}

// ---------------------------------------------------------------------
// Normal-direction reconstruction at one face from its six-cell stencil
// c[0..5] (the face lies between c[2] and c[3]).
//   characteristic: weno.rs:385-454 (basis from the Roe state of c2|c3,
//                   projected WENO5, back-projection, positivity limiter);
//   component-wise: weno.rs:654-717 (no projection, no limiter).
// ---------------------------------------------------------------------
template <typename T> WX_HD void recon_normal(int dir, bool cons, const T c[6][4], T gam,
                                              T ql[4], T qr[4])
{
    if (cons) {
#pragma unroll
        for (int k = 0; k < 4; ++k) weno5_pair(c[0][k], c[1][k], c[2][k], c[3][k], c[4][k], c[5][k], ql[k], qr[k]);
        return;
    }
    Roe<T> s = roe_avg(c[2], c[3], gam);
    T w[6][4];
#if WX_FAST
    const LeftRows<T> L = left_rows(dir, s, gam);
#pragma unroll
    for (int t = 0; t < 6; ++t) left_project(dir, L, c[t], w[t]);
#else
    T L[4][4], R[4][4];
    eig_left(dir, s, gam, L);
    eig_right(dir, s, R);
#pragma unroll
    for (int t = 0; t < 6; ++t) mat4_vec4(L, c[t], w[t]);
#endif
    T chl[4], chr[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) weno5_pair(w[0][k], w[1][k], w[2][k], w[3][k], w[4][k], w[5][k], chl[k], chr[k]);
#if WX_FAST
    const RightCols<T> R = right_cols(dir, s);
    right_project(dir, R, chl, ql);
    right_project(dir, R, chr, qr);
#else
    mat4_vec4(R, chl, ql);
    mat4_vec4(R, chr, qr);
#endif
    pp_limit(ql, c[2], gam);
    pp_limit(qr, c[3], gam);
}

// ---------------------------------------------------------------------
// FP32 reconstruction on increments (the FP32-split path).
//
// WENO5 and the Gauss-point pass are exactly translation invariant (the
// smoothness indicators see differences only; every candidate value is an
// affine combination with coefficients summing to one), so
//     weno(c_0..c_4) = c_2 + weno(c_0 - c_2, ..., c_4 - c_2).
// The increments are formed in FP64 and rounded to FP32, the WENO arithmetic
// runs in FP32, and the result is added back to the FP64 centre value: the
// FP32 rounding error scales with the local variation |c_j - c_2| instead
// of |c|, which keeps the flux differences of smooth regions at FP64 level.
//
// Characteristic variant: with w = L c the reconstructed face state is
// R weno(L c_j) = R L c_2 + R weno(L (c_j - c_2)).  The reference's L is not
// the inverse of R (rows 0 and 3 carry 1/(2 rho c) in the normal-momentum
// column where R^-1 has 1/(2c), weno.rs:88,103 / :149,164); the product is
//     R L c = c - ((rho~ - 1)/rho~) c_n (0, 1, 0, u~)   (x faces)
//     R L c = c - ((rho~ - 1)/rho~) c_n (0, 0, 1, v~)   (y faces)
// with c_n the normal momentum and rho~, u~, v~ the Roe state, evaluated in
// FP64.  L and R for the increments are built in FP32 from the rounded
// Roe state.
// ---------------------------------------------------------------------
//
// The FP32 error of a reconstruction scales with the increments it is given,
// so a face whose stencils span a large relative jump is computed in FP64 by
// the sweep: a jump is an increment of density or energy above tau times
// the centre value (tau: the solver's split-jump threshold), in the normal stencil of the face or of one of the
// four faces beside it that feed its Gauss points (recon_normal_dev's
// result), or in the face values feeding its Gauss points (weno5_gq_dev's
// result for those components).  Smooth flow stays below it: the relative
// increments are a few grid spacings times the relative gradient.
// ---------------------------------------------------------------------
template <bool CHECK, typename T> WX_HD bool weno5_gq_dev(T q0, T q1, T q2, T q3, T q4, T& ga, T& gb, float tau)
{
    float a, b;
    const float i0 = (float)(q0 - q2), i1 = (float)(q1 - q2), i3 = (float)(q3 - q2), i4 = (float)(q4 - q2);
    weno5_gq<float>(i0, i1, 0.0f, i3, i4, a, b);
    ga = q2 + (T)a;
    gb = q2 + (T)b;
    if constexpr (!CHECK) return false;
    const float m = fmaxf(fmaxf(fabsf(i0), fabsf(i1)), fmaxf(fabsf(i3), fabsf(i4)));
    return !(m <= tau * fabsf((float)q2));
}

WX_HD bool recon_normal_dev(int dir, bool cons, const double c[6][4], double gam, double ql[4], double qr[4],
                            float tau)
{
    float d[6][4];
#pragma unroll
    for (int t = 0; t < 6; ++t)
#pragma unroll
        for (int k = 0; k < 4; ++k) d[t][k] = t == 2 ? 0.0f : (float)(c[t][k] - c[2][k]);
    // relative jump of density or energy across the stencil
    bool jump = false;
#pragma unroll
    for (int k = 0; k < 4; k += 3) {
        const float m = fmaxf(fmaxf(fmaxf(fabsf(d[0][k]), fabsf(d[1][k])), fmaxf(fabsf(d[3][k]), fabsf(d[4][k]))),
                              fabsf(d[5][k]));
        jump = jump || !(m <= tau * fabsf((float)c[2][k]));
    }
    if (cons) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            ql[k] = c[2][k] + (double)weno5_left<float>(d[0][k], d[1][k], 0.0f, d[3][k], d[4][k]);
            qr[k] = c[2][k] + (double)weno5_right<float>(d[1][k], 0.0f, d[3][k], d[4][k], d[5][k]);
        }
        return jump;
    }
    Roe<double> s = roe_avg(c[2], c[3], gam);
    // FP64 base R L c_2
    double base[4] = {c[2][0], c[2][1], c[2][2], c[2][3]};
    {
        const int n = dir == DIR_X ? 1 : 2;
        double kap = (s.rho - 1.0) * rcp(s.rho) * c[2][n];
        base[n] -= kap;
        base[3] -= kap * (dir == DIR_X ? s.u : s.v);
    }
    Roe<float> sf{(float)s.rho, (float)s.u, (float)s.v, (float)s.h, (float)s.c, (float)s.ic, (float)s.irho};
    const LeftRows<float> L = left_rows<float>(dir, sf, (float)gam);
    const RightCols<float> R = right_cols<float>(dir, sf);
    float w[6][4];
#pragma unroll
    for (int t = 0; t < 6; ++t) {
        if (t == 2) {
#pragma unroll
            for (int k = 0; k < 4; ++k) w[t][k] = 0.0f;
        } else {
            left_project<float>(dir, L, d[t], w[t]);
        }
    }
    float chl[4], chr[4], el[4], er[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        chl[k] = weno5_left<float>(w[0][k], w[1][k], 0.0f, w[3][k], w[4][k]);
        chr[k] = weno5_right<float>(w[1][k], 0.0f, w[3][k], w[4][k], w[5][k]);
    }
    right_project<float>(dir, R, chl, el);
    right_project<float>(dir, R, chr, er);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        ql[k] = base[k] + (double)el[k];
        qr[k] = base[k] + (double)er[k];
    }
    pp_limit(ql, c[2], gam);
    pp_limit(qr, c[3], gam);
    return jump;
}

// ---------------------------------------------------------------------
// Riemann solvers (riemann.rs:6-349).  Both use Roe-average wave speeds
// s_l = ũ_n - ã, s_r = ũ_n + ã with ã^2 floored at 0, and the guard
// |d| < 1e-12 -> copysign(1e-12, d) + 1e-12.
// ---------------------------------------------------------------------
template <typename T> struct FaceState {
    T ql0, ql1, ql2, ql3, qr0, qr1, qr2, qr3;   // floored densities
    T u_l, v_l, u_r, v_r, p_l, p_r;
    T s_l, s_r;
    T il, ir;                                   // 1 / floored densities (fast variant)
    T s_lc, s_rc;                               // HLLC's speeds (twin: no max(0, .) under the root)
};

template <typename T> WX_HD FaceState<T> face_speeds(int dir, const T ql[4], const T qr[4], T gam)
{
    const T eps = K<T>::eps;
    FaceState<T> s;
    s.ql0 = vmax(ql[0], eps); s.ql1 = ql[1]; s.ql2 = ql[2]; s.ql3 = ql[3];
    s.qr0 = vmax(qr[0], eps); s.qr1 = qr[1]; s.qr2 = qr[2]; s.qr3 = qr[3];
#if WX_FAST
    RoeW<T> rw = roe_weights(s.ql0, s.qr0);
    T il = rw.ia, ir = rw.ib;
    s.il = il;
    s.ir = ir;
    s.u_l = s.ql1 * il; s.v_l = s.ql2 * il;
    s.u_r = s.qr1 * ir; s.v_r = s.qr2 * ir;
#else
    s.u_l = s.ql1 / s.ql0; s.v_l = s.ql2 / s.ql0;
    s.u_r = s.qr1 / s.qr0; s.v_r = s.qr2 / s.qr0;
#endif
    T p_l = (gam - T(1)) * (s.ql3 - T(0.5) * s.ql0 * (s.u_l * s.u_l + s.v_l * s.v_l));
    T p_r = (gam - T(1)) * (s.qr3 - T(0.5) * s.qr0 * (s.u_r * s.u_r + s.v_r * s.v_r));
    s.p_l = vmax(p_l, eps);
    s.p_r = vmax(p_r, eps);
#if WX_FAST
    T h_l = (s.ql3 + s.p_l) * il;
    T h_r = (s.qr3 + s.p_r) * ir;
    T idr = rcp(T(1) + rw.k);
    T ut = (s.u_l + rw.k * s.u_r) * idr;
    T vt = (s.v_l + rw.k * s.v_r) * idr;
    T ht = (h_l + rw.k * h_r) * idr;
    T tmp = ht - T(0.5) * (ut * ut + vt * vt);
    T at = sqrt_z(vmax((gam - T(1)) * tmp, T(0)));
#else
    T sl = sqrt(s.ql0), sr = sqrt(s.qr0);
    T dr = sl + sr;
    T h_l = (s.ql3 + s.p_l) / s.ql0;
    T h_r = (s.qr3 + s.p_r) / s.qr0;
    T ut = (sl * s.u_l + sr * s.u_r) / dr;
    T vt = (sl * s.v_l + sr * s.v_r) / dr;
    T ht = (sl * h_l + sr * h_r) / dr;
    T tmp = ht - T(0.5) * (ut * ut + vt * vt);
    T at = sqrt(vmax((gam - T(1)) * tmp, T(0)));
#endif
    T wt = dir == DIR_X ? ut : vt;
    s.s_l = wt - at;
    s.s_r = wt + at;
#if WX_TWIN
    const T atc = sqrt((gam - T(1)) * tmp);
    s.s_lc = wt - atc;
    s.s_rc = wt + atc;
#else
    s.s_lc = s.s_l;
    s.s_rc = s.s_r;
#endif
    return s;
}

template <typename T> WX_HD T guard_denom(T d)
{
    const T eps = K<T>::eps;
    if (fabs(d) < eps) d = copysign(eps, d) + eps;
    return d;
}

#if WX_TWIN
// WENO_HLLC_2d.py:435-574: HLLC of the NumPy twin (see the file header).
template <typename T> WX_HD void hllc_from(int dir, const FaceState<T>& s, const T ql[4], const T fl[4],
                                           const T qr[4], const T fr[4], T out[4])
{
    const T wl = dir == DIR_X ? s.u_l : s.v_l;
    const T wr = dir == DIR_X ? s.u_r : s.v_r;
    const T sl = s.s_lc, sr = s.s_rc;
    T den = s.ql0 * (sl - wl) - s.qr0 * (sr - wr);
    // numpy's guard: eps * sign(d) + eps, sign(+-0) = 0
    if (fabs(den) < K<T>::eps) den = (den > T(0) ? K<T>::eps : (den < T(0) ? -K<T>::eps : T(0))) + K<T>::eps;
    const T s_st = (s.p_r - s.p_l + s.ql0 * wl * (sl - wl) - s.qr0 * wr * (sr - wr)) / den;
#pragma unroll
    for (int k = 0; k < 4; ++k) out[k] = T(NAN);   // no region: the twin leaves F unset
    if (sl >= T(0)) {
#pragma unroll
        for (int k = 0; k < 4; ++k) out[k] = fl[k];
    }
    if (sl < T(0) && s_st >= T(0)) {
        T cf = s.ql0 * (sl - wl) / (sl - s_st);
        T e = s.ql3 / s.ql0 + (s_st - wl) * (s_st + s.p_l / (s.ql0 * (sl - wl)));
        T st[4] = {cf, dir == DIR_X ? cf * s_st : cf * s.u_l, dir == DIR_X ? cf * s.v_l : cf * s_st, cf * e};
#pragma unroll
        for (int k = 0; k < 4; ++k) out[k] = fl[k] + sl * (st[k] - ql[k]);
    }
    if (s_st < T(0) && sr > T(0)) {
        T cf = s.qr0 * (s_st - wr) / (sr - s_st);   // the twin's right-star coefficient
        T e = s.qr3 / s.qr0 + (s_st - wr) * (s_st + s.p_r / (s.qr0 * (sr - wr)));
        T st[4] = {cf, dir == DIR_X ? cf * s_st : cf * s.u_r, dir == DIR_X ? cf * s.v_r : cf * s_st, cf * e};
#pragma unroll
        for (int k = 0; k < 4; ++k) out[k] = fr[k] + sr * (st[k] - qr[k]);
    }
    if (sr <= T(0)) {
#pragma unroll
        for (int k = 0; k < 4; ++k) out[k] = fr[k];
    }
}
#else
// riemann.rs:6-96 (x) / :99-189 (y): HLLC.  ql/qr are the raw (unfloored)
// face states; the star-state difference uses them as given.
template <typename T> WX_HD void hllc_from(int dir, const FaceState<T>& s, const T ql[4], const T fl[4],
                                           const T qr[4], const T fr[4], T out[4])
{
    T wl = dir == DIR_X ? s.u_l : s.v_l;
    T wr = dir == DIR_X ? s.u_r : s.v_r;
    T den = guard_denom(s.ql0 * (s.s_l - wl) - s.qr0 * (s.s_r - wr));
#if WX_FAST
    // reciprocals keep the quotients' zero-divisor results (inf / NaN)
#define WX_QUOT(n, d) ((n) * rcp_z(d))
#define WX_OVER_RHO(n, side) ((n) * (side == 0 ? s.il : s.ir))
#define WX_P_TERM(p, rho, irho, sw) ((p) * (irho) * rcp_z(sw))
#else
#define WX_QUOT(n, d) ((n) / (d))
#define WX_OVER_RHO(n, side) ((n) / (side == 0 ? s.ql0 : s.qr0))
#define WX_P_TERM(p, rho, irho, sw) ((p) / ((rho) * (sw)))
#endif
    T s_st = WX_QUOT(s.p_r - s.p_l + s.ql0 * wl * (s.s_l - wl) - s.qr0 * wr * (s.s_r - wr), den);
    if (s.s_l >= T(0)) {
#pragma unroll
        for (int k = 0; k < 4; ++k) out[k] = fl[k];
    } else if (s.s_l < T(0) && s_st >= T(0)) {
        T cf = WX_QUOT(s.ql0 * (s.s_l - wl), s.s_l - s_st);
        T e = WX_OVER_RHO(s.ql3, 0) + (s_st - wl) * (s_st + WX_P_TERM(s.p_l, s.ql0, s.il, s.s_l - wl));
        T st[4];
        st[0] = cf;
        st[1] = dir == DIR_X ? cf * s_st : cf * s.u_l;
        st[2] = dir == DIR_X ? cf * s.v_l : cf * s_st;
        st[3] = cf * e;
#pragma unroll
        for (int k = 0; k < 4; ++k) out[k] = fl[k] + s.s_l * (st[k] - ql[k]);
    } else if (s_st < T(0) && s.s_r > T(0)) {
        T cf = WX_QUOT(s.qr0 * (s.s_r - wr), s.s_r - s_st);
        T e = WX_OVER_RHO(s.qr3, 1) + (s_st - wr) * (s_st + WX_P_TERM(s.p_r, s.qr0, s.ir, s.s_r - wr));
        T st[4];
        st[0] = cf;
        st[1] = dir == DIR_X ? cf * s_st : cf * s.u_r;
        st[2] = dir == DIR_X ? cf * s.v_r : cf * s_st;
        st[3] = cf * e;
#pragma unroll
        for (int k = 0; k < 4; ++k) out[k] = fr[k] + s.s_r * (st[k] - qr[k]);
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) out[k] = fr[k];
    }
#undef WX_QUOT
#undef WX_OVER_RHO
#undef WX_P_TERM
}

#endif

// riemann.rs:191-269 (x) / :271-349 (y): HLLE with floored densities.
// ql / qr are the raw face states (used by the twin build only).
template <typename T> WX_HD void hlle_from(const FaceState<T>& s, const T ql[4], const T fl[4], const T qr[4],
                                           const T fr[4], T out[4])
{
#if WX_TWIN
    // WENO_HLLC_2d.py:576-669: masks SL >= 0, then SR <= 0 (assigned later,
    // so it wins where both hold), middle over the raw SR - SL with the
    // unfloored states
    if (s.s_r <= T(0)) {
#pragma unroll
        for (int k = 0; k < 4; ++k) out[k] = fr[k];
    } else if (s.s_l >= T(0)) {
#pragma unroll
        for (int k = 0; k < 4; ++k) out[k] = fl[k];
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k)
            out[k] = (s.s_r * fl[k] - s.s_l * fr[k] + s.s_r * s.s_l * (qr[k] - ql[k])) / (s.s_r - s.s_l);
    }
#else
    (void)ql;
    (void)qr;
    T den = guard_denom(s.s_r - s.s_l);
    if (s.s_l >= T(0)) {
#pragma unroll
        for (int k = 0; k < 4; ++k) out[k] = fl[k];
    } else if (s.s_r < T(0)) {
#pragma unroll
        for (int k = 0; k < 4; ++k) out[k] = fr[k];
    } else {
        const T a[4] = {s.ql0, s.ql1, s.ql2, s.ql3};
        const T b[4] = {s.qr0, s.qr1, s.qr2, s.qr3};
#pragma unroll
        for (int k = 0; k < 4; ++k)
#if WX_FAST
        {
            const T rden = rcp_z(den);   // one reciprocal for the four components
            out[k] = (s.s_r * fl[k] - s.s_l * fr[k] + s.s_r * s.s_l * (b[k] - a[k])) * rden;
        }
#else
            out[k] = (s.s_r * fl[k] - s.s_l * fr[k] + s.s_r * s.s_l * (b[k] - a[k])) / den;
#endif
    }
#endif
}

// riemann.rs:389-407  blend weight from the sensor value.  inv_span =
// 1 / (hi - lo) is used by the fast variant only.
template <typename T> WX_HD T sensor_phi(T dpdx, T dpdy, T pbar, T lo, T hi, T inv_span)
{
#if WX_FAST
    T grad = sqrt_z(dpdx * dpdx + dpdy * dpdy);
    T jp = grad * rcp(vmax(pbar, K<T>::eps));
#else
    T grad = sqrt(dpdx * dpdx + dpdy * dpdy);
    T jp = grad / vmax(pbar, K<T>::eps);
#endif
    if (jp <= lo) return T(1);
    if (jp >= hi) return T(0);
#if WX_FAST
    return T(1) - (jp - lo) * inv_span;
#else
    (void)inv_span;
    return T(1) - (jp - lo) / (hi - lo);
#endif
}

}  // namespace wx
