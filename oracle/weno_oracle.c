/*
 * weno_oracle.c -- CPU oracle for the 2D WENO5 / HLLC right-hand side.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a plain-C FP64 restatement of the
 * reference Rust crate `weno_ext` (weno_ext/src/utils.rs, weno_ext/src/weno.rs,
 * weno_ext/src/riemann.rs).  It is the checker that the GPU path is compared
 * against and the "port" CPU baseline that bench.py times.  Nothing in the
 * product path links, loads or calls it.
 *
 * Single-threaded by default, like the reference.  orc_set_threads(n) splits
 * the outer loops of the heavy passes over n OpenMP threads (a multi-core CPU
 * comparator); every output element is computed by the same arithmetic in
 * either case, so the results do not depend on n.
 *
 * Semantics that are kept on purpose (they decide bit-level parity):
 *   - same passes, loop bounds, temporaries and operation order as the Rust;
 *     compiled with -ffp-contract=off so no multiply-add is fused (Rust never
 *     contracts);
 *   - Rust f64::max / f64::min return the non-NaN operand: fmax / fmin;
 *   - x.powi(2) is x*x;
 *   - the truncated constant SQRT3_12 = 0.1443375673 (utils.rs:4).
 *
 * orc_set_twin(1) switches to the reference's pure-NumPy twin
 * WENO_HLLC_2d.py where the two differ: the exact sqrt(3)/12, and the
 * twin's HLLC and HLLE (hllc_twin_pt / hlle_twin_pt below).  Its
 * component-wise reconstruction, fluxes, sensor and blend are those of the
 * Rust crate's component-wise path operation for operation (for finite
 * inputs: numpy's maximum propagates NaN where fmax does not).
 *
 * Each function names the reference lines it follows.  Functions whose result
 * no independent check pins are listed as "parity unpinned" in the test files.
 *
 * Shapes follow the reference: u is (m, n, c>=4) channel-last with arbitrary
 * element strides; every output is a new C-contiguous array.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef ORC_FP32
/* The FP32 build (libweno_oracle_f32.so, test instrumentation): every
 * floating-point value and operation in single precision, so a test can
 * measure how far FP32 arithmetic in one stage of the scheme moves the
 * results.  Arrays passed to it are float32. */
#define double float
#define fmax fmaxf
#define fmin fminf
#define sqrt sqrtf
#define fabs fabsf
#define copysign copysignf
#endif

/* threads of the outer loops (1 = the reference's single thread) */
static int orc_nthreads = 1;
void orc_set_threads(int n) { orc_nthreads = n > 0 ? n : 1; }
#ifdef _OPENMP
#define ORC_PAR _Pragma("omp parallel for num_threads(orc_nthreads) schedule(static)")
#else
#define ORC_PAR
#endif

typedef struct {
    const double *p;
    int64_t d0, d1, d2;     /* extents */
    int64_t s0, s1, s2;     /* element strides */
} orc_view3;

#define V3(v, i, j, k) ((v)->p[(int64_t)(i) * (v)->s0 + (int64_t)(j) * (v)->s1 + (int64_t)(k) * (v)->s2])
#define C3(a, n1, i, j, k) ((a)[(((int64_t)(i) * (n1)) + (int64_t)(j)) * 4 + (k)])

/* utils.rs:3-6 */
static const double ORC_EPS = 1e-12;
static const double ORC_SQRT3_12 = 0.1443375673;
static const double ORC_RHO_MIN = 1e-12;
static const double ORC_P_MIN = 1e-12;

/* Twin semantics (orc_set_twin(1)): the reference's pure-NumPy twin
 * WENO_HLLC_2d.py where it differs from the Rust crate -- the exact
 * np.sqrt(3)/12 in the Gauss-point pass (:5), and its own HLLC (:435-574)
 * and HLLE (:576-669).  Off by default. */
static int orc_twin = 0;
static double orc_sqrt3_12 = 0.1443375673;
void orc_set_twin(int on)
{
    orc_twin = on != 0;
    orc_sqrt3_12 = orc_twin ? sqrt(3.0) / 12.0 : ORC_SQRT3_12;
}

static inline double sq(double x) { return x * x; }

static double *zalloc(int64_t count)
{
    if (count <= 0) return (double *)calloc(1, sizeof(double));
    return (double *)calloc((size_t)count, sizeof(double));
}

/* ------------------------------------------------------------------ */
/* utils.rs:74-108  physical fluxes                                    */
/* ------------------------------------------------------------------ */
static inline void flux_x_pt(const double q[4], double g, double f[4])
{
    double rho = fmax(q[0], ORC_EPS);
    double u = q[1] / rho;
    double v = q[2] / rho;
    double p = (g - 1.0) * (q[3] - 0.5 * rho * (u * u + v * v));
    f[0] = q[1];
    f[1] = q[1] * u + p;
    f[2] = q[2] * u;
    f[3] = u * (q[3] + p);
}

static inline void flux_y_pt(const double q[4], double g, double f[4])
{
    double rho = fmax(q[0], ORC_EPS);
    double u = q[1] / rho;
    double v = q[2] / rho;
    double p = (g - 1.0) * (q[3] - 0.5 * rho * (u * u + v * v));
    f[0] = q[2];
    f[1] = q[1] * v;
    f[2] = q[2] * v + p;
    f[3] = v * (q[3] + p);
}

void orc_flux(const orc_view3 *q, double gamma, int dir, double *out)
{
    for (int64_t i = 0; i < q->d0; ++i)
        for (int64_t j = 0; j < q->d1; ++j) {
            double c[4], f[4];
            for (int k = 0; k < 4; ++k) c[k] = V3(q, i, j, k);
            if (dir == 0) flux_x_pt(c, gamma, f); else flux_y_pt(c, gamma, f);
            for (int k = 0; k < 4; ++k) C3(out, q->d1, i, j, k) = f[k];
        }
}

/* utils.rs:9-35  conservative -> primitive (p, a, rho_safe, u, v, h) */
void orc_con2primi(const orc_view3 *q, double gamma, double *p, double *a,
                   double *rho_s, double *u, double *v, double *h)
{
    ORC_PAR
    for (int64_t i = 0; i < q->d0; ++i)
        for (int64_t j = 0; j < q->d1; ++j) {
            int64_t o = i * q->d1 + j;
            double rho = fmax(V3(q, i, j, 0), ORC_EPS);
            rho_s[o] = rho;
            double uu = V3(q, i, j, 1) / rho;
            double vv = V3(q, i, j, 2) / rho;
            u[o] = uu;
            v[o] = vv;
            double pp = fmax((gamma - 1.0) * (V3(q, i, j, 3) - 0.5 * rho * (uu * uu + vv * vv)), ORC_EPS);
            p[o] = pp;
            a[o] = sqrt(gamma * pp / rho);
            h[o] = (V3(q, i, j, 3) + pp) / rho;
        }
}

/* ------------------------------------------------------------------ */
/* weno.rs:13-63  Roe average used by the characteristic projection    */
/* ------------------------------------------------------------------ */
static void roe_avg(const double ul[4], const double ur[4], double g,
                    double *rho_o, double *u_o, double *v_o, double *h_o, double *c_o)
{
    double rho_l = fmax(ul[0], ORC_RHO_MIN);
    double rho_r = fmax(ur[0], ORC_RHO_MIN);
    double sl = sqrt(rho_l);
    double sr = sqrt(rho_r);
    double den = sl + sr;
    double u_l = ul[1] / rho_l, v_l = ul[2] / rho_l;
    double u_r = ur[1] / rho_r, v_r = ur[2] / rho_r;
    double p_l = fmax((g - 1.0) * (ul[3] - 0.5 * rho_l * (u_l * u_l + v_l * v_l)), ORC_P_MIN);
    double p_r = fmax((g - 1.0) * (ur[3] - 0.5 * rho_r * (u_r * u_r + v_r * v_r)), ORC_P_MIN);
    double h_l = (ul[3] + p_l) / rho_l;
    double h_r = (ur[3] + p_r) / rho_r;
    double u_roe = (sl * u_l + sr * u_r) / den;
    double v_roe = (sl * v_l + sr * v_r) / den;
    double h_roe = (sl * h_l + sr * h_r) / den;
    double kin = 0.5 * (u_roe * u_roe + v_roe * v_roe);
    double c2 = fmax((g - 1.0) * (h_roe - kin), ORC_P_MIN / fmin(rho_l, rho_r));
    *c_o = sqrt(c2);
    *rho_o = sl * sr;
    *u_o = u_roe;
    *v_o = v_roe;
    *h_o = h_roe;
}

/* weno.rs:76-189  left / right eigenvector matrices; dir 0 = x, 1 = y */
static void eig_left(int dir, double rho, double u, double v, double c, double g, double L[4][4])
{
    double g1 = g - 1.0;
    double kin = 0.5 * (u * u + v * v);
    double rc = rho * c;
    double inv_rc = 1.0 / rc;
    double inv_c2 = 1.0 / (c * c);
    double un = dir == 0 ? u : v;
    L[0][0] = g1 * kin / (2.0 * c * c) + un / (2.0 * c);
    L[3][0] = g1 * kin / (2.0 * c * c) - un / (2.0 * c);
    L[1][0] = 1.0 - g1 * kin * inv_c2;
    L[1][1] = g1 * u * inv_c2;
    L[1][2] = g1 * v * inv_c2;
    L[1][3] = -g1 * inv_c2;
    L[0][3] = g1 * inv_c2 / 2.0;
    L[3][3] = g1 * inv_c2 / 2.0;
    if (dir == 0) {
        L[0][1] = -g1 * u * inv_c2 / 2.0 - 0.5 * inv_rc;
        L[0][2] = -g1 * v * inv_c2 / 2.0;
        L[3][1] = -g1 * u * inv_c2 / 2.0 + 0.5 * inv_rc;
        L[3][2] = -g1 * v * inv_c2 / 2.0;
        L[2][0] = -v / rho; L[2][1] = 0.0; L[2][2] = 1.0 / rho; L[2][3] = 0.0;
    } else {
        L[0][1] = -g1 * u * inv_c2 / 2.0;
        L[0][2] = -g1 * v * inv_c2 / 2.0 - 0.5 * inv_rc;
        L[3][1] = -g1 * u * inv_c2 / 2.0;
        L[3][2] = -g1 * v * inv_c2 / 2.0 + 0.5 * inv_rc;
        L[2][0] = -u / rho; L[2][1] = 1.0 / rho; L[2][2] = 0.0; L[2][3] = 0.0;
    }
}

static void eig_right(int dir, double rho, double u, double v, double h, double c, double R[4][4])
{
    double kin = 0.5 * (u * u + v * v);
    R[0][0] = 1.0; R[0][1] = 1.0; R[0][2] = 0.0; R[0][3] = 1.0;
    if (dir == 0) {
        R[1][0] = u - c; R[1][1] = u; R[1][2] = 0.0; R[1][3] = u + c;
        R[2][0] = v; R[2][1] = v; R[2][2] = rho; R[2][3] = v;
        R[3][0] = h - u * c; R[3][1] = kin; R[3][2] = rho * v; R[3][3] = h + u * c;
    } else {
        R[1][0] = u; R[1][1] = u; R[1][2] = rho; R[1][3] = u;
        R[2][0] = v - c; R[2][1] = v; R[2][2] = 0.0; R[2][3] = v + c;
        R[3][0] = h - v * c; R[3][1] = kin; R[3][2] = rho * u; R[3][3] = h + v * c;
    }
}

/* weno.rs:197-205  out[r] accumulates from 0.0 in column order */
static void mat4_vec4(const double M[4][4], const double x[4], double y[4])
{
    for (int r = 0; r < 4; ++r) {
        double acc = 0.0;
        for (int c = 0; c < 4; ++c) acc += M[r][c] * x[c];
        y[r] = acc;
    }
}

/* weno.rs:214-233  WENO5-JS value at the right face of the centre cell u2 */
static double weno5_left(double u0, double u1, double u2, double u3, double u4)
{
    double b0 = 13.0 / 12.0 * sq(u2 - 2.0 * u3 + u4) + 0.25 * sq(3.0 * u2 - 4.0 * u3 + u4);
    double b1 = 13.0 / 12.0 * sq(u1 - 2.0 * u2 + u3) + 0.25 * sq(u1 - u3);
    double b2 = 13.0 / 12.0 * sq(u0 - 2.0 * u1 + u2) + 0.25 * sq(u0 - 4.0 * u1 + 3.0 * u2);
    double a0 = 0.3 / sq(ORC_EPS + b0);
    double a1 = 0.6 / sq(ORC_EPS + b1);
    double a2 = 0.1 / sq(ORC_EPS + b2);
    double s = a0 + a1 + a2;
    double w0 = a0 / s, w1 = a1 / s, w2 = a2 / s;
    double p0 = -u4 + 5.0 * u3 + 2.0 * u2;
    double p1 = -u1 + 5.0 * u2 + 2.0 * u3;
    double p2 = 2.0 * u0 - 7.0 * u1 + 11.0 * u2;
    return (w0 * p0 + w1 * p1 + w2 * p2) / 6.0;
}

/* weno.rs:238-257  WENO5-JS value at the left face of the centre cell u3 */
static double weno5_right(double u1, double u2, double u3, double u4, double u5)
{
    double b0 = 13.0 / 12.0 * sq(u3 - 2.0 * u4 + u5) + 0.25 * sq(3.0 * u3 - 4.0 * u4 + u5);
    double b1 = 13.0 / 12.0 * sq(u2 - 2.0 * u3 + u4) + 0.25 * sq(u2 - u4);
    double b2 = 13.0 / 12.0 * sq(u1 - 2.0 * u2 + u3) + 0.25 * sq(u1 - 4.0 * u2 + 3.0 * u3);
    double a0 = 0.1 / sq(ORC_EPS + b0);
    double a1 = 0.6 / sq(ORC_EPS + b1);
    double a2 = 0.3 / sq(ORC_EPS + b2);
    double s = a0 + a1 + a2;
    double w0 = a0 / s, w1 = a1 / s, w2 = a2 / s;
    double p0 = 2.0 * u5 - 7.0 * u4 + 11.0 * u3;
    double p1 = -u4 + 5.0 * u3 + 2.0 * u2;
    double p2 = -u1 + 5.0 * u2 + 2.0 * u3;
    return (w0 * p0 + w1 * p1 + w2 * p2) / 6.0;
}

/* weno.rs:262-303  2-point Gauss values along the face (a: -, b: + point) */
static void weno5_gq(double q0, double q1, double q2, double q3, double q4, double *ga, double *gb)
{
    double b0 = 13.0 / 12.0 * sq(q2 - 2.0 * q3 + q4) + 0.25 * sq(3.0 * q2 - 4.0 * q3 + q4);
    double b1 = 13.0 / 12.0 * sq(q1 - 2.0 * q2 + q3) + 0.25 * sq(q1 - q3);
    double b2 = 13.0 / 12.0 * sq(q0 - 2.0 * q1 + q2) + 0.25 * sq(q0 - 4.0 * q1 + 3.0 * q2);
    {
        double a0 = 0.1928406937 / sq(ORC_EPS + b0);
        double a1 = 0.6111111111 / sq(ORC_EPS + b1);
        double a2 = 0.1960481952 / sq(ORC_EPS + b2);
        double s = a0 + a1 + a2;
        double w0 = a0 / s, w1 = a1 / s, w2 = a2 / s;
        double p0 = q2 + (3.0 * q2 - 4.0 * q3 + q4) * orc_sqrt3_12;
        double p1 = q2 - (-q1 + q3) * orc_sqrt3_12;
        double p2 = q2 - (3.0 * q2 - 4.0 * q1 + q0) * orc_sqrt3_12;
        *ga = w0 * p0 + w1 * p1 + w2 * p2;
    }
    {
        double a0 = 0.1960481952 / sq(ORC_EPS + b0);
        double a1 = 0.6111111111 / sq(ORC_EPS + b1);
        double a2 = 0.1928406937 / sq(ORC_EPS + b2);
        double s = a0 + a1 + a2;
        double w0 = a0 / s, w1 = a1 / s, w2 = a2 / s;
        double p0 = q2 - (3.0 * q2 - 4.0 * q3 + q4) * orc_sqrt3_12;
        double p1 = q2 - (q1 - q3) * orc_sqrt3_12;
        double p2 = q2 + (3.0 * q2 - 4.0 * q1 + q0) * orc_sqrt3_12;
        *gb = w0 * p0 + w1 * p1 + w2 * p2;
    }
}

/* weno.rs:312-315 */
static double pressure_pp(const double q[4], double g)
{
    if (q[0] <= 0.0) return 0.0;
    return (g - 1.0) * (q[3] - 0.5 * (q[1] * q[1] + q[2] * q[2]) / q[0]);
}

/* faces where the limiter's density / pressure pass acted since the last
   orc_pp_reset() (test instrumentation: shows that a case exercises them) */
static long orc_pp_fired[2];
void orc_pp_reset(void) { orc_pp_fired[0] = orc_pp_fired[1] = 0; }
void orc_pp_counts(long out[2]) { out[0] = orc_pp_fired[0]; out[1] = orc_pp_fired[1]; }

/* weno.rs:327-352  density pass, then pressure pass, toward the cell average */
static void pp_limit(double q[4], const double ub[4], double g)
{
    if (q[0] < ORC_RHO_MIN) {
#ifdef _OPENMP
#pragma omp atomic
#endif
        orc_pp_fired[0]++;
        double rho_bar = fmax(ub[0], ORC_RHO_MIN);
        double th = fmax(fmin((rho_bar - ORC_RHO_MIN) / (rho_bar - q[0]), 1.0), 0.0);
        for (int k = 0; k < 4; ++k) q[k] = th * (q[k] - ub[k]) + ub[k];
    }
    double pq = pressure_pp(q, g);
    if (pq < ORC_P_MIN) {
#ifdef _OPENMP
#pragma omp atomic
#endif
        orc_pp_fired[1]++;
        double p_bar = fmax(pressure_pp(ub, g), ORC_P_MIN);
        double th = fmax(fmin((p_bar - ORC_P_MIN) / (p_bar - pq), 1.0), 0.0);
        for (int k = 0; k < 4; ++k) q[k] = th * (q[k] - ub[k]) + ub[k];
    }
}

/* ------------------------------------------------------------------ */
/* Directional reconstruction.                                         */
/*   dir 0: weno.rs:367-502 (characteristic) / :635-852 (component)    */
/*   dir 1: weno.rs:508-632 (characteristic) / :854-1071 (component)   */
/* Output arrays (q_l, f_l, q_r, f_r) are C-contiguous (F1, F2, 4):    */
/*   dir 0: (m-5, n-6); dir 1: (m-6, n-5).                              */
/* ------------------------------------------------------------------ */
void orc_weno(const orc_view3 *u, double gamma, int dir, int cons,
              double *ql_o, double *fl_o, double *qr_o, double *fr_o)
{
    const int64_t m = u->d0, n = u->d1;
    /* copy of the first four channels (weno.rs:374-379) */
    double *uc = zalloc(m * n * 4);
    for (int64_t i = 0; i < m; ++i)
        for (int64_t j = 0; j < n; ++j)
            for (int k = 0; k < 4; ++k) C3(uc, n, i, j, k) = V3(u, i, j, k);

    /* pass 1: normal-direction reconstruction */
    const int64_t om = dir == 0 ? m - 5 : m;
    const int64_t on = dir == 0 ? n : n - 5;
    double *pl = zalloc(om * on * 4);
    double *pr = zalloc(om * on * 4);
    ORC_PAR
    for (int64_t i = 0; i < om; ++i)
        for (int64_t j = 0; j < on; ++j) {
            double cell[6][4];
            for (int s = 0; s < 6; ++s)
                for (int k = 0; k < 4; ++k)
                    cell[s][k] = dir == 0 ? C3(uc, n, i + s, j, k) : C3(uc, n, i, j + s, k);
            double cl[4], cr[4];
            if (cons) {
                for (int k = 0; k < 4; ++k) {
                    cl[k] = weno5_left(cell[0][k], cell[1][k], cell[2][k], cell[3][k], cell[4][k]);
                    cr[k] = weno5_right(cell[1][k], cell[2][k], cell[3][k], cell[4][k], cell[5][k]);
                }
            } else {
                double rho, uu, vv, hh, cc;
                double L[4][4], R[4][4], w[6][4], chl[4], chr[4];
                /* left-state face: Roe basis of cells 2|3 */
                roe_avg(cell[2], cell[3], gamma, &rho, &uu, &vv, &hh, &cc);
                eig_left(dir, rho, uu, vv, cc, gamma, L);
                eig_right(dir, rho, uu, vv, hh, cc, R);
                for (int s = 0; s < 5; ++s) mat4_vec4(L, cell[s], w[s]);
                for (int k = 0; k < 4; ++k)
                    chl[k] = weno5_left(w[0][k], w[1][k], w[2][k], w[3][k], w[4][k]);
                mat4_vec4(R, chl, cl);
                /* right-state face: the reference recomputes the same basis (weno.rs:428-431) */
                roe_avg(cell[2], cell[3], gamma, &rho, &uu, &vv, &hh, &cc);
                eig_left(dir, rho, uu, vv, cc, gamma, L);
                eig_right(dir, rho, uu, vv, hh, cc, R);
                for (int s = 0; s < 5; ++s) mat4_vec4(L, cell[s + 1], w[s]);
                for (int k = 0; k < 4; ++k)
                    chr[k] = weno5_right(w[0][k], w[1][k], w[2][k], w[3][k], w[4][k]);
                mat4_vec4(R, chr, cr);
                pp_limit(cl, cell[2], gamma);
                pp_limit(cr, cell[3], gamma);
            }
            for (int k = 0; k < 4; ++k) {
                C3(pl, on, i, j, k) = cl[k];
                C3(pr, on, i, j, k) = cr[k];
            }
        }

    /* pass 2: transverse 2-point Gauss-quadrature WENO (weno.rs:460-490 / :590-620) */
    const int64_t m1 = dir == 0 ? om : om - 6;
    const int64_t n1 = dir == 0 ? on - 6 : on;
    double *l1 = zalloc(m1 * n1 * 4), *l2 = zalloc(m1 * n1 * 4);
    double *r1 = zalloc(m1 * n1 * 4), *r2 = zalloc(m1 * n1 * 4);
    ORC_PAR
    for (int64_t i = 0; i < m1; ++i)
        for (int64_t j = 0; j < n1; ++j)
            for (int k = 0; k < 4; ++k) {
                double a[5], b[5];
                for (int s = 0; s < 5; ++s) {
                    a[s] = dir == 0 ? C3(pl, on, i, j + 1 + s, k) : C3(pl, on, i + 1 + s, j, k);
                    b[s] = dir == 0 ? C3(pr, on, i, j + 1 + s, k) : C3(pr, on, i + 1 + s, j, k);
                }
                weno5_gq(a[0], a[1], a[2], a[3], a[4], &C3(l1, n1, i, j, k), &C3(l2, n1, i, j, k));
                weno5_gq(b[0], b[1], b[2], b[3], b[4], &C3(r1, n1, i, j, k), &C3(r2, n1, i, j, k));
            }

    /* fluxes at the Gauss points, averaged (weno.rs:492-501 / :622-631) */
    ORC_PAR
    for (int64_t i = 0; i < m1; ++i)
        for (int64_t j = 0; j < n1; ++j) {
            double fa[4], fb[4], fc[4], fd[4];
            double *ql1 = &C3(l1, n1, i, j, 0), *ql2 = &C3(l2, n1, i, j, 0);
            double *qr1 = &C3(r1, n1, i, j, 0), *qr2 = &C3(r2, n1, i, j, 0);
            if (dir == 0) {
                flux_x_pt(ql1, gamma, fa); flux_x_pt(ql2, gamma, fb);
                flux_x_pt(qr1, gamma, fc); flux_x_pt(qr2, gamma, fd);
            } else {
                flux_y_pt(ql1, gamma, fa); flux_y_pt(ql2, gamma, fb);
                flux_y_pt(qr1, gamma, fc); flux_y_pt(qr2, gamma, fd);
            }
            for (int k = 0; k < 4; ++k) {
                C3(fl_o, n1, i, j, k) = 0.5 * (fa[k] + fb[k]);
                C3(fr_o, n1, i, j, k) = 0.5 * (fd[k] + fc[k]);
                C3(ql_o, n1, i, j, k) = 0.5 * (ql1[k] + ql2[k]);
                C3(qr_o, n1, i, j, k) = 0.5 * (qr1[k] + qr2[k]);
            }
        }
    free(uc); free(pl); free(pr); free(l1); free(l2); free(r1); free(r2);
}

/* ------------------------------------------------------------------ */
/* Twin semantics: WENO_HLLC_2d.py's Riemann solvers, one face.        */
/* ------------------------------------------------------------------ */
/* con2primi (WENO_HLLC_2d.py:400-432) of one face state plus the
 * enthalpy the solvers form from it (:445-446) */
typedef struct { double rho, u, v, p, h; } orc_twin_prim;

static orc_twin_prim twin_prim(const double q[4], double g)
{
    orc_twin_prim s;
    s.rho = fmax(q[0], ORC_EPS);
    s.u = q[1] / s.rho;
    s.v = q[2] / s.rho;
    s.p = fmax((g - 1.0) * (q[3] - 0.5 * s.rho * (s.u * s.u + s.v * s.v)), ORC_EPS);
    s.h = (q[3] + s.p) / s.rho;
    return s;
}

/* Roe-weighted velocity and enthalpy (:450-452 / :524-526); the sound
 * speed is left to the caller (HLLC takes the root unguarded, HLLE of
 * max(0, .)) */
static void twin_roe(const orc_twin_prim *l, const orc_twin_prim *r, double *ut, double *vt, double *ht)
{
    double a = sqrt(l->rho), b = sqrt(r->rho);
    *ut = (a * l->u + b * r->u) / (a + b);
    *vt = (a * l->v + b * r->v) / (a + b);
    *ht = (a * l->h + b * r->h) / (a + b);
}

/* numpy guard of a denominator: eps * sign(d) + eps where |d| < eps, with
 * sign(+-0) = 0 (so a small negative d becomes 0) */
static double twin_guard(double d)
{
    if (fabs(d) < ORC_EPS) d = ORC_EPS * (d > 0.0 ? 1.0 : (d < 0.0 ? -1.0 : 0.0)) + ORC_EPS;
    return d;
}

/* HLLC_x / HLLC_y (:435-508 / :510-574).  The star states use the twin's
 * coefficients -- right: rho_R (S* - w_R) / (S_R - S*) -- and the four
 * masks are assigned in order, later ones overwriting earlier ones.  Where
 * no mask holds (NaN wave speeds) the twin leaves np.empty memory; NaN here. */
static void hllc_twin_pt(int dir, const double ql[4], const double fl[4], const double qr[4],
                         const double fr[4], double g, double out[4])
{
    orc_twin_prim L = twin_prim(ql, g), R = twin_prim(qr, g);
    double ut, vt, ht;
    twin_roe(&L, &R, &ut, &vt, &ht);
    double at = sqrt((g - 1.0) * (ht - 0.5 * (ut * ut + vt * vt)));
    double wl = dir == 0 ? L.u : L.v, wr = dir == 0 ? R.u : R.v;
    double SL = (dir == 0 ? ut : vt) - at, SR = (dir == 0 ? ut : vt) + at;
    double den = twin_guard(L.rho * (SL - wl) - R.rho * (SR - wr));
    double Ss = (R.p - L.p + L.rho * wl * (SL - wl) - R.rho * wr * (SR - wr)) / den;
    double cl = L.rho * (SL - wl) / (SL - Ss);
    double el = ql[3] / L.rho + (Ss - wl) * (Ss + L.p / (L.rho * (SL - wl)));
    double cr = R.rho * (Ss - wr) / (SR - Ss);
    double er = qr[3] / R.rho + (Ss - wr) * (Ss + R.p / (R.rho * (SR - wr)));
    double sl[4], sr[4];
    sl[0] = cl;
    sl[1] = dir == 0 ? cl * Ss : cl * L.u;
    sl[2] = dir == 0 ? cl * L.v : cl * Ss;
    sl[3] = cl * el;
    sr[0] = cr;
    sr[1] = dir == 0 ? cr * Ss : cr * R.u;
    sr[2] = dir == 0 ? cr * R.v : cr * Ss;
    sr[3] = cr * er;
    for (int k = 0; k < 4; ++k) out[k] = NAN;
    if (SL >= 0.0)
        for (int k = 0; k < 4; ++k) out[k] = fl[k];
    if (SL < 0.0 && Ss >= 0.0)
        for (int k = 0; k < 4; ++k) out[k] = fl[k] + SL * (sl[k] - ql[k]);
    if (Ss < 0.0 && SR > 0.0)
        for (int k = 0; k < 4; ++k) out[k] = fr[k] + SR * (sr[k] - qr[k]);
    if (SR <= 0.0)
        for (int k = 0; k < 4; ++k) out[k] = fr[k];
}

/* HLLE_x / HLLE_y (:576-623 / :625-669): masks SL >= 0 then SR <= 0; the
 * middle state divides by the raw SR - SL (the guarded denominator of
 * :602-603 is never used) and differences the unfloored face states */
static void hlle_twin_pt(int dir, const double ql[4], const double fl[4], const double qr[4],
                         const double fr[4], double g, double out[4])
{
    orc_twin_prim L = twin_prim(ql, g), R = twin_prim(qr, g);
    double ut, vt, ht;
    twin_roe(&L, &R, &ut, &vt, &ht);
    double at = sqrt(fmax(0.0, (g - 1.0) * (ht - 0.5 * (ut * ut + vt * vt))));
    double SL = (dir == 0 ? ut : vt) - at, SR = (dir == 0 ? ut : vt) + at;
    int ml = SL >= 0.0, mr = SR <= 0.0;
    if (ml)
        for (int k = 0; k < 4; ++k) out[k] = fl[k];
    if (mr)
        for (int k = 0; k < 4; ++k) out[k] = fr[k];
    if (!ml && !mr)
        for (int k = 0; k < 4; ++k)
            out[k] = (SR * fl[k] - SL * fr[k] + SR * SL * (qr[k] - ql[k])) / (SR - SL);
}

/* ------------------------------------------------------------------ */
/* riemann.rs:6-189  HLLC (Toro) with Roe-average wave speeds          */
/* ------------------------------------------------------------------ */
static void hllc_pt(int dir, const double qlr[4], const double flv[4], const double qrr[4],
                    const double frv[4], double g, double out[4])
{
    if (orc_twin) { hllc_twin_pt(dir, qlr, flv, qrr, frv, g, out); return; }
    double ql0 = fmax(qlr[0], ORC_EPS), ql1 = qlr[1], ql2 = qlr[2], ql3 = qlr[3];
    double qr0 = fmax(qrr[0], ORC_EPS), qr1 = qrr[1], qr2 = qrr[2], qr3 = qrr[3];
    double u_l = ql1 / ql0, v_l = ql2 / ql0;
    double u_r = qr1 / qr0, v_r = qr2 / qr0;
    double p_l = (g - 1.0) * (ql3 - 0.5 * ql0 * (u_l * u_l + v_l * v_l));
    double p_r = (g - 1.0) * (qr3 - 0.5 * qr0 * (u_r * u_r + v_r * v_r));
    p_l = fmax(p_l, ORC_EPS);
    p_r = fmax(p_r, ORC_EPS);
    double h_l = (ql3 + p_l) / ql0;
    double h_r = (qr3 + p_r) / qr0;
    double sl_ = sqrt(ql0), sr_ = sqrt(qr0);
    double den_r = sl_ + sr_;
    double ut = (sl_ * u_l + sr_ * u_r) / den_r;
    double vt = (sl_ * v_l + sr_ * v_r) / den_r;
    double ht = (sl_ * h_l + sr_ * h_r) / den_r;
    double tmp = ht - 0.5 * (ut * ut + vt * vt);
    double at = sqrt(fmax((g - 1.0) * tmp, 0.0));
    double wn_l = dir == 0 ? u_l : v_l;   /* normal velocity */
    double wn_r = dir == 0 ? u_r : v_r;
    double wt = dir == 0 ? ut : vt;
    double s_l = wt - at, s_r = wt + at;
    double den = ql0 * (s_l - wn_l) - qr0 * (s_r - wn_r);
    if (fabs(den) < ORC_EPS) den = copysign(ORC_EPS, den) + ORC_EPS;
    double s_st = (p_r - p_l + ql0 * wn_l * (s_l - wn_l) - qr0 * wn_r * (s_r - wn_r)) / den;
    double cf_l = ql0 * (s_l - wn_l) / (s_l - s_st);
    double e_l = ql3 / ql0 + (s_st - wn_l) * (s_st + p_l / (ql0 * (s_l - wn_l)));
    double cf_r = qr0 * (s_r - wn_r) / (s_r - s_st);
    double e_r = qr3 / qr0 + (s_st - wn_r) * (s_st + p_r / (qr0 * (s_r - wn_r)));
    double st_l[4], st_r[4];
    st_l[0] = cf_l; st_r[0] = cf_r;
    if (dir == 0) {
        st_l[1] = cf_l * s_st; st_l[2] = cf_l * v_l;
        st_r[1] = cf_r * s_st; st_r[2] = cf_r * v_r;
    } else {
        st_l[1] = cf_l * u_l; st_l[2] = cf_l * s_st;
        st_r[1] = cf_r * u_r; st_r[2] = cf_r * s_st;
    }
    st_l[3] = cf_l * e_l; st_r[3] = cf_r * e_r;
    if (s_l >= 0.0) {
        for (int k = 0; k < 4; ++k) out[k] = flv[k];
    } else if (s_l < 0.0 && s_st >= 0.0) {
        for (int k = 0; k < 4; ++k) out[k] = flv[k] + s_l * (st_l[k] - qlr[k]);
    } else if (s_st < 0.0 && s_r > 0.0) {
        for (int k = 0; k < 4; ++k) out[k] = frv[k] + s_r * (st_r[k] - qrr[k]);
    } else {
        for (int k = 0; k < 4; ++k) out[k] = frv[k];
    }
}

/* riemann.rs:191-349  HLLE with the same wave speeds */
static void hlle_pt(int dir, const double qlr[4], const double flv[4], const double qrr[4],
                    const double frv[4], double g, double out[4])
{
    if (orc_twin) { hlle_twin_pt(dir, qlr, flv, qrr, frv, g, out); return; }
    double ql0 = fmax(qlr[0], ORC_EPS), ql1 = qlr[1], ql2 = qlr[2], ql3 = qlr[3];
    double qr0 = fmax(qrr[0], ORC_EPS), qr1 = qrr[1], qr2 = qrr[2], qr3 = qrr[3];
    double u_l = ql1 / ql0, v_l = ql2 / ql0;
    double u_r = qr1 / qr0, v_r = qr2 / qr0;
    double p_l = (g - 1.0) * (ql3 - 0.5 * ql0 * (u_l * u_l + v_l * v_l));
    double p_r = (g - 1.0) * (qr3 - 0.5 * qr0 * (u_r * u_r + v_r * v_r));
    p_l = fmax(p_l, ORC_EPS);
    p_r = fmax(p_r, ORC_EPS);
    double h_l = (ql3 + p_l) / ql0;
    double h_r = (qr3 + p_r) / qr0;
    double sl_ = sqrt(ql0), sr_ = sqrt(qr0);
    double den_r = sl_ + sr_;
    double ut = (sl_ * u_l + sr_ * u_r) / den_r;
    double vt = (sl_ * v_l + sr_ * v_r) / den_r;
    double ht = (sl_ * h_l + sr_ * h_r) / den_r;
    double tmp = ht - 0.5 * (ut * ut + vt * vt);
    double at = sqrt(fmax((g - 1.0) * tmp, 0.0));
    double wt = dir == 0 ? ut : vt;
    double s_l = wt - at, s_r = wt + at;
    double den = s_r - s_l;
    if (fabs(den) < ORC_EPS) den = copysign(ORC_EPS, den) + ORC_EPS;
    if (s_l >= 0.0) {
        for (int k = 0; k < 4; ++k) out[k] = flv[k];
    } else if (s_r < 0.0) {
        for (int k = 0; k < 4; ++k) out[k] = frv[k];
    } else {
        double a[4] = {ql0, ql1, ql2, ql3}, b[4] = {qr0, qr1, qr2, qr3};
        for (int k = 0; k < 4; ++k)
            out[k] = (s_r * flv[k] - s_l * frv[k] + s_r * s_l * (b[k] - a[k])) / den;
    }
}

/* kind 0: HLLC, 1: HLLE.  Output sized from q_l (riemann.rs:13-14). */
void orc_riemann(const orc_view3 *ql, const orc_view3 *fl, const orc_view3 *qr,
                 const orc_view3 *fr, double gamma, int dir, int kind, double *out)
{
    ORC_PAR
    for (int64_t i = 0; i < ql->d0; ++i)
        for (int64_t j = 0; j < ql->d1; ++j) {
            double a[4], b[4], c[4], d[4], o[4];
            for (int k = 0; k < 4; ++k) {
                a[k] = V3(ql, i, j, k); b[k] = V3(fl, i, j, k);
                c[k] = V3(qr, i, j, k); d[k] = V3(fr, i, j, k);
            }
            if (kind == 0) hllc_pt(dir, a, b, c, d, gamma, o);
            else hlle_pt(dir, a, b, c, d, gamma, o);
            for (int k = 0; k < 4; ++k) C3(out, ql->d1, i, j, k) = o[k];
        }
}

/* riemann.rs:352-410  pressure-gradient shock sensor on the face grid.
 * Requires m >= 2 and n >= 2 (the reference indexes p[1] and p[m-2]). */
void orc_sensor(const double *p, int64_t m, int64_t n, double lo, double hi, double *phi)
{
    int64_t mn = m < n ? m : n;
    double dx = 1.0 / (double)(mn > 1 ? mn : 1);
    double *dpdx = zalloc(m * n), *dpdy = zalloc(m * n);
    for (int64_t i = 1; i < m - 1; ++i)
        for (int64_t j = 0; j < n; ++j)
            dpdx[i * n + j] = (p[(i + 1) * n + j] - p[(i - 1) * n + j]) / (2.0 * dx);
    for (int64_t j = 0; j < n; ++j) {
        dpdx[j] = (p[n + j] - p[j]) / (2.0 * dx);
        dpdx[(m - 1) * n + j] = (p[(m - 1) * n + j] - p[(m - 2) * n + j]) / (2.0 * dx);
    }
    for (int64_t i = 0; i < m; ++i) {
        for (int64_t j = 1; j < n - 1; ++j)
            dpdy[i * n + j] = (p[i * n + j + 1] - p[i * n + j - 1]) / (2.0 * dx);
        dpdy[i * n] = (p[i * n + 1] - p[i * n]) / (2.0 * dx);
        dpdy[i * n + n - 1] = (p[i * n + n - 1] - p[i * n + n - 2]) / (2.0 * dx);
    }
    ORC_PAR
    for (int64_t o = 0; o < m * n; ++o) {
        double grad = sqrt(sq(dpdx[o]) + sq(dpdy[o]));
        double jp = grad / fmax(p[o], ORC_EPS);
        double w;
        if (jp <= lo) w = 1.0;
        else if (jp >= hi) w = 0.0;
        else w = 1.0 - (jp - lo) / (hi - lo);
        phi[o] = w;
    }
    free(dpdx); free(dpdy);
}

/* riemann.rs:412-486  HLLC/HLLE blend driven by the sensor */
void orc_blend(const orc_view3 *ql, const orc_view3 *fl, const orc_view3 *qr,
               const orc_view3 *fr, double gamma, int dir, double lo, double hi, double *out)
{
    const int64_t m = ql->d0, n = ql->d1;
    double *t[6], *u[6];
    for (int s = 0; s < 6; ++s) { t[s] = zalloc(m * n); u[s] = zalloc(qr->d0 * qr->d1); }
    orc_con2primi(ql, gamma, t[0], t[1], t[2], t[3], t[4], t[5]);
    orc_con2primi(qr, gamma, u[0], u[1], u[2], u[3], u[4], u[5]);
    double *pb = zalloc(m * n), *phi = zalloc(m * n);
    for (int64_t i = 0; i < m; ++i)
        for (int64_t j = 0; j < n; ++j)
            pb[i * n + j] = 0.5 * (t[0][i * n + j] + u[0][i * qr->d1 + j]);
    orc_sensor(pb, m, n, lo, hi, phi);
    double *fc = zalloc(m * n * 4), *fe = zalloc(m * n * 4);
    orc_riemann(ql, fl, qr, fr, gamma, dir, 0, fc);
    orc_riemann(ql, fl, qr, fr, gamma, dir, 1, fe);
    ORC_PAR
    for (int64_t o = 0; o < m * n; ++o) {
        double w = phi[o];
        for (int k = 0; k < 4; ++k) out[o * 4 + k] = w * fc[o * 4 + k] + (1.0 - w) * fe[o * 4 + k];
    }
    for (int s = 0; s < 6; ++s) { free(t[s]); free(u[s]); }
    free(pb); free(phi); free(fc); free(fe);
}

/* riemann.rs:489-541  L(U) = -(df/dx + dg/dy) on the (m-6, n-6) interior */
void orc_l_local(const orc_view3 *u, double dx, double gamma, int force_hlle,
                 double lo, double hi, int cons, double *out)
{
    const int64_t m = u->d0, n = u->d1;
    const int64_t ax = m - 5, bx = n - 6;   /* x-face grid */
    const int64_t ay = m - 6, by = n - 5;   /* y-face grid */
    double *xq[4], *yq[4];
    for (int s = 0; s < 4; ++s) { xq[s] = zalloc(ax * bx * 4); yq[s] = zalloc(ay * by * 4); }
    orc_weno(u, gamma, 0, cons, xq[0], xq[1], xq[2], xq[3]);
    orc_weno(u, gamma, 1, cons, yq[0], yq[1], yq[2], yq[3]);
    orc_view3 vx[4], vy[4];
    for (int s = 0; s < 4; ++s) {
        vx[s] = (orc_view3){xq[s], ax, bx, 4, bx * 4, 4, 1};
        vy[s] = (orc_view3){yq[s], ay, by, 4, by * 4, 4, 1};
    }
    double *fx = zalloc(ax * bx * 4), *fy = zalloc(ay * by * 4);
    if (force_hlle) {
        orc_riemann(&vx[0], &vx[1], &vx[2], &vx[3], gamma, 0, 1, fx);
        orc_riemann(&vy[0], &vy[1], &vy[2], &vy[3], gamma, 1, 1, fy);
    } else {
        orc_blend(&vx[0], &vx[1], &vx[2], &vx[3], gamma, 0, lo, hi, fx);
        orc_blend(&vy[0], &vy[1], &vy[2], &vy[3], gamma, 1, lo, hi, fy);
    }
    const int64_t nx = ax - 1, ny = by - 1;
    double *dfdx = zalloc(nx * bx * 4), *dgdy = zalloc(ay * ny * 4);
    ORC_PAR
    for (int64_t i = 0; i < nx; ++i)
        for (int64_t j = 0; j < bx; ++j)
            for (int k = 0; k < 4; ++k)
                C3(dfdx, bx, i, j, k) = (C3(fx, bx, i + 1, j, k) - C3(fx, bx, i, j, k)) / dx;
    for (int64_t i = 0; i < ay; ++i)
        for (int64_t j = 0; j < ny; ++j)
            for (int k = 0; k < 4; ++k)
                C3(dgdy, ny, i, j, k) = (C3(fy, by, i, j + 1, k) - C3(fy, by, i, j, k)) / dx;
    for (int64_t i = 0; i < nx; ++i)
        for (int64_t j = 0; j < ny; ++j)
            for (int k = 0; k < 4; ++k)
                C3(out, ny, i, j, k) = -(C3(dfdx, bx, i, j, k) + C3(dgdy, ny, i, j, k));
    for (int s = 0; s < 4; ++s) { free(xq[s]); free(yq[s]); }
    free(fx); free(fy); free(dfdx); free(dgdy);
}

/* ------------------------------------------------------------------ */
/* Test hooks: the point functions above, exposed for the oracle's own */
/* invariant tests (eigenvectors, Roe property, limiter, WENO5/GQ).    */
/* ------------------------------------------------------------------ */
void orc_test_roe_eig(int dir, const double ul[4], const double ur[4], double gamma,
                      double roe5[5], double L[16], double R[16])
{
    double rho, u, v, h, c;
    roe_avg(ul, ur, gamma, &rho, &u, &v, &h, &c);
    double LL[4][4], RR[4][4];
    eig_left(dir, rho, u, v, c, gamma, LL);
    eig_right(dir, rho, u, v, h, c, RR);
    roe5[0] = rho; roe5[1] = u; roe5[2] = v; roe5[3] = h; roe5[4] = c;
    for (int r = 0; r < 4; ++r)
        for (int k = 0; k < 4; ++k) { L[r * 4 + k] = LL[r][k]; R[r * 4 + k] = RR[r][k]; }
}

void orc_test_pp_limit(double q[4], const double ub[4], double gamma) { pp_limit(q, ub, gamma); }

void orc_test_weno5(const double u[6], double *left, double *right)
{
    *left = weno5_left(u[0], u[1], u[2], u[3], u[4]);
    *right = weno5_right(u[1], u[2], u[3], u[4], u[5]);
}

void orc_test_gq(const double q[5], double *ga, double *gb)
{
    weno5_gq(q[0], q[1], q[2], q[3], q[4], ga, gb);
}

void orc_test_riemann_pt(int dir, int kind, const double ql[4], const double fl[4],
                         const double qr[4], const double fr[4], double gamma, double out[4])
{
    if (kind == 0) hllc_pt(dir, ql, fl, qr, fr, gamma, out);
    else hlle_pt(dir, ql, fl, qr, fr, gamma, out);
}
