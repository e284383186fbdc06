// wenox_kernels.hip -- gfx950 kernels of the 2D WENO5 / HLLC hot path.
//
// Compiled three times: -DWX_FAST=0 -ffp-contract=off (exports
// wenox::ops_exact), -DWX_FAST=1 -ffp-contract=fast (wenox::ops_fast) and
// -DWX_FAST=0 -DWX_TWIN=1 -ffp-contract=off (wenox::ops_twin, the NumPy twin's
// semantics, see wenox_numerics.h).
//
// This file holds the layer-by-layer kernels that mirror the reference's
// passes (weno_ext/src/weno.rs pass 1 / pass 2, riemann.rs solvers, sensor and
// divergence).  They serve the drop-in functions (weno_x_py, hllc_*_rs_py, ...)
// and the unfused stepper engine; the fused sweep kernels live in
// wenox_sweep.hip.  It also holds the stepper's ghost fill, RK stage combine and
// the single-pass CFL reduction.
#include "wenox_numerics.h"
#include "wenox_ops.h"

#if WX_TWIN
#define WX_NS twink
#define WX_OPS ops_twin
#elif WX_FAST
#define WX_NS fastk
#define WX_OPS ops_fast
#else
#define WX_NS exactk
#define WX_OPS ops_exact
#endif

namespace wenox {
namespace WX_NS {

using namespace wx;

constexpr int TB = 256;

static inline unsigned blocks_for(int64_t n) { return (unsigned)((n + TB - 1) / TB); }

struct Acc {
    const double* p;
    Lay l;
    __device__ __forceinline__ double operator()(int64_t i, int64_t j, int k) const
    {
        return p[i * l.is + j * l.js + k * l.ks];
    }
};
struct AccW {
    double* p;
    Lay l;
    __device__ __forceinline__ double& operator()(int64_t i, int64_t j, int k) const
    {
        return p[i * l.is + j * l.js + k * l.ks];
    }
};

// ------------------------------------------------------------------------
// utils.rs:74-108 / :9-35
// ------------------------------------------------------------------------
__global__ void k_flux(Acc q, int64_t M, int64_t N, int dir, double gam, AccW out)
{
    int64_t t = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (t >= M * N) return;
    int64_t i = t / N, j = t - i * N;
    double c[4], f[4];
    for (int k = 0; k < 4; ++k) c[k] = q(i, j, k);
    flux_pt(dir, c, gam, f);
    for (int k = 0; k < 4; ++k) out(i, j, k) = f[k];
}

__global__ void k_con2primi(Acc q, int64_t M, int64_t N, double gam, double* p, double* a,
                            double* rho, double* u, double* v, double* h)
{
    int64_t t = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (t >= M * N) return;
    int64_t i = t / N, j = t - i * N;
    double c[4];
    for (int k = 0; k < 4; ++k) c[k] = q(i, j, k);
    Prim<double> r = con2primi_pt(c, gam);
    p[t] = r.p; a[t] = r.a; rho[t] = r.rho; u[t] = r.u; v[t] = r.v; h[t] = r.h;
}

// Snapshot rows [row0, row0 + rows) of the padded SoA4 state as the reference
// drivers' (m, n, 8) array: conserved channels, then rho_safe, u, v, p of
// con2primi (Rie_prob_2d_rs.py:157-163), one cell (64 B) per thread.
__global__ void k_snapshot8(const double* U, int64_t P, int64_t n, int64_t ld, double gam, int64_t row0,
                            int64_t cells, double* out)
{
    int64_t t = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (t >= cells) return;
    const int64_t row = t / n;
    const int64_t o = (row0 + row) * ld + (t - row * n);
    double c[4];
    for (int k = 0; k < 4; ++k) c[k] = U[k * P + o];
    Prim<double> r = con2primi_pt(c, gam);
    double* d = out + t * 8;
    d[0] = c[0]; d[1] = c[1]; d[2] = c[2]; d[3] = c[3];
    d[4] = r.rho; d[5] = r.u; d[6] = r.v; d[7] = r.p;
}

// ------------------------------------------------------------------------
// weno.rs:385-454 (characteristic) / :654-717 (component-wise): pass 1.
// x: face i between cells i+2 and i+3 of column j, (m-5, n) faces.
// y: face j between cells j+2 and j+3 of row i, (m, n-5) faces.
// ------------------------------------------------------------------------
__global__ void k_pass1(Acc u, int64_t om, int64_t on, int dir, int cons, double gam, AccW pl,
                        AccW pr)
{
    int64_t t = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (t >= om * on) return;
    int64_t i = t / on, j = t - i * on;
    double c[6][4];
#pragma unroll
    for (int s = 0; s < 6; ++s)
#pragma unroll
        for (int k = 0; k < 4; ++k) c[s][k] = dir == DIR_X ? u(i + s, j, k) : u(i, j + s, k);
    double ql[4], qr[4];
    recon_normal(dir, cons != 0, c, gam, ql, qr);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        pl(i, j, k) = ql[k];
        pr(i, j, k) = qr[k];
    }
}

// weno.rs:460-501 / :590-631: pass 2 (transverse 2-point Gauss WENO) and the
// fluxes at both Gauss points, averaged.
__global__ void k_pass2(Acc pl, Acc pr, int64_t m1, int64_t n1, int dir, double gam, AccW ql,
                        AccW fl, AccW qr, AccW fr)
{
    int64_t t = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (t >= m1 * n1) return;
    int64_t i = t / n1, j = t - i * n1;
    double l1[4], l2[4], r1[4], r2[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double a[5], b[5];
#pragma unroll
        for (int s = 0; s < 5; ++s) {
            a[s] = dir == DIR_X ? pl(i, j + 1 + s, k) : pl(i + 1 + s, j, k);
            b[s] = dir == DIR_X ? pr(i, j + 1 + s, k) : pr(i + 1 + s, j, k);
        }
        weno5_gq(a[0], a[1], a[2], a[3], a[4], l1[k], l2[k]);
        weno5_gq(b[0], b[1], b[2], b[3], b[4], r1[k], r2[k]);
    }
    double fa[4], fb[4], fc[4], fd[4];
    flux_pt(dir, l1, gam, fa);
    flux_pt(dir, l2, gam, fb);
    flux_pt(dir, r1, gam, fc);
    flux_pt(dir, r2, gam, fd);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        fl(i, j, k) = 0.5 * (fa[k] + fb[k]);
        fr(i, j, k) = 0.5 * (fd[k] + fc[k]);
        ql(i, j, k) = 0.5 * (l1[k] + l2[k]);
        qr(i, j, k) = 0.5 * (r1[k] + r2[k]);
    }
}

// ------------------------------------------------------------------------
// riemann.rs:6-349 per face; kind 0 HLLC, 1 HLLE
// ------------------------------------------------------------------------
__device__ __forceinline__ void load4(Acc a, int64_t i, int64_t j, double o[4])
{
#pragma unroll
    for (int k = 0; k < 4; ++k) o[k] = a(i, j, k);
}

__global__ void k_riemann(Acc QL, Acc FL, Acc QR, Acc FR, int64_t F1, int64_t F2, int dir,
                          int kind, double gam, AccW out)
{
    int64_t t = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (t >= F1 * F2) return;
    int64_t i = t / F2, j = t - i * F2;
    double ql[4], fl[4], qr[4], fr[4], o[4];
    load4(QL, i, j, ql); load4(FL, i, j, fl); load4(QR, i, j, qr); load4(FR, i, j, fr);
    FaceState<double> s = face_speeds(dir, ql, qr, gam);
    if (kind == 0) hllc_from(dir, s, ql, fl, qr, fr, o);
    else hlle_from(s, ql, fl, qr, fr, o);
#pragma unroll
    for (int k = 0; k < 4; ++k) out(i, j, k) = o[k];
}

// riemann.rs:421-430: p̄ = 0.5 (p_l + p_r) from floored pressures
__global__ void k_pbar(Acc QL, Acc QR, int64_t F1, int64_t F2, double gam, double* pbar)
{
    int64_t t = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (t >= F1 * F2) return;
    int64_t i = t / F2, j = t - i * F2;
    double ql[4], qr[4];
    load4(QL, i, j, ql); load4(QR, i, j, qr);
    pbar[t] = 0.5 * (pressure_floored(ql, gam) + pressure_floored(qr, gam));
}

// riemann.rs:352-410 (sensor on the face grid, one-sided at its edges) and
// :433-447 (phi*HLLC + (1-phi)*HLLE)
__device__ __forceinline__ double sensor_at(const double* p, int64_t F1, int64_t F2, int64_t i,
                                            int64_t j, double two_h, double lo, double hi)
{
    double dpdx, dpdy;
    if (i == 0) dpdx = (p[F2 + j] - p[j]) / two_h;
    else if (i == F1 - 1) dpdx = (p[i * F2 + j] - p[(i - 1) * F2 + j]) / two_h;
    else dpdx = (p[(i + 1) * F2 + j] - p[(i - 1) * F2 + j]) / two_h;
    if (j == 0) dpdy = (p[i * F2 + 1] - p[i * F2]) / two_h;
    else if (j == F2 - 1) dpdy = (p[i * F2 + j] - p[i * F2 + j - 1]) / two_h;
    else dpdy = (p[i * F2 + j + 1] - p[i * F2 + j - 1]) / two_h;
    return sensor_phi(dpdx, dpdy, p[i * F2 + j], lo, hi, 1.0 / (hi - lo));
}

__global__ void k_blend(Acc QL, Acc FL, Acc QR, Acc FR, const double* pbar, int64_t F1, int64_t F2,
                        int dir, double gam, double two_h, double lo, double hi, AccW out)
{
    int64_t t = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (t >= F1 * F2) return;
    int64_t i = t / F2, j = t - i * F2;
    double ql[4], fl[4], qr[4], fr[4], fc[4], fe[4];
    load4(QL, i, j, ql); load4(FL, i, j, fl); load4(QR, i, j, qr); load4(FR, i, j, fr);
    FaceState<double> s = face_speeds(dir, ql, qr, gam);
    hllc_from(dir, s, ql, fl, qr, fr, fc);
    hlle_from(s, ql, fl, qr, fr, fe);
    double w = sensor_at(pbar, F1, F2, i, j, two_h, lo, hi);
#pragma unroll
    for (int k = 0; k < 4; ++k) out(i, j, k) = w * fc[k] + (1.0 - w) * fe[k];
}

// riemann.rs:513-540
__global__ void k_divergence(Acc fx, Acc fy, int64_t Nx, int64_t Ny, double dx, AccW out)
{
    int64_t t = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (t >= Nx * Ny) return;
    int64_t i = t / Ny, j = t - i * Ny;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double dfdx = (fx(i + 1, j, k) - fx(i, j, k)) / dx;
        double dgdy = (fy(i, j + 1, k) - fy(i, j, k)) / dx;
        out(i, j, k) = -(dfdx + dgdy);
    }
}

// ------------------------------------------------------------------------
// host launchers of the drop-in pieces
// ------------------------------------------------------------------------
static void h_flux(const double* q, Lay lq, int64_t M, int64_t N, int dir, double gam, double* out,
                   Lay lo, hipStream_t s)
{
    if (M * N == 0) return;
    k_flux<<<blocks_for(M * N), TB, 0, s>>>(Acc{q, lq}, M, N, dir, gam, AccW{out, lo});
}

static void h_con2primi(const double* q, Lay lq, int64_t M, int64_t N, double gam, double* p,
                        double* a, double* rho, double* u, double* v, double* h, hipStream_t s)
{
    if (M * N == 0) return;
    k_con2primi<<<blocks_for(M * N), TB, 0, s>>>(Acc{q, lq}, M, N, gam, p, a, rho, u, v, h);
}

static void h_snapshot8(const double* U, int64_t m, int64_t n, int64_t ld, double gam, int64_t row0,
                        int64_t rows, double* out, hipStream_t s)
{
    const int64_t cells = rows * n;
    if (cells <= 0) return;
    k_snapshot8<<<blocks_for(cells), TB, 0, s>>>(U, m * ld, n, ld, gam, row0, cells, out);
}

static void h_weno(const double* u, Lay lu, int64_t m, int64_t n, int dir, int cons, double gam,
                   double* pl, double* pr, double* ql, double* fl, double* qr, double* fr, Lay lo,
                   hipStream_t s)
{
    int64_t om = dir == DIR_X ? m - 5 : m;
    int64_t on = dir == DIR_X ? n : n - 5;
    Lay lp = lay_soa4(om, on);
    if (om * on > 0)
        k_pass1<<<blocks_for(om * on), TB, 0, s>>>(Acc{u, lu}, om, on, dir, cons, gam,
                                                    AccW{pl, lp}, AccW{pr, lp});
    int64_t m1 = dir == DIR_X ? om : om - 6;
    int64_t n1 = dir == DIR_X ? on - 6 : on;
    if (m1 * n1 > 0)
        k_pass2<<<blocks_for(m1 * n1), TB, 0, s>>>(Acc{pl, lp}, Acc{pr, lp}, m1, n1, dir, gam,
                                                    AccW{ql, lo}, AccW{fl, lo}, AccW{qr, lo},
                                                    AccW{fr, lo});
}

static void h_riemann(const double* ql, const double* fl, const double* qr, const double* fr,
                      Lay li, int64_t F1, int64_t F2, int dir, int kind, double gam, double lo,
                      double hi, double* pbar, double* out, Lay lout, int64_t h_min, hipStream_t s)
{
    if (F1 * F2 == 0) return;
    Acc QL{ql, li}, FL{fl, li}, QR{qr, li}, FR{fr, li};
    if (kind < 2) {
        k_riemann<<<blocks_for(F1 * F2), TB, 0, s>>>(QL, FL, QR, FR, F1, F2, dir, kind, gam,
                                                      AccW{out, lout});
        return;
    }
    int64_t mn = h_min > 0 ? h_min : (F1 < F2 ? F1 : F2);
    double h = 1.0 / (double)(mn > 1 ? mn : 1);
    k_pbar<<<blocks_for(F1 * F2), TB, 0, s>>>(QL, QR, F1, F2, gam, pbar);
    k_blend<<<blocks_for(F1 * F2), TB, 0, s>>>(QL, FL, QR, FR, pbar, F1, F2, dir, gam, 2.0 * h,
                                                lo, hi, AccW{out, lout});
}

static void h_divergence(const double* fx, Lay lfx, const double* fy, Lay lfy, int64_t Nx,
                         int64_t Ny, double dx, double* out, Lay lout, hipStream_t s)
{
    if (Nx * Ny == 0) return;
    k_divergence<<<blocks_for(Nx * Ny), TB, 0, s>>>(Acc{fx, lfx}, Acc{fy, lfy}, Nx, Ny, dx,
                                                    AccW{out, lout});
}

#include "wenox_state.h"

static void h_bc(double* U, double* UT, int64_t nx, int64_t ny, int64_t ld, int64_t ldT, BcParams bp,
                 const StepCtl* ctl, int stage, hipStream_t s)
{
    launch_bc(StF64{U, (nx + 6) * ld}, StF64{UT, (ny + 6) * ldT}, UT != nullptr, nx, ny, ld, ldT, bp, ctl,
              stage, s);
}

static void h_cfl(const double* U, int64_t nx, int64_t ny, int64_t ld, double gam, StepCtl* ctl, hipStream_t s)
{
    launch_cfl(StF64{const_cast<double*>(U), (nx + 6) * ld}, nx, ny, ld, gam, ctl, s);
}

__global__ void k_advance(StepCtl* ctl)
{
    if (threadIdx.x == 0 && blockIdx.x == 0 && ctl->active) {
        ctl->T += ctl->dt;
        ctl->step += 1;
    }
}

static void h_advance(StepCtl* ctl, hipStream_t s) { k_advance<<<1, 64, 0, s>>>(ctl); }

static void h_cfl_part(const double* U, int64_t nx, int64_t ny, int64_t ld, double gam, StepCtl* ctl,
                       hipStream_t s)
{
    launch_cfl(StF64{const_cast<double*>(U), (nx + 6) * ld}, nx, ny, ld, gam, ctl, s, false);
}

static void h_step_dt(StepCtl* ctl, hipStream_t s) { k_step_dt<<<1, 1024, 0, s>>>(ctl); }

static void h_settle(StepCtl* ctl, hipStream_t s) { k_settle<<<1, 64, 0, s>>>(ctl); }

// WENO_HLLC_2d.py:813-837 (con_var2primi_pfloor): where the pressure of a cell
// (density floored at 1e-12 for the velocities and the kinetic energy) falls
// below pmin, raise its total energy so that p = pmin.  Every padded cell of
// the M x N planes (row pitch ld), after an active step.  With cfl != 0 the
// interior cells also give the next step's wave speed as the twin's loop
// forms it from the primitive channels this function writes (:895-897):
// sqrt(g p / max(rho, rho_floor)) + sqrt(u^2 + v^2) with p = pmin where the
// floor acted; the maximum (NaN-propagating, as np.max) goes to
// ctl->floor_bits by a bit-pattern atomicMax (values are >= 0).
__global__ void k_pfloor(double* U, int64_t M, int64_t N, int64_t ld, double gam, double pmin, StepCtl* ctl,
                         int cfl)
{
    if (!ctl->active) return;
    const int64_t t = (int64_t)blockIdx.x * TB + threadIdx.x;
    double w = 0.0;
    if (t < M * N) {
        const int64_t i = t / N, j = t - i * N, o = i * ld + j, P = M * ld;
        const double rho = fmax(U[o], 1e-12);
        const double u = U[P + o] / rho, v = U[2 * P + o] / rho;
        const double kinetic = 0.5 * rho * (u * u + v * v);
        double p = (gam - 1.0) * (U[3 * P + o] - kinetic);
        if (p < pmin) {
            U[3 * P + o] = kinetic + pmin / (gam - 1.0);
            p = pmin;
        }
        if (cfl && i >= 3 && i < M - 3 && j >= 3 && j < N - 3)
            w = sqrt(gam * p / fmax(U[o], ctl->rho_floor)) + sqrt(u * u + v * v);
    }
    if (!cfl) return;
    w = wave_max(w);
    if ((threadIdx.x & 63) == 0) atomicMax(&ctl->floor_bits, (unsigned long long)__double_as_longlong(w));
    if (t == 0) ctl->floor_valid = 1;
}

static void h_pfloor(double* U, int64_t M, int64_t N, int64_t ld, double gam, double pmin, StepCtl* ctl, int cfl,
                     hipStream_t s)
{
    if (M * N == 0) return;
    k_pfloor<<<blocks_for(M * N), TB, 0, s>>>(U, M, N, ld, gam, pmin, ctl, cfl);
}

// ------------------------------------------------------------------------
// Unfused RHS for the stepper: the reference's pass structure on SoA4.
// work layout: pl, pr (x pass 1), then ql, fl, qr, fr (x faces), fx; then the
// same for y; then pbar.
// ------------------------------------------------------------------------
static void h_rhs_unfused(const double* U, int64_t nx, int64_t ny, int64_t ld, double dx, double gam, int cons,
                          int force_hlle, double lo, double hi, double* work, double* L,
                          const StepCtl* /*ctl*/, hipStream_t s)
{
    const int64_t m = nx + 6, n = ny + 6;
    const Lay lu = Lay{ld, 1, m * ld};
    const int64_t p1 = 4 * (m - 5) * n > 4 * m * (n - 5) ? 4 * (m - 5) * n : 4 * m * (n - 5);
    const int64_t fxs = 4 * (nx + 1) * ny, fys = 4 * nx * (ny + 1);
    double* pl = work;
    double* pr = pl + p1;
    double* f[4];
    f[0] = pr + p1;
    for (int k = 1; k < 4; ++k) f[k] = f[k - 1] + (fxs > fys ? fxs : fys);
    double* fx = f[3] + (fxs > fys ? fxs : fys);
    double* fy = fx + fxs;
    double* pbar = fy + fys;
    const int kind = force_hlle ? 1 : 2;
    Lay lx = lay_soa4(nx + 1, ny);
    h_weno(U, lu, m, n, DIR_X, cons, gam, pl, pr, f[0], f[1], f[2], f[3], lx, s);
    h_riemann(f[0], f[1], f[2], f[3], lx, nx + 1, ny, DIR_X, kind, gam, lo, hi, pbar, fx, lx, 0, s);
    Lay ly = lay_soa4(nx, ny + 1);
    h_weno(U, lu, m, n, DIR_Y, cons, gam, pl, pr, f[0], f[1], f[2], f[3], ly, s);
    h_riemann(f[0], f[1], f[2], f[3], ly, nx, ny + 1, DIR_Y, kind, gam, lo, hi, pbar, fy, ly, 0, s);
    h_divergence(fx, lx, fy, ly, nx, ny, dx, L, lay_soa4(nx, ny), s);
}

// ------------------------------------------------------------------------
// SSP-RK3 stage combine (Rie_prob_2d_rs.py:183-202):
//   1: q + dt*l      2: (0.75 q + 0.25 q1) + (0.25 dt) l
//   3: ((1/3) q + (2/3) q2) + ((2/3) dt) l
// ------------------------------------------------------------------------
__global__ void k_rk(const double* Un, const double* Us, const double* L, double* out, int64_t nx,
                     int64_t ny, int64_t ld, int stage, const StepCtl* ctl)
{
    if (!ctl->active) return;
    int64_t t = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (t >= nx * ny) return;
    const int64_t P = (nx + 6) * ld, PL = nx * ny;
    int64_t i = t / ny, j = t - i * ny;
    int64_t o = (i + 3) * ld + (j + 3);
    const double dt = ctl->dt;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double q = Un[k * P + o];
        double l = L[k * PL + t];
        double r;
        if (stage == 1) {
            r = q + dt * l;
        } else if (stage == 2) {
            r = 0.75 * q + 0.25 * Us[k * P + o] + (0.25 * dt) * l;
        } else {
            r = (1.0 / 3.0) * q + (2.0 / 3.0) * Us[k * P + o] + ((2.0 / 3.0) * dt) * l;
        }
        out[k * P + o] = r;
    }
}

static void h_rk(const double* Un, const double* Us, const double* L, double* out, int64_t nx,
                 int64_t ny, int64_t ld, int stage, const StepCtl* ctl, hipStream_t s)
{
    if (nx * ny == 0) return;
    k_rk<<<blocks_for(nx * ny), TB, 0, s>>>(Un, Us, L, out, nx, ny, ld, stage, ctl);
}

}  // namespace WX_NS

#ifndef __HIP_DEVICE_COMPILE__
const Ops WX_OPS = {
    WX_NS::h_flux, WX_NS::h_con2primi, WX_NS::h_snapshot8, WX_NS::h_weno, WX_NS::h_riemann, WX_NS::h_divergence,
    WX_NS::h_bc,   WX_NS::h_cfl,       WX_NS::h_advance, WX_NS::h_cfl_part, WX_NS::h_step_dt,
    WX_NS::h_settle, WX_NS::h_rhs_unfused, WX_NS::h_rk, WX_NS::h_pfloor,
};

#if !WX_FAST && !WX_TWIN
size_t rhs_unfused_scratch(int64_t nx, int64_t ny)
{
    const int64_t m = nx + 6, n = ny + 6;
    const int64_t p1 = 4 * (m - 5) * n > 4 * m * (n - 5) ? 4 * (m - 5) * n : 4 * m * (n - 5);
    const int64_t fxs = 4 * (nx + 1) * ny, fys = 4 * nx * (ny + 1);
    const int64_t fm = fxs > fys ? fxs : fys;
    return (size_t)(2 * p1 + 4 * fm + fxs + fys + (nx + 1) * (ny + 1));
}
#endif
#endif  // !__HIP_DEVICE_COMPILE__

}  // namespace wenox
