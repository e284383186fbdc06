// wenox_sweep.hip -- fused direction sweeps of the device-resident solver.
//
// Compiled twice like wenox_kernels.hip (-DWX_FAST=0 / 1); each build exports
// one SweepOps table.
//
// k_sweep<DIR> computes, for every interior cell r of a chunk, the normal
// flux difference D = (F[r+1] - F[r]) / dx of one direction, where the
// normal direction is the SLOW axis of the padded SoA4 array it reads:
//   DIR = x: the state U itself ((nx+6) x (ny+6) planes);
//   DIR = y: the transposed state U^T ((ny+6) x (nx+6) planes), evaluated with
//            the reference's y-direction formulas (no channel swap), so both
//            sweeps read and write with unit stride along the lanes.
//
// One 64-lane wavefront per workgroup.  Lane l owns padded transverse column
// c0 + l; the wave marches along the normal axis over the faces a of its
// chunk, keeping the six-cell stencil of face a (rows a..a+5) in registers:
//   pass 1  (weno.rs:385-454 / :654-717) on every lane -> LDS,
//   pass 2  (weno.rs:460-501) from the LDS values of lanes l-2..l+2, the
//           averaged Gauss-point fluxes, HLLC + HLLE and p̄ (riemann.rs:6-349,
//           :421-430) on lanes 2..61,
//   sensor  (riemann.rs:352-410) one face behind: normal neighbours from the
//           lane's registers, transverse neighbours from a 2-slot LDS ring;
//           one-sided at the face grid's edges, as the reference;
//   blend   (riemann.rs:433-447) and the flux difference of the cell between
//           the last two finished faces.
// Face arrays never reach HBM.  58 output columns per wave (lanes 3..60).
//
// k_combine applies the SSP-RK3 stage combination with L = -(Dx + Dy^T)
// (riemann.rs:536; Rie_prob_2d_rs.py:183-202) through an LDS tile transpose,
// writing the next state in both orientations.
#include "wenox_numerics.h"
#include "wenox_ops.h"
#include "wenox_sweep.h"

#if WX_FAST
#define WX_NS fasts
#define WX_SWEEP_OPS sweep_ops_fast
#else
#define WX_NS exacts
#define WX_SWEEP_OPS sweep_ops_exact
#endif

namespace wenox {
namespace WX_NS {

using namespace wx;

constexpr int WAVE = 64;
constexpr int OUT_COLS = WAVE - 6;   // lanes 3..60

struct SweepArgs {
    const double* U;      // padded SoA4, M x N per plane, normal = rows
    int64_t M, N;         // padded extents (nn + 6, nt + 6)
    int64_t nn, nt;       // interior cells along normal / transverse
    int chunk;            // interior rows per workgroup
    double gam, dx, two_h, lo, hi;
    double* D;            // interior SoA4 nn x nt
    const StepCtl* ctl;
};

template <int DIR, bool CONS, bool BLEND>
__global__ void __launch_bounds__(WAVE) k_sweep(SweepArgs A)
{
    if (A.ctl && !A.ctl->active) return;
    __shared__ double s_p1[8][WAVE];     // pass-1 q_l (0..3), q_r (4..7) of the current face
    __shared__ double s_pb[2][WAVE];     // p̄ ring (face parity)

    const int l = threadIdx.x;
    const int64_t c0 = (int64_t)blockIdx.x * OUT_COLS;      // padded column of lane 0
    const int64_t pc = c0 + l;                              // this lane's padded column
    const int64_t t = pc - 3;                               // interior transverse index
    const bool col_ok = pc < A.N;
    const int64_t pcl = col_ok ? pc : A.N - 1;              // clamped for loads
    const bool gauss_lane = l >= 2 && l < WAVE - 2;
    const bool out_lane = l >= 3 && l < WAVE - 3 && t >= 0 && t < A.nt;
    const int lm = l > 0 ? l - 1 : 0, lp = l < WAVE - 1 ? l + 1 : WAVE - 1;   // sensor neighbours

    const int64_t r0 = (int64_t)blockIdx.y * A.chunk;
    const int64_t r1 = r0 + A.chunk < A.nn ? r0 + A.chunk : A.nn;   // cells [r0, r1), faces [r0, r1]
    const int64_t a0 = BLEND ? (r0 > 0 ? r0 - 1 : 0) : r0;
    const int64_t a1 = BLEND ? (r1 < A.nn ? r1 + 1 : A.nn) : r1;     // inclusive
    const int64_t P = A.M * A.N;
    const double* U0 = A.U + pcl;

    // stencil window rows a..a+5 (padded), 4 conserved components
    double w[6][4];
#pragma unroll
    for (int s = 0; s < 5; ++s)
#pragma unroll
        for (int k = 0; k < 4; ++k) w[s + 1][k] = U0[k * P + (a0 + s) * A.N];

    double fc_prev[4], fe_prev[4], F_prev[4];
    double pb_m2 = 0.0, pb_m1 = 0.0;    // p̄ of faces a-2, a-1 (this column)
#pragma unroll
    for (int k = 0; k < 4; ++k) { fc_prev[k] = 0.0; fe_prev[k] = 0.0; F_prev[k] = 0.0; }
    const int64_t Pd = A.nn * A.nt;

    auto finish_face = [&](int64_t f, const double F[4]) {
        // cell f-1 lies between faces f-1 and f
        if (out_lane && f - 1 >= r0 && f - 1 < r1) {
            double* d = A.D + (f - 1) * A.nt + t;
#pragma unroll
            for (int k = 0; k < 4; ++k) d[k * Pd] = (F[k] - F_prev[k]) / A.dx;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) F_prev[k] = F[k];
    };

    for (int64_t a = a0; a <= a1; ++a) {
        // slide the window and load row a+5
#pragma unroll
        for (int s = 0; s < 5; ++s)
#pragma unroll
            for (int k = 0; k < 4; ++k) w[s][k] = w[s + 1][k];
#pragma unroll
        for (int k = 0; k < 4; ++k) w[5][k] = U0[k * P + (a + 5) * A.N];

        // pass 1 (normal reconstruction) on every lane
        {
            double ql[4], qr[4];
            recon_normal(DIR, CONS, w, A.gam, ql, qr);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                s_p1[k][l] = ql[k];
                s_p1[4 + k][l] = qr[k];
            }
        }
        __syncthreads();

        double fc[4], fe[4], pb = 1.0;
        if (gauss_lane) {
            // pass 2: Gauss points along the face from lanes l-2..l+2
            double l1[4], l2[4], q1[4], q2[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                weno5_gq(s_p1[k][l - 2], s_p1[k][l - 1], s_p1[k][l], s_p1[k][l + 1], s_p1[k][l + 2],
                         l1[k], l2[k]);
                weno5_gq(s_p1[4 + k][l - 2], s_p1[4 + k][l - 1], s_p1[4 + k][l], s_p1[4 + k][l + 1],
                         s_p1[4 + k][l + 2], q1[k], q2[k]);
            }
            double fa[4], fb[4], fcc[4], fd[4];
            flux_pt(DIR, l1, A.gam, fa);
            flux_pt(DIR, l2, A.gam, fb);
            flux_pt(DIR, q1, A.gam, fcc);
            flux_pt(DIR, q2, A.gam, fd);
            double QL[4], FL[4], QR[4], FR[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                FL[k] = 0.5 * (fa[k] + fb[k]);
                FR[k] = 0.5 * (fd[k] + fcc[k]);
                QL[k] = 0.5 * (l1[k] + l2[k]);
                QR[k] = 0.5 * (q1[k] + q2[k]);
            }
            FaceState<double> st = face_speeds(DIR, QL, QR, A.gam);
            hlle_from(st, FL, FR, fe);
            if (BLEND) {
                hllc_from(DIR, st, QL, FL, QR, FR, fc);
                pb = 0.5 * (pressure_floored(QL, A.gam) + pressure_floored(QR, A.gam));
            }
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) { fc[k] = 0.0; fe[k] = 0.0; }
        }

        if (!BLEND) {
            finish_face(a, fe);
        } else {
            s_pb[a & 1][l] = pb;
            __syncthreads();
            if (a > a0) {
                // sensor and blend of face f = a-1 (riemann.rs:360-407)
                const int64_t f = a - 1;
                double dn = f == 0 ? (pb - pb_m1) / A.two_h : (pb - pb_m2) / A.two_h;
                const double* ring = s_pb[f & 1];
                double dt_;
                if (t <= 0) dt_ = (ring[lp] - ring[l]) / A.two_h;
                else if (t >= A.nt - 1) dt_ = (ring[l] - ring[lm]) / A.two_h;
                else dt_ = (ring[lp] - ring[lm]) / A.two_h;
                double phi = DIR == DIR_X ? sensor_phi(dn, dt_, pb_m1, A.lo, A.hi)
                                          : sensor_phi(dt_, dn, pb_m1, A.lo, A.hi);
                double F[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) F[k] = phi * fc_prev[k] + (1.0 - phi) * fe_prev[k];
                if (f >= r0) finish_face(f, F);
            }
            pb_m2 = pb_m1;
            pb_m1 = pb;
#pragma unroll
            for (int k = 0; k < 4; ++k) { fc_prev[k] = fc[k]; fe_prev[k] = fe[k]; }
        }
        __syncthreads();
    }

    if (BLEND && a1 == A.nn) {
        // last face of the grid: one-sided normal difference, no face a1+1
        const int64_t f = a1;
        double dn = (pb_m1 - pb_m2) / A.two_h;
        const double* ring = s_pb[f & 1];
        double dt_;
        if (t <= 0) dt_ = (ring[lp] - ring[l]) / A.two_h;
        else if (t >= A.nt - 1) dt_ = (ring[l] - ring[lm]) / A.two_h;
        else dt_ = (ring[lp] - ring[lm]) / A.two_h;
        double phi = DIR == DIR_X ? sensor_phi(dn, dt_, pb_m1, A.lo, A.hi)
                                  : sensor_phi(dt_, dn, pb_m1, A.lo, A.hi);
        double F[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) F[k] = phi * fc_prev[k] + (1.0 - phi) * fe_prev[k];
        finish_face(f, F);
    }
}

// ------------------------------------------------------------------------
// RK stage combine with L = -(Dx + Dy^T), both orientations written.
// Tile of 32 x 32 interior cells, 256 threads; Dy^T and the transposed
// output go through LDS.
// ------------------------------------------------------------------------
constexpr int TT = 32;

__global__ void __launch_bounds__(256) k_combine(const double* __restrict__ Dx, const double* __restrict__ DyT,
                                                 const double* Un, const double* Us, double* out,
                                                 double* outT, int64_t nx, int64_t ny, int stage,
                                                 const StepCtl* ctl)
{
    if (ctl && !ctl->active) return;
    __shared__ double tile[TT][TT + 1];
    const int tx = threadIdx.x & (TT - 1);   // fast index
    const int ty = threadIdx.x / TT;         // 0..7
    const int64_t i0 = (int64_t)blockIdx.y * TT, j0 = (int64_t)blockIdx.x * TT;
    const int64_t m = nx + 6, n = ny + 6, P = m * n, PI = nx * ny;
    const double dt = ctl ? ctl->dt : 0.0;
    for (int k = 0; k < 4; ++k) {
        // Dy^T[k][j][i]: rows j, contiguous i
        for (int r = ty; r < TT; r += 8) {
            int64_t j = j0 + r, i = i0 + tx;
            if (j < ny && i < nx) tile[r][tx] = DyT[k * PI + j * nx + i];
        }
        __syncthreads();
        double res[TT / 8];
        for (int r = ty, q = 0; r < TT; r += 8, ++q) {
            int64_t i = i0 + r, j = j0 + tx;
            res[q] = 0.0;
            if (i < nx && j < ny) {
                double dgdy = tile[tx][r];
                double l = -(Dx[k * PI + i * ny + j] + dgdy);
                int64_t o = (i + 3) * n + (j + 3);
                double qn = Un[k * P + o];
                double v;
                if (stage == 1) v = qn + dt * l;
                else if (stage == 2) v = 0.75 * qn + 0.25 * Us[k * P + o] + (0.25 * dt) * l;
                else v = (1.0 / 3.0) * qn + (2.0 / 3.0) * Us[k * P + o] + ((2.0 / 3.0) * dt) * l;
                out[k * P + o] = v;
                res[q] = v;
            }
        }
        __syncthreads();
        for (int r = ty, q = 0; r < TT; r += 8, ++q) tile[r][tx] = res[q];
        __syncthreads();
        // outT[k][j+3][i+3], contiguous i
        for (int r = ty; r < TT; r += 8) {
            int64_t j = j0 + r, i = i0 + tx;
            if (j < ny && i < nx) outT[k * P + (j + 3) * m + (i + 3)] = tile[tx][r];
        }
        __syncthreads();
    }
}

// Plain RHS from the two sweeps (for tests): L = -(Dx + Dy^T), SoA4 interior.
__global__ void k_rhs_out(const double* Dx, const double* DyT, double* L, int64_t nx, int64_t ny)
{
    int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= nx * ny) return;
    int64_t i = t / ny, j = t - i * ny;
    for (int k = 0; k < 4; ++k) L[k * nx * ny + t] = -(Dx[k * nx * ny + t] + DyT[k * nx * ny + j * nx + i]);
}

// Padded SoA4 transpose: out[k][j][i] = in[k][i][j] for an M x N plane set.
__global__ void __launch_bounds__(256) k_transpose(const double* in, double* out, int64_t M, int64_t N)
{
    __shared__ double tile[TT][TT + 1];
    const int tx = threadIdx.x & (TT - 1), ty = threadIdx.x / TT;
    const int64_t i0 = (int64_t)blockIdx.y * TT, j0 = (int64_t)blockIdx.x * TT;
    const int64_t P = M * N;
    for (int k = 0; k < 4; ++k) {
        for (int r = ty; r < TT; r += 8) {
            int64_t i = i0 + r, j = j0 + tx;
            if (i < M && j < N) tile[r][tx] = in[k * P + i * N + j];
        }
        __syncthreads();
        for (int r = ty; r < TT; r += 8) {
            int64_t j = j0 + r, i = i0 + tx;
            if (i < M && j < N) out[k * P + j * M + i] = tile[tx][r];
        }
        __syncthreads();
    }
}

template <int DIR>
static void launch_sweep(const SweepArgs& a, int cons, int blend, hipStream_t s)
{
    dim3 grid((unsigned)((a.nt + OUT_COLS - 1) / OUT_COLS), (unsigned)((a.nn + a.chunk - 1) / a.chunk));
    if (cons) {
        if (blend) k_sweep<DIR, true, true><<<grid, WAVE, 0, s>>>(a);
        else k_sweep<DIR, true, false><<<grid, WAVE, 0, s>>>(a);
    } else {
        if (blend) k_sweep<DIR, false, true><<<grid, WAVE, 0, s>>>(a);
        else k_sweep<DIR, false, false><<<grid, WAVE, 0, s>>>(a);
    }
}

static void h_sweeps(const SweepPlan& p, const double* Us, const double* UsT, double* Dx, double* DyT,
                     const StepCtl* ctl, hipStream_t s)
{
    const int blend = p.force_hlle ? 0 : 1;
    SweepArgs ax{Us, p.nx + 6, p.ny + 6, p.nx, p.ny, p.chunk_x, p.gamma, p.dx, p.two_h_x,
                 p.jp_lo, p.jp_hi, Dx, ctl};
    launch_sweep<DIR_X>(ax, p.cons, blend, s);
    SweepArgs ay{UsT, p.ny + 6, p.nx + 6, p.ny, p.nx, p.chunk_y, p.gamma, p.dx, p.two_h_y,
                 p.jp_lo, p.jp_hi, DyT, ctl};
    launch_sweep<DIR_Y>(ay, p.cons, blend, s);
}

static void h_stage(const SweepPlan& p, const double* Us, const double* UsT, const double* Un,
                    double* Dx, double* DyT, double* out, double* outT, int stage, const StepCtl* ctl,
                    hipStream_t s)
{
    h_sweeps(p, Us, UsT, Dx, DyT, ctl, s);
    dim3 grid((unsigned)((p.ny + TT - 1) / TT), (unsigned)((p.nx + TT - 1) / TT));
    k_combine<<<grid, 256, 0, s>>>(Dx, DyT, Un, Us, out, outT, p.nx, p.ny, stage, ctl);
}

static void h_rhs(const SweepPlan& p, const double* U, const double* UT, double* Dx, double* DyT,
                  double* L, const StepCtl* ctl, hipStream_t s)
{
    h_sweeps(p, U, UT, Dx, DyT, ctl, s);
    int64_t nc = p.nx * p.ny;
    k_rhs_out<<<(unsigned)((nc + 255) / 256), 256, 0, s>>>(Dx, DyT, L, p.nx, p.ny);
}

static void h_transpose(const double* in, double* out, int64_t M, int64_t N, hipStream_t s)
{
    dim3 grid((unsigned)((N + TT - 1) / TT), (unsigned)((M + TT - 1) / TT));
    k_transpose<<<grid, 256, 0, s>>>(in, out, M, N);
}

}  // namespace WX_NS

#ifndef __HIP_DEVICE_COMPILE__
const SweepOps WX_SWEEP_OPS = {WX_NS::h_stage, WX_NS::h_rhs, WX_NS::h_transpose};

#if !WX_FAST
// Chunk length: enough waves to cover the 1024 SIMDs twice where the grid
// allows it, within [8, 64] rows (a chunk recomputes 2 halo faces for the
// sensor, so longer is cheaper per cell).
static int pick_chunk(int64_t nn, int64_t nt)
{
    int64_t cols = (nt + WX_NS::OUT_COLS - 1) / WX_NS::OUT_COLS;
    int64_t c = nn * cols / 2048;
    if (c < 8) c = 8;
    if (c > 64) c = 64;
    return (int)c;
}

SweepPlan make_sweep_plan(int64_t nx, int64_t ny, double dx, double gamma, int cons,
                          int force_hlle, double jp_lo, double jp_hi)
{
    SweepPlan p{};
    p.nx = nx; p.ny = ny; p.dx = dx; p.gamma = gamma;
    p.cons = cons; p.force_hlle = force_hlle; p.jp_lo = jp_lo; p.jp_hi = jp_hi;
    int64_t mx = (nx + 1) < ny ? (nx + 1) : ny;   // x-face grid (nx+1, ny)
    int64_t my = nx < (ny + 1) ? nx : (ny + 1);   // y-face grid (nx, ny+1)
    p.two_h_x = 2.0 * (1.0 / (double)(mx > 1 ? mx : 1));
    p.two_h_y = 2.0 * (1.0 / (double)(my > 1 ? my : 1));
    p.chunk_x = pick_chunk(nx, ny);
    p.chunk_y = pick_chunk(ny, nx);
    return p;
}
#endif
#endif

}  // namespace wenox
