// wenox_sweep.hip -- fused direction sweeps of the device-resident solver.
//
// Compiled twice like wenox_kernels.hip (-DWX_FAST=0 / 1); each build exports
// one SweepOps table.  The FP32-split kernels exist in the fast build only.
//
// k_sweep computes, for every interior cell r of a chunk, the normal flux
// difference D = (F[r+1] - F[r]) / dx of one direction, where the normal
// direction is the SLOW axis of the padded plane set it reads:
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
// Template parameters: T compute type; IN the state encoding read (FP64
// planes, FP32 hi planes, FP32 hi+lo planes read as FP64); DT the type of D;
// MODE plain or DETECT (FP32 arithmetic on increments; a face row where a
// Gauss-point state has rho <= 0 or p <= 0 or is not finite is recomputed
// in FP64 by its wavefront on the spot).
//
// k_combine applies the SSP-RK3 stage combination with L = -(Dx + Dy^T)
// (riemann.rs:536; Rie_prob_2d_rs.py:183-202) through an LDS tile transpose,
// writing the next state in both orientations.  The exact build keeps the
// reference's operation order; the fast build's FP64 step lets the x sweep
// form a Un + b Us - (c dt) Dx in place of Dx (FOLD), so the combine reads
// two arrays instead of four.
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

// Phase markers for static instruction counts (phase_counts.py builds with
// -DWX_PHASE_MARKS=1): a scheduling barrier and an assembly comment at each
// phase boundary of the sweep's march.  Off in the library build.
#if WX_PHASE_MARKS
#define WX_PHASE(name)                                 \
    do {                                               \
        __builtin_amdgcn_sched_barrier(0);             \
        asm volatile(";@@phase " #name ::: "memory");  \
    } while (0)
#else
#define WX_PHASE(name) \
    do {               \
    } while (0)
#endif

namespace wenox {
namespace WX_NS {

using namespace wx;

#include "wenox_state.h"

constexpr int WAVE = 64;
constexpr int OUT_COLS = WAVE - 6;   // lanes 3..60

enum { IN_F64 = 0, IN_F32 = 1, IN_SPLIT = 2 };
enum { MODE_PLAIN = 0, MODE_DETECT = 1 };
// What a sweep writes (its FOLD): the flux differences D of its cells
// (FOLD_NONE); E = a Un + b Us - (c dt) D, the x sweep's part of the stage
// combination (FOLD_E: row layout, interior planes; FOLD_BE: blocked padded
// planes); or, as the y sweep of the blocked step, the stage's new state
// E - (c dt) D (FOLD_BO: blocked padded planes, E read where Un is).
enum { FOLD_NONE = 0, FOLD_E = 1, FOLD_BE = 2, FOLD_BO = 3 };

struct SweepArgs {
    const void* U;        // padded planes, M x N each, normal = rows
    int64_t M, N;         // padded extents (nn + 6, nt + 6)
    int64_t ld, ldd;      // row pitches (elements) of U's planes and of D's
    int64_t nn, nt;       // interior cells along normal / transverse
    int chunk;            // interior rows per wavefront
    int waves;            // wavefronts per workgroup (consecutive chunks)
    double gam, dx, two_h, lo, hi;
    void* D;              // interior planes nn x nt (DT), row pitch ldd
    int* flags;           // one per wavefront: DETECT records whether it recomputed a face row in FP64
    const StepCtl* ctl;
    const void* Un = nullptr;   // step-start state, planes like U (FOLD)
    int stage = 0;              // RK stage 1..3 (FOLD)
    // the launch's workgroups cover column blocks [bx0, bx0 + ncols) and row
    // groups [by0, by0 + nrows) (ncols, nrows = 0: to the grid's end)
    int64_t bx0 = 0, by0 = 0, ncols = 0, nrows = 0;
    // FOLD_BE / FOLD_BO: U, Un and D are blocked states of the padded grid
    // (blk_row / blk_col with nb lines per line row, components adjacent;
    // for the y sweep rows are y and lanes x, as in the transposed copy)
    int64_t nb = 0;
    int cfl_out = 0;            // FOLD_BO, stage 3: leave the CFL maxima in ctl->part
    // FOLD_BO: also write the ghost cells anchored in the workgroup's tile by
    // the rule bp (ghost_in_ysweep holds: every source in the same tile)
    int ghost = 0;
    BcParams bp{};
    float split_jump = SPLIT_JUMP_DEFAULT;   // DETECT: the FP64 threshold on relative jumps
};

// Shu-Osher coefficients of stage s: U_{s+1} = a Un + b Us + (c dt) L
// (Rie_prob_2d_rs.py:183-202; in stage 1 Us = Un, so a = 0, b = 1).
__host__ __device__ inline void stage_coeffs(int stage, double& a, double& b, double& c)
{
    if (stage == 1) { a = 0.0; b = 1.0; c = 1.0; }
    else if (stage == 2) { a = 0.75; b = 0.25; c = 0.25; }
    else { a = 1.0 / 3.0; b = 2.0 / 3.0; c = 2.0 / 3.0; }
}

template <typename T, int IN>
__device__ __forceinline__ T ld_state(const void* U, int64_t P, int k, int64_t idx)
{
    if constexpr (IN == IN_F64) {
        return (T)((const double*)U)[k * P + idx];
    } else if constexpr (IN == IN_F32) {
        return (T)((const float*)U)[k * P + idx];
    } else {
        const float* f = (const float*)U;
        return (T)((double)f[k * P + idx] + (double)f[(k + 4) * P + idx]);
    }
}

template <typename T> __device__ __forceinline__ bool bad_state(const T q[4], T gam)
{
    if (!(q[0] > T(0))) return true;
    T p = (gam - T(1)) * (q[3] - T(0.5) * (q[1] * q[1] + q[2] * q[2]) / q[0]);
    return !(p > T(0)) || !(p < T(3.0e38));
}

// Lanes of one wavefront exchange values through LDS inside the march.  The
// DS instructions of a wavefront execute in program order, so the only
// requirement is that the compiler keeps the accesses on their side of this
// point (the same three builtins as HIP's __syncwarp).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A lane of a packed sweep (sweep_body's PACK): its instance's blocked-state
// offset, its padded column in that instance, the instance's dt and whether
// the instance is active.
struct LanePack {
    int64_t off = 0, col = 0;
    double dt = 0.0;
    bool active = false;
};

// LDS of one sweep workgroup (declared by the kernel, so a kernel running
// both sweeps holds one copy).
template <typename T, int W> struct SweepLds {
    T p1[W][8][WAVE];     // pass-1 q_l (0..3), q_r (4..7) of the current face
    T pb[W][2][WAVE];     // p̄ ring (face parity)
    T xpb[W][2][WAVE];    // p̄ of the first [0] and last [1] owned face
    T xf[W][12][WAVE];    // deferred first face: HLLC (0..3), HLLE (4..7), F of the next face (8..11)
    T xF[W][4][WAVE];     // blended flux of the first owned face
    double wm[W];         // FOLD_BO, stage 3: per-wave CFL maxima
};

// FP64 recompute of one face row of a wavefront (the FP32-split sweep's
// fallback): pass 1 from the lane's stencil c on every lane, through the
// wavefront's pass-1 LDS rows p1, then pass 2 on the Gauss lanes whose face
// is flagged (take); g holds the lane's Gauss-point states l1, l2, q1, q2 (4
// each) and receives the FP64 ones where take is set.
template <typename T, int DIR, bool CONS>
__device__ __attribute__((noinline)) void face_row_fp64(const T (*c)[4], T gam, T (*p1)[WAVE], int l,
                                                        bool gauss_lane, bool take, T* g)
{
    wave_sync();
    T ql[4], qr[4];
    recon_normal(DIR, CONS, c, gam, ql, qr);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        p1[k][l] = ql[k];
        p1[4 + k][l] = qr[k];
    }
    wave_sync();
    if (gauss_lane && take) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            weno5_gq(p1[k][l - 2], p1[k][l - 1], p1[k][l], p1[k][l + 1], p1[k][l + 2], g[k], g[4 + k]);
            weno5_gq(p1[4 + k][l - 2], p1[4 + k][l - 1], p1[4 + k][l], p1[4 + k][l + 1], p1[4 + k][l + 2], g[8 + k],
                     g[12 + k]);
        }
    }
}

// Cells of one workgroup of W waves (see sweep_body), and the first cell of
// wave k of the group (k = W: the group's end).
__host__ __device__ inline int64_t group_cells(int W, int chunk, bool blend)
{
    return W > 1 ? (int64_t)W * chunk - (blend ? 3 : 1) : (int64_t)chunk;
}
__host__ __device__ inline int64_t wave_start(int W, int chunk, bool blend, int k)
{
    if (k <= 0) return 0;
    if (k >= W) return group_cells(W, chunk, blend);
    return (int64_t)k * chunk - (blend ? 1 : 0);
}

// One workgroup: column block bx, row group by of W consecutive chunks, one
// wavefront per chunk; `bid` indexes the flags (W per workgroup).
//
// Wave w owns the cells [r0, r1) of its chunk and computes the faces it owns,
// [r0, r1) plus, for the group's last wave, face r1.  The sensor of a face
// needs the p̄ of both normal neighbours, so the group's first wave also
// computes the halo face r0 - 1 and its last wave the halo face r1 + 1 (the
// reference's one-sided differences apply at the grid's first and last face
// instead).  Inside the group the neighbours' p̄ come through LDS: a wave
// defers the blend of its first face (it needs the previous wave's last p̄)
// and of its last face (it needs the next wave's first p̄) to the end of the
// march, and the flux differences of its first and last cell follow them.
// Halo cost per group: 2 faces.  With W > 1 the cells are split so that
// every wave marches the same number of faces: `chunk` cells for the middle
// waves, one fewer for the first (it adds the halo face r0 - 1) and, for the
// last, as many fewer as the faces it adds (r1, and the halo face r1 + 1 with
// the blend), so a group covers W * chunk - 3 cells (W * chunk - 1 without
// the blend); group_cells gives the count.  The waves of a group run on
// separate SIMDs of one CU, and equal marches keep each of them busy until
// the group's end.  Needs chunk >= 3 when W > 1 (the host enforces it).
//
// FOLD_E / FOLD_BE (x sweep of the fast FP64 step, which needs A.ctl for
// dt): D receives, in place of Dx, the part of the stage's combination this
// sweep can form, E = a Un + b Us - (c dt) Dx (stage_coeffs), so the combine
// (or the blocked y sweep) reads E instead of Dx, Un and Us.  Us of a cell
// finished inside the march is a row of the stencil window and Un is fetched
// one face ahead; the cells finished after the march reload both.
// FOLD_BO (y sweep of the blocked step): E comes through the Un path (A.Un)
// and D receives the new state fma(-(c dt), Dy, E), the operation the
// combine performs, so both steps give the same state bit for bit; in stage
// 3 the workgroup also leaves the maximum of c + |V| over its new cells in
// ctl->part and marks the step's advance pending, as the stage-3 combine.
// PACK (blocked folds, one wave, no CFL output): the lanes belong to the
// instances of a pack (BatchPack); the caller gives each lane its instance's
// blocked-state offset, its padded column in that instance (clamped into the
// pack), the instance's dt and whether it is active, in pk.  A.U, A.Un and
// A.D are then the batch's blocked states, A.N, A.nt the members' widths.
template <typename T, int IN, typename DT, int MODE, int DIR, bool CONS, bool BLEND, int W, int FOLD = FOLD_NONE,
          bool PACK = false>
__device__ __forceinline__ void sweep_body(const SweepArgs& A, int64_t bx, int64_t by, int64_t bid,
                                           SweepLds<T, W>& S, const LanePack& pk = LanePack{})
{
    static_assert(FOLD == FOLD_NONE || (IN == IN_F64 && MODE == MODE_PLAIN), "fold: FP64 sweeps only");
    static_assert(FOLD == FOLD_NONE || (DIR == DIR_X) == (FOLD != FOLD_BO), "fold: E from the x sweep, state from the y sweep");
    constexpr bool BLK = FOLD == FOLD_BE || FOLD == FOLD_BO;
    // the fast FP64 Gauss-point pass shares first differences between lanes
    constexpr bool GQ_DIFF = WX_FAST && MODE == MODE_PLAIN && sizeof(T) == 8;
    static_assert(!PACK || (BLK && W == 1), "packs: blocked folds, one wave per workgroup");
    if (!PACK && A.ctl && !A.ctl->active) return;
    auto& s_p1 = S.p1;
    auto& s_pb = S.pb;
    auto& s_xpb = S.xpb;
    auto& s_xf = S.xf;
    auto& s_xF = S.xF;

    const T gam = (T)A.gam, two_h = (T)A.two_h, jlo = (T)A.lo, jhi = (T)A.hi, dx = (T)A.dx;
#if WX_FAST
    // the fast variant multiplies by host-formed reciprocals of the constant divisors
    const T inv_two_h = (T)(1.0 / A.two_h), inv_dx = (T)(1.0 / A.dx);
    const T inv_span = (T)(1.0 / (A.hi - A.lo));
#define WX_DIV_2H(x) ((x) * inv_two_h)
#define WX_DIV_DX(x) ((x) * inv_dx)
#else
#define WX_DIV_2H(x) ((x) / two_h)
#define WX_DIV_DX(x) ((x) / dx)
    const T inv_span = T(0);   // unused by the exact variant
#endif
    // the wave index is uniform across the wave: keep it (and every row bound
    // derived from it) in scalar registers
    const int wv = W > 1 ? __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE)) : 0;
    const int l = (int)(threadIdx.x % WAVE);
    const int64_t c0 = bx * OUT_COLS;                       // padded column of lane 0
    const int64_t pc = PACK ? pk.col : c0 + l;              // this lane's padded column
    const int64_t t = pc - 3;                               // interior transverse index
    const int64_t pcl = PACK || pc < A.N ? pc : A.N - 1;    // clamped for loads
    const bool gauss_lane = l >= 2 && l < WAVE - 2;
    const bool out_lane = l >= 3 && l < WAVE - 3 && t >= 0 && t < A.nt && (!PACK || pk.active);
    const int lm = l > 0 ? l - 1 : 0, lp = l < WAVE - 1 ? l + 1 : WAVE - 1;   // sensor neighbours

    const int64_t g0 = by * group_cells(W, A.chunk, BLEND);
    const int64_t r0 = g0 + wave_start(W, A.chunk, BLEND, wv);
    const int64_t e1 = g0 + wave_start(W, A.chunk, BLEND, wv + 1);
    const int64_t r1 = e1 < A.nn ? e1 : A.nn;                         // cells [r0, r1)
    const bool empty = r0 >= A.nn;
    const bool last = wv == W - 1 || r1 == A.nn;                    // owns face r1
    const bool dfirst = BLEND && wv > 0;                             // first face blended at the end
    const int64_t fa = (BLEND && wv == 0 && r0 > 0) ? r0 - 1 : r0;
    const int64_t fb = last ? ((BLEND && r1 < A.nn) ? r1 + 1 : r1) : r1 - 1;   // inclusive
    // component stride: planes, or adjacent components in the blocked layout
    const int64_t P = BLK ? 1 : A.M * A.ld;
    const int64_t Pd = A.nn * A.ldd;
    DT* Dout = (DT*)A.D;
    bool bad = false;
    // element index of padded row r (normal) at this lane's (clamped) column:
    // row layout r ld + pcl; blocked, with rows x (x sweep) or y (y sweep)
    const int64_t lane_off = !BLK ? pcl : (PACK ? pk.off : 0) + (DIR == DIR_X ? blk_col(pcl) : blk_row(pcl, A.nb));
    auto at = [&](int64_t r) -> int64_t {
        if constexpr (!BLK) return r * A.ld + pcl;
        else return (DIR == DIR_X ? blk_row(r, A.nb) : blk_col(r)) + lane_off;
    };
    double wmax = 0.0;   // FOLD_BO, stage 3: max of c + |V| over this lane's new cells
    const bool cfl_max = !PACK && FOLD == FOLD_BO && A.stage == 3 && A.cfl_out;

    T F_prev[4], fc_prev[4], fe_prev[4];
    T pb_m2 = T(0), pb_m1 = T(0), pb_r0n = T(0);   // p̄ of faces a-2, a-1; p̄(r0 + 1)
#pragma unroll
    for (int k = 0; k < 4; ++k) { fc_prev[k] = T(0); fe_prev[k] = T(0); F_prev[k] = T(0); }

    double fa_c = 0.0, fb_c = 0.0, fc_c = 0.0;   // FOLD: a, b and c dt of the stage
    if constexpr (FOLD != FOLD_NONE) {
        stage_coeffs(A.stage, fa_c, fb_c, fc_c);
        fc_c *= PACK ? pk.dt : A.ctl->dt;
    }
    // Un read: the x fold from stage 2 on (stage 1: a = 0), E always
    const bool fold_un = FOLD == FOLD_BO || (FOLD != FOLD_NONE && A.stage > 1);
    // FOLD: Us and Un (FOLD_BO: E) of interior cell `cell` from memory
    auto load_cell = [&](int64_t cell, T us[4], T un[4]) {
        if constexpr (FOLD != FOLD_NONE) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                us[k] = FOLD == FOLD_BO ? T(0) : ld_state<T, IN>(A.U, P, k, at(cell + 3));
                un[k] = fold_un ? ld_state<T, IN>(A.Un, P, k, at(cell + 3)) : T(0);
            }
        }
    };
    // cell's flux difference; with FOLD its part of the stage combination
    // (us, un: the cell's Us and Un), in explicit fused operations so that
    // every cell rounds alike wherever the march finishes it
    auto put_d = [&](int64_t cell, const T Fr[4], const T Fl[4], const T* us, const T* un) {
        if (out_lane) {
            if constexpr (BLK) {
                // the cell's padded element (out lanes are never clamped)
                DT* d = Dout + at(cell + 3);
                double v[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const T dv = WX_DIV_DX(Fr[k] - Fl[k]);
                    if constexpr (FOLD == FOLD_BE) v[k] = fma(-T(fc_c), dv, fma(T(fa_c), un[k], T(fb_c) * us[k]));
                    else v[k] = fma(-T(fc_c), dv, un[k]);
                    d[k * P] = (DT)v[k];
                }
                if constexpr (FOLD == FOLD_BO && !PACK) {
                    if (cfl_max) wmax = nan_max(wmax, cfl_speed(v, A.gam, A.ctl->rho_floor));
                }
            } else {
                DT* d = Dout + cell * A.ldd + t;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const T dv = WX_DIV_DX(Fr[k] - Fl[k]);
                    if constexpr (FOLD == FOLD_NONE) d[k * Pd] = (DT)dv;
                    else d[k * Pd] = (DT)fma(-T(fc_c), dv, fma(T(fa_c), un[k], T(fb_c) * us[k]));
                }
            }
        }
    };
    // face f finished with flux F: cell f-1 lies between faces f-1 and f
    // (us, un: that cell's Us and Un when it is finished here)
    auto finish_face = [&](int64_t f, const T F[4], const T* us, const T* un) {
        if (f - 1 >= r0 && f - 1 < r1) {
            if (dfirst && f - 1 == r0) {
#pragma unroll
                for (int k = 0; k < 4; ++k) s_xf[wv][8 + k][l] = F[k];
            } else {
                put_d(f - 1, F, F_prev, us, un);
            }
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) F_prev[k] = F[k];
    };
    auto sensor = [&](const T* ring, T dn, T pbf) -> T {
        T dt_;
        if (t <= 0) dt_ = WX_DIV_2H(ring[lp] - ring[l]);
        else if (t >= A.nt - 1) dt_ = WX_DIV_2H(ring[l] - ring[lm]);
        else dt_ = WX_DIV_2H(ring[lp] - ring[lm]);
        // the reference's dpdx is along x: normal for x faces, transverse for y faces
        return DIR == DIR_X ? sensor_phi(dn, dt_, pbf, jlo, jhi, inv_span)
                            : sensor_phi(dt_, dn, pbf, jlo, jhi, inv_span);
    };

    if (!empty) {
        // stencil window rows a..a+5 (padded), 4 conserved components
        T w[6][4], nxt[4];
#pragma unroll
        for (int s = 0; s < 5; ++s)
#pragma unroll
            for (int k = 0; k < 4; ++k) w[s + 1][k] = ld_state<T, IN>(A.U, P, k, at(fa + s));
#pragma unroll
        for (int k = 0; k < 4; ++k) nxt[k] = ld_state<T, IN>(A.U, P, k, at(fa + 5));
        // FOLD: Un of the cell finished at face a, a - 2 with the blend (one
        // face behind) and a - 1 without, fetched one face ahead like the
        // window's next row (measured 0.4 % faster than at the top of the same
        // face); rows fa - 2 .. fb - 1 lie inside the padded planes
        constexpr int FIN = BLEND ? 2 : 1;
        T un_cur[4], un_nxt[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            un_cur[k] = T(0);
            un_nxt[k] = fold_un ? ld_state<T, IN>(A.Un, P, k, at(fa - FIN + 3)) : T(0);
        }

        for (int64_t a = fa; a <= fb; ++a) {
            WX_PHASE(window);
            // slide the window; row a+5 was fetched one face ahead, fetch row a+6
#pragma unroll
            for (int s = 0; s < 5; ++s)
#pragma unroll
                for (int k = 0; k < 4; ++k) w[s][k] = w[s + 1][k];
#pragma unroll
            for (int k = 0; k < 4; ++k) w[5][k] = nxt[k];
            if (a < fb) {
#pragma unroll
                for (int k = 0; k < 4; ++k) nxt[k] = ld_state<T, IN>(A.U, P, k, at(a + 6));
            }
            if (fold_un) {
#pragma unroll
                for (int k = 0; k < 4; ++k) un_cur[k] = un_nxt[k];
                if (a < fb) {
#pragma unroll
                    for (int k = 0; k < 4; ++k)
                        un_nxt[k] = ld_state<T, IN>(A.Un, P, k, at(a + 1 - FIN + 3));
                }
            }

            // pass 1 (normal reconstruction) on every lane
            WX_PHASE(pass1);
            T pv[8];        // this lane's q_l (0..3), q_r (4..7)
            bool jn = false;   // FP32-split: a large jump across this lane's normal stencil
            if constexpr (MODE == MODE_DETECT) jn = recon_normal_dev(DIR, CONS, w, gam, pv, pv + 4, A.split_jump);
            else recon_normal(DIR, CONS, w, gam, pv, pv + 4);
#pragma unroll
            for (int k = 0; k < 8; ++k) s_p1[wv][k][l] = pv[k];
            wave_sync();

            T fc[4], fe[4], pb = T(1);
            T l1[4], l2[4], q1[4], q2[4];
            bool jt = false;   // FP32-split: a large jump across this face's Gauss-point stencil
            WX_PHASE(pass2);
            if constexpr (GQ_DIFF) {
                // pass 2 from first differences (weno5_gq_d): each lane forms
                // q(l+1) - q(l) of its eight values and leaves it in its own
                // pass-1 slot (every lane's read of its right neighbour's slot
                // is issued before any lane's store: DS instructions of a
                // wavefront run in program order), then reads the differences
                // of lanes l-2, l-1 and l+1
                T (*sp)[WAVE] = s_p1[wv];
                const int lr = l < WAVE - 1 ? l + 1 : l;
                T dq[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) dq[k] = sp[k][lr] - pv[k];
                wave_sync();
#pragma unroll
                for (int k = 0; k < 8; ++k) sp[k][l] = dq[k];
                wave_sync();
                if (gauss_lane) {
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        weno5_gq_d(pv[k], sp[k][l - 2], sp[k][l - 1], dq[k], sp[k][l + 1], l1[k], l2[k]);
                        weno5_gq_d(pv[4 + k], sp[4 + k][l - 2], sp[4 + k][l - 1], dq[4 + k], sp[4 + k][l + 1], q1[k],
                                   q2[k]);
                    }
                }
            } else if (gauss_lane) {
                // pass 2: Gauss points along the face from lanes l-2..l+2
                const T(*sp)[WAVE] = s_p1[wv];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    if constexpr (MODE == MODE_DETECT) {
                        // density and energy also report their relative jump
                        if (k == 0 || k == 3) {
                            jt |= weno5_gq_dev<true>(sp[k][l - 2], sp[k][l - 1], sp[k][l], sp[k][l + 1], sp[k][l + 2],
                                                     l1[k], l2[k], A.split_jump);
                            jt |= weno5_gq_dev<true>(sp[4 + k][l - 2], sp[4 + k][l - 1], sp[4 + k][l],
                                                     sp[4 + k][l + 1], sp[4 + k][l + 2], q1[k], q2[k], A.split_jump);
                        } else {
                            weno5_gq_dev<false>(sp[k][l - 2], sp[k][l - 1], sp[k][l], sp[k][l + 1], sp[k][l + 2],
                                                l1[k], l2[k], A.split_jump);
                            weno5_gq_dev<false>(sp[4 + k][l - 2], sp[4 + k][l - 1], sp[4 + k][l], sp[4 + k][l + 1],
                                                sp[4 + k][l + 2], q1[k], q2[k], A.split_jump);
                        }
                    } else {
                        weno5_gq(sp[k][l - 2], sp[k][l - 1], sp[k][l], sp[k][l + 1], sp[k][l + 2], l1[k], l2[k]);
                        weno5_gq(sp[4 + k][l - 2], sp[4 + k][l - 1], sp[4 + k][l], sp[4 + k][l + 1],
                                 sp[4 + k][l + 2], q1[k], q2[k]);
                    }
                }
            }
            WX_PHASE(detect);
            if constexpr (MODE == MODE_DETECT) {
                // A Gauss-point state with rho <= 0, p <= 0 or not finite on any
                // lane: the wavefront recomputes this face row in FP64 (pass 1
                // from the registers' stencil, then pass 2), and the flagged
                // faces take the FP64 states, so that every face's result is
                // the same whatever tiling puts it in a row with others.
                // The same holds for a face whose Gauss-point stencil, or the
                // normal stencil of a face among lanes l-2..l+2 that feed its
                // Gauss points, spans a large jump (recon_normal_dev,
                // weno5_gq_dev).
                const uint64_t jm = __ballot(jn ? 1 : 0);
                const bool jnear = ((jm >> (gauss_lane ? l - 2 : 0)) & 0x1full) != 0ull;
                const bool badf = gauss_lane && (jnear || jt || bad_state(l1, gam) || bad_state(l2, gam) ||
                                                 bad_state(q1, gam) || bad_state(q2, gam));
                if (__ballot(badf ? 1 : 0) != 0ull) {
                    bad = true;
                    // only the flagged faces take the FP64 states: a face's
                    // choice depends on its own states alone (a neighbour's
                    // flag at the wave's edge lanes is not known to the wave,
                    // and the sensor reads the edge lanes' p̄)
                    const bool take = badf;
                    // an out-of-line call with copies of the stencil and the
                    // results, so the FP32 path's registers are not shared
                    // with the FP64 recompute
                    T wc[6][4], g[16];
#pragma unroll
                    for (int r = 0; r < 6; ++r)
#pragma unroll
                        for (int k = 0; k < 4; ++k) wc[r][k] = w[r][k];
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        g[k] = l1[k]; g[4 + k] = l2[k]; g[8 + k] = q1[k]; g[12 + k] = q2[k];
                    }
                    face_row_fp64<T, DIR, CONS>(wc, gam, s_p1[wv], l, gauss_lane, take, g);
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        l1[k] = g[k]; l2[k] = g[4 + k]; q1[k] = g[8 + k]; q2[k] = g[12 + k];
                    }
                }
            }
            WX_PHASE(fluxes);
            if (gauss_lane) {
                T fa_[4], fb_[4], fcc[4], fd[4];
                flux_pt2(DIR, l1, l2, gam, fa_, fb_);
                flux_pt2(DIR, q1, q2, gam, fcc, fd);
                T QL[4], FL[4], QR[4], FR[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    FL[k] = T(0.5) * (fa_[k] + fb_[k]);
                    FR[k] = T(0.5) * (fd[k] + fcc[k]);
                    QL[k] = T(0.5) * (l1[k] + l2[k]);
                    QR[k] = T(0.5) * (q1[k] + q2[k]);
                }
                WX_PHASE(riemann);
                FaceState<T> st = face_speeds(DIR, QL, QR, gam);
                hlle_from(st, QL, FL, QR, FR, fe);
                if (BLEND) {
                    hllc_from(DIR, st, QL, FL, QR, FR, fc);
                    // floored pressures of QL and QR (riemann.rs:421-430), as formed by face_speeds
                    pb = T(0.5) * (st.p_l + st.p_r);
                }
            } else {
#pragma unroll
                for (int k = 0; k < 4; ++k) { fc[k] = T(0); fe[k] = T(0); }
            }

            WX_PHASE(sensor_blend);
            if (!BLEND) {
                if (W > 1 && a == r0 && wv > 0) {
#pragma unroll
                    for (int k = 0; k < 4; ++k) s_xF[wv][k][l] = fe[k];
                }
                finish_face(a, fe, w[2], un_cur);
            } else {
                s_pb[wv][a & 1][l] = pb;
                if (dfirst && a == r0) {
                    s_xpb[wv][0][l] = pb;
#pragma unroll
                    for (int k = 0; k < 4; ++k) { s_xf[wv][k][l] = fc[k]; s_xf[wv][4 + k][l] = fe[k]; }
                }
                if (dfirst && a == r0 + 1) pb_r0n = pb;
                wave_sync();
                if (a > fa) {
                    // sensor and blend of face f = a-1 (riemann.rs:360-407)
                    const int64_t f = a - 1;
                    if (!(dfirst && f == r0)) {
                        T dn = f == 0 ? WX_DIV_2H(pb - pb_m1) : WX_DIV_2H(pb - pb_m2);
                        T phi = sensor(s_pb[wv][f & 1], dn, pb_m1);
                        T F[4];
#pragma unroll
                        for (int k = 0; k < 4; ++k) F[k] = phi * fc_prev[k] + (T(1) - phi) * fe_prev[k];
                        if (f >= r0) finish_face(f, F, w[1], un_cur);
                    }
                }
                pb_m2 = pb_m1;
                pb_m1 = pb;
#pragma unroll
                for (int k = 0; k < 4; ++k) { fc_prev[k] = fc[k]; fe_prev[k] = fe[k]; }
            }
            wave_sync();
            WX_PHASE(end);
        }

        if (BLEND && fb == A.nn) {
            // last face of the grid: one-sided normal difference, no face nn+1
            const int64_t f = fb;
            T dn = WX_DIV_2H(pb_m1 - pb_m2);
            T phi = sensor(s_pb[wv][f & 1], dn, pb_m1);
            T F[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) F[k] = phi * fc_prev[k] + (T(1) - phi) * fe_prev[k];
            T us[4], un[4];
            load_cell(f - 1, us, un);
            finish_face(f, F, us, un);
        }
        if (BLEND && !last) s_xpb[wv][1][l] = pb_m1;     // p̄ of face r1 - 1
    }

    if constexpr (W > 1) {
        // FOLD: Us and Un of the cells finished below, fetched before the
        // barrier so that the wait covers their latency
        T us0[4], un0[4], us1[4], un1[4], us2[4], un2[4];
        if (!empty) {
            if (dfirst) load_cell(r0, us0, un0);
            if (BLEND && !last) load_cell(r1 - 2, us1, un1);
            if (!last) load_cell(r1 - 1, us2, un2);
        }
        __syncthreads();
        T FLast[4];
        if (!empty && BLEND) {
            if (dfirst) {
                // face r0: previous wave's last p̄ below, p̄(r0 + 1) above
                T dn = WX_DIV_2H(pb_r0n - s_xpb[wv - 1][1][l]);
                T pb0 = s_xpb[wv][0][l];
                T phi = sensor(s_xpb[wv][0], dn, pb0);
                T F0[4], F1[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    F0[k] = phi * s_xf[wv][k][l] + (T(1) - phi) * s_xf[wv][4 + k][l];
                    F1[k] = s_xf[wv][8 + k][l];
                    s_xF[wv][k][l] = F0[k];
                }
                put_d(r0, F1, F0, us0, un0);
            }
            if (!last) {
                // face r1 - 1: p̄(r1 - 2) below, next wave's first p̄ above
                T dn = WX_DIV_2H(s_xpb[wv + 1][0][l] - pb_m2);
                T phi = sensor(s_pb[wv][(r1 - 1) & 1], dn, pb_m1);
#pragma unroll
                for (int k = 0; k < 4; ++k) FLast[k] = phi * fc_prev[k] + (T(1) - phi) * fe_prev[k];
                put_d(r1 - 2, FLast, F_prev, us1, un1);
            }
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) FLast[k] = F_prev[k];
        }
        __syncthreads();
        if (!empty && !last) {
            // cell r1 - 1 closes with the next wave's first face
            T Fn[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) Fn[k] = s_xF[wv + 1][k][l];
            put_d(r1 - 1, Fn, FLast, us2, un2);
        }
    }
    if constexpr (FOLD == FOLD_BO && !PACK) {
        if (A.ghost) {
            // The ghost cells anchored in this workgroup's tile (x rows
            // [x0, x1) on the lanes, y cells [y0, y1) of the march): its own
            // padded range widened by the 3 ghost layers on the grid's edges,
            // as the RK combine's ghost-on-store; each from its source cell of
            // the new state, which this workgroup has just written (the
            // barrier makes every wave's stores visible to the others).
            __syncthreads();
            const int64_t nx = A.nt, ny = A.nn;
            const int64_t x0 = c0, x1 = c0 + OUT_COLS, y0 = g0, y1 = g0 + group_cells(W, A.chunk, BLEND);
            const bool e0 = x0 == 0, e1 = x1 >= nx, f0 = y0 == 0, f1 = y1 >= ny;
            if (e0 || e1 || f0 || f1) {
                const int64_t ia = e0 ? 0 : x0 + 3, ib = e1 ? nx + 5 : x1 + 2;
                const int64_t ja = f0 ? 0 : y0 + 3, jb = f1 ? ny + 5 : y1 + 2;
                const int64_t ci = ib - ia + 1, cj = jb - ja + 1;
                const int64_t ri = e0 ? 3 : 0, rb = e1 ? 3 : 0;       // ghost rows at the low / high x end
                const int64_t gj = (f0 ? 3 : 0) + (f1 ? 3 : 0);       // ghost columns
                // ghost rows over the tile's whole column range, then ghost
                // columns over its interior rows
                const int64_t nrow = (ri + rb) * cj, n = nrow + (ci - ri - rb) * gj;
                const double ts = A.bp.kind == BC_DMR_STD ? bc_time(A.ctl, A.bp, A.stage) : 0.0;
                double* const St = (double*)A.D;
                for (int64_t q = threadIdx.x; q < n; q += WAVE * W) {
                    int64_t i, j;
                    if (q < nrow) {
                        const int64_t r = q / cj;
                        i = r < ri ? ia + r : nx + 3 + (r - ri);
                        j = ja + q % cj;
                    } else {
                        const int64_t u = q - nrow, c = u % gj;
                        i = ia + ri + u / gj;
                        j = (f0 && c < 3) ? c : ny + 3 + (c - (f0 ? 3 : 0));
                    }
                    const BcSrc b = bc_source(i, j, nx, ny, A.bp, ts);
                    const int64_t so = blk_row(b.si, A.nb) + blk_col(b.sj), go = blk_row(i, A.nb) + blk_col(j);
                    double v[4];
                    bc_value(b, A.bp, [&](int k) { return St[so + k]; }, v);
#pragma unroll
                    for (int k = 0; k < 4; ++k) St[go + k] = v[k];
                }
            }
        }
        if (cfl_max) {
            // the workgroup's maximum into ctl->part[bid] (the launch's grid
            // has gridDim.x * gridDim.y workgroups, bid their row-major index)
            const double m = wave_max(wmax);
            StepCtl* cw = const_cast<StepCtl*>(A.ctl);
            if constexpr (W > 1) {
                if (l == 0) S.wm[wv] = m;
                __syncthreads();
            }
            if (threadIdx.x == 0) {
                double g = m;
                if constexpr (W > 1) {
#pragma unroll
                    for (int v = 1; v < W; ++v) g = nan_max(g, S.wm[v]);
                }
                cw->part[bid] = g;
                if (bid == 0) {
                    cw->nparts = (int)(gridDim.x * gridDim.y);
                    cw->pending = 1;
                }
            }
        }
    }
    if (MODE == MODE_DETECT) {
        if (l == 0) A.flags[bid * W + wv] = bad ? 1 : 0;
    }
}

#undef WX_DIV_2H
#undef WX_DIV_DX

// Two wavefronts per SIMD (at most 256 VGPRs): the sweeps are VALU-bound and
// need the second wave to cover LDS and memory latency.
template <typename T, int IN, typename DT, int MODE, int DIR, bool CONS, bool BLEND, int W, int FOLD = FOLD_NONE>
__global__ void __launch_bounds__(WAVE * W) __attribute__((amdgpu_waves_per_eu(2))) k_sweep(SweepArgs A)
{
    __shared__ SweepLds<T, W> S;
    sweep_body<T, IN, DT, MODE, DIR, CONS, BLEND, W, FOLD>(A, A.bx0 + blockIdx.x, A.by0 + blockIdx.y,
                                                           (int64_t)blockIdx.y * gridDim.x + blockIdx.x, S);
}

// Both sweeps of a stage in one launch: workgroups [0, nbx) run the x sweep,
// the rest the y sweep (they read different planes and write different D
// arrays).  The two halves fill each other's tail, which on grids of about
// one wave round per sweep is most of the gain.  Flag slots follow the
// workgroup number, so the x and y halves use disjoint ranges.  FOLD applies
// to the x half.
template <typename T, int IN, typename DT, int MODE, bool CONS, bool BLEND, int W, bool FOLD = false>
__global__ void __launch_bounds__(WAVE * W) __attribute__((amdgpu_waves_per_eu(2)))
k_sweep_xy(SweepArgs Ax, SweepArgs Ay, unsigned nbx, unsigned colsx, unsigned colsy)
{
    __shared__ SweepLds<T, W> S;
    const unsigned b = blockIdx.x;
    if (b < nbx) {
        sweep_body<T, IN, DT, MODE, DIR_X, CONS, BLEND, W, FOLD>(Ax, Ax.bx0 + b % colsx, Ax.by0 + b / colsx, b, S);
    } else {
        const unsigned c = b - nbx;
        sweep_body<T, IN, DT, MODE, DIR_Y, CONS, BLEND, W>(Ay, Ay.bx0 + c % colsy, Ay.by0 + c / colsy, b, S);
    }
}

// ------------------------------------------------------------------------
// RK stage combine with L = -(Dx + Dy^T), both orientations written.
// Tile of 32 x 32 interior cells, 256 threads; Dy^T and the transposed
// output go through LDS.
// ------------------------------------------------------------------------
constexpr int TT = 32;

// One tile of TT rows (x) by TJ columns (y), TJ = 32 or 64: tile row ti,
// tile column tj.  The four components' Dy^T tiles are staged in LDS
// together, so each thread has all loads of its cells in flight at once.
// A thread overwrites in LDS exactly the Dy^T entries it read with the new
// state, so the tile needs two barriers; the transposed state then leaves
// LDS along x.  Wider tiles give the x-major streams 512 B rows.
// Row pitches: ld of Un, Us, out; ldT of outT; ldx of Dx; ldyT of DyT.
// CFL (the final stage of a step on the fused FP64 engine): also reduce the
// next step's wave speed c + |V| over the tile's new cells into ctl->part[t]
// (t the tile's index among ntiles) and mark the step's advance pending; the
// next step's first launch finishes both.
// FOLD: Dx holds the x sweep's part E of the combination (the sweep's FOLD);
// the new state is E - (c dt) Dy.
// GHOST (ghost_on_store): a tile on the grid's edge also writes, in both
// orientations, the ghost cells whose nearest interior cell it holds, from
// the new states in LDS by the solver's rule bp (bc_source / bc_value, as
// the ghost-fill kernel).
template <class St, typename DT, int TJ, bool CFL = false, bool FOLD = false, bool GHOST = false>
__device__ __forceinline__ void combine_tile(const DT* __restrict__ Dx, const DT* __restrict__ DyT, St Un,
                                             St Us, St out, St outT, int64_t nx, int64_t ny, int64_t ld,
                                             int64_t ldT, int64_t ldx, int64_t ldyT, int stage,
                                             const StepCtl* ctl, int64_t ti, int64_t tj, double gam = 0.0,
                                             int64_t t = 0, int64_t ntiles = 0, BcParams bp = BcParams{})
{
    if (ctl && !ctl->active) return;
    double wmax = 0.0;
    __shared__ double wpart[4];   // per-wave CFL maxima (CFL only)
    __shared__ double tile[4][TJ][TT + 1];   // [k][j - j0][i - i0]
    const int64_t i0 = ti * TT, j0 = tj * TJ;
    const int64_t PX = nx * ldx, PY = ny * ldyT;
    const double dt = ctl ? ctl->dt : 0.0;
    double fa_c, fb_c, fc_c;   // FOLD: c dt of the stage
    stage_coeffs(stage, fa_c, fb_c, fc_c);
    fc_c *= dt;
    // Dy^T[k][j][i]: rows j, contiguous i (32 threads per row)
    {
        const int c = threadIdx.x & (TT - 1), r0 = threadIdx.x / TT;
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int r = r0; r < TJ; r += 256 / TT) {
                const int64_t j = j0 + r, i = i0 + c;
                if (j < ny && i < nx) tile[k][r][c] = (double)DyT[k * PY + j * ldyT + i];
            }
    }
    __syncthreads();
    // x-major cells: TJ threads per row, contiguous j
    {
        const int tx = threadIdx.x & (TJ - 1), ty = threadIdx.x / TJ;
#pragma unroll
        for (int r = ty; r < TT; r += 256 / TJ) {
            const int64_t i = i0 + r, j = j0 + tx;
            if (i < nx && j < ny) {
                const int64_t o = (i + 3) * ld + (j + 3);
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    double v;
                    if constexpr (!FOLD) {
                        const double l = -((double)Dx[k * PX + i * ldx + j] + tile[k][tx][r]);
                        const double qn = Un.ld(k, o);
                        if (stage == 1) v = qn + dt * l;
                        else if (stage == 2) v = 0.75 * qn + 0.25 * Us.ld(k, o) + (0.25 * dt) * l;
                        else v = (1.0 / 3.0) * qn + (2.0 / 3.0) * Us.ld(k, o) + ((2.0 / 3.0) * dt) * l;
                    } else {
                        v = fma(-fc_c, tile[k][tx][r], (double)Dx[k * PX + i * ldx + j]);
                    }
                    out.st(k, o, v);
                    tile[k][tx][r] = v;
                }
            }
        }
    }
    __syncthreads();
    // outT[k][j+3][i+3], contiguous i
    {
        const int c = threadIdx.x & (TT - 1), r0 = threadIdx.x / TT;
        if constexpr (CFL) {
            // the transposed stores visit every new cell once: take its c + |V| here,
            // away from the x-major loop's loads and stores
#pragma unroll
            for (int r = r0; r < TJ; r += 256 / TT) {
                const int64_t j = j0 + r, i = i0 + c;
                if (j < ny && i < nx) {
                    double qv[4];
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        outT.st(k, (j + 3) * ldT + (i + 3), tile[k][r][c]);
                        qv[k] = St::stored(tile[k][r][c]);   // the state as stored (FP32-split: hi + lo)
                    }
                    wmax = nan_max(wmax, cfl_speed(qv, gam, ctl->rho_floor));
                }
            }
            wmax = wave_max(wmax);
            if ((threadIdx.x & 63) == 0) wpart[threadIdx.x >> 6] = wmax;
            __syncthreads();
            if (threadIdx.x == 0) {
                StepCtl* cw = const_cast<StepCtl*>(ctl);
                cw->part[t] = nan_max(nan_max(wpart[0], wpart[1]), nan_max(wpart[2], wpart[3]));
                if (t == 0) {
                    cw->nparts = (int)ntiles;
                    cw->pending = 1;
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k)
#pragma unroll
                for (int r = r0; r < TJ; r += 256 / TT) {
                    const int64_t j = j0 + r, i = i0 + c;
                    if (j < ny && i < nx) outT.st(k, (j + 3) * ldT + (i + 3), tile[k][r][c]);
                }
        }
    }
    if constexpr (GHOST) {
        // padded rows / columns of the ghost cells anchored in this tile: its
        // own, widened by the 3 ghost layers on the grid's edges (the tile is
        // read-only in LDS from the barrier above on)
        const bool e0 = ti == 0, e1 = i0 + TT >= nx, f0 = tj == 0, f1 = j0 + TJ >= ny;
        if (e0 || e1 || f0 || f1) {
            const int ia = e0 ? 0 : (int)i0 + 3, ib = e1 ? (int)nx + 5 : (int)i0 + TT + 2;
            const int ja = f0 ? 0 : (int)j0 + 3, jb = f1 ? (int)ny + 5 : (int)j0 + TJ + 2;
            // the ghost cells only: ghost rows over the full column range
            // (top, bottom), then ghost columns over the interior rows (left,
            // right)
            const int ci = ib - ia + 1, cj = jb - ja + 1;
            const int ri = ((int)i0 + 3 > ia ? 3 : 0), rb = (ib > (int)nx + 2 ? 3 : 0);   // ghost rows above / below
            const int gj = (f0 ? 3 : 0) + (f1 ? 3 : 0);                                    // ghost columns
            const int nrow = (ri + rb) * cj, n = nrow + (ci - ri - rb) * gj;
            const double ts = bp.kind == BC_DMR_STD ? bc_time(ctl, bp, stage) : 0.0;
            for (int q = threadIdx.x; q < n; q += 256) {
                int64_t i, j;
                if (q < nrow) {
                    const int r = q / cj;
                    i = r < ri ? ia + r : (int64_t)nx + 3 + (r - ri);
                    j = ja + q % cj;
                } else {
                    const int gs = gj > 0 ? gj : 1;   // (gj = 0 leaves this branch unused)
                    const int u = q - nrow, c = u % gs;
                    i = ia + ri + u / gs;
                    j = (f0 && c < 3) ? c : (int64_t)ny + 3 + (c - (f0 ? 3 : 0));
                }
                const BcSrc b = bc_source(i, j, nx, ny, bp, ts);
                double v[4];
                bc_value(b, bp, [&](int k) { return tile[k][b.sj - 3 - j0][b.si - 3 - i0]; }, v);
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    out.st(k, i * ld + j, v[k]);
                    outT.st(k, j * ldT + i, v[k]);
                }
            }
        }
    }
}

// Workgroups go to the 8 XCDs round-robin in launch order.  Workgroup
// b = 8 k + x takes tile k of band x, where the row-major tile order is cut
// into 8 contiguous bands (the first ntiles % 8 of them one tile longer): each
// XCD's L2 then holds a contiguous band, and the cache lines two neighbouring
// tiles share (rows are 32 doubles, 3 off a line boundary) are fetched once.
template <class St, typename DT, int TJ, bool CFL = false, bool FOLD = false, bool GHOST = false>
__global__ void __launch_bounds__(256) k_combine(const DT* __restrict__ Dx, const DT* __restrict__ DyT,
                                                 St Un, St Us, St out, St outT, int64_t nx, int64_t ny,
                                                 int64_t ld, int64_t ldT, int64_t ldx, int64_t ldyT,
                                                 int stage, const StepCtl* ctl, double gam, BcParams bp,
                                                 int64_t ti0, int64_t tiles_i)
{
    // this launch holds tile rows [ti0, ti0 + gridDim.y) of the tiles_i rows
    const int64_t tiles_j = gridDim.x, nloc = (int64_t)gridDim.x * gridDim.y;
    const int64_t b = (int64_t)blockIdx.y * tiles_j + blockIdx.x;
    const int64_t q = nloc / 8, r = nloc % 8, x = b % 8, k = b / 8;
    const int64_t t = ti0 * tiles_j + x * q + (x < r ? x : r) + k;
    combine_tile<St, DT, TJ, CFL, FOLD, GHOST>(Dx, DyT, Un, Us, out, outT, nx, ny, ld, ldT, ldx, ldyT, stage, ctl,
                                               t / tiles_j, t % tiles_j, gam, t, tiles_i * tiles_j, bp);
}

// cfl: the step's last combine also leaves the next step's CFL maxima (one
// per tile, combine_tiles(p) of them) in ctl->part and the advance pending.
// FOLD: Dx holds the x sweep's part of the combination (the sweep's FOLD).
// ghost: the combine also writes the ghost cells (ghost_on_store(p)).
// Tile rows [ti0, ti1) (ti1 < 0: to the last row).
template <class St, typename DT, int TJ, bool CFL, bool FOLD, bool GHOST>
static void launch_combine_t(const SweepPlan& p, const DT* Dx, const DT* DyT, St Un, St Us, St out, St outT,
                             int stage, const StepCtl* ctl, hipStream_t s, int64_t ti0, int64_t ti1)
{
    const int64_t tiles_i = (p.nx + TT - 1) / TT;
    if (ti1 < 0) ti1 = tiles_i;
    if (ti1 <= ti0) return;
    const dim3 g((unsigned)((p.ny + TJ - 1) / TJ), (unsigned)(ti1 - ti0));
    k_combine<St, DT, TJ, CFL, FOLD, GHOST><<<g, 256, 0, s>>>(Dx, DyT, Un, Us, out, outT, p.nx, p.ny, p.ld, p.ldT,
                                                             p.ldx, p.ldyT, stage, ctl, p.gamma, p.bp, ti0,
                                                             tiles_i);
}

template <class St, typename DT, int TJ, bool FOLD>
static void launch_combine_f(const SweepPlan& p, const DT* Dx, const DT* DyT, St Un, St Us, St out, St outT,
                             int stage, const StepCtl* ctl, hipStream_t s, bool cfl, bool ghost, int64_t ti0,
                             int64_t ti1)
{
    if (cfl) {
        if (ghost) launch_combine_t<St, DT, TJ, true, FOLD, true>(p, Dx, DyT, Un, Us, out, outT, stage, ctl, s, ti0, ti1);
        else launch_combine_t<St, DT, TJ, true, FOLD, false>(p, Dx, DyT, Un, Us, out, outT, stage, ctl, s, ti0, ti1);
    } else {
        if (ghost) launch_combine_t<St, DT, TJ, false, FOLD, true>(p, Dx, DyT, Un, Us, out, outT, stage, ctl, s, ti0, ti1);
        else launch_combine_t<St, DT, TJ, false, FOLD, false>(p, Dx, DyT, Un, Us, out, outT, stage, ctl, s, ti0, ti1);
    }
}

template <class St, typename DT, bool FOLD = false>
static void launch_combine(const SweepPlan& p, const DT* Dx, const DT* DyT, St Un, St Us, St out, St outT,
                           int stage, const StepCtl* ctl, hipStream_t s, bool cfl = false, bool ghost = false,
                           int64_t ti0 = 0, int64_t ti1 = -1)
{
    if (p.combine_cols == 64)
        launch_combine_f<St, DT, 64, FOLD>(p, Dx, DyT, Un, Us, out, outT, stage, ctl, s, cfl, ghost, ti0, ti1);
    else
        launch_combine_f<St, DT, 32, FOLD>(p, Dx, DyT, Un, Us, out, outT, stage, ctl, s, cfl, ghost, ti0, ti1);
}

// Plain RHS from the two sweeps (for tests): L = -(Dx + Dy^T), SoA4 interior
// (unpitched) from Dx (row pitch ldx) and Dy^T (row pitch ldyT).
template <typename DT>
__global__ void k_rhs_out(const DT* Dx, const DT* DyT, double* L, int64_t nx, int64_t ny, int64_t ldx,
                          int64_t ldyT)
{
    int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= nx * ny) return;
    int64_t i = t / ny, j = t - i * ny;
    for (int k = 0; k < 4; ++k)
        L[k * nx * ny + t] = -((double)Dx[k * nx * ldx + i * ldx + j] + (double)DyT[k * ny * ldyT + j * ldyT + i]);
}

// Padded plane-set transpose: out[k][j][i] = in[k][i][j] for NP planes of
// M x N (in: row pitch ld_in, plane M ld_in; out: row pitch ld_out, plane
// N ld_out).
template <typename E, int NP>
__global__ void __launch_bounds__(256) k_transpose(const E* in, E* out, int64_t M, int64_t N, int64_t ld_in,
                                                   int64_t ld_out)
{
    __shared__ E tile[TT][TT + 1];
    const int tx = threadIdx.x & (TT - 1), ty = threadIdx.x / TT;
    const int64_t i0 = (int64_t)blockIdx.y * TT, j0 = (int64_t)blockIdx.x * TT;
    const int64_t Pi = M * ld_in, Po = N * ld_out;
    for (int k = 0; k < NP; ++k) {
        for (int r = ty; r < TT; r += 8) {
            int64_t i = i0 + r, j = j0 + tx;
            if (i < M && j < N) tile[r][tx] = in[k * Pi + i * ld_in + j];
        }
        __syncthreads();
        for (int r = ty; r < TT; r += 8) {
            int64_t j = j0 + r, i = i0 + tx;
            if (i < M && j < N) out[k * Po + j * ld_out + i] = tile[tx][r];
        }
        __syncthreads();
    }
}

// Column blocks and row groups of a sweep launch (its ncols / nrows, else
// the rest of the grid from bx0 / by0).
static int64_t sweep_cols(const SweepArgs& a)
{
    return a.ncols ? a.ncols : (a.nt + OUT_COLS - 1) / OUT_COLS - a.bx0;
}
static int64_t sweep_rows(const SweepArgs& a, int W, bool blend)
{
    const int64_t rows = group_cells(W, a.chunk, blend);
    return a.nrows ? a.nrows : (a.nn + rows - 1) / rows - a.by0;
}

template <typename T, int IN, typename DT, int MODE, int DIR, int W, int FOLD = FOLD_NONE>
static void launch_sweep_w(const SweepArgs& a, int cons, int blend, hipStream_t s)
{
    dim3 grid((unsigned)sweep_cols(a), (unsigned)sweep_rows(a, W, blend != 0));
    if (cons) {
        if (blend) k_sweep<T, IN, DT, MODE, DIR, true, true, W, FOLD><<<grid, WAVE * W, 0, s>>>(a);
        else k_sweep<T, IN, DT, MODE, DIR, true, false, W, FOLD><<<grid, WAVE * W, 0, s>>>(a);
    } else {
        if (blend) k_sweep<T, IN, DT, MODE, DIR, false, true, W, FOLD><<<grid, WAVE * W, 0, s>>>(a);
        else k_sweep<T, IN, DT, MODE, DIR, false, false, W, FOLD><<<grid, WAVE * W, 0, s>>>(a);
    }
}

template <typename T, int IN, typename DT, int MODE, int DIR, int FOLD = FOLD_NONE>
static void launch_sweep(const SweepArgs& a, int cons, int blend, hipStream_t s)
{
    if (a.waves == 4) launch_sweep_w<T, IN, DT, MODE, DIR, 4, FOLD>(a, cons, blend, s);
    else if (a.waves == 2) launch_sweep_w<T, IN, DT, MODE, DIR, 2, FOLD>(a, cons, blend, s);
    else launch_sweep_w<T, IN, DT, MODE, DIR, 1, FOLD>(a, cons, blend, s);
}

template <typename T, int IN, typename DT, int MODE, int W, bool FOLD = false>
static void launch_sweep_xy_w(const SweepArgs& ax, const SweepArgs& ay, int cons, int blend, hipStream_t s)
{
    const unsigned colsx = (unsigned)sweep_cols(ax), colsy = (unsigned)sweep_cols(ay);
    const unsigned nbx = colsx * (unsigned)sweep_rows(ax, W, blend != 0);
    const unsigned nby = colsy * (unsigned)sweep_rows(ay, W, blend != 0);
    const dim3 grid(nbx + nby);
    if (cons) {
        if (blend)
            k_sweep_xy<T, IN, DT, MODE, true, true, W, FOLD><<<grid, WAVE * W, 0, s>>>(ax, ay, nbx, colsx, colsy);
        else
            k_sweep_xy<T, IN, DT, MODE, true, false, W, FOLD><<<grid, WAVE * W, 0, s>>>(ax, ay, nbx, colsx, colsy);
    } else {
        if (blend)
            k_sweep_xy<T, IN, DT, MODE, false, true, W, FOLD><<<grid, WAVE * W, 0, s>>>(ax, ay, nbx, colsx, colsy);
        else
            k_sweep_xy<T, IN, DT, MODE, false, false, W, FOLD><<<grid, WAVE * W, 0, s>>>(ax, ay, nbx, colsx, colsy);
    }
}

// x then y sweep of one stage: one launch when both use the same number of
// wavefronts per workgroup, else two.  FOLD applies to the x sweep.
template <typename T, int IN, typename DT, int MODE, bool FOLD = false>
static void launch_sweeps(const SweepArgs& ax, const SweepArgs& ay, int cons, int blend, hipStream_t s)
{
    if (ax.waves != ay.waves) {
        launch_sweep<T, IN, DT, MODE, DIR_X, FOLD>(ax, cons, blend, s);
        launch_sweep<T, IN, DT, MODE, DIR_Y>(ay, cons, blend, s);
    } else if (ax.waves == 4) {
        launch_sweep_xy_w<T, IN, DT, MODE, 4, FOLD>(ax, ay, cons, blend, s);
    } else if (ax.waves == 2) {
        launch_sweep_xy_w<T, IN, DT, MODE, 2, FOLD>(ax, ay, cons, blend, s);
    } else {
        launch_sweep_xy_w<T, IN, DT, MODE, 1, FOLD>(ax, ay, cons, blend, s);
    }
}

static SweepArgs args_x(const SweepPlan& p, const void* U, void* D, int* flags, const StepCtl* ctl)
{
    SweepArgs a{U,     p.nx + 6,   p.ny + 6, p.ld,    p.ldx,     p.nx,    p.ny,  p.chunk_x,
                p.waves_x, p.gamma, p.dx,     p.two_h_x, p.jp_lo, p.jp_hi, D,     flags, ctl};
    a.split_jump = p.split_jump;
    return a;
}

static SweepArgs args_y(const SweepPlan& p, const void* UT, void* D, int* flags, const StepCtl* ctl)
{
    SweepArgs a{UT,    p.ny + 6,   p.nx + 6, p.ldT,   p.ldyT,    p.ny,    p.nx,  p.chunk_y,
                p.waves_y, p.gamma, p.dx,     p.two_h_y, p.jp_lo, p.jp_hi, D,     flags, ctl};
    a.split_jump = p.split_jump;
    return a;
}

// ---- FP64 ---------------------------------------------------------------
static void h_sweeps(const SweepPlan& p, const double* Us, const double* UsT, double* Dx, double* DyT,
                     const StepCtl* ctl, hipStream_t s)
{
    const int blend = p.force_hlle ? 0 : 1;
    launch_sweeps<double, IN_F64, double, MODE_PLAIN>(args_x(p, Us, Dx, nullptr, ctl),
                                                      args_y(p, UsT, DyT, nullptr, ctl), p.cons, blend, s);
}

#if WX_FAST
// Both sweeps of a stage with the x sweep's fold (needs ctl for dt), then the
// combine that completes the stage.
static void h_stage_fold(const SweepPlan& p, const double* Us, const double* UsT, const double* Un, double* Dx,
                         double* DyT, double* out, double* outT, int stage, const StepCtl* ctl, hipStream_t s,
                         bool sweeps, bool combine)
{
    const int blend = p.force_hlle ? 0 : 1;
    const int64_t P = p.P();
    SweepArgs ax = args_x(p, Us, Dx, nullptr, ctl);
    ax.Un = Un;
    ax.stage = stage;
    const SweepArgs ay = args_y(p, UsT, DyT, nullptr, ctl);
    const StF64 un{const_cast<double*>(Un), P}, us{const_cast<double*>(Us), P}, o{out, P}, oT{outT, p.PT()};
    if (sweeps && combine && p.parts > 1 && p.side) {
        // Slabs of x rows: the sweeps of slab k run on s while the combine of
        // the slabs finished before runs on the side stream (the sweeps are
        // VALU-bound, the combine HBM-bound).  Slab k's sweeps cover x-sweep
        // row groups [gx0, gx1) and y-sweep column blocks [cy0, cy1); the
        // combine takes the tile rows whose cells both have finished.
        const int wx = ax.waves;
        const int64_t gcx = group_cells(wx, ax.chunk, blend != 0);
        const int64_t Gx = (p.nx + gcx - 1) / gcx, Cy = (p.nx + OUT_COLS - 1) / OUT_COLS;
        const int K = (int)std::min<int64_t>(std::min<int64_t>(p.parts, Gx), Cy);
        int64_t t_done = 0;
        for (int k = 0; k < K; ++k) {
            SweepArgs bx = ax, by = ay;
            bx.by0 = k * Gx / K;
            bx.nrows = (k + 1) * Gx / K - bx.by0;
            by.bx0 = k * Cy / K;
            by.ncols = (k + 1) * Cy / K - by.bx0;
            launch_sweeps<double, IN_F64, double, MODE_PLAIN, true>(bx, by, p.cons, blend, s);
            const int64_t x_done = std::min((bx.by0 + bx.nrows) * gcx, (by.bx0 + by.ncols) * OUT_COLS);
            const int64_t t1 = k == K - 1 ? (p.nx + TT - 1) / TT : std::min(x_done, p.nx) / TT;
            if (t1 > t_done) {
                (void)hipEventRecord(p.ev[k], s);
                (void)hipStreamWaitEvent(p.side, p.ev[k], 0);
                launch_combine<StF64, double, true>(p, Dx, DyT, un, us, o, oT, stage, ctl, p.side, stage == 3,
                                                    ghost_on_store(p), t_done, t1);
                t_done = t1;
            }
        }
        // the next launch on s needs every slab's combine
        (void)hipEventRecord(p.ev[K], p.side);
        (void)hipStreamWaitEvent(s, p.ev[K], 0);
        return;
    }
    if (sweeps) launch_sweeps<double, IN_F64, double, MODE_PLAIN, true>(ax, ay, p.cons, blend, s);
    if (combine)
        launch_combine<StF64, double, true>(p, Dx, DyT, un, us, o, oT, stage, ctl, s, stage == 3, ghost_on_store(p));
}
#endif

static void h_stage(const SweepPlan& p, const double* Us, const double* UsT, const double* Un,
                    double* Dx, double* DyT, double* out, double* outT, int stage, const StepCtl* ctl,
                    hipStream_t s)
{
#if WX_FAST
    if (ctl) {
        h_stage_fold(p, Us, UsT, Un, Dx, DyT, out, outT, stage, ctl, s, true, true);
        return;
    }
#endif
    h_sweeps(p, Us, UsT, Dx, DyT, ctl, s);
    const int64_t P = p.P();
    launch_combine<StF64, double>(p, Dx, DyT, StF64{const_cast<double*>(Un), P}, StF64{const_cast<double*>(Us), P},
                                  StF64{out, P}, StF64{outT, p.PT()}, stage, ctl, s, stage == 3, ghost_on_store(p));
}

static void h_rhs(const SweepPlan& p, const double* U, const double* UT, double* Dx, double* DyT,
                  double* L, const StepCtl* ctl, hipStream_t s)
{
    h_sweeps(p, U, UT, Dx, DyT, ctl, s);
    int64_t nc = p.nx * p.ny;
    k_rhs_out<double><<<(unsigned)((nc + 255) / 256), 256, 0, s>>>(Dx, DyT, L, p.nx, p.ny, p.ldx, p.ldyT);
}

static void h_transpose(const double* in, double* out, int64_t M, int64_t N, int64_t ld_in, int64_t ld_out,
                        hipStream_t s)
{
    dim3 grid((unsigned)((N + TT - 1) / TT), (unsigned)((M + TT - 1) / TT));
    k_transpose<double, 4><<<grid, 256, 0, s>>>(in, out, M, N, ld_in, ld_out);
}

static void h_sweep1(const SweepPlan& p, int dir, int split, const void* U, void* D, int* flags,
                     const StepCtl* ctl, hipStream_t s)
{
    const int blend = p.force_hlle ? 0 : 1;
#if WX_FAST
    if (split) {
        if (dir == 0) {
            launch_sweep<double, IN_SPLIT, float, MODE_DETECT, DIR_X>(args_x(p, U, D, flags, ctl), p.cons, blend, s);
        } else {
            launch_sweep<double, IN_SPLIT, float, MODE_DETECT, DIR_Y>(args_y(p, U, D, flags, ctl), p.cons, blend, s);
        }
        return;
    }
#endif
    if (dir == 0)
        launch_sweep<double, IN_F64, double, MODE_PLAIN, DIR_X>(args_x(p, U, D, nullptr, ctl), p.cons, blend, s);
    else
        launch_sweep<double, IN_F64, double, MODE_PLAIN, DIR_Y>(args_y(p, U, D, nullptr, ctl), p.cons, blend, s);
}

static void h_sweep2(const SweepPlan& p, int split, const void* U, const void* UT, void* Dx, void* DyT,
                     int* flags, const StepCtl* ctl, hipStream_t s)
{
    const int blend = p.force_hlle ? 0 : 1;
#if WX_FAST
    if (split) {
        const SweepArgs ax = args_x(p, U, Dx, flags, ctl), ay = args_y(p, UT, DyT, flags, ctl);
        launch_sweeps<double, IN_SPLIT, float, MODE_DETECT>(ax, ay, p.cons, blend, s);
        return;
    }
    if (ctl) {
        // as a stage-2 step launches them (U standing in for Un)
        h_stage_fold(p, (const double*)U, (const double*)UT, (const double*)U, (double*)Dx, (double*)DyT, nullptr,
                     nullptr, 2, ctl, s, true, false);
        return;
    }
#endif
    launch_sweeps<double, IN_F64, double, MODE_PLAIN>(args_x(p, U, Dx, nullptr, ctl), args_y(p, UT, DyT, nullptr, ctl),
                                                      p.cons, blend, s);
}

static void h_combine1(const SweepPlan& p, int split, const void* Dx, const void* DyT, const void* Un,
                       const void* Us, void* out, void* outT, const StepCtl* ctl, hipStream_t s)
{
    const int64_t P = p.P(), PT = p.PT();
#if WX_FAST
    if (split) {
        launch_combine<StSplit, float>(p, (const float*)Dx, (const float*)DyT, StSplit{(float*)Un, P},
                                       StSplit{(float*)Us, P}, StSplit{(float*)out, P}, StSplit{(float*)outT, PT},
                                       2, ctl, s, false, ghost_on_store(p));
        return;
    }
    if (ctl) {
        h_stage_fold(p, (const double*)Us, nullptr, (const double*)Un, (double*)Dx, (double*)DyT, (double*)out,
                     (double*)outT, 2, ctl, s, false, true);
        return;
    }
#endif
    launch_combine<StF64, double>(p, (const double*)Dx, (const double*)DyT, StF64{(double*)Un, P},
                                  StF64{(double*)Us, P}, StF64{(double*)out, P}, StF64{(double*)outT, PT}, 2, ctl,
                                  s, false, ghost_on_store(p));
}

#if WX_FAST
// ---- blocked FP64 step ------------------------------------------------------
// The x sweep writes E, the y sweep reads the same blocked states with lanes
// along x and finishes the stage; no transposed copy, no combine launch.
static SweepArgs args_blk(const SweepPlan& p, int dir, const double* U, const double* Un, double* D, int stage,
                          const StepCtl* ctl)
{
    SweepArgs a = dir == 0 ? args_x(p, U, D, nullptr, ctl) : args_y(p, U, D, nullptr, ctl);
    a.Un = Un;
    a.stage = stage;
    a.nb = p.nb;
    a.cfl_out = 1;
    if (dir == 1 && ghost_in_ysweep(p)) {
        a.ghost = 1;
        a.bp = p.bp;
    }
    return a;
}

static void h_sweep1_blk(const SweepPlan& p, int dir, const double* Us, const double* Un, double* E, double* out,
                         int stage, const StepCtl* ctl, hipStream_t s)
{
    const int blend = p.force_hlle ? 0 : 1;
    if (dir == 0)
        launch_sweep<double, IN_F64, double, MODE_PLAIN, DIR_X, FOLD_BE>(args_blk(p, 0, Us, Un, E, stage, ctl),
                                                                       p.cons, blend, s);
    else
        launch_sweep<double, IN_F64, double, MODE_PLAIN, DIR_Y, FOLD_BO>(args_blk(p, 1, Us, E, out, stage, ctl),
                                                                       p.cons, blend, s);
}

static void h_stage_blk(const SweepPlan& p, const double* Us, const double* Un, double* E, double* out, int stage,
                        const StepCtl* ctl, hipStream_t s)
{
    h_sweep1_blk(p, 0, Us, Un, E, out, stage, ctl, s);
    h_sweep1_blk(p, 1, Us, Un, E, out, stage, ctl, s);
}

__global__ void k_bc_blk(StF64 U, int64_t nx, int64_t ny, int64_t nb, BcParams bp, const StepCtl* ctl, int stage)
{
    auto at = [&](int64_t i, int64_t j) { return blk_row(i, nb) + blk_col(j); };
    bc_cell_at(U, nx, ny, bp, ctl, stage, (int64_t)blockIdx.x * ST_TB + threadIdx.x, at,
               [&](int64_t i, int64_t j, const double* v) {
#pragma unroll
                   for (int k = 0; k < 4; ++k) U.st(k, at(i, j), v[k]);
               });
}

static void h_bc_blk(const SweepPlan& p, double* U, const StepCtl* ctl, int stage, hipStream_t s)
{
    const int64_t count = 6 * (p.ny + 6) + 6 * p.nx;
    k_bc_blk<<<(unsigned)((count + ST_TB - 1) / ST_TB), ST_TB, 0, s>>>(StF64{U, 1}, p.nx, p.ny, p.nb, p.bp, ctl,
                                                                      stage);
}

// Row-major padded planes (row pitch ld, plane P) to the blocked state or back.
template <bool TO_BLK>
__global__ void k_blk_copy(const double* in, double* out, int64_t M, int64_t N, int64_t ld, int64_t P, int64_t nb)
{
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= M * N) return;
    const int64_t i = t / N, j = t - i * N;
    const int64_t r = i * ld + j, b = blk_row(i, nb) + blk_col(j);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if constexpr (TO_BLK) out[b + k] = in[k * P + r];
        else out[k * P + r] = in[b + k];
    }
}

static void h_to_blk(const SweepPlan& p, const double* U, double* Ub, hipStream_t s)
{
    const int64_t n = (p.nx + 6) * (p.ny + 6);
    k_blk_copy<true><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(U, Ub, p.nx + 6, p.ny + 6, p.ld, p.P(), p.nb);
}

static void h_from_blk(const SweepPlan& p, const double* Ub, double* U, hipStream_t s)
{
    const int64_t n = (p.nx + 6) * (p.ny + 6);
    k_blk_copy<false><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(Ub, U, p.nx + 6, p.ny + 6, p.ld, p.P(), p.nb);
}
#endif

#if WX_FAST
// ---- FP32-split -----------------------------------------------------------
// Bulk sweep in FP32 on increments of the hi+lo values; face rows with a
// non-positive or non-finite Gauss-point state are recomputed in FP64.
static void h_sweeps_split(const SweepPlan& p, const float* Us, const float* UsT, float* Dx,
                           float* DyT, int* flags, const StepCtl* ctl, hipStream_t s)
{
    const int blend = p.force_hlle ? 0 : 1;
    const SweepArgs ax = args_x(p, Us, Dx, flags, ctl), ay = args_y(p, UsT, DyT, flags, ctl);
    launch_sweeps<double, IN_SPLIT, float, MODE_DETECT>(ax, ay, p.cons, blend, s);
}

static void h_stage_split(const SweepPlan& p, const float* Us, const float* UsT, const float* Un,
                          float* Dx, float* DyT, float* out, float* outT, int* flags, int stage,
                          const StepCtl* ctl, hipStream_t s)
{
    h_sweeps_split(p, Us, UsT, Dx, DyT, flags, ctl, s);
    const int64_t P = p.P();
    launch_combine<StSplit, float>(p, Dx, DyT, StSplit{const_cast<float*>(Un), P}, StSplit{const_cast<float*>(Us), P},
                                   StSplit{out, P}, StSplit{outT, p.PT()}, stage, ctl, s, stage == 3, ghost_on_store(p));
}

static void h_rhs_split(const SweepPlan& p, const float* U, const float* UT, float* Dx, float* DyT,
                        int* flags, double* L, const StepCtl* ctl, hipStream_t s)
{
    h_sweeps_split(p, U, UT, Dx, DyT, flags, ctl, s);
    int64_t nc = p.nx * p.ny;
    k_rhs_out<float><<<(unsigned)((nc + 255) / 256), 256, 0, s>>>(Dx, DyT, L, p.nx, p.ny, p.ldx, p.ldyT);
}

static void h_transpose_split(const float* in, float* out, int64_t M, int64_t N, int64_t ld_in, int64_t ld_out,
                              hipStream_t s)
{
    dim3 grid((unsigned)((N + TT - 1) / TT), (unsigned)((M + TT - 1) / TT));
    k_transpose<float, 8><<<grid, 256, 0, s>>>(in, out, M, N, ld_in, ld_out);
}

__global__ void k_to_split(const double* in, float* out, int64_t P)
{
    int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= P) return;
    StSplit o{out, P};
    for (int k = 0; k < 4; ++k) o.st(k, t, in[k * P + t]);
}

__global__ void k_from_split(const float* in, double* out, int64_t P)
{
    int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= P) return;
    StSplit q{const_cast<float*>(in), P};
    for (int k = 0; k < 4; ++k) out[k * P + t] = q.ld(k, t);
}

static void h_to_split(const double* in, float* out, int64_t P, hipStream_t s)
{
    k_to_split<<<(unsigned)((P + 255) / 256), 256, 0, s>>>(in, out, P);
}

static void h_from_split(const float* in, double* out, int64_t P, hipStream_t s)
{
    k_from_split<<<(unsigned)((P + 255) / 256), 256, 0, s>>>(in, out, P);
}

static void h_bc_split(float* U, float* UT, int64_t nx, int64_t ny, int64_t ld, int64_t ldT, BcParams bp,
                       const StepCtl* ctl, int stage, hipStream_t s)
{
    launch_bc(StSplit{U, (nx + 6) * ld}, StSplit{UT, (ny + 6) * ldT}, UT != nullptr, nx, ny, ld, ldT, bp, ctl,
              stage, s);
}

static void h_cfl_split(const float* U, int64_t nx, int64_t ny, int64_t ld, double gam, StepCtl* ctl,
                        hipStream_t s, int finish)
{
    launch_cfl(StSplit{const_cast<float*>(U), (nx + 6) * ld}, nx, ny, ld, gam, ctl, s, finish != 0);
}
#endif

// ------------------------------------------------------------------------
// Batched engine: the same bodies, workgroups mapped to instances by table.
// ------------------------------------------------------------------------
template <class St> __device__ __forceinline__ St make_st(void* base, int64_t soff, int64_t P);
template <> __device__ __forceinline__ StF64 make_st<StF64>(void* base, int64_t soff, int64_t P)
{
    return StF64{(double*)base + soff, P};
}
template <> __device__ __forceinline__ StSplit make_st<StSplit>(void* base, int64_t soff, int64_t P)
{
    return StSplit{(float*)base + 2 * soff, P};
}

// Both sweeps of a stage in one launch: workgroups [0, n_x) run the x sweep,
// the rest the y sweep (their tails fill each other, as k_sweep_xy).
template <typename T, int IN, typename DT, int MODE, bool CONS, bool BLEND, bool FOLD = false>
__global__ void __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(2)))
k_sweep_batch_xy(BatchTables B, BatchCommon C, const void* Ubase, const void* UTbase, void* Dxbase, void* DyTbase,
                 int* flags, StepCtl* ctls, const void* Unbase, int stage)
{
    const int b = blockIdx.x;
    const bool xdir = b < B.n_x;
    const int i = xdir ? B.blk_x[b] : B.blk_y[b - B.n_x];
    const BatchInst I = B.inst[i];
    SweepArgs A;
    A.U = (const char*)(xdir ? Ubase : UTbase) + I.soff * 8;
    A.M = xdir ? I.nx + 6 : I.ny + 6;
    A.N = xdir ? I.ny + 6 : I.nx + 6;
    A.ld = A.N;                                   // instances are stored unpitched
    A.ldd = xdir ? I.ny : I.nx;
    A.nn = xdir ? I.nx : I.ny;
    A.nt = xdir ? I.ny : I.nx;
    A.chunk = xdir ? I.chunk_x : I.chunk_y;
    A.waves = 1;
    A.gam = C.gam;
    A.dx = I.dx;
    A.two_h = xdir ? I.two_h_x : I.two_h_y;
    A.lo = C.jp_lo;
    A.split_jump = C.split_jump;
    A.hi = C.jp_hi;
    A.D = (DT*)(xdir ? Dxbase : DyTbase) + I.ioff;
    A.flags = flags;
    A.ctl = &ctls[i];
    A.Un = FOLD && xdir ? (const char*)Unbase + I.soff * 8 : nullptr;
    A.stage = stage;
    __shared__ SweepLds<T, 1> S;
    if (xdir) {
        const int local = b - I.blk0_x;
        sweep_body<T, IN, DT, MODE, DIR_X, CONS, BLEND, 1, FOLD>(A, local % I.gx_x, local / I.gx_x, b, S);
    } else {
        const int local = b - B.n_x - I.blk0_y;
        sweep_body<T, IN, DT, MODE, DIR_Y, CONS, BLEND, 1>(A, local % I.gx_y, local / I.gx_y, b, S);
    }
}

template <typename T, int IN, typename DT, int MODE, bool FOLD = false>
static void launch_sweep_batch_xy(const BatchTables& t, const BatchCommon& c, const void* U, const void* UT,
                                  void* Dx, void* DyT, int* flags, StepCtl* ctls, hipStream_t s,
                                  const void* Un = nullptr, int stage = 0)
{
    const unsigned n = (unsigned)(t.n_x + t.n_y);
    if (!n) return;
    const bool blend = !c.force_hlle;
#define WX_KB(CO, BL) k_sweep_batch_xy<T, IN, DT, MODE, CO, BL, FOLD><<<n, WAVE, 0, s>>>(t, c, U, UT, Dx, DyT, flags, \
                                                                                      ctls, Un, stage)
    if (c.cons) {
        if (blend) WX_KB(true, true);
        else WX_KB(true, false);
    } else {
        if (blend) WX_KB(false, true);
        else WX_KB(false, false);
    }
#undef WX_KB
}

template <class St, typename DT, bool FOLD = false>
__global__ void __launch_bounds__(256) k_combine_batch(BatchTables B, const DT* Dx, const DT* DyT, void* Un,
                                                       void* Us, void* out, void* outT, int stage, StepCtl* ctls)
{
    const int b = blockIdx.x;
    const int i = B.blk_c[b];
    const BatchInst I = B.inst[i];
    const int64_t P = (I.nx + 6) * (I.ny + 6);
    const int local = b - I.tile0;
    combine_tile<St, DT, TT, false, FOLD>(Dx + I.ioff, DyT + I.ioff, make_st<St>(Un, I.soff, P),
                                          make_st<St>(Us, I.soff, P), make_st<St>(out, I.soff, P),
                                          make_st<St>(outT, I.soff, P), I.nx, I.ny, I.ny + 6, I.nx + 6, I.ny, I.nx,
                                          stage, &ctls[i], local / I.tiles_j, local % I.tiles_j);
}

template <class St>
__global__ void k_bc_batch(BatchTables B, BcParams bp, void* U, void* UT, StepCtl* ctls, int stage)
{
    const int b = blockIdx.x;
    const int i = B.blk_g[b];
    const BatchInst I = B.inst[i];
    const int64_t P = (I.nx + 6) * (I.ny + 6);
    bc_cell<St>(make_st<St>(U, I.soff, P), make_st<St>(UT, I.soff, P), true, I.nx, I.ny, I.ny + 6, I.nx + 6, bp,
                &ctls[i], stage, (int64_t)(b - I.ghost_blk0) * ST_TB + threadIdx.x);
}

template <class St>
__global__ void __launch_bounds__(ST_TB) k_cfl_batch(BatchTables B, void* U, double gam, StepCtl* ctls)
{
    const int b = blockIdx.x;
    const int i = B.blk_f[b];
    const BatchInst I = B.inst[i];
    const int64_t P = (I.nx + 6) * (I.ny + 6);
    cfl_block<St>(make_st<St>(U, I.soff, P), I.nx, I.ny, I.ny + 6, gam, &ctls[i], b - I.cfl_blk0, I.cfl_nblk);
}

__global__ void k_advance_batch(StepCtl* ctls, int n)
{
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n && ctls[i].active) {
        ctls[i].T += ctls[i].dt;
        ctls[i].step += 1;
    }
}

static void hb_stage(const BatchTables& t, const BatchCommon& c, int split, const void* Us, const void* UsT,
                     const void* Un, void* Dx, void* DyT, void* out, void* outT, int* flags, int stage,
                     StepCtl* ctls, hipStream_t s)
{
#if WX_FAST
    if (split) {
        launch_sweep_batch_xy<double, IN_SPLIT, float, MODE_DETECT>(t, c, Us, UsT, Dx, DyT, flags, ctls, s);
        k_combine_batch<StSplit, float><<<(unsigned)t.n_c, 256, 0, s>>>(
            t, (const float*)Dx, (const float*)DyT, (void*)Un, (void*)Us, out, outT, stage, ctls);
        return;
    }
#endif
#if WX_FAST
    // the x sweep forms part of the stage combination, as in the single-grid
    // step; both sweeps in one launch
    launch_sweep_batch_xy<double, IN_F64, double, MODE_PLAIN, true>(t, c, Us, UsT, Dx, DyT, nullptr, ctls, s, Un,
                                                                     stage);
    k_combine_batch<StF64, double, true><<<(unsigned)t.n_c, 256, 0, s>>>(
        t, (const double*)Dx, (const double*)DyT, (void*)Un, (void*)Us, out, outT, stage, ctls);
#else
    launch_sweep_batch_xy<double, IN_F64, double, MODE_PLAIN>(t, c, Us, UsT, Dx, DyT, nullptr, ctls, s);
    k_combine_batch<StF64, double><<<(unsigned)t.n_c, 256, 0, s>>>(
        t, (const double*)Dx, (const double*)DyT, (void*)Un, (void*)Us, out, outT, stage, ctls);
#endif
}

static void hb_bc(const BatchTables& t, BcParams bp, int split, void* U, void* UT, StepCtl* ctls, int stage,
                  hipStream_t s)
{
#if WX_FAST
    if (split) {
        k_bc_batch<StSplit><<<(unsigned)t.n_g, ST_TB, 0, s>>>(t, bp, U, UT, ctls, stage);
        return;
    }
#endif
    k_bc_batch<StF64><<<(unsigned)t.n_g, ST_TB, 0, s>>>(t, bp, U, UT, ctls, stage);
}

static void hb_cfl(const BatchTables& t, int split, const void* U, double gam, StepCtl* ctls, hipStream_t s)
{
#if WX_FAST
    if (split) {
        k_cfl_batch<StSplit><<<(unsigned)t.n_f, ST_TB, 0, s>>>(t, (void*)U, gam, ctls);
        return;
    }
#endif
    k_cfl_batch<StF64><<<(unsigned)t.n_f, ST_TB, 0, s>>>(t, (void*)U, gam, ctls);
}

static void hb_advance(StepCtl* ctls, int n, hipStream_t s)
{
    k_advance_batch<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(ctls, n);
}

#if WX_FAST
// ---- batched blocked FP64 step ----------------------------------------------
// One direction's sweeps of every instance: workgroup b of the launch is
// block b - blk0 of instance blk_x[b] (x) or blk_y[b] (y), one wavefront each.
template <int DIR, int FOLD, bool CONS, bool BLEND>
__global__ void __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(2)))
k_sweep_batch_blk(BatchTables B, BatchCommon C, const double* Ub, const double* Unb, double* Db, StepCtl* ctls,
                  int stage)
{
    const int b = blockIdx.x;
    const int i = DIR == DIR_X ? B.blk_x[b] : B.blk_y[b];
    const BatchInst I = B.inst[i];
    SweepArgs A;
    A.U = Ub + I.boff;
    A.Un = Unb + I.boff;
    A.D = Db + I.boff;
    A.M = DIR == DIR_X ? I.nx + 6 : I.ny + 6;
    A.N = DIR == DIR_X ? I.ny + 6 : I.nx + 6;
    A.ld = A.N;
    A.ldd = DIR == DIR_X ? I.ny : I.nx;
    A.nn = DIR == DIR_X ? I.nx : I.ny;
    A.nt = DIR == DIR_X ? I.ny : I.nx;
    A.chunk = DIR == DIR_X ? I.chunk_x : I.chunk_y;
    A.waves = 1;
    A.gam = C.gam;
    A.dx = I.dx;
    A.two_h = DIR == DIR_X ? I.two_h_x : I.two_h_y;
    A.lo = C.jp_lo;
    A.split_jump = C.split_jump;
    A.hi = C.jp_hi;
    A.flags = nullptr;
    A.ctl = &ctls[i];
    A.stage = stage;
    A.nb = I.nb;
    __shared__ SweepLds<double, 1> S;
    const int local = b - (DIR == DIR_X ? I.blk0_x : I.blk0_y);
    const int gx = DIR == DIR_X ? I.gx_x : I.gx_y;
    sweep_body<double, IN_F64, double, MODE_PLAIN, DIR, CONS, BLEND, 1, FOLD>(A, local % gx, local / gx, b, S);
}

template <int DIR, int FOLD>
static void launch_sweep_batch_blk(const BatchTables& t, const BatchCommon& c, const double* U, const double* Un,
                                   double* D, StepCtl* ctls, int stage, hipStream_t s)
{
    const unsigned n = (unsigned)(DIR == DIR_X ? t.n_x : t.n_y);
    if (!n) return;
#define WX_KB(CO, BL) k_sweep_batch_blk<DIR, FOLD, CO, BL><<<n, WAVE, 0, s>>>(t, c, U, Un, D, ctls, stage)
    if (c.cons) {
        if (!c.force_hlle) WX_KB(true, true);
        else WX_KB(true, false);
    } else {
        if (!c.force_hlle) WX_KB(false, true);
        else WX_KB(false, false);
    }
#undef WX_KB
}

static void hb_stage_blk(const BatchTables& t, const BatchCommon& c, const double* Us, const double* Un, double* E,
                         double* out, int stage, StepCtl* ctls, hipStream_t s)
{
    launch_sweep_batch_blk<DIR_X, FOLD_BE>(t, c, Us, Un, E, ctls, stage, s);
    launch_sweep_batch_blk<DIR_Y, FOLD_BO>(t, c, Us, E, out, ctls, stage, s);
}

__global__ void k_bc_batch_blk(BatchTables B, BcParams bp, double* U, StepCtl* ctls, int stage)
{
    const int b = blockIdx.x;
    const int i = B.blk_g[b];
    const BatchInst I = B.inst[i];
    const StF64 Ui{U + I.boff, 1};
    auto at = [&](int64_t r, int64_t c) { return blk_row(r, I.nb) + blk_col(c); };
    bc_cell_at(Ui, I.nx, I.ny, bp, &ctls[i], stage, (int64_t)(b - I.ghost_blk0) * ST_TB + threadIdx.x, at,
               [&](int64_t r, int64_t c, const double* v) {
#pragma unroll
                   for (int k = 0; k < 4; ++k) Ui.st(k, at(r, c), v[k]);
               });
}

__global__ void __launch_bounds__(ST_TB) k_cfl_batch_blk(BatchTables B, const double* U, double gam, StepCtl* ctls)
{
    const int b = blockIdx.x;
    const int i = B.blk_f[b];
    const BatchInst I = B.inst[i];
    cfl_block_at(StF64{const_cast<double*>(U) + I.boff, 1}, I.nx, I.ny,
                 [&](int64_t r, int64_t c) { return blk_row(r, I.nb) + blk_col(c); }, gam, &ctls[i],
                 b - I.cfl_blk0, I.cfl_nblk);
}

static void hb_bc_blk(const BatchTables& t, BcParams bp, double* U, StepCtl* ctls, int stage, hipStream_t s)
{
    k_bc_batch_blk<<<(unsigned)t.n_g, ST_TB, 0, s>>>(t, bp, U, ctls, stage);
}

static void hb_cfl_blk(const BatchTables& t, const double* U, double gam, StepCtl* ctls, hipStream_t s)
{
    k_cfl_batch_blk<<<(unsigned)t.n_f, ST_TB, 0, s>>>(t, U, gam, ctls);
}

// ---- persistent batched step (persist_item describes the work items) ------
struct PersistArgs {
    double *Ub, *Ab, *Bb, *Eb;
    StepCtl* ctls;
    const int* items;   // the queues' item lists
    int* cnt;           // n_inst * PC_N counters, then the PW_N words
    int steps;
    int spin_limit;
    PersistQueue q[PERSIST_QUEUES];
};

// The items of one instance run on one XCD and exchange data through its L2,
// which all its CUs share.  A finished item waits until its stores have
// reached the L2 (they pass the write-through L1), then counts itself with an
// atomic; a waiting item polls the count with atomic loads, then acquires at
// agent scope, which invalidates its CU's L1 (and no line of local memory in
// the L2).  Measured on the 512-instance batch (ms per step, launch per phase
// 10.5): agent-scope release and acquire at every item, 12.7 -- the release
// writes the XCD's whole L2 back each time; the acquire alone, 9.7.  A
// workgroup-scope L1 invalidation in place of the acquire left stale lines
// (states differed after one step).  Every word is reached one way only: the
// counters, tickets, flags and the CFL maximum (cmax_bits) through atomics and
// atomic loads, which act at the device's point of coherence; the states and
// the other step-control fields through plain loads and stores, which stay in
// the XCD's L2 until the launch ends.

// Lane 0 waits until *c >= target (at most spin_limit polls); then the
// wavefront acquires what the counted items wrote.  False when the wait gave
// up (error flag set, by this wait or an earlier one).
__device__ __forceinline__ bool persist_wait(const int* c, int target, int* err, int spin_limit)
{
    int ok = 1;
    if (threadIdx.x == 0) {
        int spins = 0;
        while (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                ok = 0;
                break;
            }
            if (++spins > spin_limit) {
                __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(4);
        }
    }
    ok = __builtin_amdgcn_readfirstlane(ok);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return ok != 0;
}

// The wavefront's stores complete (in the XCD's L2), then *c += 1; returns
// the previous count to every lane.
__device__ __forceinline__ int persist_signal(int* c)
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0);
    int prev = 0;
    if (threadIdx.x == 0) prev = __hip_atomic_fetch_add(c, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return __builtin_amdgcn_readfirstlane(prev);
}

template <bool CONS, bool BLEND>
__global__ void __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(2)))
k_batch_persist(BatchTables B, BatchCommon C, BcParams bp, PersistArgs P)
{
    __shared__ SweepLds<double, 1> S;
    int* const words = P.cnt + (int64_t)B.n_inst * PC_N;
    int* const err = words + PW_ERR;
    const int l = (int)threadIdx.x;
    // the queue of the XCD this workgroup runs on
    const int xcc = (int)(__builtin_amdgcn_s_getreg((15 << 11) | (0 << 6) | 20) & 0xffff);   // HW_REG_XCC_ID
    const bool served = xcc < PERSIST_QUEUES;
    PersistQueue Q{0, 0, 0, 0, 0};
    if (served) Q = P.q[xcc];
    int* const ticket = words + PW_TICKET + (served ? xcc : 0);
    const int total = persist_items_per_step(Q) * P.steps;
    for (; served;) {
        int t = 0;
        if (l == 0) t = __hip_atomic_fetch_add(ticket, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        t = __builtin_amdgcn_readfirstlane(t);
        if (t >= total) break;
        const PersistItem it = persist_item(t, Q, P.items, B);
        int* const cnt = P.cnt + (int64_t)it.pack * PC_N;
        const bool ok = it.wait_ctr < 0 || persist_wait(cnt + it.wait_ctr, it.wait_target, err, P.spin_limit);
        const BatchPack K = B.packs[it.pack];
        if (ok && (it.kind == PK_X || it.kind == PK_Y)) {
            // a column block of the pack's sweep: lane l's column of the
            // members' padded columns laid side by side (clamped into the
            // pack), its member, that member's offset, dt and activity (plain
            // loads, after the acquire of the wait)
            const bool xd = it.kind == PK_X;
            const BatchInst I = B.inst[B.pack_inst[K.first]];   // the members' common shape
            const double* const src = it.stage == 1 ? P.Ub : (it.stage == 2 ? P.Ab : P.Bb);
            double* const dst = it.stage == 1 ? P.Ab : (it.stage == 2 ? P.Bb : P.Ub);
            const int nm = (int)(xd ? I.ny + 6 : I.nx + 6);   // padded columns per member
            const int gx = xd ? K.gx_x : K.gx_y;
            const int bx = it.local % gx, by = it.local / gx;
            int v = bx * OUT_COLS + l;
            if (v > K.k * nm - 1) v = K.k * nm - 1;
            const int m = v / nm;
            const int mi = B.pack_inst[K.first + m];
            LanePack lp;
            lp.off = B.inst[mi].boff;
            lp.col = v - m * nm;
            lp.dt = P.ctls[mi].dt;
            lp.active = P.ctls[mi].active != 0;
            SweepArgs A;
            A.U = src;
            A.Un = xd ? P.Ub : P.Eb;
            A.D = xd ? P.Eb : dst;
            A.M = xd ? I.nx + 6 : I.ny + 6;
            A.N = nm;
            A.ld = A.N;
            A.ldd = xd ? I.ny : I.nx;
            A.nn = xd ? I.nx : I.ny;
            A.nt = xd ? I.ny : I.nx;
            A.chunk = xd ? I.chunk_x : I.chunk_y;
            A.waves = 1;
            A.gam = C.gam;
            A.dx = I.dx;
            A.two_h = xd ? I.two_h_x : I.two_h_y;
            A.lo = C.jp_lo;
            A.split_jump = C.split_jump;
            A.hi = C.jp_hi;
            A.flags = nullptr;
            A.ctl = nullptr;
            A.stage = it.stage;
            A.nb = I.nb;
            if (xd)
                sweep_body<double, IN_F64, double, MODE_PLAIN, DIR_X, CONS, BLEND, 1, FOLD_BE, true>(A, bx, by, t, S,
                                                                                                     lp);
            else
                sweep_body<double, IN_F64, double, MODE_PLAIN, DIR_Y, CONS, BLEND, 1, FOLD_BO, true>(A, bx, by, t, S,
                                                                                                     lp);
        } else if (ok) {
            const BatchInst I = B.inst[it.inst];
            StepCtl* const ctl = P.ctls + it.inst;
            // the step-control fields the item reads (plain loads, after the
            // acquire of its wait)
            StepCtl lc{};
            lc.active = ctl->active;
            lc.dt = ctl->dt;
            lc.T = ctl->T;
            lc.rho_floor = ctl->rho_floor;
            if (it.kind == PK_CFL) {
                // rows local, local + nblk, ... of the interior, lanes along the row
                const StF64 Ui{P.Ub + I.boff, 1};
                double vm = 0.0;
                for (int64_t r = it.local; r < I.nx; r += I.cfl_nblk) {
                    for (int64_t j = l; j < I.ny; j += WAVE) {
                        const int64_t o = blk_row(r + 3, I.nb) + blk_col(j + 3);
                        const double q[4] = {Ui.ld(0, o), Ui.ld(1, o), Ui.ld(2, o), Ui.ld(3, o)};
                        vm = nan_max(vm, cfl_speed(q, C.gam, lc.rho_floor));
                    }
                }
                vm = wave_max(vm);
                if (l == 0) atomicMax(&ctl->cmax_bits, (unsigned long long)__double_as_longlong(vm));
            } else if (it.kind == PK_G) {
                double* const dst = it.stage == 1 ? P.Ab : (it.stage == 2 ? P.Bb : P.Ub);
                const StF64 Ui{dst + I.boff, 1};
                auto at = [&](int64_t r, int64_t c) { return blk_row(r, I.nb) + blk_col(c); };
#pragma unroll
                for (int q = 0; q < GHOST_TB / WAVE; ++q) {
                    bc_cell_at(Ui, I.nx, I.ny, bp, &lc, it.stage, (int64_t)it.local * GHOST_TB + q * WAVE + l, at,
                               [&](int64_t r, int64_t c, const double* v) {
#pragma unroll
                                   for (int k = 0; k < 4; ++k) Ui.st(k, at(r, c), v[k]);
                               });
                }
            }
        }
        const int prev = persist_signal(cnt + it.sig_ctr);
        if (it.last_count > 0 && prev + 1 == it.last_count) {
            // every other block of the pack's step has signalled: finish it
            // for every member
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            if (l == 0) {
                for (int m = 0; m < K.k; ++m) {
                    StepCtl* const ctl = P.ctls + B.pack_inst[K.first + m];
                    if (it.kind == PK_CFL) {
                        const unsigned long long bits =
                            __hip_atomic_load(&ctl->cmax_bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        cfl_finish(ctl, __longlong_as_double((long long)bits), false);
                        __hip_atomic_store(&ctl->cmax_bits, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    } else if (ctl->active) {
                        ctl->T += ctl->dt;
                        ctl->step += 1;
                    }
                }
            }
            persist_signal(cnt + it.done_ctr);
        }
    }
    // the last workgroup to leave checks that every queue was served (a queue
    // whose XCD ran none of the workgroups keeps its items)
    if (l == 0) {
        const int gone = __hip_atomic_fetch_add(words + PW_EXIT, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (gone == (int)gridDim.x - 1) {
            for (int q = 0; q < PERSIST_QUEUES; ++q) {
                const int need = persist_items_per_step(P.q[q]) * P.steps;
                if (__hip_atomic_load(words + PW_TICKET + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need)
                    __hip_atomic_store(err, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
}

__global__ void k_fill_int(int* p, int n, int v)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

// Each workgroup marks the XCD it runs on: seen[id] = 1 for ids below 31,
// seen[31] for any other.
__global__ void k_xcc_probe(int* seen)
{
    if (threadIdx.x == 0) {
        const int xcc = (int)(__builtin_amdgcn_s_getreg((15 << 11) | (0 << 6) | 20) & 0xffff);   // HW_REG_XCC_ID
        seen[xcc < 31 ? xcc : 31] = 1;
    }
}

// Bit i set when a workgroup of a `grid`-workgroup launch ran on XCD i (bit
// 31: an id of 31 or more); words: 32 device ints.  0 on a HIP error.
static unsigned hb_xcc_mask(int grid, int* words, hipStream_t s)
{
    int seen[32] = {};
    k_fill_int<<<1, 256, 0, s>>>(words, 32, 0);
    k_xcc_probe<<<(unsigned)(grid > 0 ? grid : 1), WAVE, 0, s>>>(words);
    if (hipMemcpyAsync(seen, words, sizeof seen, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return 0u;
    unsigned m = 0;
    for (int i = 0; i < 32; ++i)
        if (seen[i]) m |= 1u << i;
    return m;
}

template <bool CONS, bool BLEND> static int persist_grid_of()
{
    static int grid = 0;   // per scheme; the library serves one device type
    if (!grid) {
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_persist<CONS, BLEND>, WAVE, 0) !=
                hipSuccess)
            return 0;
        grid = cus * (per_cu > 0 ? per_cu : 1);
    }
    return grid;
}

static int hb_persist_grid(const BatchCommon& c)
{
    if (c.cons) return c.force_hlle ? persist_grid_of<true, false>() : persist_grid_of<true, true>();
    return c.force_hlle ? persist_grid_of<false, false>() : persist_grid_of<false, true>();
}

static void hb_persist(const BatchTables& t, const BatchCommon& c, BcParams bp, double* Ub, double* Ab, double* Bb,
                       double* Eb, StepCtl* ctls, const PersistQueue* queues, const int* items, int* cnt, int steps,
                       int grid, int spin_limit, hipStream_t s)
{
    // counters, tickets and the exit count (not the error flag) start at zero
    const int n = t.n_inst * PC_N + PW_ERR;
    k_fill_int<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(cnt, n, 0);
    PersistArgs P{Ub, Ab, Bb, Eb, ctls, items, cnt, steps, spin_limit, {}};
    for (int q = 0; q < PERSIST_QUEUES; ++q) P.q[q] = queues[q];
    if (grid < 1) return;
#define WX_KP(CO, BL) k_batch_persist<CO, BL><<<(unsigned)grid, WAVE, 0, s>>>(t, c, bp, P)
    if (c.cons) {
        if (!c.force_hlle) WX_KP(true, true);
        else WX_KP(true, false);
    } else {
        if (!c.force_hlle) WX_KP(false, true);
        else WX_KP(false, false);
    }
#undef WX_KP
}
#endif

}  // namespace WX_NS

#ifndef __HIP_DEVICE_COMPILE__
#if WX_FAST
const BatchOps batch_ops_fast = {WX_NS::hb_stage,     WX_NS::hb_bc,     WX_NS::hb_cfl,     WX_NS::hb_advance,
                                 WX_NS::hb_stage_blk, WX_NS::hb_bc_blk, WX_NS::hb_cfl_blk, WX_NS::hb_persist,
                                 WX_NS::hb_persist_grid, WX_NS::hb_xcc_mask};
#else
const BatchOps batch_ops_exact = {WX_NS::hb_stage, WX_NS::hb_bc, WX_NS::hb_cfl, WX_NS::hb_advance,
                                  nullptr,         nullptr,      nullptr,       nullptr,
                                  nullptr,         nullptr};
#endif
#if WX_FAST
const SweepOps WX_SWEEP_OPS = {WX_NS::h_stage,       WX_NS::h_rhs,        WX_NS::h_transpose,
                               WX_NS::h_stage_split, WX_NS::h_rhs_split,  WX_NS::h_transpose_split,
                               WX_NS::h_to_split,    WX_NS::h_from_split, WX_NS::h_bc_split,
                               WX_NS::h_cfl_split,   WX_NS::h_sweep1,     WX_NS::h_combine1,
                               WX_NS::h_sweep2,      WX_NS::h_stage_blk,  WX_NS::h_sweep1_blk,
                               WX_NS::h_bc_blk,      WX_NS::h_to_blk,     WX_NS::h_from_blk};
#else
const SweepOps WX_SWEEP_OPS = {WX_NS::h_stage,  WX_NS::h_rhs,      WX_NS::h_transpose, nullptr, nullptr,
                               nullptr,         nullptr,           nullptr,            nullptr, nullptr,
                               WX_NS::h_sweep1, WX_NS::h_combine1, WX_NS::h_sweep2,    nullptr, nullptr,
                               nullptr,         nullptr,           nullptr};

// Default tiling from a round model of the 256 CUs: two wavefronts per SIMD
// are resident (2048 waves, 2048 / W workgroups of W waves); a sweep takes
// rounds x (faces one wave computes) face-times, where a wave computes its
// chunk plus its share of the group's two halo faces (and, with W > 1, about
// half a face of deferred work).  The autotuner refines this on the device.
static void pick_tiling(int64_t nn, int64_t nt, int& chunk, int& waves)
{
    const int64_t cols = (nt + WX_NS::OUT_COLS - 1) / WX_NS::OUT_COLS;
    double best = 1e300;
    chunk = 8;
    waves = 1;
    for (int w : {1, 2, 4}) {
        const int cmin = w == 1 ? 8 : (w == 2 ? 4 : 3), cmax = w == 1 ? 64 : (w == 2 ? 24 : 16);
        for (int c = cmin; c <= cmax; ++c) {
            if (w > 1 && WX_NS::group_cells(w, c, true) > nn) continue;
            const int64_t gc = WX_NS::group_cells(w, c, true);
            const int64_t groups = (nn + gc - 1) / gc;
            const int64_t wg = groups * cols, slots = 2048 / w;
            const int64_t rounds = (wg + slots - 1) / slots;
            // faces one wave marches: c + 3 alone; c (balanced) + deferred work in a group
            const double cost = (double)rounds * (w > 1 ? c + 0.5 : c + 3.0);
            if (cost < best - 1e-9) { best = cost; chunk = c; waves = w; }
        }
    }
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
    pick_tiling(nx, ny, p.chunk_x, p.waves_x);
    pick_tiling(ny, nx, p.chunk_y, p.waves_y);
    p.combine_cols = 32;   // measured faster than 64-wide on 1920x480 and no slower at 4096^2
    p.ld = row_pitch(ny + 6);
    p.ldT = row_pitch(nx + 6);
    p.ldx = row_pitch(ny);
    p.ldyT = row_pitch(nx);
    p.nb = (ny + 6 + 1) / 2;
    p.Pb = (nx + 6 + 1) / 2 * p.nb * 16;
    return p;
}

int64_t sweep_flag_count(const SweepPlan& p)
{
    const int64_t c = WX_NS::OUT_COLS;
    // the blend's groups are the smaller ones: enough flags either way
    const int64_t rx = WX_NS::group_cells(p.waves_x, p.chunk_x, true);
    const int64_t ry = WX_NS::group_cells(p.waves_y, p.chunk_y, true);
    int64_t fx = ((p.ny + c - 1) / c) * ((p.nx + rx - 1) / rx) * p.waves_x;
    int64_t fy = ((p.nx + c - 1) / c) * ((p.ny + ry - 1) / ry) * p.waves_y;
    // the one-launch sweep pair numbers its workgroups across both halves
    return fx + fy;
}
#endif
#endif

}  // namespace wenox
