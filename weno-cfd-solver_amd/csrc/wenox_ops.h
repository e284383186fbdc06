// wenox_ops.h -- host-side launch table of the device kernels.
//
// wenox_kernels.hip is compiled twice (arithmetic variants "exact" and
// "fast", see wenox_numerics.h); each build exports one Ops table.  All
// pointers are device pointers; every launch is asynchronous on `s`.
//
// Array conventions:
//   AoS4  (M, N, 4) C-contiguous, the numpy layout of the reference's arrays;
//   SoA4  four planes [4][M][N], j (y) fastest -- the stepper's HBM layout.
//
// The stepper stores each padded plane (M = nx+6 rows of N = ny+6 cells)
// with a row pitch ld = row_pitch(N) >= N and plane stride M * ld, and the
// plane set begins STATE_SHIFT elements into its allocation: padded column 3
// (the first interior cell) of every row then starts on a 256 B (FP64) or
// 128 B (FP32 planes) boundary, so 32-cell row segments of the interior are
// whole cache lines.  The transposed copy uses row_pitch(M) the same way, and
// the interior flux-difference arrays (nx x ny, ny x nx) use row_pitch of
// their row length without a shift.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wenox {

constexpr int64_t ROW_ALIGN = 32;                // elements
constexpr int64_t STATE_SHIFT = ROW_ALIGN - 3;   // elements before padded column 0 of row 0
inline int64_t row_pitch(int64_t n) { return (n + ROW_ALIGN - 1) / ROW_ALIGN * ROW_ALIGN; }

// Device-resident time control of one solve (read by every step kernel).
// Slots of the CFL reduction's per-block maxima (StepCtl::part).
constexpr int CFL_SLOTS = 1024;

struct StepCtl {
    double T;              // simulated time
    double dt;             // dt of the current step
    double t_final;        // < 0: step-count controlled
    double cfl_dx;         // CFL * dx (host product, as the reference forms it)
    double cmax0;          // first-step wave speed from the caller's primitives
    double rho_floor;      // CFL: c = sqrt(g p / max(rho, rho_floor))   (1e-14; the 1-D driver 1e-10)
    double cmax_eps;       // CFL: dt = cfl_dx / (c_max + cmax_eps)      (1e-16; the 1-D driver 0)
    long long step;        // steps taken
    long long max_steps;   // safety cap / step count
    int active;            // current step runs (loop condition held)
    int use_cmax0;         // next CFL uses cmax0 instead of the state
    unsigned long long cmax_bits;   // running max of c+|V| (bit pattern, >= 0)
    unsigned int blocks_done;       // last-block detection of the reduction
    int floor_valid;                // floor_bits holds the next step's c_max (pressure floor ran)
    unsigned long long floor_bits;  // max of c+|V| from the floored primitives (bit pattern)
    int pending;                    // a finished step's T += dt, step += 1 not yet applied
    int nparts;                     // valid entries of part (the last writer's block count)
    double* part;                   // per-block maxima of the CFL reduction (nullptr: atomics)
};

enum BcKind : int {
    BC_MIRROR = 0,      // even mirror, y then x (Rie_prob_2d_rs.py:123-152)
    BC_PERIODIC = 1,
    BC_DMR_REF = 2,     // unitest.py:207-227 (final state of its write sequence)
    BC_DMR_STD = 3,     // standard double Mach reflection with exact moving shock
    BC_COPY = 4,        // constant extrapolation: ghost <- nearest interior (Rie_prob_1D_rs.py:59-76)
};

struct BcParams {
    int kind;
    int dmr_bound;      // BC_DMR_REF: `bound` of unitest.init_refac
    double dx, dy;      // BC_DMR_STD geometry
    double x0;          // BC_DMR_STD wall start (1/6)
    double gamma;
    double t0;          // BC_DMR_STD ghost time of a fill launched without StepCtl
};

// Strides (in elements) of a 3-index array (i, j, k): element at
// i*is + j*js + k*ks.  AoS4 (M,N,4): {4N, 4, 1}.  SoA4 [4][M][N]: {N, 1, M*N}.
struct Lay {
    int64_t is, js, ks;
};
inline Lay lay_aos4(int64_t M, int64_t N) { return Lay{4 * N, 4, 1}; }
inline Lay lay_soa4(int64_t M, int64_t N) { return Lay{N, 1, M * N}; }

struct Ops {
    // ---- drop-in building blocks (reference shapes, any layout) ---------
    void (*flux)(const double* q, Lay lq, int64_t M, int64_t N, int dir, double gam, double* out,
                 Lay lo, hipStream_t s);
    void (*con2primi)(const double* q, Lay lq, int64_t M, int64_t N, double gam, double* p,
                      double* a, double* rho, double* u, double* v, double* h, hipStream_t s);
    // rows [row0, row0 + rows) of a padded SoA4 state (m x n planes, row
    // pitch ld) as (rows, n, 8): conserved channels, then rho_safe, u, v, p
    void (*snapshot8)(const double* U, int64_t m, int64_t n, int64_t ld, double gam, int64_t row0,
                      int64_t rows, double* out, hipStream_t s);
    // pass 1 (normal) into pl/pr, then pass 2 (transverse Gauss points) and the
    // averaged fluxes into ql/fl/qr/fr.  u is (m, n); pl/pr are SoA4 of
    // (m-5, n) for x or (m, n-5) for y; outputs are (F1, F2) in layout lo.
    void (*weno)(const double* u, Lay lu, int64_t m, int64_t n, int dir, int cons, double gam,
                 double* pl, double* pr, double* ql, double* fl, double* qr, double* fr, Lay lo,
                 hipStream_t s);
    // kind 0 HLLC, 1 HLLE, 2 blend (pbar: scratch of F1*F2 doubles); the
    // blend's sensor takes h = 1/h_min (riemann.rs:360-364 with the face
    // grid's min(F1, F2) when h_min = 0; the whole grid's when (F1, F2) is a
    // band of its rows)
    void (*riemann)(const double* ql, const double* fl, const double* qr, const double* fr,
                    Lay li, int64_t F1, int64_t F2, int dir, int kind, double gam, double lo,
                    double hi, double* pbar, double* out, Lay lout, int64_t h_min, hipStream_t s);
    // out (Nx, Ny) = -((fx[i+1]-fx[i])/dx + (fy[j+1]-fy[j])/dx)
    void (*divergence)(const double* fx, Lay lfx, const double* fy, Lay lfy, int64_t Nx,
                       int64_t Ny, double dx, double* out, Lay lout, hipStream_t s);

    // ---- stepper (SoA4 padded state, (nx+6) x (ny+6) per plane) ---------
    // Row pitches: ld of U's planes, ldT of UT's (plane strides (nx+6) ld and
    // (ny+6) ldT).
    // ghost fill of U, and of its transposed copy UT when UT != nullptr
    void (*bc)(double* U, double* UT, int64_t nx, int64_t ny, int64_t ld, int64_t ldT, BcParams bp,
               const StepCtl* ctl, int stage, hipStream_t s);
    void (*cfl)(const double* U, int64_t nx, int64_t ny, int64_t ld, double gam, StepCtl* ctl, hipStream_t s);
    void (*advance)(StepCtl* ctl, hipStream_t s);
    // the per-block maxima of the CFL reduction only (into ctl->part)
    void (*cfl_part)(const double* U, int64_t nx, int64_t ny, int64_t ld, double gam, StepCtl* ctl,
                     hipStream_t s);
    // a step's first launch on the fused engine: apply the previous step's
    // pending advance, reduce ctl->part (written by the previous step's last
    // RK combine, or by cfl_part), form dt and the loop condition
    void (*step_dt)(StepCtl* ctl, hipStream_t s);
    // apply a pending advance (before the host reads T and the step count)
    void (*settle)(StepCtl* ctl, hipStream_t s);
    // Unfused RHS: L (interior Nx x Ny, SoA4) from U (padded SoA4), through
    // face arrays in `work` (rhs_unfused_scratch(nx, ny) doubles).
    void (*rhs_unfused)(const double* U, int64_t nx, int64_t ny, int64_t ld, double dx, double gam, int cons,
                        int force_hlle, double lo, double hi, double* work, double* L,
                        const StepCtl* ctl, hipStream_t s);
    // RK stage combine on the SoA4 interior (stage 1..3), skipped when !ctl->active.
    void (*rk)(const double* Un, const double* Us, const double* L, double* out, int64_t nx,
               int64_t ny, int64_t ld, int stage, const StepCtl* ctl, hipStream_t s);
    // pressure floor of the NumPy twin's time loop on every cell of M x N
    // planes (row pitch ld), skipped when !ctl->active.  With cfl != 0 it also
    // reduces max(sqrt(g p / max(rho, rho_floor)) + |V|) over the interior
    // (three ghost layers) from the floored primitives into ctl->floor_bits,
    // which the next step's dt then uses (the twin's CFL reads them).
    void (*pfloor)(double* U, int64_t M, int64_t N, int64_t ld, double gam, double pmin, StepCtl* ctl,
                   int cfl, hipStream_t s);
};

size_t rhs_unfused_scratch(int64_t nx, int64_t ny);

extern const Ops ops_exact;
extern const Ops ops_fast;
extern const Ops ops_twin;

}  // namespace wenox
