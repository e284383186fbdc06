// wenox_sweep.h -- fused direction-sweep engine of the device-resident solver.
//
// One RK stage = an x sweep over the state, a y sweep over its transposed copy
// (each: reconstruction, Gauss pass, Riemann solve, sensor, blend and the
// normal flux difference of every cell, with no face array in HBM), then one
// combine kernel forming L = -(Dx + Dy) and the SSP-RK3 stage, which writes
// the next state in both orientations.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wenox_ops.h"

namespace wenox {

struct SweepPlan {
    int64_t nx, ny;
    double dx, gamma;
    int cons, force_hlle;
    double jp_lo, jp_hi;
    double two_h_x, two_h_y;   // 2 / max(min(F1, F2), 1) of the x and y face grids (riemann.rs:360-361)
    int chunk_x, chunk_y;      // interior rows marched per wavefront
    int waves_x, waves_y;      // wavefronts per workgroup along the normal (1, 2 or 4)
    int combine_cols;          // y extent of an RK-combine tile (32 or 64)
    int ghost;                 // the step may let the combine write the ghost cells (ghost_on_store)
    BcParams bp;               // ghost-fill rule of the solver
    // row pitches (elements): U planes, U^T planes, Dx rows, Dy^T rows
    int64_t ld, ldT, ldx, ldyT;
    // Slab pipelining of the fast FP64 stage: with parts > 1 the sweeps run in
    // `parts` slabs of x rows on the stage's stream and each finished slab's
    // RK combine on `side`, ordered through the events ev[0 .. parts]
    // (PIPE_MAX_PARTS + 1 of them).
    int parts = 1;
    hipStream_t side = nullptr;
    hipEvent_t* ev = nullptr;
    // Blocked layout of the fast FP64 step without transposed copies
    // (SweepOps::stage_blk): the four conserved components of a cell side by
    // side (32 B), 2 x 2 cells per 128 B line, lines row-major; nb lines per
    // line row, Pb doubles per state (blk_row / blk_col).
    int64_t nb = 0, Pb = 0;
    int64_t P() const { return (nx + 6) * ld; }       // plane stride of U
    int64_t PT() const { return (ny + 6) * ldT; }     // plane stride of U^T
    int64_t PDx() const { return nx * ldx; }          // plane stride of Dx
    int64_t PDyT() const { return ny * ldyT; }        // plane stride of Dy^T
};

// Index of component 0 of padded cell (i, j) in a blocked state:
// blk_row(i, nb) + blk_col(j); component k follows at + k.  A 2 x 2 group
// of cells is one 128 B line and a cell one 32 B sector of it, so a lane
// reads or writes its cell's four components as 32 contiguous bytes, a
// wavefront's 64 lanes along either axis touch 32 lines, and the next row of
// its march uses the same lines.  (With each component in its own plane of
// 4 x 4-cell lines, the lines of a wave stayed in use over four rows of its
// march and the L2 did not hold them: the 4096^2 step read 30 GB per step
// from beyond the L2, against 16.7 GB for the row-major step.)
__host__ __device__ inline int64_t blk_row(int64_t i, int64_t nb) { return (i >> 1) * nb * 16 + ((i & 1) << 3); }
__host__ __device__ inline int64_t blk_col(int64_t j) { return ((j >> 1) << 4) + ((j & 1) << 2); }

struct SweepOps {
    // stage s: Us/UsT stage input (ghost-filled), Un step start; Dx (4 planes
    // of nx rows, pitch ldx), DyT (4 planes of ny rows, pitch ldyT) scratch;
    // out/outT receive the next state's interior.  State buffers use the
    // plan's pitches (ld, ldT).
    void (*stage)(const SweepPlan& p, const double* Us, const double* UsT, const double* Un,
                  double* Dx, double* DyT, double* out, double* outT, int stage,
                  const StepCtl* ctl, hipStream_t s);
    // L(U) (SoA4 interior) only.
    void (*rhs)(const SweepPlan& p, const double* U, const double* UT, double* Dx, double* DyT,
                double* L, const StepCtl* ctl, hipStream_t s);
    // out[k][j][i] = in[k][i][j] for an M x N padded plane set (row pitch
    // ld_in of in, ld_out of out)
    void (*transpose)(const double* in, double* out, int64_t M, int64_t N, int64_t ld_in, int64_t ld_out,
                      hipStream_t s);

    // ---- FP32-split state: float planes [8][M][N], value = hi[k] + lo[k+4] ----
    // (fast build only; nullptr in the exact build)
    void (*stage_split)(const SweepPlan& p, const float* Us, const float* UsT, const float* Un,
                        float* Dx, float* DyT, float* out, float* outT, int* flags, int stage,
                        const StepCtl* ctl, hipStream_t s);
    void (*rhs_split)(const SweepPlan& p, const float* U, const float* UT, float* Dx, float* DyT,
                      int* flags, double* L, const StepCtl* ctl, hipStream_t s);
    void (*transpose_split)(const float* in, float* out, int64_t M, int64_t N, int64_t ld_in, int64_t ld_out,
                            hipStream_t s);
    void (*to_split)(const double* in, float* out, int64_t P, hipStream_t s);
    void (*from_split)(const float* in, double* out, int64_t P, hipStream_t s);
    void (*bc_split)(float* U, float* UT, int64_t nx, int64_t ny, int64_t ld, int64_t ldT, BcParams bp,
                     const StepCtl* ctl, int stage, hipStream_t s);
    // CFL maxima of a split state into ctl->part; finish: also form dt (as
    // the step's first launch does from the last combine's maxima)
    void (*cfl_split)(const float* U, int64_t nx, int64_t ny, int64_t ld, double gam, StepCtl* ctl, hipStream_t s,
                      int finish);

    // Single kernels, for per-kernel timing: one direction's sweep of the
    // stage input U (U^T for dir 1) into D, and the stage-2 combine.
    void (*sweep1)(const SweepPlan& p, int dir, int split, const void* U, void* D, int* flags,
                   const StepCtl* ctl, hipStream_t s);
    void (*combine1)(const SweepPlan& p, int split, const void* Dx, const void* DyT, const void* Un,
                     const void* Us, void* out, void* outT, const StepCtl* ctl, hipStream_t s);
    // Both sweeps of a stage as the step launches them (one kernel when the
    // two tilings use the same number of wavefronts per workgroup).
    void (*sweep2)(const SweepPlan& p, int split, const void* U, const void* UT, void* Dx, void* DyT,
                   int* flags, const StepCtl* ctl, hipStream_t s);

    // ---- blocked FP64 step (fast build only; nullptr in the exact build) ----
    // Stage s on blocked states (p.nb, p.Pb): the x sweep of Us writes E = a Un
    // + b Us - (c dt) Dx, then the y sweep of Us finishes the stage, out = E -
    // (c dt) Dy, and in stage 3 leaves the next step's CFL maxima (one per
    // workgroup, at most blk_part_slots(p)) and the advance pending.  out
    // receives the interior; Us must hold its ghost cells.
    void (*stage_blk)(const SweepPlan& p, const double* Us, const double* Un, double* E, double* out,
                      int stage, const StepCtl* ctl, hipStream_t s);
    // one of the two sweeps alone (dir 0: x into E; 1: y into out), for timing
    void (*sweep1_blk)(const SweepPlan& p, int dir, const double* Us, const double* Un, double* E,
                       double* out, int stage, const StepCtl* ctl, hipStream_t s);
    void (*bc_blk)(const SweepPlan& p, double* U, const StepCtl* ctl, int stage, hipStream_t s);
    // row-major padded planes (row pitch p.ld) <-> blocked state
    void (*to_blk)(const SweepPlan& p, const double* U, double* Ub, hipStream_t s);
    void (*from_blk)(const SweepPlan& p, const double* Ub, double* U, hipStream_t s);
};


// ---- batched engine: many independent problems in each launch ----------------
// Every state buffer holds the instances back to back: instance i's padded
// planes start `soff` 8-byte words in (FP64 planes [4][M][N] or FP32-split
// planes [8][M][N], the same bytes); Dx / Dy^T hold its interior planes at
// element offset `ioff`.  Workgroups of each launch map to instances through
// the blk_* tables.
struct BatchInst {
    int64_t nx, ny;
    int64_t soff, ioff;
    double dx, two_h_x, two_h_y;
    int chunk_x, chunk_y;
    int gx_x, gx_y;        // column blocks of the x / y sweeps
    int blk0_x, blk0_y;    // first workgroup of this instance in the x / y sweep grids
    int tiles_j, tile0;    // combine: tile columns, first tile
    int ghost_blk0;        // first ghost-fill block
    int cfl_blk0, cfl_nblk;
    int64_t boff, nb;      // blocked step: offset (doubles) of the instance's blocked state, its nb
};

struct BatchTables {       // device pointers and grid sizes
    const BatchInst* inst;
    const int* blk_x;
    const int* blk_y;
    const int* blk_c;
    const int* blk_g;
    const int* blk_f;
    int n_x, n_y, n_c, n_g, n_f, n_inst;
};

struct BatchCommon {
    double gam, jp_lo, jp_hi;
    int cons, force_hlle;
};

struct BatchOps {
    // one RK stage of every instance (sweeps, FP64 fix-up when split, combine)
    void (*stage)(const BatchTables& t, const BatchCommon& c, int split, const void* Us,
                  const void* UsT, const void* Un, void* Dx, void* DyT, void* out, void* outT,
                  int* flags, int stage, StepCtl* ctls, hipStream_t s);
    void (*bc)(const BatchTables& t, BcParams bp, int split, void* U, void* UT, StepCtl* ctls,
               int stage, hipStream_t s);
    void (*cfl)(const BatchTables& t, int split, const void* U, double gam, StepCtl* ctls,
                hipStream_t s);
    void (*advance)(StepCtl* ctls, int n, hipStream_t s);
    // ---- blocked FP64 fast step (fast build only; nullptr in the exact build) ----
    // the instances' blocked states start at BatchInst::boff: per stage the x
    // sweeps write E, the y sweeps finish the stage (SweepOps::stage_blk, the
    // CFL maxima left to cfl_blk); the ghost fill and the CFL reduction of the
    // blocked states
    void (*stage_blk)(const BatchTables& t, const BatchCommon& c, const double* Us, const double* Un, double* E,
                      double* out, int stage, StepCtl* ctls, hipStream_t s);
    void (*bc_blk)(const BatchTables& t, BcParams bp, double* U, StepCtl* ctls, int stage, hipStream_t s);
    void (*cfl_blk)(const BatchTables& t, const double* U, double gam, StepCtl* ctls, hipStream_t s);
};

extern const BatchOps batch_ops_exact;
extern const BatchOps batch_ops_fast;
constexpr int PIPE_MAX_PARTS = 8;
// workgroups per sweep column block (the wave's output columns)
constexpr int SWEEP_OUT_COLS = 58;

// Entries of ctl->part the blocked step's y sweep may write (one per
// workgroup of any tiling: chunk >= 1).
inline int64_t blk_part_slots(const SweepPlan& p)
{
    return ((p.nx + SWEEP_OUT_COLS - 1) / SWEEP_OUT_COLS) * (p.ny > 0 ? p.ny : 1);
}

constexpr int COMBINE_TILE = 32;
constexpr int GHOST_TB = 256;

// The RK combine writes the ghost cells of the state it forms, in place of a
// ghost-fill launch, when every ghost cell's source lies in the combine tile
// that holds the ghost's nearest interior cell: the local rules (mirror,
// copy, both double-Mach-reflection fills; periodic takes cells from the far
// side) on grids whose edge tiles are at least 3 cells (the ghost depth) wide.
inline bool ghost_on_store(const SweepPlan& p)
{
    if (!p.ghost || p.bp.kind == BC_PERIODIC) return false;
    const int64_t rx = p.nx % COMBINE_TILE, ry = p.ny % p.combine_cols;
    return p.nx >= 3 && p.ny >= 3 && rx != 1 && rx != 2 && ry != 1 && ry != 2;
}

extern const SweepOps sweep_ops_exact;
extern const SweepOps sweep_ops_fast;

SweepPlan make_sweep_plan(int64_t nx, int64_t ny, double dx, double gamma, int cons,
                          int force_hlle, double jp_lo, double jp_hi);

// Flag slots of the FP32-split sweeps: the workgroups of both sweeps of a stage.
int64_t sweep_flag_count(const SweepPlan& p);

}  // namespace wenox
