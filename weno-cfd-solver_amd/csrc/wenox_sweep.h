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

// Default split-jump threshold of the FP32-split sweeps (SweepPlan::split_jump).
constexpr float SPLIT_JUMP_DEFAULT = 0.25f;

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
    // FP32-split: a face whose stencils span a relative jump of density or
    // energy above this takes FP64 states (recon_normal_dev, weno5_gq_dev)
    float split_jump = SPLIT_JUMP_DEFAULT;
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

// A pack of the persistent step: k instances of one shape (nx, ny, dx) whose
// padded columns its sweeps lay side by side along the wavefront's lanes.
// The x sweep sees k (ny + 6) padded columns, member m's padded column c at
// lane column m (ny + 6) + c, and cuts them into column blocks of 58 output
// lanes as it cuts one grid; a lane computes its own column's faces from its
// own member's cells, so only ghost columns fall on the lanes where two
// members meet.  A 64-wide instance alone fills two wavefronts with 64
// outputs; five packed fill six.  Likewise the y sweep over k (nx + 6).
struct BatchPack {
    int first, k;          // members: pack_inst[first .. first + k)
    int gx_x, gx_y;        // column blocks of the x / y sweeps over the members' padded columns
    int px0, npx;          // x-sweep items: blk_px[px0 .. px0 + npx)
    int py0, npy;          // y-sweep items: blk_py[py0 .. py0 + npy)
    int cfl_nblk, g_nblk;  // CFL and ghost-fill blocks of all members
};

// Column blocks of a sweep over k members of `nt` interior transverse cells
// (k (nt + 6) padded columns, the outputs on lanes 3..60 of each block).
__host__ __device__ inline int pack_blocks(int64_t nt, int k, int out_cols)
{
    return (int)((k * (nt + 6) - 6 + out_cols - 1) / out_cols);
}

struct BatchTables {       // device pointers and grid sizes
    const BatchInst* inst;
    const int* blk_x;
    const int* blk_y;
    const int* blk_c;
    const int* blk_g;
    const int* blk_f;
    int n_x, n_y, n_c, n_g, n_f, n_inst;
    // the persistent step's packs (n_packs = 0 where it does not run):
    // inst_pack[i] the pack of instance i, blk_px / blk_py the pack of each
    // x / y sweep item
    const BatchPack* packs;
    const int* pack_inst;
    const int* inst_pack;
    const int* blk_px;
    const int* blk_py;
    int n_packs, n_px, n_py;
};

struct BatchCommon {
    double gam, jp_lo, jp_hi;
    int cons, force_hlle;
    float split_jump = SPLIT_JUMP_DEFAULT;   // FP32-split (SweepPlan::split_jump)
};

constexpr int GHOST_TB = 256;

// ---- persistent batched step (blocked FP64 fast step) -------------------------
// One launch advances every instance `steps` steps.  The instances are split
// over PERSIST_QUEUES work queues, one per XCD (each with its own L2), and
// within a queue into packs (BatchPack): a workgroup (one wavefront) takes
// work items, in ticket order, from the queue of the XCD it runs on and runs
// each to its end, so every item of an instance runs on one XCD and the items
// hand data to each other through that XCD's L2.  An item first waits, on a
// counter of its pack, for the items whose results it reads.  The items of
// step k of a queue, in ticket order: the CFL blocks of its instances (n_f),
// then for each stage s = 1..3 the x-sweep blocks of its packs (n_x), their
// y-sweep blocks (n_y) and the ghost-fill blocks of its instances (n_g).
// Waits, all within one pack:
//   CFL block, step k > 0   on the previous step's end    (PC_STEP >= k)
//   x sweep, stage 1        on the step's dt              (PC_CFL >= k + 1)
//   x sweep, stage s > 1    on the ghost fill of stage s - 1
//   y sweep                 on the x sweep of its stage
//   ghost fill              on the y sweep of its stage
// (counts of done items of the pack, summed over the launch's steps).  The
// last CFL block of a pack's step forms every member's dt, the last stage-3
// ghost-fill block advances every member's T and step count.  Every item
// waits only for items of its queue with smaller tickets, so the oldest
// unfinished item of a queue can always run; persist_item is shared with a
// host walk of every queue (build_tables) that checks this for the packs.
enum PersistCounter : int {
    PC_FARR = 0,    // CFL blocks arrived (the last of a step forms dt)
    PC_CFL = 1,     // dt formed (one per step)
    PC_SX = 2,      // + 3 (s - 1): x-sweep blocks of stage s done
    PC_SY = 3,      // + 3 (s - 1): y-sweep blocks of stage s done
    PC_SG = 4,      // + 3 (s - 1): ghost-fill blocks of stage s done (stage 3: arrived)
    PC_STEP = 11,   // step ended (T, step advanced)
    PC_N = 12       // counters per pack
};
enum PersistKind : int { PK_CFL = 0, PK_X = 1, PK_Y = 2, PK_G = 3 };
constexpr int PERSIST_STEPS = 8;   // steps per launch of the captured multi-step graph
// Members per pack, at most (wenox_batch_set_pack_max): measured on the
// 512-instance batch of sizes {64..384}^2 with the packs' lane-row choice.
constexpr int PACK_MAX_DEFAULT = 8;
// Spin limit of a dependency wait: about a second, far beyond any wait the
// item order allows (the items an item waits for have smaller tickets and are
// already running); passing it sets the error flag, which also ends every
// other wait, so the launch always drains.
constexpr int PERSIST_SPIN_LIMIT = 1 << 22;

struct PersistItem {
    int kind, step, stage;
    int pack;                    // the pack whose counters the item waits on and signals
    int inst;                    // CFL and ghost-fill blocks: their instance (sweeps: -1)
    int local;                   // block of the instance (CFL, ghost fill) or of the pack (sweeps)
    int wait_ctr, wait_target;   // wait_ctr < 0: no wait
    int sig_ctr;                 // counter the finished item increments
    int last_count;              // > 0: the increment that reaches it finishes the step's
    int done_ctr;                //   CFL (dt) or the step (advance), then increments done_ctr
};

__host__ __device__ inline int batch_items_g(const BatchInst& I)
{
    return (int)((6 * (I.ny + 6) + 6 * I.nx + GHOST_TB - 1) / GHOST_TB);
}
// One work queue: the global indices (into blk_f, blk_px, blk_py, blk_g) of
// its items of one step, items[off ..]: n_f CFL blocks, n_x x-sweep blocks,
// n_y y-sweep blocks, n_g ghost-fill blocks (the sweep and ghost lists serve
// each of the three stages).
struct PersistQueue {
    int n_f, n_x, n_y, n_g, off;
};
constexpr int PERSIST_QUEUES = 8;   // XCDs of the MI355X

__host__ __device__ inline int persist_items_per_step(const PersistQueue& Q)
{
    return Q.n_f + 3 * (Q.n_x + Q.n_y + Q.n_g);
}

// Work item of ticket t of queue Q (0 <= t < steps * persist_items_per_step(Q)).
__host__ __device__ inline PersistItem persist_item(int t, const PersistQueue& Q, const int* items,
                                                    const BatchTables& B)
{
    PersistItem it{};
    const int* const L = items + Q.off;
    const int per_stage = Q.n_x + Q.n_y + Q.n_g;
    const int per_step = Q.n_f + 3 * per_stage;
    it.step = t / per_step;
    int u = t - it.step * per_step;
    const int k1 = it.step + 1;
    it.inst = -1;
    it.wait_ctr = -1;
    it.wait_target = 0;
    it.last_count = 0;
    it.done_ctr = -1;
    if (u < Q.n_f) {
        const int g = L[u];
        it.kind = PK_CFL;
        it.stage = 0;
        it.inst = B.blk_f[g];
        it.pack = B.inst_pack[it.inst];
        it.local = g - B.inst[it.inst].cfl_blk0;
        if (it.step > 0) {
            it.wait_ctr = PC_STEP;
            it.wait_target = it.step;
        }
        it.sig_ctr = PC_FARR;
        it.last_count = k1 * B.packs[it.pack].cfl_nblk;
        it.done_ctr = PC_CFL;
        return it;
    }
    u -= Q.n_f;
    const int s = u / per_stage;   // stage - 1
    u -= s * per_stage;
    it.stage = s + 1;
    if (u < Q.n_x) {
        const int g = L[Q.n_f + u];
        it.kind = PK_X;
        it.pack = B.blk_px[g];
        const BatchPack& K = B.packs[it.pack];
        it.local = g - K.px0;
        if (s == 0) {
            it.wait_ctr = PC_CFL;
            it.wait_target = k1;
        } else {
            it.wait_ctr = PC_SG + 3 * (s - 1);
            it.wait_target = k1 * K.g_nblk;
        }
        it.sig_ctr = PC_SX + 3 * s;
    } else if (u < Q.n_x + Q.n_y) {
        const int g = L[Q.n_f + u];
        it.kind = PK_Y;
        it.pack = B.blk_py[g];
        const BatchPack& K = B.packs[it.pack];
        it.local = g - K.py0;
        it.wait_ctr = PC_SX + 3 * s;
        it.wait_target = k1 * K.npx;
        it.sig_ctr = PC_SY + 3 * s;
    } else {
        const int g = L[Q.n_f + u];
        it.kind = PK_G;
        it.inst = B.blk_g[g];
        it.pack = B.inst_pack[it.inst];
        const BatchPack& K = B.packs[it.pack];
        it.local = g - B.inst[it.inst].ghost_blk0;
        it.wait_ctr = PC_SY + 3 * s;
        it.wait_target = k1 * K.npy;
        it.sig_ctr = PC_SG + 3 * s;
        if (s == 2) {
            it.last_count = k1 * K.g_nblk;
            it.done_ctr = PC_STEP;
        }
    }
    return it;
}

// Device words after the n_inst * PC_N counters (packs use the first
// n_packs * PC_N of them): the queues' tickets, the
// count of exited workgroups, the error flag (1: a wait passed its spin
// limit; 2: a queue was left with items, no workgroup running on its XCD).
constexpr int PW_TICKET = 0, PW_EXIT = PERSIST_QUEUES, PW_ERR = PERSIST_QUEUES + 1, PW_N = PERSIST_QUEUES + 2;

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
    // `steps` whole steps of every instance in one persistent launch (see
    // persist_item) on the blocked states Ub (U^n), Ab, Bb (stage states) and
    // Eb, on `grid` workgroups; queues[PERSIST_QUEUES] with their item lists
    // in items (device); cnt holds n_inst * PC_N counters and the PW_N words
    // (tickets, exit count, error flag), all but the error flag cleared by a
    // fill kernel before the launch.
    void (*persist)(const BatchTables& t, const BatchCommon& c, BcParams bp, double* Ub, double* Ab, double* Bb,
                    double* Eb, StepCtl* ctls, const PersistQueue* queues, const int* items, int* cnt, int steps,
                    int grid, int spin_limit, hipStream_t s);
    // resident workgroups of the persistent kernel on the current device
    // (occupancy x compute units), for c's scheme
    int (*persist_grid)(const BatchCommon& c);
    // the XCDs a launch of `grid` workgroups runs on (bit i: XCC id i; bit
    // 31: ids of 31 and more), through `words` (32 device ints); 0 on error
    unsigned (*xcc_mask)(int grid, int* words, hipStream_t s);
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

// The blocked step's y sweep writes the ghost cells of the state it forms,
// in place of a ghost-fill launch, when every ghost cell's source lies in the
// sweep tile that holds the ghost's nearest interior cell: the local rules
// (not periodic) on grids whose edge tiles are at least 3 cells wide along x
// (58-row blocks of the lanes) and along y (the groups of the march).
inline bool ghost_in_ysweep(const SweepPlan& p)
{
    if (!p.ghost || p.bp.kind == BC_PERIODIC || p.nx < 3 || p.ny < 3) return false;
    const int64_t w = p.waves_y, c = p.chunk_y;
    const int64_t g = w > 1 ? w * c - (p.force_hlle ? 1 : 3) : c;   // cells per workgroup along y
    if (g < 3) return false;
    const int64_t rx = p.nx % SWEEP_OUT_COLS, ry = p.ny % g;
    return rx != 1 && rx != 2 && ry != 1 && ry != 2;
}

extern const SweepOps sweep_ops_exact;
extern const SweepOps sweep_ops_fast;

SweepPlan make_sweep_plan(int64_t nx, int64_t ny, double dx, double gamma, int cons,
                          int force_hlle, double jp_lo, double jp_hi);

// Flag slots of the FP32-split sweeps: the workgroups of both sweeps of a stage.
int64_t sweep_flag_count(const SweepPlan& p);

}  // namespace wenox
