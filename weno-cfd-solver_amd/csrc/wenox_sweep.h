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
    int chunk_x, chunk_y;      // interior rows marched per workgroup
};

struct SweepOps {
    // stage s: Us/UsT stage input (ghost-filled), Un step start; Dx (nx*ny*4),
    // DyT (ny*nx*4) scratch; out/outT receive the next state's interior.
    void (*stage)(const SweepPlan& p, const double* Us, const double* UsT, const double* Un,
                  double* Dx, double* DyT, double* out, double* outT, int stage,
                  const StepCtl* ctl, hipStream_t s);
    // L(U) (SoA4 interior) only.
    void (*rhs)(const SweepPlan& p, const double* U, const double* UT, double* Dx, double* DyT,
                double* L, const StepCtl* ctl, hipStream_t s);
    // out[k][j][i] = in[k][i][j] for an M x N padded plane set
    void (*transpose)(const double* in, double* out, int64_t M, int64_t N, hipStream_t s);
};

extern const SweepOps sweep_ops_exact;
extern const SweepOps sweep_ops_fast;

SweepPlan make_sweep_plan(int64_t nx, int64_t ny, double dx, double gamma, int cons,
                          int force_hlle, double jp_lo, double jp_hi);

}  // namespace wenox
