/*
 * wenox.h -- C ABI of the MI355X (gfx950) WENO5 / HLLC hot path.
 *
 * Two groups of entry points:
 *
 *  1. Drop-in functions.  One per function of the reference Python extension
 *     module `weno_ext` (PyO3, weno_ext/src/lib.rs:156-170), with the same
 *     argument meaning, shapes and results.  Inputs are host arrays described
 *     by wenox_view3 (any element strides, like the reference's
 *     PyReadonlyArray3 views); outputs are caller-allocated C-contiguous host
 *     arrays of the shapes given below.  The computation runs on the current
 *     HIP device; these calls copy in, compute and copy out, so they are
 *     PCIe-bound at large sizes.
 *
 *  2. Device-resident solver.  Holds the padded conserved state in HBM and
 *     advances it with the reference's SSP-RK3 + ghost fill + CFL loop
 *     (Rie_prob_2d_rs.py:169-207,311-345; double_mach_reflection.py:97-134,
 *     229-260), hipGraph replays of one or eight time steps, dt computed on the device.
 *
 * Every function returns WENOX_OK (0) on success, a negative value for an
 * argument or shape error (the cases where the reference raises), or a
 * positive value for a HIP runtime failure; wenox_last_error() then returns a
 * message (thread-local, valid until the next call on the same thread).
 * There is no CPU fallback: without a usable gfx950 device every compute call
 * fails with WENOX_E_DEVICE.
 */
#ifndef WENOX_H
#define WENOX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WENOX_OK 0
#define WENOX_E_ARG (-1)    /* invalid argument (null pointer, bad enum, c < 4) */
#define WENOX_E_SHAPE (-2)  /* a shape the reference rejects with a Rust panic */
#define WENOX_E_DEVICE 1    /* HIP runtime error or no usable device */

/* Arithmetic variant of the device kernels. */
#define WENOX_EXACT 0  /* the reference's operation order, no FMA contraction */
#define WENOX_FAST 1   /* FMA contraction and shared reciprocals (rounding-level differences) */
#define WENOX_TWIN 2   /* the reference's NumPy twin WENO_HLLC_2d.py where it differs from the Rust:
                          exact sqrt(3)/12, its HLLC right-star coefficient and wave speeds, its HLLE
                          middle state (drop-in functions and the unfused engine only) */

/* A read-only host array view: extents and element (not byte) strides. */
typedef struct wenox_view3 {
    const double *data;
    int64_t dim[3];
    int64_t stride[3];
} wenox_view3;

const char *wenox_last_error(void);
/* Library version string. */
const char *wenox_version(void);
/* Selects the arithmetic variant used by the drop-in functions (default WENOX_EXACT;
 * WENOX_TWIN gives the NumPy twin's flux functions, WENO_HLLC_2d.py:435-757). */
int wenox_set_variant(int variant);
/* Number of visible HIP devices (0 when none); does not create a context. */
int wenox_device_count(void);

/* ---- 1. drop-in functions ------------------------------------------------ */

/* lib.rs:18-24 flux_x_py(q, gamma) -> (M, N, 4); q is (M, N, c>=4). */
int wenox_flux_x(const wenox_view3 *q, double gamma, double *out);
/* lib.rs:25-31 flux_y_py(q, gamma) -> (M, N, 4). */
int wenox_flux_y(const wenox_view3 *q, double gamma, double *out);
/* lib.rs:32-39 con2primi_py(q, gamma) -> (p, a, rho_safe, u, v, h), each (M, N). */
int wenox_con2primi(const wenox_view3 *q, double gamma, double *p, double *a, double *rho,
                    double *u, double *v, double *h);
/* lib.rs:40-46 weno_x_py(u, gamma) -> (q_l, f_l, q_r, f_r), each (m-5, n-6, 4).
 * cons_switch = 0 is the reference binding (characteristic path,
 * weno.rs:367-502); 1 selects the component-wise path (weno.rs:635-852). */
int wenox_weno_x(const wenox_view3 *u, double gamma, int cons_switch, double *q_l, double *f_l,
                 double *q_r, double *f_r);
/* lib.rs:47-53 weno_y_py(u, gamma) -> (q_l, f_l, q_r, f_r), each (m-6, n-5, 4). */
int wenox_weno_y(const wenox_view3 *u, double gamma, int cons_switch, double *q_l, double *f_l,
                 double *q_r, double *f_r);
/* lib.rs:54-64 hllc_x_rs_py(q_l, f_l, q_r, f_r, gamma) -> (F1, F2, 4), F1 x F2 from q_l. */
int wenox_hllc_x(const wenox_view3 *q_l, const wenox_view3 *f_l, const wenox_view3 *q_r,
                 const wenox_view3 *f_r, double gamma, double *out);
/* lib.rs:65-75 hllc_y_rs_py */
int wenox_hllc_y(const wenox_view3 *q_l, const wenox_view3 *f_l, const wenox_view3 *q_r,
                 const wenox_view3 *f_r, double gamma, double *out);
/* lib.rs:76-86 hlle_x_rs_py */
int wenox_hlle_x(const wenox_view3 *q_l, const wenox_view3 *f_l, const wenox_view3 *q_r,
                 const wenox_view3 *f_r, double gamma, double *out);
/* lib.rs:87-97 hlle_y_rs_py */
int wenox_hlle_y(const wenox_view3 *q_l, const wenox_view3 *f_l, const wenox_view3 *q_r,
                 const wenox_view3 *f_r, double gamma, double *out);
/* lib.rs:98-115 hllc_hlle_x_rs_py(q_l, f_l, q_r, f_r, gamma, jp_lo, jp_hi, force_hlle) */
int wenox_hllc_hlle_x(const wenox_view3 *q_l, const wenox_view3 *f_l, const wenox_view3 *q_r,
                      const wenox_view3 *f_r, double gamma, double jp_lo, double jp_hi,
                      int force_hlle, double *out);
/* lib.rs:116-133 hllc_hlle_y_rs_py */
int wenox_hllc_hlle_y(const wenox_view3 *q_l, const wenox_view3 *f_l, const wenox_view3 *q_r,
                      const wenox_view3 *f_r, double gamma, double jp_lo, double jp_hi,
                      int force_hlle, double *out);
/* lib.rs:134-153 l_local_py(u, dx, gamma, force_hlle, jp_cri, cons_switch)
 * -> (m-6, n-6, 4) = -(dF/dx + dG/dy) (riemann.rs:489-541). */
int wenox_l_local(const wenox_view3 *u, double dx, double gamma, int force_hlle, double jp_lo,
                  double jp_hi, int cons_switch, double *out);

/* l_local on device arrays, allocating nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * hipGraph capture, tight serving loops, harnesses that forbid hipMalloc).
 * u: a view of a (m, n, c>=4) array in device memory (any element strides);
 * out: a C-contiguous (m-6, n-6, 4) device array; stream: a hipStream_t (NULL:
 * the default stream).  Same results as wenox_l_local (the variant set by
 * wenox_set_variant).  The kernels are enqueued on `stream` and the call
 * returns without waiting.  The face arrays live in one 64 MiB device array
 * of the library (per device, held until the process ends), so grids are
 * limited to those whose face arrays fit (about 500 x 500 interior cells;
 * larger ones fail with WENOX_E_SHAPE).  Every call on a device shares that
 * array: calls that overlap on one device (other streams, other threads)
 * race.  The first call on a device looks the array up
 * (hipGetSymbolAddress); make one call before a graph capture if the capture
 * must not include it. */
int wenox_l_local_noalloc(const wenox_view3 *u, double dx, double gamma, int force_hlle, double jp_lo,
                          double jp_hi, int cons_switch, double *out, void *stream);

/* ---- 2. device-resident solver --------------------------------------------- */

#define WENOX_BC_MIRROR 0    /* Rie_prob_2d_rs.py:123-152 apply_bc_zero_gradient (even mirror) */
#define WENOX_BC_PERIODIC 1
#define WENOX_BC_DMR_REF 2   /* unitest.py:207-227 apply_bc_refac */
#define WENOX_BC_DMR_STD 3   /* standard double Mach reflection, exact moving shock on top
                                (domain [0, 4] x [0, 1]: ny * dx must be 1) */
#define WENOX_BC_COPY 4      /* Rie_prob_1D_rs.py:59-76 apply_bc_corrected (constant extrapolation) */

#define WENOX_ENGINE_UNFUSED 0  /* the reference's pass structure, face arrays in HBM */
#define WENOX_ENGINE_FUSED 1    /* one fused kernel per direction sweep per RK stage */

#define WENOX_PREC_FP64 0
#define WENOX_PREC_FP32SPLIT 1  /* FP32 bulk arithmetic, state as FP32 hi+lo pairs */

typedef struct wenox_solver_desc {
    int64_t nx, ny;      /* interior cells; the state is (nx+6, ny+6) with 3 ghosts */
    double dx;           /* cell size (the reference uses dx for both directions) */
    double gamma;
    double cfl;
    int cons_switch;     /* 0 characteristic, 1 component-wise reconstruction */
    int force_hlle;      /* 1: HLLE everywhere; 0: HLLC/HLLE blend (sensor jp_lo, jp_hi) */
    double jp_lo, jp_hi;
    int bc;              /* WENOX_BC_* */
    int dmr_bound;       /* WENOX_BC_DMR_REF: `bound` from unitest.init_refac */
    double dmr_x0;       /* WENOX_BC_DMR_STD: wall start (1/6) */
    int engine;          /* WENOX_ENGINE_* */
    int variant;         /* WENOX_EXACT / WENOX_FAST / WENOX_TWIN (unfused engine, FP64) */
    int precision;       /* WENOX_PREC_* */
    int use_graph;       /* capture one time step in a hipGraph and replay it */
} wenox_solver_desc;

typedef struct wenox_solver wenox_solver;

int wenox_solver_create(const wenox_solver_desc *desc, wenox_solver **out);
void wenox_solver_destroy(wenox_solver *s);
/* Upload q (nx+6, ny+6, c>=4) host array (channels 0-3 conserved; when c>=8
 * channels 4-7 are the caller's primitives (rho, u, v, p), used for the first
 * CFL step as the reference does).  Applies the solver's ghost fill and sets
 * the simulated time to t0. */
int wenox_solver_upload(wenox_solver *s, const wenox_view3 *q, double t0);
/* Download the state as (nx+6, ny+6, 8): conserved channels and the
 * primitives (rho_safe, u, v, p) of con2primi. */
int wenox_solver_download(wenox_solver *s, double *q8);
/* Advance: at most max_steps steps, and while T < t_final - 1e-12 when
 * t_final >= 0 (Rie_prob_2d_rs.py:335).  Reports the time and steps taken. */
int wenox_solver_run(wenox_solver *s, int64_t max_steps, double t_final, double *t_out,
                     int64_t *steps_out);
/* Compensated global sums of the current state over the interior cells:
 * out[0..3] = sum of rho, rho u, rho v, E (the conserved channels 0-3 of the
 * reference's state array, Rie_prob_2d_rs.py:28-32), in FP64 with TwoSum error
 * compensation (Ogita-Rump-Oishi Sum2) for FP64 and FP32-split states, in an
 * order fixed by the grid size; with n_out = 8, out[4..7] = the compensated
 * sums of their magnitudes.  Multiply by dx^2 for the totals.  No reference
 * counterpart (the reference forms no global sums); conservation checks. */
int wenox_solver_conserved_sums(wenox_solver *s, double *out, int n_out);
/* Evaluate L(U) of the current state into out (nx, ny, 4) host array. */
int wenox_solver_rhs(wenox_solver *s, double *out);
/* Average device time per step of the last run (ms), from HIP events. */
double wenox_solver_last_ms_per_step(wenox_solver *s);
/* Per-kernel device times (ms, HIP events on the solver's stream, averaged
 * over reps launches of each kernel alone), fused engine only:
 * out_ms[0] CFL reduction, [1] x sweep, [2] y sweep, [3] RK combine,
 * [4] ghost fill, [5] their sum over one step (3 stages), and with
 * n_out >= 7 [6] both sweeps launched as the step launches them ([5] then
 * uses [6] for the sweeps), with n_out >= 8 [7] one whole RK stage as the
 * step launches it (sweeps and combine, pipelined slabs included; [5] then
 * uses [7]).  With the blocked layout [1], [2], [6] and [7] are averages
 * over the three stages.  n_out >= 6. */
int wenox_solver_profile(wenox_solver *s, int reps, double *out_ms, int n_out);
/* CFL rule dt = CFL dx / (max(sqrt(g p / max(rho, rho_floor)) + |V|) + cmax_eps):
 * (1e-14, 1e-16) by default as Rie_prob_2d_rs.py:311-319; the 1-D driver
 * Rie_prob_1D_rs.py:165-168 uses (1e-10, 0). */
int wenox_solver_set_cfl_rule(wenox_solver *s, double rho_floor, double cmax_eps);
/* Pressure floor after every step (pmin > 0; 0 switches it off): where a
 * cell's pressure is below pmin its total energy is raised to give p = pmin,
 * over every padded cell.  Replaces con_var2primi_pfloor of the NumPy twin's
 * time loop (WENO_HLLC_2d.py:813-837, called at :903); FP64 states only. */
int wenox_solver_set_pfloor(wenox_solver *s, double pmin);
/* Tiling of the fused engine's sweeps: rows marched per wavefront (chunk) and
 * wavefronts per workgroup along the normal (waves: 1, 2 or 4; consecutive
 * chunks of one workgroup share their sensor halo through LDS, chunk >= 3
 * then).  No reference counterpart (the reference has no device tiling). */
int wenox_solver_plan(wenox_solver *s, int *chunk_x, int *chunk_y);
int wenox_solver_set_chunks(wenox_solver *s, int chunk_x, int chunk_y);
int wenox_solver_tiling(wenox_solver *s, int *chunk_x, int *chunk_y, int *waves_x, int *waves_y);
int wenox_solver_set_tiling(wenox_solver *s, int chunk_x, int chunk_y, int waves_x, int waves_y);
/* y extent (32 or 64 cells) of the fused engine's RK-combine tiles (32 rows
 * in x).  No reference counterpart. */
int wenox_solver_combine_cols(wenox_solver *s, int *cols);
int wenox_solver_set_combine_cols(wenox_solver *s, int cols);

/* Slab pipelining of the fused engine's FP64 fast step: with parts > 1 (at
 * most 8) each RK stage runs its sweeps in `parts` slabs of x rows on the
 * solver's stream while a second stream runs the RK combine of every slab
 * already swept (the sweeps are bound by the FP64 VALU, the combine by HBM).
 * Results do not depend on parts.  1 (the default) launches each stage's
 * sweeps and combine once.  No reference counterpart. */
int wenox_solver_pipeline(wenox_solver *s, int *parts);
int wenox_solver_set_pipeline(wenox_solver *s, int parts);

/* FP32-split precision: besides a face whose FP32 Gauss-point states are
 * inadmissible (rho <= 0, p <= 0, not finite), a face takes FP64 states when
 * its normal stencil, the normal stencils of the four faces beside it that
 * feed its Gauss points, or the face values feeding its Gauss points span an
 * increment of density or energy above tau times the centre value (default
 * 0.25; inf disables the test).  Smaller tau: more faces in FP64, smaller
 * FP32 error, slower.  The choice is per face, so the state does not depend
 * on the tiling.  No reference counterpart. */
int wenox_solver_split_jump(wenox_solver *s, double *tau);
int wenox_solver_set_split_jump(wenox_solver *s, double tau);

/* State layout of the fused engine's fast FP64 step.  0: row-major planes
 * with a transposed copy for the y sweep, and per RK stage the x and y
 * sweeps and an RK-combine launch.  1 (the default): a blocked state (a
 * cell's four components side by side, 2 x 2 cells per 128 B line) that
 * both sweeps read; per stage the x sweep writes
 * its part of the stage combination and the y sweep finishes the stage
 * (no transposed copy, no combine launch), then a ghost-fill launch.  The
 * states are the same bit for bit.  The other calls (upload, download,
 * sums, RHS, objective) work with either.  Not with the FP32-split
 * precision, the exact variant or the pressure floor.  No reference
 * counterpart. */
int wenox_solver_layout(wenox_solver *s, int *blocked);
int wenox_solver_set_layout(wenox_solver *s, int blocked);

/* ---- 3. batched solver: many independent problems of uneven sizes ---------- */
/* The n instances share the numerics of desc (nx, ny, dx of desc are ignored;
 * bc must be MIRROR, PERIODIC or COPY) and each has its own grid, dx, time and
 * CFL dt.  Step structures: every phase of a step (CFL, x and y sweeps, RK
 * combine or stage finish, ghost fill) one launch over all instances, a step
 * one hipGraph replay; or, for the blocked fast FP64 batch, the persistent
 * step (wenox_batch_set_persistent), one launch per step or per eight steps.
 * chunk: rows marched per sweep workgroup (<= 0: 16).  No counterpart in the
 * reference, which runs one problem per process; this serves its batched
 * workload. */
typedef struct wenox_batch wenox_batch;
int wenox_batch_create(const wenox_solver_desc *desc, int n, const int64_t *nx, const int64_t *ny,
                       const double *dx, int chunk, wenox_batch **out);
void wenox_batch_destroy(wenox_batch *b);
int wenox_batch_upload(wenox_batch *b, int i, const wenox_view3 *q, double t0);
int wenox_batch_download(wenox_batch *b, int i, double *q8);
/* Every instance advances at most max_steps steps (and while T < t_final - 1e-12
 * when t_final >= 0); *steps_out = steps summed over instances. */
int wenox_batch_run(wenox_batch *b, int64_t max_steps, double t_final, int64_t *steps_out);
int wenox_batch_status(wenox_batch *b, double *T, int64_t *steps);
double wenox_batch_last_ms_per_step(wenox_batch *b);
/* Rows marched per sweep workgroup (x and y sweeps) of every instance; set
 * rebuilds the launch tables.  No reference counterpart (autotuning). */
int wenox_batch_chunks(wenox_batch *b, int *chunk_x, int *chunk_y);
int wenox_batch_set_chunks(wenox_batch *b, int chunk_x, int chunk_y);
/* Device time (ms, HIP events) of one RK stage of all instances as a step
 * launches it, averaged over reps; the state is left as it was. */
int wenox_batch_profile(wenox_batch *b, int reps, double *stage_ms);
/* State layout of the batched fast FP64 step, as wenox_solver_set_layout:
 * 1 (the default for the fast FP64 batch): each instance's blocked state,
 * the y sweeps finishing every stage; 0: row-major planes with transposed
 * copies and an RK-combine launch.  The states are the same bit for bit.
 * No reference counterpart. */
int wenox_batch_layout(wenox_batch *b, int *blocked);
int wenox_batch_set_layout(wenox_batch *b, int blocked);
/* Persistent step of the blocked fast FP64 batch (on = 1, its default; the
 * getter reports the structure in effect, 0 for the row-major batch): one launch advances
 * every instance one step, or eight in the graph that wenox_batch_run replays
 * while at least eight steps remain.  Its workgroups take work items (CFL,
 * x-sweep, y-sweep and ghost-fill blocks of each pack of instances and
 * stage) in order from a device counter; an item waits only for the earlier
 * items of its own pack whose results it reads, so packs move through the
 * phases of a step independently.  Same states bit for bit as the
 * launch-per-phase step (on = 0).  A wait that passes its spin limit (about a
 * second) makes wenox_batch_run return WENOX_E_DEVICE.  No reference
 * counterpart. */
int wenox_batch_persistent(wenox_batch *b, int *on);
int wenox_batch_set_persistent(wenox_batch *b, int on);
/* Packs of the persistent step: within each work queue, instances of one
 * shape (nx, ny, dx) whose padded columns its sweeps lay side by side along
 * a wavefront's 64 lanes, so narrow instances stop leaving lanes idle (a
 * 64-wide instance alone fills two wavefronts with 64 outputs, five packed
 * fill six).  Each shape's instances form packs of at most pack_max members
 * (default 8), the size chosen for the fewest lane-rows; 1 turns packing off.
 * Same states bit for bit for every pack_max.  n_packs: packs in effect (0
 * when the persistent step does not run).  No reference counterpart. */
int wenox_batch_packs(wenox_batch *b, int *pack_max, int *n_packs);
int wenox_batch_set_pack_max(wenox_batch *b, int pack_max);
/* XCDs (bit i: XCC id i) the persistent step's work queues serve.  probed:
 * those a launch of the persistent grid reached when the batch was created
 * (all eight on an MI355X in its one-partition mode; fewer on a partitioned
 * device, where the persistent step then queues work only to those); in_use:
 * the ones the queues use, by default every probed one below 8.  set: a
 * non-empty subset of them (0: the default); the states are the same bit for
 * bit for every subset.  No reference counterpart. */
int wenox_batch_persist_xcds(wenox_batch *b, unsigned *in_use, unsigned *probed);
int wenox_batch_set_persist_xcds(wenox_batch *b, unsigned mask);
/* Polls of a persistent step's dependency wait before it gives up and sets
 * the error flag (default 4194304, about a second).  A small value forces the
 * error path, for its tests; the launch still drains.  No reference
 * counterpart. */
int wenox_batch_set_spin_limit(wenox_batch *b, int spins);
/* The FP32-split threshold of wenox_solver_set_split_jump for every
 * instance.  No reference counterpart. */
int wenox_batch_set_split_jump(wenox_batch *b, double tau);
/* Device time (ms, HIP events) per step of `steps` whole steps of every
 * instance as wenox_batch_run launches them (for the autotuner); the states
 * and step controls are restored afterwards.  No reference counterpart. */
int wenox_batch_profile_steps(wenox_batch *b, int steps, double *ms_per_step);

/* ---- 4. jp_cri tuning objective --------------------------------------------- */
/* A read-only 2-D host array view: extents and element strides. */
typedef struct wenox_view2 {
    const double *data;
    int64_t dim[2];
    int64_t stride[2];
} wenox_view2;

/* metrics.py:1-77 compute_2d_objective(rho, u, v, p, dx, dy, weights), the
 * objective opt_jp_cro.py:41-47 minimises, on the GPU.  rho, u, v: (nx, ny)
 * views of one shape (p is not read, as in the reference).  weights: {I, E,
 * O, TV} (NULL: {1, 1, 50, 1}).  *J receives the objective; metrics (8
 * values, may be NULL) receives I_curve, E_perp, O, TV, cI, cE, cO, cTV.
 * The shock mask is thinned with the Zhang-Suen algorithm, which
 * skimage.morphology.skeletonize implements for 2-D images. */
int wenox_objective_2d(const wenox_view2 *rho, const wenox_view2 *u, const wenox_view2 *v, double dx,
                       double dy, const double *weights, double *J, double *metrics);
/* The same objective of a solver's current state (FP64 states): rho the
 * conserved density, u and v the con2primi velocities of the interior, as
 * opt_jp_cro.py:31-35 takes them from main_rs's final array, with no copy of
 * the state to the host. */
int wenox_solver_objective(wenox_solver *s, double dx, double dy, const double *weights, double *J,
                           double *metrics);
/* Intermediates of the objective of rho (metrics.py:13,16,22,49): shock mask
 * S, skeleton C, band (nx*ny bytes each, 0 or 1) and the 7 x 7 median of rho
 * (nx*ny doubles); any output may be NULL.  For inspection and tests. */
int wenox_objective_fields(const wenox_view2 *rho, double dx, double dy, uint8_t *S, uint8_t *C,
                           uint8_t *band, double *median);

#ifdef __cplusplus
}
#endif

#endif /* WENOX_H */
