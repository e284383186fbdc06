// wenox_state.h -- storage policies of the padded state and the kernels that
// touch every cell generically: the ghost fill and the single-pass CFL
// reduction.  Included inside namespace wenox::WX_NS by each kernel
// translation unit (names stay distinct per arithmetic variant).
//
// Storage policies (element (k, idx) of plane k at flat index idx):
//   StF64   double planes [4][M][N]
//   StSplit float planes [8][M][N]: value = (double)hi[k] + (double)lo[k+4],
//           written back as hi = (float)v, lo = (float)(v - hi)   (FP32-split)
#pragma once

struct StF64 {
    double* p;
    int64_t P;
    __device__ __forceinline__ double ld(int k, int64_t idx) const { return p[k * P + idx]; }
    __device__ __forceinline__ void st(int k, int64_t idx, double v) const { p[k * P + idx] = v; }
    // the value st(v) leaves in memory, as ld reads it back
    static __device__ __forceinline__ double stored(double v) { return v; }
};

struct StSplit {
    float* p;
    int64_t P;
    __device__ __forceinline__ double ld(int k, int64_t idx) const
    {
        return (double)p[k * P + idx] + (double)p[(k + 4) * P + idx];
    }
    __device__ __forceinline__ void st(int k, int64_t idx, double v) const
    {
        float h = (float)v;
        p[k * P + idx] = h;
        p[(k + 4) * P + idx] = (float)(v - (double)h);
    }
    static __device__ __forceinline__ double stored(double v)
    {
        const float h = (float)v;
        return (double)h + (double)(float)(v - (double)h);
    }
};

constexpr int ST_TB = 256;

// ------------------------------------------------------------------------
// Stepper: ghost fill on the padded SoA4 state
// ------------------------------------------------------------------------
// Ghost cells of an (nx+6) x (ny+6) array: 3 full rows at each x end plus 3
// columns at each y end of the interior rows.
__device__ __forceinline__ void ghost_index(int64_t t, int64_t nx, int64_t ny, int64_t& i, int64_t& j)
{
    const int64_t n = ny + 6;
    const int64_t rows = 6 * n;                 // x-ghost rows, all columns
    if (t < rows) {
        int64_t r = t / n;
        j = t - r * n;
        i = r < 3 ? r : nx + r;                 // 0,1,2, nx+3, nx+4, nx+5
    } else {
        t -= rows;
        int64_t r = t / 6;                      // interior row
        int64_t c = t - r * 6;
        i = 3 + r;
        j = c < 3 ? c : ny + c;                 // 0,1,2, ny+3, ny+4, ny+5
    }
}

// Even mirror (Rie_prob_2d_rs.py:123-152): ghost g <- 5-g at the low end,
// ghost N+5-g <- N+g at the high end; the y pass runs first over all rows and
// the x pass over all columns, so a corner takes the doubly mirrored cell.
__device__ __forceinline__ int64_t mirror_src(int64_t g, int64_t N)
{
    if (g < 3) return 5 - g;
    if (g >= N + 3) return 2 * N + 5 - g;
    return g;
}

__device__ __forceinline__ int64_t periodic_src(int64_t g, int64_t N)
{
    if (g < 3) return g + N;
    if (g >= N + 3) return g - N;
    return g;
}

// unitest.py:207-227, written as the final value of each ghost cell after the
// reference's sequence of row/column assignments (valid for nx+2 >= bound+10,
// ny >= 3): left ghost rows = post-shock row; top ghosts copy column ny+2;
// right ghost rows copy row nx+2; bottom ghosts of rows <= bound are post-shock,
// rows bound+1..bound+9 copy column 3, the rest mirror with rho*v negated.
__constant__ double c_dmr_post[4] = {8.0, 33.0, -57.1576, 563.499};

// Where the value of ghost cell (i, j) (padded indices) comes from: the
// interior cell (si, sj) (src 0, rho v negated when flip) or the post-shock
// (src 1) or pre-shock (src 2) state of the DMR rules.  ts is the stage time
// of the standard DMR's moving shock.
struct BcSrc {
    int64_t si, sj;
    int src;
    bool flip;
};

__device__ __forceinline__ BcSrc bc_source(int64_t i, int64_t j, int64_t nx, int64_t ny, const BcParams& bp,
                                           double ts)
{
    BcSrc r{i, j, 0, false};
    if (bp.kind == BC_COPY) {
        // x rows first, then y columns over all rows: ghost (i, j) <- (clamp i, clamp j)
        r.si = i < 3 ? 3 : (i > nx + 2 ? nx + 2 : i);
        r.sj = j < 3 ? 3 : (j > ny + 2 ? ny + 2 : j);
    } else if (bp.kind == BC_MIRROR || bp.kind == BC_PERIODIC) {
        r.si = bp.kind == BC_MIRROR ? mirror_src(i, nx) : periodic_src(i, nx);
        r.sj = bp.kind == BC_MIRROR ? mirror_src(j, ny) : periodic_src(j, ny);
    } else if (bp.kind == BC_DMR_REF) {
        const int64_t B = bp.dmr_bound;
        if (i < 3) {
            r.src = 1;
        } else {
            if (i >= nx + 3) r.si = nx + 2;
            if (j >= ny + 3) {
                r.sj = ny + 2;
            } else if (j < 3) {
                if (i >= nx + 3) {
                    r.sj = 5 - j; r.flip = true;
                } else if (i <= B) {
                    r.src = 1;
                } else if (i <= B + 9) {
                    r.sj = 3;
                } else {
                    r.sj = 5 - j; r.flip = true;
                }
            }
        }
    } else {
        // Standard double Mach reflection (Woodward & Colella 1984): left
        // inflow at the post-shock state; right outflow copies the last
        // interior row; bottom: post-shock for x < x0, reflecting wall
        // (rho*v negated) beyond; top: post-shock where x lies behind the
        // exact shock trace x0 + (1 + 20 t) / sqrt(3) at the stage time, else
        // the resting pre-shock gas.  Ghost columns of the right ghost rows
        // follow the same rules at the last interior row.
        const int64_t ie = i > nx + 2 ? nx + 2 : i;
        const double xc = ((double)(ie - 3) + 0.5) * bp.dx;
        r.si = ie;
        if (i < 3) {
            r.src = 1;
        } else if (j >= ny + 3) {
            r.src = xc < bp.x0 + (1.0 + 20.0 * ts) / 1.7320508075688772935 ? 1 : 2;
        } else if (j < 3) {
            if (xc < bp.x0) r.src = 1;
            else { r.sj = 5 - j; r.flip = true; }
        }
    }
    return r;
}

// The value of a ghost cell from its source: rd(idx of the interior cell in
// the (si, sj) frame, k) reads the state.
template <class Rd>
__device__ __forceinline__ void bc_value(const BcSrc& b, const BcParams& bp, Rd rd, double v[4])
{
    if (b.src == 0) {
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = rd(k);
        if (b.flip) v[2] = -v[2];
    } else if (bp.kind == BC_DMR_REF) {
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = c_dmr_post[k];
    } else {
        const double g = bp.gamma;
        const double rho1 = 8.0, u1 = 8.25 * 0.86602540378443864676, v1 = -8.25 * 0.5, p1 = 116.5;
        if (b.src == 1) {
            v[0] = rho1; v[1] = rho1 * u1; v[2] = rho1 * v1;
            v[3] = p1 / (g - 1.0) + 0.5 * rho1 * (u1 * u1 + v1 * v1);
        } else {
            v[0] = 1.4; v[1] = 0.0; v[2] = 0.0; v[3] = 1.0 / (g - 1.0);
        }
    }
}

// Ghost time of a state written by RK stage `stage` (0: the step's input):
// T, T + dt, T + dt/2, T + dt.  Without ctl: bp.t0.
__device__ __forceinline__ double bc_time(const StepCtl* ctl, const BcParams& bp, int stage)
{
    if (!ctl) return bp.t0;
    return ctl->T + (stage == 2 ? 0.5 * ctl->dt : (stage == 0 ? 0.0 : ctl->dt));
}

// Ghost cell number t (0 <= t < 6 (ny+6) + 6 nx) of one state U, whose
// padded cell (i, j) is element at(i, j) of its planes: the value by the rule
// bp, handed to put(i, j, v).
template <class St, class At, class Put>
__device__ __forceinline__ void bc_cell_at(St U, int64_t nx, int64_t ny, BcParams bp, const StepCtl* ctl,
                                           int stage, int64_t t, At at, Put put)
{
    if (ctl && !ctl->active) return;
    const int64_t count = 6 * (ny + 6) + 6 * nx;
    if (t >= count) return;
    int64_t i, j;
    ghost_index(t, nx, ny, i, j);
    const double ts = bp.kind == BC_DMR_STD ? bc_time(ctl, bp, stage) : 0.0;
    const BcSrc b = bc_source(i, j, nx, ny, bp, ts);
    double v[4];
    bc_value(b, bp, [&](int k) { return U.ld(k, at(b.si, b.sj)); }, v);
    put(i, j, v);
}

// Row-major planes: ld and ldT are the row pitches of U and its transposed
// copy UT (written too when has_t).
template <class St>
__device__ __forceinline__ void bc_cell(St U, St UT, bool has_t, int64_t nx, int64_t ny, int64_t ld, int64_t ldT,
                                        BcParams bp, const StepCtl* ctl, int stage, int64_t t)
{
    bc_cell_at(
        U, nx, ny, bp, ctl, stage, t, [&](int64_t i, int64_t j) { return i * ld + j; },
        [&](int64_t i, int64_t j, const double* v) {
#pragma unroll
            for (int k = 0; k < 4; ++k) U.st(k, i * ld + j, v[k]);
            if (has_t) {
#pragma unroll
                for (int k = 0; k < 4; ++k) UT.st(k, j * ldT + i, v[k]);
            }
        });
}

// ------------------------------------------------------------------------
// Single-pass CFL reduction (Rie_prob_2d_rs.py:311-319):
//   c_max = max over the interior of sqrt(g p / max(rho, 1e-14)) + sqrt(u^2+v^2)
// with (rho, u, v, p) the synced primitives of con2primi (utils.rs:9-35).
// Blocks reduce with wave shuffles and store one value each; a one-wave
// kernel takes the maximum and forms dt.  (The batched engine's few blocks per
// instance instead publish with atomicMax on the bit pattern -- a
// non-negative double orders like its bits -- and the last block finishes.)
// max is exact and order-free, so the result does not depend on the split.
// ------------------------------------------------------------------------
constexpr int CFL_BLOCKS = CFL_SLOTS;

// max that propagates NaN like numpy.max
__device__ __forceinline__ double nan_max(double a, double b)
{
    return (a != a) ? a : ((b > a || b != b) ? b : a);
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = nan_max(v, __shfl_xor(v, o, 64));
    return v;
}

template <class St>
__global__ void k_bc(St U, St UT, bool has_t, int64_t nx, int64_t ny, int64_t ld, int64_t ldT, BcParams bp,
                     const StepCtl* ctl, int stage)
{
    bc_cell<St>(U, UT, has_t, nx, ny, ld, ldT, bp, ctl, stage, (int64_t)blockIdx.x * ST_TB + threadIdx.x);
}

// dt and the loop condition from the step's wave-speed maximum
// (Rie_prob_2d_rs.py:335-336 loop condition, then :315-318).  The wave speed
// is, in order of precedence: the caller's primitives at upload (first
// step), the pressure floor's reduction after the previous step, the state's
// own reduction.  A step that does not run leaves the first two for the next
// one that does.
// c + |V| of one cell as the CFL rule forms it (Rie_prob_2d_rs.py:312-315):
// con2primi velocities and pressure, sound speed over max(rho, rho_floor)
__device__ __forceinline__ double cfl_speed(const double q[4], double gam, double rho_floor)
{
#if WX_FAST
    // the fast variant: reciprocals and roots from v_rcp_f64 / v_rsq_f64 with
    // their Newton / Goldschmidt steps (rounding-level differences)
    const double rho = fmax(q[0], K<double>::eps);
    const double ir = rcp(rho);
    const double u = q[1] * ir, v = q[2] * ir;
    const double p = fmax((gam - 1.0) * (q[3] - 0.5 * rho * (u * u + v * v)), K<double>::eps);
    return sqrt_z(gam * p * rcp(fmax(q[0], rho_floor))) + sqrt_z(u * u + v * v);
#else
    Prim<double> r = con2primi_pt(q, gam);
    double spd = sqrt(r.u * r.u + r.v * r.v);
    double c = sqrt(gam * r.p / fmax(q[0], rho_floor));
    return c + spd;
#endif
}

__device__ __forceinline__ void apply_pending(StepCtl* ctl)
{
    if (ctl->pending) {
        ctl->T += ctl->dt;
        ctl->step += 1;
        ctl->pending = 0;
    }
}

// reset_bits: clear the atomics' words (cmax_bits, blocks_done) with plain
// stores (callers that update them through atomics only clear them their way)
__device__ __forceinline__ void cfl_finish(StepCtl* ctl, double cmax_state, bool reset_bits = true)
{
    apply_pending(ctl);
    if (reset_bits) {
        ctl->cmax_bits = 0ull;
        ctl->blocks_done = 0u;
    }
    double T = ctl->T;
    bool run = ctl->step < ctl->max_steps;
    if (ctl->t_final >= 0.0) run = run && (T < ctl->t_final - 1e-12);
    ctl->active = run ? 1 : 0;
    if (!run) return;
    double cmax = ctl->use_cmax0 ? ctl->cmax0
                : (ctl->floor_valid ? __longlong_as_double((long long)ctl->floor_bits) : cmax_state);
    ctl->use_cmax0 = 0;
    ctl->floor_valid = 0;
    ctl->floor_bits = 0ull;
    double c_max = cmax + ctl->cmax_eps;
    double dt = ctl->cfl_dx / c_max;
    if (ctl->t_final >= 0.0) {
        double rem = ctl->t_final - T;
        dt = rem < dt ? rem : dt;
    }
    ctl->dt = dt;
}

// Block `blk` of `nblk` blocks reducing one state's interior, whose padded
// cell (i, j) is element at(i, j) of its planes.  With a per-block array
// (ctl->part) the block stores its maximum there and k_cfl_dt finishes;
// without one (the batched engine, a few blocks per instance) blocks combine
// with atomics and the last one finishes.
template <class St, class At>
__device__ __forceinline__ void cfl_block_at(St U, int64_t nx, int64_t ny, At at, double gam, StepCtl* ctl,
                                             int64_t blk, int64_t nblk)
{
    __shared__ double part[ST_TB / 64];
    __shared__ bool last;
    double vmaxl = 0.0;
    // rows blk, blk + nblk, ...; the block's threads stride along the row
    for (int64_t i = blk; i < nx; i += nblk) {
        for (int64_t j = threadIdx.x; j < ny; j += ST_TB) {
            const int64_t o = at(i + 3, j + 3);
            double q[4] = {U.ld(0, o), U.ld(1, o), U.ld(2, o), U.ld(3, o)};
            vmaxl = nan_max(vmaxl, cfl_speed(q, gam, ctl->rho_floor));
        }
    }
    vmaxl = wave_max(vmaxl);
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) part[wv] = vmaxl;
    __syncthreads();
    double* const slots = ctl->part;
    if (threadIdx.x == 0) {
        double b = part[0];
        for (int w = 1; w < ST_TB / 64; ++w) b = nan_max(b, part[w]);
        if (slots) {
            slots[blk] = b;
            if (blk == 0) ctl->nparts = (int)nblk;
        } else {
            atomicMax(&ctl->cmax_bits, (unsigned long long)__double_as_longlong(b));
            __threadfence();
            unsigned prev = atomicAdd(&ctl->blocks_done, 1u);
            last = (prev == (unsigned)(nblk - 1));
        }
    }
    if (slots) return;
    __syncthreads();
    if (!last || threadIdx.x != 0) return;
    __threadfence();
    unsigned long long bits = atomicAdd(&ctl->cmax_bits, 0ull);
    cfl_finish(ctl, __longlong_as_double((long long)bits));
}

// Row-major planes with row pitch ld.
template <class St>
__device__ __forceinline__ void cfl_block(St U, int64_t nx, int64_t ny, int64_t ld, double gam, StepCtl* ctl,
                                          int64_t blk, int64_t nblk)
{
    cfl_block_at(U, nx, ny, [&](int64_t i, int64_t j) { return i * ld + j; }, gam, ctl, blk, nblk);
}

template <class St>
__global__ void __launch_bounds__(ST_TB) k_cfl(St U, int64_t nx, int64_t ny, int64_t ld, double gam, StepCtl* ctl)
{
    cfl_block<St>(U, nx, ny, ld, gam, ctl, blockIdx.x, gridDim.x);
}

// One wavefront: maximum of the nblk per-block values, then dt.
__global__ void __launch_bounds__(64) k_cfl_dt(StepCtl* ctl, int nblk)
{
    double v = 0.0;
    for (int b = (int)threadIdx.x; b < nblk; b += 64) v = nan_max(v, ctl->part[b]);
    v = wave_max(v);
    if (threadIdx.x == 0) cfl_finish(ctl, v);
}

// The same over ctl->nparts values (up to one per RK-combine tile: 16384 at
// 4096^2), 1024 threads.
__global__ void __launch_bounds__(1024) k_step_dt(StepCtl* ctl)
{
    __shared__ double part[16];
    const int n = ctl->nparts;
    double v = 0.0;
    for (int b = (int)threadIdx.x; b < n; b += 1024) v = nan_max(v, ctl->part[b]);
    v = wave_max(v);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = part[0];
        for (int w = 1; w < 16; ++w) m = nan_max(m, part[w]);
        cfl_finish(ctl, m);
    }
}

__global__ void k_settle(StepCtl* ctl)
{
    if (threadIdx.x == 0) apply_pending(ctl);
}

template <class St>
static void launch_bc(St U, St UT, bool has_t, int64_t nx, int64_t ny, int64_t ld, int64_t ldT, BcParams bp,
                      const StepCtl* ctl, int stage, hipStream_t s)
{
    int64_t count = 6 * (ny + 6) + 6 * nx;
    k_bc<St><<<(unsigned)((count + ST_TB - 1) / ST_TB), ST_TB, 0, s>>>(U, UT, has_t, nx, ny, ld, ldT, bp, ctl,
                                                                       stage);
}

template <class St>
static void launch_cfl(St U, int64_t nx, int64_t ny, int64_t ld, double gam, StepCtl* ctl, hipStream_t s,
                       bool finish = true)
{
    // whole rows per block; ctl->part holds CFL_BLOCKS values
    int64_t nb = nx < CFL_BLOCKS ? nx : CFL_BLOCKS;
    if (nb < 1) nb = 1;
    k_cfl<St><<<(unsigned)nb, ST_TB, 0, s>>>(U, nx, ny, ld, gam, ctl);
    if (finish) k_cfl_dt<<<1, 64, 0, s>>>(ctl, (int)nb);
}
