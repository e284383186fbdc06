// wenox_batch.cpp -- C ABI of the batched engine: many independent 2D problems
// of uneven sizes advanced together, each with its own dt (per-instance CFL),
// every phase of a step one launch over all instances, the step one hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/wenox.h"
#include "wenox_ops.h"
#include "wenox_sweep.h"

using namespace wenox;

namespace wenox {
int abi_fail(int code, const std::string& msg);   // wenox_abi.cpp
}

namespace {

#define BCHK(expr)                                                                                  \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            return abi_fail(WENOX_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));    \
    } while (0)

int ceil_div(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

}  // namespace

struct wenox_batch {
    wenox_solver_desc d{};
    int n = 0;
    int split = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<BatchInst> inst;
    BatchTables tab{};
    BatchCommon com{};
    BcParams bp{};
    const BatchOps* bops = nullptr;
    const SweepOps* sops = nullptr;
    const Ops* ops = nullptr;
    void* dev_tables = nullptr;
    double *U = nullptr, *A = nullptr, *B = nullptr, *UT = nullptr, *AT = nullptr, *BT = nullptr;
    double *Dx = nullptr, *DyT = nullptr, *work = nullptr;
    int* flags = nullptr;
    StepCtl* ctls = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double last_ms = 0.0;
    int chunk_x = 16, chunk_y = 16;
    size_t flags_n = 0;
    // blocked FP64 fast step (the default where it applies): the instances'
    // blocked states (BatchInst::boff, nb) in Ub (U^n), Ab, Bb (stage states)
    // and Eb (the x sweeps' part of each stage), one allocation; the
    // row-major U holds an instance's state only after sync_rows
    bool blk = false;
    bool rows_stale = false;
    double* blk_mem = nullptr;
    double *Ub = nullptr, *Ab = nullptr, *Bb = nullptr, *Eb = nullptr;
};

namespace {

void batch_step(wenox_batch* b)
{
    hipStream_t st = b->stream;
    if (b->blk) {
        double* src[3] = {b->Ub, b->Ab, b->Bb};
        double* dst[3] = {b->Ab, b->Bb, b->Ub};
        b->bops->cfl_blk(b->tab, b->Ub, b->d.gamma, b->ctls, st);
        for (int s = 1; s <= 3; ++s) {
            b->bops->stage_blk(b->tab, b->com, src[s - 1], b->Ub, b->Eb, dst[s - 1], s, b->ctls, st);
            b->bops->bc_blk(b->tab, b->bp, dst[s - 1], b->ctls, s, st);
        }
        b->bops->advance(b->ctls, b->n, st);
        return;
    }
    void* src[3] = {b->U, b->A, b->B};
    void* dst[3] = {b->A, b->B, b->U};
    void* srcT[3] = {b->UT, b->AT, b->BT};
    void* dstT[3] = {b->AT, b->BT, b->UT};
    b->bops->cfl(b->tab, b->split, b->U, b->d.gamma, b->ctls, st);
    for (int s = 1; s <= 3; ++s) {
        b->bops->stage(b->tab, b->com, b->split, src[s - 1], srcT[s - 1], b->U, b->Dx, b->DyT, dst[s - 1],
                       dstT[s - 1], b->flags, s, b->ctls, st);
        b->bops->bc(b->tab, b->bp, b->split, dst[s - 1], dstT[s - 1], b->ctls, s, st);
    }
    b->bops->advance(b->ctls, b->n, st);
}

// Instance table and workgroup maps for sweep chunks (chunk_x, chunk_y),
// uploaded to the device (the previous tables are freed); the instances'
// offsets and sizes stay.
int build_tables(wenox_batch* b, int chunk_x, int chunk_y)
{
    std::vector<int> bx, by, bc, bg, bf;
    for (int i = 0; i < b->n; ++i) {
        BatchInst& I = b->inst[i];
        I.chunk_x = chunk_x;
        I.chunk_y = chunk_y;
        I.gx_x = ceil_div(I.ny, SWEEP_OUT_COLS);
        I.gx_y = ceil_div(I.nx, SWEEP_OUT_COLS);
        I.blk0_x = (int)bx.size();
        for (int k = 0; k < I.gx_x * ceil_div(I.nx, chunk_x); ++k) bx.push_back(i);
        I.blk0_y = (int)by.size();
        for (int k = 0; k < I.gx_y * ceil_div(I.ny, chunk_y); ++k) by.push_back(i);
        I.tiles_j = ceil_div(I.ny, COMBINE_TILE);
        I.tile0 = (int)bc.size();
        for (int k = 0; k < I.tiles_j * ceil_div(I.nx, COMBINE_TILE); ++k) bc.push_back(i);
        I.ghost_blk0 = (int)bg.size();
        for (int k = 0; k < ceil_div(6 * (I.ny + 6) + 6 * I.nx, GHOST_TB); ++k) bg.push_back(i);
        I.cfl_blk0 = (int)bf.size();
        I.cfl_nblk = std::min(ceil_div(I.nx * I.ny, GHOST_TB), 64);
        for (int k = 0; k < I.cfl_nblk; ++k) bf.push_back(i);
    }
    const size_t tbytes = sizeof(BatchInst) * b->n +
                          sizeof(int) * (bx.size() + by.size() + bc.size() + bg.size() + bf.size());
    if (b->dev_tables) BCHK(hipFree(b->dev_tables));
    b->dev_tables = nullptr;
    BCHK(hipMalloc(&b->dev_tables, tbytes));
    // one flag per workgroup of a stage's sweep launch (x and y together)
    const size_t nf = bx.size() + by.size();
    if (nf > b->flags_n) {
        if (b->flags) BCHK(hipFree(b->flags));
        b->flags = nullptr;
        BCHK(hipMalloc(&b->flags, sizeof(int) * nf));
        b->flags_n = nf;
    }
    std::vector<char> host(tbytes);
    char* h = host.data();
    char* dbase = (char*)b->dev_tables;
    auto put = [&](const void* src, size_t bytes) {
        std::memcpy(h, src, bytes);
        char* dptr = dbase + (h - host.data());
        h += bytes;
        return dptr;
    };
    b->tab.inst = (const BatchInst*)put(b->inst.data(), sizeof(BatchInst) * b->n);
    b->tab.blk_x = (const int*)put(bx.data(), sizeof(int) * bx.size());
    b->tab.blk_y = (const int*)put(by.data(), sizeof(int) * by.size());
    b->tab.blk_c = (const int*)put(bc.data(), sizeof(int) * bc.size());
    b->tab.blk_g = (const int*)put(bg.data(), sizeof(int) * bg.size());
    b->tab.blk_f = (const int*)put(bf.data(), sizeof(int) * bf.size());
    b->tab.n_x = (int)bx.size(); b->tab.n_y = (int)by.size(); b->tab.n_c = (int)bc.size();
    b->tab.n_g = (int)bg.size(); b->tab.n_f = (int)bf.size(); b->tab.n_inst = b->n;
    BCHK(hipMemcpy(b->dev_tables, host.data(), tbytes, hipMemcpyHostToDevice));
    b->chunk_x = chunk_x;
    b->chunk_y = chunk_y;
    // the captured step holds the old grids
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    return WENOX_OK;
}

// Geometry of instance i for the layout conversions (unpitched row-major
// planes, the instance's blocked geometry).
SweepPlan inst_plan(const BatchInst& I)
{
    SweepPlan p{};
    p.nx = I.nx;
    p.ny = I.ny;
    p.ld = I.ny + 6;
    p.nb = I.nb;
    return p;
}

// After blocked steps: every instance's row-major state from its blocked one
// (a download converts its own instance only).
void sync_rows(wenox_batch* b)
{
    if (!b->blk || !b->rows_stale) return;
    for (const BatchInst& I : b->inst) b->sops->from_blk(inst_plan(I), b->Ub + I.boff, b->U + I.soff, b->stream);
    b->rows_stale = false;
}

int read_ctls(wenox_batch* b, std::vector<StepCtl>& h)
{
    h.resize(b->n);
    BCHK(hipMemcpyAsync(h.data(), b->ctls, sizeof(StepCtl) * b->n, hipMemcpyDeviceToHost, b->stream));
    BCHK(hipStreamSynchronize(b->stream));
    return WENOX_OK;
}

int write_ctls(wenox_batch* b, const std::vector<StepCtl>& h)
{
    BCHK(hipMemcpyAsync(b->ctls, h.data(), sizeof(StepCtl) * b->n, hipMemcpyHostToDevice, b->stream));
    BCHK(hipStreamSynchronize(b->stream));
    return WENOX_OK;
}

}  // namespace

extern "C" void wenox_batch_destroy(wenox_batch* b)
{
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (double* p : {b->U, b->A, b->B, b->UT, b->AT, b->BT, b->Dx, b->DyT, b->work, b->blk_mem})
        if (p) (void)hipFree(p);
    if (b->flags) (void)hipFree(b->flags);
    if (b->ctls) (void)hipFree(b->ctls);
    if (b->dev_tables) (void)hipFree(b->dev_tables);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

extern "C" int wenox_batch_create(const wenox_solver_desc* desc, int n, const int64_t* nx, const int64_t* ny,
                                  const double* dx, int chunk, wenox_batch** out)
{
    if (!desc || !out || n < 1 || !nx || !ny || !dx) return abi_fail(WENOX_E_ARG, "wenox_batch_create: bad argument");
    *out = nullptr;
    const auto& d = *desc;
    if (d.bc != WENOX_BC_MIRROR && d.bc != WENOX_BC_PERIODIC && d.bc != WENOX_BC_COPY)
        return abi_fail(WENOX_E_ARG, "the batched engine supports the mirror, periodic and copy ghost fills");
    if (d.precision != WENOX_PREC_FP64 && d.precision != WENOX_PREC_FP32SPLIT)
        return abi_fail(WENOX_E_ARG, "unknown precision");
    for (int i = 0; i < n; ++i)
        if (nx[i] < 3 || ny[i] < 3 || !(dx[i] > 0)) return abi_fail(WENOX_E_SHAPE, "each instance needs nx, ny >= 3, dx > 0");
    if (chunk < 1) chunk = 16;
    if (d.variant == WENOX_TWIN) return abi_fail(WENOX_E_ARG, "the twin variant runs on the single-problem unfused engine");
    int dev_count = 0;
    if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0)
        return abi_fail(WENOX_E_DEVICE, "no HIP device visible: the wenox kernels need an MI355X (gfx950)");
    auto* b = new wenox_batch();
    b->d = d;
    b->n = n;
    b->split = d.precision == WENOX_PREC_FP32SPLIT;
    const bool fast = b->split || d.variant == WENOX_FAST;
    b->bops = fast ? &batch_ops_fast : &batch_ops_exact;
    b->sops = fast ? &sweep_ops_fast : &sweep_ops_exact;
    b->ops = fast ? &ops_fast : &ops_exact;
    (void)hipGetDevice(&b->device);
    auto cleanup = [&](int code) {
        wenox_batch_destroy(b);
        return code;
    };
    // instance layout (back to back, unpitched; blocked states back to back)
    int64_t soff = 0, ioff = 0, maxP = 0, boff = 0;
    b->inst.resize(n);
    for (int i = 0; i < n; ++i) {
        BatchInst& I = b->inst[i];
        I = BatchInst{};
        I.nx = nx[i]; I.ny = ny[i]; I.dx = dx[i];
        I.soff = soff; I.ioff = ioff;
        const int64_t P = (I.nx + 6) * (I.ny + 6);
        soff += 4 * P;
        ioff += 4 * I.nx * I.ny;
        maxP = std::max(maxP, P);
        SweepPlan p = make_sweep_plan(I.nx, I.ny, I.dx, d.gamma, d.cons_switch, d.force_hlle, d.jp_lo, d.jp_hi);
        I.two_h_x = p.two_h_x; I.two_h_y = p.two_h_y;
        I.boff = boff;
        I.nb = p.nb;
        boff += p.Pb;
    }
    // the fast FP64 step runs on the blocked states
    b->blk = !b->split && d.variant == WENOX_FAST;
    bool ok = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreate(&b->ev0) == hipSuccess && hipEventCreate(&b->ev1) == hipSuccess;
    for (double** p : {&b->U, &b->A, &b->B, &b->UT, &b->AT, &b->BT})
        ok = ok && hipMalloc(p, (size_t)soff * 8) == hipSuccess;
    ok = ok && hipMalloc(&b->Dx, (size_t)ioff * 8) == hipSuccess && hipMalloc(&b->DyT, (size_t)ioff * 8) == hipSuccess &&
         hipMalloc(&b->work, (size_t)maxP * 10 * 8) == hipSuccess &&
         hipMalloc(&b->ctls, sizeof(StepCtl) * n) == hipSuccess;
    if (ok && b->blk) {
        ok = hipMalloc(&b->blk_mem, (size_t)boff * 4 * 8) == hipSuccess;
        if (ok) {
            b->Ub = b->blk_mem;
            b->Ab = b->Ub + boff;
            b->Bb = b->Ab + boff;
            b->Eb = b->Bb + boff;
        }
    }
    if (!ok) return cleanup(abi_fail(WENOX_E_DEVICE, "batch allocation failed (" + std::to_string((6 * soff + 2 * ioff + (b->blk ? 4 * boff : 0)) * 8 >> 20) + " MiB)"));
    int rct = build_tables(b, chunk, chunk);
    if (rct) return cleanup(rct);
    b->com = BatchCommon{d.gamma, d.jp_lo, d.jp_hi, d.cons_switch, d.force_hlle};
    b->bp.kind = d.bc;
    b->bp.gamma = d.gamma;
    std::vector<StepCtl> c(n);
    for (int i = 0; i < n; ++i) {
        c[i] = StepCtl{};
        c[i].cfl_dx = d.cfl * dx[i];
        c[i].rho_floor = 1e-14;   // Rie_prob_2d_rs.py:314-315
        c[i].cmax_eps = 1e-16;
        c[i].t_final = -1.0;
        c[i].active = 1;
    }
    int rc = write_ctls(b, c);
    if (rc) return cleanup(rc);
    *out = b;
    return WENOX_OK;
}

extern "C" int wenox_batch_upload(wenox_batch* b, int i, const wenox_view3* q, double t0)
{
    if (!b || !q || i < 0 || i >= b->n) return abi_fail(WENOX_E_ARG, "wenox_batch_upload: bad argument");
    const BatchInst& I = b->inst[i];
    const int64_t m = I.nx + 6, n = I.ny + 6, P = m * n;
    if (q->dim[0] != m || q->dim[1] != n || q->dim[2] < 4) return abi_fail(WENOX_E_SHAPE, "instance state must be (nx+6, ny+6, c>=4)");
    BCHK(hipSetDevice(b->device));
    std::vector<double> pin((size_t)P * 4);
    const int64_t s0 = q->stride[0], s1 = q->stride[1], s2 = q->stride[2];
    for (int64_t a = 0; a < m; ++a)
        for (int64_t c = 0; c < n; ++c)
            for (int k = 0; k < 4; ++k) pin[k * P + a * n + c] = q->data[a * s0 + c * s1 + k * s2];
    std::vector<StepCtl> ctl;
    int rc = read_ctls(b, ctl);
    if (rc) return rc;
    StepCtl& h = ctl[i];
    h.T = t0; h.step = 0; h.active = 1; h.use_cmax0 = 0; h.cmax_bits = 0; h.blocks_done = 0;
    if (q->dim[2] >= 8) {
        double cm = 0.0;
        for (int64_t a = 3; a < 3 + I.nx; ++a)
            for (int64_t c = 3; c < 3 + I.ny; ++c) {
                const double* v = q->data + a * s0 + c * s1;
                double w = std::sqrt(b->d.gamma * v[7 * s2] / std::max(v[0], h.rho_floor)) +
                           std::sqrt(v[5 * s2] * v[5 * s2] + v[6 * s2] * v[6 * s2]);
                if (w != w || w > cm) cm = w;
            }
        h.cmax0 = cm;
        h.use_cmax0 = 1;
    }
    if ((rc = write_ctls(b, ctl))) return rc;
    double* Ui = b->U + I.soff;
    double* UTi = b->UT + I.soff;
    StepCtl* ci = b->ctls + i;
    BcParams bp = b->bp;
    bp.dx = I.dx;
    if (b->split) {
        BCHK(hipMemcpy(b->work, pin.data(), pin.size() * 8, hipMemcpyHostToDevice));
        b->sops->to_split(b->work, (float*)Ui, P, b->stream);
        b->sops->transpose_split((const float*)Ui, (float*)UTi, m, n, n, m, b->stream);
        b->sops->bc_split((float*)Ui, (float*)UTi, I.nx, I.ny, n, m, bp, ci, 0, b->stream);
    } else {
        BCHK(hipMemcpy(Ui, pin.data(), pin.size() * 8, hipMemcpyHostToDevice));
        b->sops->transpose(Ui, UTi, m, n, n, m, b->stream);
        b->ops->bc(Ui, UTi, I.nx, I.ny, n, m, bp, ci, 0, b->stream);
        if (b->blk) b->sops->to_blk(inst_plan(I), Ui, b->Ub + I.boff, b->stream);
    }
    BCHK(hipStreamSynchronize(b->stream));
    BCHK(hipGetLastError());
    return WENOX_OK;
}

extern "C" int wenox_batch_download(wenox_batch* b, int i, double* q8)
{
    if (!b || !q8 || i < 0 || i >= b->n) return abi_fail(WENOX_E_ARG, "wenox_batch_download: bad argument");
    const BatchInst& I = b->inst[i];
    const int64_t m = I.nx + 6, n = I.ny + 6, P = m * n;
    BCHK(hipSetDevice(b->device));
    // this instance's row-major state from its blocked one
    if (b->blk) b->sops->from_blk(inst_plan(I), b->Ub + I.boff, b->U + I.soff, b->stream);
    double* w = b->work;
    const double* state = b->U + I.soff;
    if (b->split) {
        b->sops->from_split((const float*)(b->U + I.soff), w + 6 * P, P, b->stream);
        state = w + 6 * P;
    }
    b->ops->con2primi(state, lay_soa4(m, n), m, n, b->d.gamma, w, w + P, w + 2 * P, w + 3 * P, w + 4 * P,
                      w + 5 * P, b->stream);
    std::vector<double> cons((size_t)P * 4), prim((size_t)P * 6);
    BCHK(hipMemcpyAsync(cons.data(), state, P * 4 * 8, hipMemcpyDeviceToHost, b->stream));
    BCHK(hipMemcpyAsync(prim.data(), w, P * 6 * 8, hipMemcpyDeviceToHost, b->stream));
    BCHK(hipStreamSynchronize(b->stream));
    for (int64_t c = 0; c < P; ++c) {
        double* o = q8 + c * 8;
        for (int k = 0; k < 4; ++k) o[k] = cons[k * P + c];
        o[4] = prim[2 * P + c];
        o[5] = prim[3 * P + c];
        o[6] = prim[4 * P + c];
        o[7] = prim[0 * P + c];
    }
    return WENOX_OK;
}

extern "C" int wenox_batch_run(wenox_batch* b, int64_t max_steps, double t_final, int64_t* steps_out)
{
    if (!b || max_steps < 0) return abi_fail(WENOX_E_ARG, "wenox_batch_run: bad argument");
    BCHK(hipSetDevice(b->device));
    std::vector<StepCtl> c;
    int rc = read_ctls(b, c);
    if (rc) return rc;
    std::vector<long long> step0(b->n);
    for (int i = 0; i < b->n; ++i) {
        step0[i] = c[i].step;
        c[i].max_steps = c[i].step + max_steps;
        c[i].t_final = t_final;
        c[i].active = 1;
    }
    if ((rc = write_ctls(b, c))) return rc;
    // the steps write the blocked states only
    if (b->blk) b->rows_stale = true;
    if (b->d.use_graph && !b->exec) {
        BCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        batch_step(b);
        BCHK(hipStreamEndCapture(b->stream, &b->graph));
        BCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    auto launch = [&](int64_t k) -> int {
        for (int64_t s = 0; s < k; ++s) {
            if (b->exec) BCHK(hipGraphLaunch(b->exec, b->stream));
            else batch_step(b);
        }
        return WENOX_OK;
    };
    BCHK(hipEventRecord(b->ev0, b->stream));
    if (t_final < 0.0) {
        if ((rc = launch(max_steps))) return rc;
    } else {
        int64_t launched = 0;
        while (launched < max_steps) {
            int64_t chunk = std::min<int64_t>(16, max_steps - launched);
            if ((rc = launch(chunk))) return rc;
            launched += chunk;
            if ((rc = read_ctls(b, c))) return rc;
            bool any = false;
            for (const auto& x : c) any = any || (x.active && x.step < x.max_steps && x.T < t_final - 1e-12);
            if (!any) break;
        }
    }
    BCHK(hipEventRecord(b->ev1, b->stream));
    BCHK(hipStreamSynchronize(b->stream));
    BCHK(hipGetLastError());
    float ms = 0.f;
    BCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    if ((rc = read_ctls(b, c))) return rc;
    int64_t total = 0, most = 0;
    for (int i = 0; i < b->n; ++i) {
        total += c[i].step - step0[i];
        most = std::max<int64_t>(most, c[i].step - step0[i]);
    }
    b->last_ms = most > 0 ? ms / (double)most : 0.0;
    if (steps_out) *steps_out = total;
    return WENOX_OK;
}

extern "C" int wenox_batch_status(wenox_batch* b, double* T, int64_t* steps)
{
    if (!b) return abi_fail(WENOX_E_ARG, "wenox_batch_status: null batch");
    std::vector<StepCtl> c;
    int rc = read_ctls(b, c);
    if (rc) return rc;
    for (int i = 0; i < b->n; ++i) {
        if (T) T[i] = c[i].T;
        if (steps) steps[i] = c[i].step;
    }
    return WENOX_OK;
}

extern "C" double wenox_batch_last_ms_per_step(wenox_batch* b) { return b ? b->last_ms : 0.0; }

extern "C" int wenox_batch_chunks(wenox_batch* b, int* chunk_x, int* chunk_y)
{
    if (!b) return abi_fail(WENOX_E_ARG, "wenox_batch_chunks: null batch");
    if (chunk_x) *chunk_x = b->chunk_x;
    if (chunk_y) *chunk_y = b->chunk_y;
    return WENOX_OK;
}

extern "C" int wenox_batch_set_chunks(wenox_batch* b, int chunk_x, int chunk_y)
{
    if (!b || chunk_x < 1 || chunk_y < 1) return abi_fail(WENOX_E_ARG, "wenox_batch_set_chunks: chunks must be >= 1");
    BCHK(hipSetDevice(b->device));
    BCHK(hipStreamSynchronize(b->stream));
    return build_tables(b, chunk_x, chunk_y);
}

extern "C" int wenox_batch_profile(wenox_batch* b, int reps, double* stage_ms)
{
    if (!b || reps < 1 || !stage_ms) return abi_fail(WENOX_E_ARG, "wenox_batch_profile: bad argument");
    BCHK(hipSetDevice(b->device));
    std::vector<StepCtl> h;
    int rc = read_ctls(b, h);
    if (rc) return rc;
    std::vector<StepCtl> act = h;
    for (auto& c : act) {
        c.active = 1;
        c.dt = 0.0;
    }
    if ((rc = write_ctls(b, act))) return rc;
    // one RK stage (stage 2, from U into B / B^T: the step overwrites B) as a step launches it
    auto stage = [&] {
        if (b->blk)
            b->bops->stage_blk(b->tab, b->com, b->Ub, b->Ub, b->Eb, b->Bb, 2, b->ctls, b->stream);
        else
            b->bops->stage(b->tab, b->com, b->split, b->U, b->UT, b->U, b->Dx, b->DyT, b->B, b->BT, b->flags, 2,
                           b->ctls, b->stream);
    };
    stage();
    BCHK(hipEventRecord(b->ev0, b->stream));
    for (int r = 0; r < reps; ++r) stage();
    BCHK(hipEventRecord(b->ev1, b->stream));
    BCHK(hipEventSynchronize(b->ev1));
    float ms = 0.f;
    BCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    *stage_ms = ms / reps;
    return write_ctls(b, h);
}

extern "C" int wenox_batch_layout(wenox_batch* b, int* blocked)
{
    if (!b || !blocked) return abi_fail(WENOX_E_ARG, "wenox_batch_layout: null argument");
    *blocked = b->blk ? 1 : 0;
    return WENOX_OK;
}

extern "C" int wenox_batch_set_layout(wenox_batch* b, int blocked)
{
    if (!b || (blocked != 0 && blocked != 1)) return abi_fail(WENOX_E_ARG, "wenox_batch_set_layout: blocked is 0 or 1");
    if ((blocked != 0) == b->blk) return WENOX_OK;
    if (blocked && !b->blk_mem)
        return abi_fail(WENOX_E_ARG, "the blocked layout serves the batched fast FP64 step");
    BCHK(hipSetDevice(b->device));
    if (blocked) {
        // the row-major states hold every instance (rows are current when not blocked)
        for (const BatchInst& I : b->inst) b->sops->to_blk(inst_plan(I), b->U + I.soff, b->Ub + I.boff, b->stream);
        b->rows_stale = false;
        b->blk = true;
    } else {
        sync_rows(b);
        // the row-major step also reads the transposed copies
        for (const BatchInst& I : b->inst)
            b->sops->transpose(b->U + I.soff, b->UT + I.soff, I.nx + 6, I.ny + 6, I.ny + 6, I.nx + 6, b->stream);
        b->blk = false;
    }
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    BCHK(hipStreamSynchronize(b->stream));
    BCHK(hipGetLastError());
    return WENOX_OK;
}
