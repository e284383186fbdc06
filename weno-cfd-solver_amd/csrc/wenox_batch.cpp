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
    // persistent blocked step (BatchOps::persist): one launch per step, or per
    // PERSIST_STEPS steps (graph8); pcnt: n * PC_N counters, ticket, error flag
    bool persist = false;
    bool persist_fits = true;   // the tickets of a PERSIST_STEPS-step launch fit in an int
    int pgrid = 0;
    int pspin = PERSIST_SPIN_LIMIT;
    int* pcnt = nullptr;
    PersistQueue pq[PERSIST_QUEUES] = {};
    int* pitems = nullptr;   // the queues' item lists (device)
    int pack_max = PACK_MAX_DEFAULT;   // members per pack of the persistent step
    int packs_n = 0;                   // packs of the current tables
    // XCDs (bit i: XCC id i) a launch reached at creation, and those the
    // queues use (a subset of the probed ones below PERSIST_QUEUES)
    unsigned xcd_probed = 0, xcd_use = 0;
    hipGraph_t graph8 = nullptr;
    hipGraphExec_t exec8 = nullptr;
};

namespace {

bool persistent(const wenox_batch* b) { return b->blk && b->persist && b->persist_fits && b->xcd_use != 0; }

void persist_steps(wenox_batch* b, int steps)
{
    b->bops->persist(b->tab, b->com, b->bp, b->Ub, b->Ab, b->Bb, b->Eb, b->ctls, b->pq, b->pitems, b->pcnt, steps,
                     b->pgrid, b->pspin, b->stream);
}

void drop_graphs(wenox_batch* b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->exec8) (void)hipGraphExecDestroy(b->exec8);
    if (b->graph8) (void)hipGraphDestroy(b->graph8);
    b->exec = nullptr;
    b->graph = nullptr;
    b->exec8 = nullptr;
    b->graph8 = nullptr;
}

void batch_step(wenox_batch* b)
{
    hipStream_t st = b->stream;
    if (persistent(b)) {
        persist_steps(b, 1);
        return;
    }
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

// The persistent step's queues and packs.  Instances go to the queues of the
// XCDs in `xcds` (bit q: the queue of XCC id q, q < PERSIST_QUEUES) largest
// first, each to the queue with the least work so far (cells).  In a queue, the instances of one shape (nx,
// ny, dx) form packs of k members, k (at most pack_max) chosen for the fewest
// lane-rows of both sweeps (column blocks times rows marched).  A queue lists,
// per phase, the blocks of its packs (CFL and ghost-fill blocks of their
// members, sweep blocks of the packs) in pack order.
struct PackPlan {
    std::vector<BatchPack> packs;
    std::vector<int> pack_inst, inst_pack, blk_px, blk_py, items;
    PersistQueue Q[PERSIST_QUEUES] = {};
};

// Lane-rows of the two sweeps of one pack of k members of instance I's shape.
int64_t pack_cost(const BatchInst& I, int k)
{
    return (int64_t)pack_blocks(I.ny, k, SWEEP_OUT_COLS) * I.nx + (int64_t)pack_blocks(I.nx, k, SWEEP_OUT_COLS) * I.ny;
}

// Members per pack for c instances of I's shape: packs of k or k - 1 members.
std::vector<int> pack_sizes(const BatchInst& I, int c, int pack_max)
{
    const int64_t widest = std::max(I.nx, I.ny) + 6;
    int best_k = 1;
    int64_t best = -1;
    for (int k = 1; k <= std::min(c, pack_max) && k * widest < ((int64_t)1 << 30); ++k) {
        const int np = (c + k - 1) / k;
        int64_t cost = 0;
        for (int j = 0; j < np; ++j) cost += pack_cost(I, c / np + (j < c % np ? 1 : 0));
        if (best < 0 || cost < best) { best = cost; best_k = k; }
    }
    const int np = (c + best_k - 1) / best_k;
    std::vector<int> sz(np);
    for (int j = 0; j < np; ++j) sz[j] = c / np + (j < c % np ? 1 : 0);
    return sz;
}

PackPlan plan_packs(const std::vector<BatchInst>& inst, int pack_max, unsigned xcds)
{
    PackPlan P;
    const int n = (int)inst.size();
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
        return inst[a].nx * inst[a].ny > inst[c].nx * inst[c].ny;
    });
    std::vector<std::vector<int>> members(PERSIST_QUEUES);
    std::vector<int64_t> load(PERSIST_QUEUES, 0);
    for (int i : order) {
        int q = -1;
        for (int c = 0; c < PERSIST_QUEUES; ++c)
            if (((xcds >> c) & 1u) && (q < 0 || load[c] < load[q])) q = c;
        members[q].push_back(i);
        load[q] += inst[i].nx * inst[i].ny;
    }
    P.inst_pack.assign(n, -1);
    auto same = [&](int a, int c) {
        return inst[a].nx == inst[c].nx && inst[a].ny == inst[c].ny && inst[a].dx == inst[c].dx;
    };
    std::vector<std::vector<int>> qpacks(PERSIST_QUEUES);   // the packs of each queue
    for (int q = 0; q < PERSIST_QUEUES; ++q) {
        // shape groups in order of first appearance (largest first)
        std::vector<std::vector<int>> groups;
        for (int i : members[q]) {
            auto g = std::find_if(groups.begin(), groups.end(), [&](const std::vector<int>& v) { return same(v[0], i); });
            if (g == groups.end()) groups.push_back({i});
            else g->push_back(i);
        }
        for (const auto& g : groups) {
            size_t at = 0;
            for (int k : pack_sizes(inst[g[0]], (int)g.size(), pack_max)) {
                BatchPack K{};
                K.first = (int)P.pack_inst.size();
                K.k = k;
                const BatchInst& I = inst[g[at]];
                K.gx_x = pack_blocks(I.ny, k, SWEEP_OUT_COLS);
                K.gx_y = pack_blocks(I.nx, k, SWEEP_OUT_COLS);
                K.npx = K.gx_x * (int)((I.nx + I.chunk_x - 1) / I.chunk_x);
                K.npy = K.gx_y * (int)((I.ny + I.chunk_y - 1) / I.chunk_y);
                K.px0 = (int)P.blk_px.size();
                K.py0 = (int)P.blk_py.size();
                const int id = (int)P.packs.size();
                for (int j = 0; j < k; ++j) {
                    const int i = g[at + j];
                    P.pack_inst.push_back(i);
                    P.inst_pack[i] = id;
                    K.cfl_nblk += inst[i].cfl_nblk;
                    K.g_nblk += batch_items_g(inst[i]);
                }
                at += k;
                P.blk_px.insert(P.blk_px.end(), K.npx, id);
                P.blk_py.insert(P.blk_py.end(), K.npy, id);
                P.packs.push_back(K);
                qpacks[q].push_back(id);
            }
        }
    }
    for (int q = 0; q < PERSIST_QUEUES; ++q) {
        PersistQueue& Q = P.Q[q];
        auto& it = P.items;
        Q.off = (int)it.size();
        for (int id : qpacks[q]) {
            const BatchPack& K = P.packs[id];
            for (int j = 0; j < K.k; ++j) {
                const BatchInst& I = inst[P.pack_inst[K.first + j]];
                for (int b = 0; b < I.cfl_nblk; ++b) it.push_back(I.cfl_blk0 + b);
            }
        }
        Q.n_f = (int)it.size() - Q.off;
        size_t m = it.size();
        for (int id : qpacks[q])
            for (int b = 0; b < P.packs[id].npx; ++b) it.push_back(P.packs[id].px0 + b);
        Q.n_x = (int)(it.size() - m);
        m = it.size();
        for (int id : qpacks[q])
            for (int b = 0; b < P.packs[id].npy; ++b) it.push_back(P.packs[id].py0 + b);
        Q.n_y = (int)(it.size() - m);
        m = it.size();
        for (int id : qpacks[q]) {
            const BatchPack& K = P.packs[id];
            for (int j = 0; j < K.k; ++j) {
                const BatchInst& I = inst[P.pack_inst[K.first + j]];
                for (int b = 0; b < batch_items_g(I); ++b) it.push_back(I.ghost_blk0 + b);
            }
        }
        Q.n_g = (int)(it.size() - m);
    }
    return P;
}

// Every queue's tickets of a PERSIST_STEPS-step persistent launch walked in
// order, as one workgroup would run them: every item's wait must already hold
// from the items before it -- then the oldest unfinished item of a queue can
// always run, whatever the number of workgroups -- and every pack ends
// PERSIST_STEPS steps.
bool persist_order_holds(const BatchTables& th, const PersistQueue* Q, const int* items)
{
    const int np = th.n_packs;
    std::vector<int> c((size_t)np * PC_N, 0);
    for (int q = 0; q < PERSIST_QUEUES; ++q) {
        const int total = persist_items_per_step(Q[q]) * PERSIST_STEPS;
        for (int t = 0; t < total; ++t) {
            const PersistItem it = persist_item(t, Q[q], items, th);
            if (it.pack < 0 || it.pack >= np) return false;
            int* ci = c.data() + (size_t)it.pack * PC_N;
            if (it.wait_ctr >= 0 && ci[it.wait_ctr] < it.wait_target) return false;
            const int prev = ci[it.sig_ctr]++;
            if (it.last_count > 0 && prev + 1 == it.last_count) ci[it.done_ctr]++;
        }
    }
    for (int p = 0; p < np; ++p)
        if (c[(size_t)p * PC_N + PC_STEP] != PERSIST_STEPS || c[(size_t)p * PC_N + PC_CFL] != PERSIST_STEPS)
            return false;
    return true;
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
    // the persistent step's packs and queues over these blocks
    PackPlan pp;
    if (b->pcnt) {
        pp = plan_packs(b->inst, b->pack_max, b->xcd_use ? b->xcd_use : 1u);
        // tickets are ints: a batch whose launch would need more runs phase by phase
        int64_t most = 0;
        for (const PersistQueue& Q : pp.Q) most = std::max<int64_t>(most, persist_items_per_step(Q));
        b->persist_fits = most * PERSIST_STEPS < ((int64_t)1 << 30);
    }
    const size_t tbytes = sizeof(BatchInst) * b->n + sizeof(BatchPack) * pp.packs.size() +
                          sizeof(int) * (bx.size() + by.size() + bc.size() + bg.size() + bf.size() +
                                         pp.pack_inst.size() + pp.inst_pack.size() + pp.blk_px.size() +
                                         pp.blk_py.size());
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
        if (bytes) std::memcpy(h, src, bytes);
        char* dptr = dbase + (h - host.data());
        h += bytes;
        return dptr;
    };
    // BatchInst and BatchPack first: their 8-byte fields stay aligned
    b->tab.inst = (const BatchInst*)put(b->inst.data(), sizeof(BatchInst) * b->n);
    b->tab.packs = (const BatchPack*)put(pp.packs.data(), sizeof(BatchPack) * pp.packs.size());
    b->tab.blk_x = (const int*)put(bx.data(), sizeof(int) * bx.size());
    b->tab.blk_y = (const int*)put(by.data(), sizeof(int) * by.size());
    b->tab.blk_c = (const int*)put(bc.data(), sizeof(int) * bc.size());
    b->tab.blk_g = (const int*)put(bg.data(), sizeof(int) * bg.size());
    b->tab.blk_f = (const int*)put(bf.data(), sizeof(int) * bf.size());
    b->tab.pack_inst = (const int*)put(pp.pack_inst.data(), sizeof(int) * pp.pack_inst.size());
    b->tab.inst_pack = (const int*)put(pp.inst_pack.data(), sizeof(int) * pp.inst_pack.size());
    b->tab.blk_px = (const int*)put(pp.blk_px.data(), sizeof(int) * pp.blk_px.size());
    b->tab.blk_py = (const int*)put(pp.blk_py.data(), sizeof(int) * pp.blk_py.size());
    b->tab.n_x = (int)bx.size(); b->tab.n_y = (int)by.size(); b->tab.n_c = (int)bc.size();
    b->tab.n_g = (int)bg.size(); b->tab.n_f = (int)bf.size(); b->tab.n_inst = b->n;
    b->tab.n_packs = (int)pp.packs.size();
    b->tab.n_px = (int)pp.blk_px.size();
    b->tab.n_py = (int)pp.blk_py.size();
    BCHK(hipMemcpy(b->dev_tables, host.data(), tbytes, hipMemcpyHostToDevice));
    b->chunk_x = chunk_x;
    b->chunk_y = chunk_y;
    // the queues' order walked on the host, then their item lists uploaded
    if (b->pcnt && b->persist_fits) {
        BatchTables th = b->tab;
        th.inst = b->inst.data();
        th.blk_x = bx.data(); th.blk_y = by.data(); th.blk_c = bc.data(); th.blk_g = bg.data(); th.blk_f = bf.data();
        th.packs = pp.packs.data();
        th.pack_inst = pp.pack_inst.data();
        th.inst_pack = pp.inst_pack.data();
        th.blk_px = pp.blk_px.data();
        th.blk_py = pp.blk_py.data();
        if (!persist_order_holds(th, pp.Q, pp.items.data()))
            return abi_fail(WENOX_E_DEVICE, "persistent batched step: an item would wait for a later ticket");
        for (int q = 0; q < PERSIST_QUEUES; ++q) b->pq[q] = pp.Q[q];
        if (b->pitems) BCHK(hipFree(b->pitems));
        b->pitems = nullptr;
        BCHK(hipMalloc(&b->pitems, sizeof(int) * std::max<size_t>(pp.items.size(), 1)));
        BCHK(hipMemcpy(b->pitems, pp.items.data(), sizeof(int) * pp.items.size(), hipMemcpyHostToDevice));
        b->packs_n = (int)pp.packs.size();
    }
    // the captured steps hold the old grids
    drop_graphs(b);
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

// The persistent step's error flag: set when a dependency wait passed its spin
// limit (the launch still drains; its results are not valid).  Cleared on read.
int persist_error(wenox_batch* b)
{
    if (!b->pcnt) return WENOX_OK;
    int* flag = b->pcnt + (size_t)b->n * PC_N + PW_ERR;
    int e = 0;
    BCHK(hipMemcpyAsync(&e, flag, sizeof(int), hipMemcpyDeviceToHost, b->stream));
    BCHK(hipStreamSynchronize(b->stream));
    if (!e) return WENOX_OK;
    BCHK(hipMemsetAsync(flag, 0, sizeof(int), b->stream));
    BCHK(hipStreamSynchronize(b->stream));
    if (e == 2)
        return abi_fail(WENOX_E_DEVICE, "persistent batched step: a work queue was left with items (no workgroup "
                                        "ran on its XCD); the states of this run are not valid");
    return abi_fail(WENOX_E_DEVICE, "persistent batched step: a dependency wait passed its spin limit; "
                                    "the states of this run are not valid");
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
    drop_graphs(b);
    if (b->pcnt) (void)hipFree(b->pcnt);
    if (b->pitems) (void)hipFree(b->pitems);
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
    if (ok && b->blk) {
        // counters (at least the 32 words of the XCD probe)
        const size_t nc = std::max<size_t>((size_t)n * PC_N + PW_N, 32);
        ok = hipMalloc(&b->pcnt, nc * sizeof(int)) == hipSuccess;
        b->pgrid = ok ? b->bops->persist_grid(BatchCommon{d.gamma, d.jp_lo, d.jp_hi, d.cons_switch, d.force_hlle}) : 0;
        ok = ok && b->pgrid > 0;
        // the XCDs the persistent launch's workgroups reach: a partitioned
        // device (fewer XCDs) gets queues only where its workgroups run
        b->xcd_probed = ok ? b->bops->xcc_mask(b->pgrid, b->pcnt, b->stream) : 0u;
        b->xcd_use = b->xcd_probed & ((1u << PERSIST_QUEUES) - 1);
        ok = ok && hipMemset(b->pcnt, 0, nc * sizeof(int)) == hipSuccess;
        b->persist = true;   // the default step structure of the blocked fast FP64 batch
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
    const bool multi = persistent(b);   // PERSIST_STEPS steps per launch where they fit
    if (b->d.use_graph && !b->exec) {
        BCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        batch_step(b);
        BCHK(hipStreamEndCapture(b->stream, &b->graph));
        BCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    if (b->d.use_graph && multi && !b->exec8) {
        BCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        persist_steps(b, PERSIST_STEPS);
        BCHK(hipStreamEndCapture(b->stream, &b->graph8));
        BCHK(hipGraphInstantiate(&b->exec8, b->graph8, nullptr, nullptr, 0));
    }
    auto launch = [&](int64_t k) -> int {
        int64_t s = 0;
        for (; multi && s + PERSIST_STEPS <= k; s += PERSIST_STEPS) {
            if (b->exec8) BCHK(hipGraphLaunch(b->exec8, b->stream));
            else persist_steps(b, PERSIST_STEPS);
        }
        for (; s < k; ++s) {
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
    if (multi && (rc = persist_error(b))) return rc;
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

extern "C" int wenox_batch_persist_xcds(wenox_batch* b, unsigned* in_use, unsigned* probed)
{
    if (!b) return abi_fail(WENOX_E_ARG, "wenox_batch_persist_xcds: null batch");
    if (in_use) *in_use = b->xcd_use;
    if (probed) *probed = b->xcd_probed;
    return WENOX_OK;
}

extern "C" int wenox_batch_set_persist_xcds(wenox_batch* b, unsigned mask)
{
    if (!b) return abi_fail(WENOX_E_ARG, "wenox_batch_set_persist_xcds: null batch");
    const unsigned avail = b->xcd_probed & ((1u << PERSIST_QUEUES) - 1);
    if (mask == 0) mask = avail;
    if ((mask & ~avail) != 0 || mask == 0)
        return abi_fail(WENOX_E_ARG, "wenox_batch_set_persist_xcds: the mask must be a non-empty subset of the XCDs "
                                     "the probe launch reached (" + std::to_string(avail) + ")");
    if (mask == b->xcd_use) return WENOX_OK;
    BCHK(hipSetDevice(b->device));
    BCHK(hipStreamSynchronize(b->stream));
    b->xcd_use = mask;
    return build_tables(b, b->chunk_x, b->chunk_y);
}

extern "C" int wenox_batch_packs(wenox_batch* b, int* pack_max, int* n_packs)
{
    if (!b) return abi_fail(WENOX_E_ARG, "wenox_batch_packs: null batch");
    if (pack_max) *pack_max = b->pack_max;
    if (n_packs) *n_packs = persistent(b) ? b->packs_n : 0;
    return WENOX_OK;
}

extern "C" int wenox_batch_set_pack_max(wenox_batch* b, int pack_max)
{
    if (!b || pack_max < 1) return abi_fail(WENOX_E_ARG, "wenox_batch_set_pack_max: pack_max must be >= 1");
    if (pack_max == b->pack_max) return WENOX_OK;
    BCHK(hipSetDevice(b->device));
    BCHK(hipStreamSynchronize(b->stream));
    b->pack_max = pack_max;
    return build_tables(b, b->chunk_x, b->chunk_y);
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
    drop_graphs(b);
    BCHK(hipStreamSynchronize(b->stream));
    BCHK(hipGetLastError());
    return WENOX_OK;
}

extern "C" int wenox_batch_persistent(wenox_batch* b, int* on)
{
    if (!b || !on) return abi_fail(WENOX_E_ARG, "wenox_batch_persistent: null argument");
    *on = persistent(b) ? 1 : 0;
    return WENOX_OK;
}

extern "C" int wenox_batch_set_persistent(wenox_batch* b, int on)
{
    if (!b || (on != 0 && on != 1)) return abi_fail(WENOX_E_ARG, "wenox_batch_set_persistent: on is 0 or 1");
    if (on && !b->pcnt)
        return abi_fail(WENOX_E_ARG, "the persistent step serves the batched fast FP64 step (blocked layout)");
    if (on && !b->persist_fits)
        return abi_fail(WENOX_E_ARG, "the persistent step: too many work items per launch for this batch and chunk");
    if (on && !b->xcd_use)
        return abi_fail(WENOX_E_DEVICE, "the persistent step: no XCD with a work queue was reached by the probe launch");
    if ((on != 0) == b->persist) return WENOX_OK;
    BCHK(hipSetDevice(b->device));
    BCHK(hipStreamSynchronize(b->stream));
    b->persist = on != 0;
    drop_graphs(b);
    return WENOX_OK;
}

extern "C" int wenox_batch_set_split_jump(wenox_batch* b, double tau)
{
    if (!b || !(tau > 0.0)) return abi_fail(WENOX_E_ARG, "wenox_batch_set_split_jump: tau must be > 0 (inf: no jump test)");
    BCHK(hipSetDevice(b->device));
    BCHK(hipStreamSynchronize(b->stream));
    b->com.split_jump = (float)tau;
    drop_graphs(b);   // the captured launches hold the old threshold
    return WENOX_OK;
}

extern "C" int wenox_batch_set_spin_limit(wenox_batch* b, int spins)
{
    if (!b || spins < 0) return abi_fail(WENOX_E_ARG, "wenox_batch_set_spin_limit: spins must be >= 0");
    BCHK(hipSetDevice(b->device));
    BCHK(hipStreamSynchronize(b->stream));
    b->pspin = spins;
    drop_graphs(b);   // the captured launches hold the old limit
    return WENOX_OK;
}

extern "C" int wenox_batch_profile_steps(wenox_batch* b, int steps, double* ms_per_step)
{
    if (!b || steps < 1 || !ms_per_step) return abi_fail(WENOX_E_ARG, "wenox_batch_profile_steps: bad argument");
    BCHK(hipSetDevice(b->device));
    // the state the steps overwrite (blocked: Ub; row-major: U and U^T) is
    // saved in a scratch allocation and restored with the step controls
    std::vector<StepCtl> h;
    int rc = read_ctls(b, h);
    if (rc) return rc;
    const size_t sbytes = (size_t)b->inst.back().soff * 8 + (size_t)4 * (b->inst.back().nx + 6) * (b->inst.back().ny + 6) * 8;
    const size_t bbytes = b->blk ? (size_t)(b->Ab - b->Ub) * 8 : 0;
    const size_t nbytes = b->blk ? bbytes : 2 * sbytes;
    char* save = nullptr;
    BCHK(hipMalloc(&save, nbytes));
    auto copy_state = [&](bool restore) -> int {
        auto cp = [&](void* dst, const void* src, size_t n) {
            return hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToDevice, b->stream);
        };
        hipError_t e;
        if (b->blk) e = restore ? cp(b->Ub, save, bbytes) : cp(save, b->Ub, bbytes);
        else {
            e = restore ? cp(b->U, save, sbytes) : cp(save, b->U, sbytes);
            if (e == hipSuccess) e = restore ? cp(b->UT, save + sbytes, sbytes) : cp(save + sbytes, b->UT, sbytes);
        }
        if (e != hipSuccess) return abi_fail(WENOX_E_DEVICE, std::string("state copy: ") + hipGetErrorString(e));
        return WENOX_OK;
    };
    std::vector<StepCtl> act = h;
    for (auto& c : act) {
        c.active = 1;
        c.max_steps = c.step + steps + 1;
        c.t_final = -1.0;
    }
    const bool multi = persistent(b);
    auto go = [&](int k) {
        int s = 0;
        for (; multi && s + PERSIST_STEPS <= k; s += PERSIST_STEPS) persist_steps(b, PERSIST_STEPS);
        for (; s < k; ++s) batch_step(b);
    };
    float ms = 0.f;
    rc = copy_state(false);
    if (!rc) rc = write_ctls(b, act);
    if (!rc) {
        go(1);
        hipError_t e = hipEventRecord(b->ev0, b->stream);
        go(steps);
        if (e == hipSuccess) e = hipEventRecord(b->ev1, b->stream);
        if (e == hipSuccess) e = hipEventSynchronize(b->ev1);
        if (e == hipSuccess) e = hipGetLastError();
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, b->ev0, b->ev1);
        if (e != hipSuccess) rc = abi_fail(WENOX_E_DEVICE, std::string("profile_steps: ") + hipGetErrorString(e));
    }
    if (!rc && multi) rc = persist_error(b);
    const int rc2 = copy_state(true);
    (void)hipStreamSynchronize(b->stream);
    (void)hipFree(save);
    if (rc) return rc;
    if (rc2) return rc2;
    *ms_per_step = ms / steps;
    return write_ctls(b, h);
}
