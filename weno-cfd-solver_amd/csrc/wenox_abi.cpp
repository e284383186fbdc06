// wenox_abi.cpp -- C ABI (include/wenox.h): the drop-in functions of the
// reference module weno_ext and the device-resident solver.
//
// Host code only; every computation is a gfx950 kernel launched through the
// Ops tables of wenox_kernels.hip / wenox_sweep.hip.  There is no CPU compute
// path: without a device every compute entry point returns WENOX_E_DEVICE.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <functional>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/wenox.h"
#include "wenox_hostpool.h"
#include "wenox_objective.h"
#include "wenox_sums.h"
#include "wenox_ops.h"
#include "wenox_sweep.h"

using namespace wenox;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(WENOX_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));   \
    } while (0)

int g_variant = WENOX_EXACT;

const Ops* ops_for(int variant)
{
    return variant == WENOX_FAST ? &ops_fast : (variant == WENOX_TWIN ? &ops_twin : &ops_exact);
}

// Checks that a gfx950 device is current; the code objects are built for it only.
// The device count, and each device's architecture, are looked up once per
// process (the queries cost hundreds of microseconds, as much as a drop-in
// call's kernels).
int require_device()
{
    static std::once_flag once;
    static int count = 0;
    std::call_once(once, [] {
        if (hipGetDeviceCount(&count) != hipSuccess) count = 0;
    });
    if (count == 0) return fail(WENOX_E_DEVICE, "no HIP device visible: the wenox kernels need an MI355X (gfx950)");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    static std::mutex mu;
    static std::vector<std::string> arch(count);
    std::lock_guard<std::mutex> lk(mu);
    if (dev < 0 || dev >= count) return fail(WENOX_E_DEVICE, "current HIP device out of range");
    if (arch[dev].empty()) {
        hipDeviceProp_t prop;
        HIPCHK(hipGetDeviceProperties(&prop, dev));
        arch[dev] = prop.gcnArchName[0] ? prop.gcnArchName : "?";
    }
    if (arch[dev].compare(0, 6, "gfx950") != 0)
        return fail(WENOX_E_DEVICE, "device is " + arch[dev] + "; the wenox kernels are built for gfx950 only");
    return WENOX_OK;
}

// ------------------------------------------------------------------------
// Drop-in workspace: one stream, a growing device arena and pinned staging.
// Calls are serialised (the reference holds the GIL for the whole call).
// ------------------------------------------------------------------------
// Slabs of a pipelined drop-in l_local (see wenox_l_local).
constexpr int DROPIN_SLABS = 4;

struct Dropin {
    std::mutex mu;
    hipStream_t stream = nullptr;
    // the pipelined l_local: input rows copied in and results copied out on
    // streams of their own; an event per slab for uploaded, computed and
    // copied out
    hipStream_t h2d = nullptr, d2h = nullptr;
    hipEvent_t ev[3][DROPIN_SLABS] = {};
    double* dev = nullptr;
    size_t dev_cap = 0;
    double* pin = nullptr;
    size_t pin_cap = 0;
};
Dropin g_drop;

int dropin_reserve(size_t dev_doubles, size_t pin_doubles)
{
    int rc = require_device();
    if (rc) return rc;
    if (!g_drop.stream) HIPCHK(hipStreamCreateWithFlags(&g_drop.stream, hipStreamNonBlocking));
    if (!g_drop.h2d) HIPCHK(hipStreamCreateWithFlags(&g_drop.h2d, hipStreamNonBlocking));
    if (!g_drop.d2h) HIPCHK(hipStreamCreateWithFlags(&g_drop.d2h, hipStreamNonBlocking));
    for (auto& row : g_drop.ev)
        for (hipEvent_t& e : row)
            if (!e) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    if (dev_doubles > g_drop.dev_cap) {
        if (g_drop.dev) HIPCHK(hipFree(g_drop.dev));
        g_drop.dev = nullptr;
        size_t cap = std::max(dev_doubles, g_drop.dev_cap * 3 / 2);
        HIPCHK(hipMalloc(&g_drop.dev, std::max<size_t>(cap, 1) * sizeof(double)));
        g_drop.dev_cap = cap;
    }
    if (pin_doubles > g_drop.pin_cap) {
        if (g_drop.pin) HIPCHK(hipHostFree(g_drop.pin));
        g_drop.pin = nullptr;
        size_t cap = std::max(pin_doubles, g_drop.pin_cap * 3 / 2);
        HIPCHK(hipHostMalloc(&g_drop.pin, std::max<size_t>(cap, 1) * sizeof(double)));
        g_drop.pin_cap = cap;
    }
    return WENOX_OK;
}

HostPool& host_pool()
{
    static HostPool* pool = new HostPool;
    return *pool;
}

// Host threads for fn over rows [0, M) holding `work` doubles in all (at
// least PAR_MIN per thread).
int par_threads(int64_t M, int64_t work)
{
    constexpr int64_t PAR_MIN = (int64_t)1 << 15;
    return (int)std::min<int64_t>({(int64_t)host_pool().threads(), M, std::max<int64_t>(work / PAR_MIN, 1)});
}

// fn(i0, i1) over row ranges [i0, i1) of [0, M), on several host threads when
// the rows hold enough doubles in all (`work`).
template <class Fn> void par_rows(int64_t M, int64_t work, Fn fn)
{
    const int T = par_threads(M, work);
    if (T <= 1) {
        fn((int64_t)0, M);
        return;
    }
    host_pool().run(M, T, std::function<void(int64_t, int64_t)>(fn));
}

// Packing of channels 0..3 of rows r0 + [a, b) x cols [0, N) of a strided
// view into the same rows of an AoS4 (M, N, 4) contiguous buffer, as a job
// over row ranges [a, b).
std::function<void(int64_t, int64_t)> pack4_job(const wenox_view3* v, int64_t r0, int64_t N, double* dst)
{
    const int64_t s0 = v->stride[0], s1 = v->stride[1], s2 = v->stride[2];
    const double* p = v->data;
    const bool dense = s2 == 1 && s1 == 4 && s0 == 4 * N;
    return [=](int64_t a, int64_t b) {
        const int64_t i0 = r0 + a, i1 = r0 + b;
        if (dense) {
            std::memcpy(dst + i0 * N * 4, p + i0 * N * 4, sizeof(double) * (size_t)((i1 - i0) * N * 4));
            return;
        }
        for (int64_t i = i0; i < i1; ++i)
            for (int64_t j = 0; j < N; ++j) {
                const double* s = p + i * s0 + j * s1;
                double* d = dst + (i * N + j) * 4;
                d[0] = s[0]; d[1] = s[s2]; d[2] = s[2 * s2]; d[3] = s[3 * s2];
            }
    };
}

void pack4(const wenox_view3* v, int64_t M, int64_t N, double* dst)
{
    par_rows(M, M * N * 4, pack4_job(v, 0, N, dst));
}

bool bad_view(const wenox_view3* v)
{
    return !v || (!v->data && v->dim[0] * v->dim[1] * v->dim[2] != 0) || v->dim[0] < 0 ||
           v->dim[1] < 0 || v->dim[2] < 0;
}

int shape_err(const char* fn, const std::string& why)
{
    return fail(WENOX_E_SHAPE, std::string(fn) + ": " + why);
}

std::string dims(const wenox_view3* v)
{
    char b[96];
    std::snprintf(b, sizeof b, "(%lld, %lld, %lld)", (long long)v->dim[0], (long long)v->dim[1],
                  (long long)v->dim[2]);
    return b;
}

// The device result into the caller's (pageable) host array: large results
// through the pinned staging buffer and a copy on several host threads, small
// ones directly.
int sync_copy_out(double* host, const double* dev, size_t n)
{
    if (n == 0) return WENOX_OK;
    if (n < ((size_t)1 << 20) || n > g_drop.pin_cap) {
        HIPCHK(hipMemcpyAsync(host, dev, n * sizeof(double), hipMemcpyDeviceToHost, g_drop.stream));
        return WENOX_OK;
    }
    double* pin = g_drop.pin;
    HIPCHK(hipMemcpyAsync(pin, dev, n * sizeof(double), hipMemcpyDeviceToHost, g_drop.stream));
    HIPCHK(hipStreamSynchronize(g_drop.stream));
    const int64_t rows = (int64_t)((n + 4095) / 4096);
    par_rows(rows, (int64_t)n, [=](int64_t r0, int64_t r1) {
        const size_t b = (size_t)r0 * 4096, e = std::min(n, (size_t)r1 * 4096);
        if (e > b) std::memcpy(host + b, pin + b, (e - b) * sizeof(double));
    });
    return WENOX_OK;
}

int upload4(const wenox_view3* v, int64_t M, int64_t N, double* dev, double* pin)
{
    size_t n = (size_t)(M * N * 4);
    if (n == 0) return WENOX_OK;
    pack4(v, M, N, pin);
    HIPCHK(hipMemcpyAsync(dev, pin, n * sizeof(double), hipMemcpyHostToDevice, g_drop.stream));
    return WENOX_OK;
}

int finish()
{
    HIPCHK(hipStreamSynchronize(g_drop.stream));
    HIPCHK(hipGetLastError());
    return WENOX_OK;
}

int flux_common(const char* fn, const wenox_view3* q, double gamma, int dir, double* out)
{
    if (bad_view(q) || (!out && q->dim[0] * q->dim[1] > 0)) return fail(WENOX_E_ARG, std::string(fn) + ": null array");
    const int64_t M = q->dim[0], N = q->dim[1];
    if (M * N > 0 && q->dim[2] < 4) return shape_err(fn, "q needs at least 4 channels, got " + dims(q));
    std::lock_guard<std::mutex> lk(g_drop.mu);
    const size_t n = (size_t)(M * N * 4);
    int rc = dropin_reserve(2 * n, n);
    if (rc) return rc;
    double* dq = g_drop.dev;
    double* dout = dq + n;
    if ((rc = upload4(q, M, N, dq, g_drop.pin))) return rc;
    if (M * N > 0)
        ops_for(g_variant)->flux(dq, lay_aos4(M, N), M, N, dir, gamma, dout, lay_aos4(M, N), g_drop.stream);
    if ((rc = sync_copy_out(out, dout, n))) return rc;
    return finish();
}

int weno_common(const char* fn, const wenox_view3* u, double gamma, int cons, int dir, double* ql,
                double* fl, double* qr, double* fr)
{
    if (bad_view(u)) return fail(WENOX_E_ARG, std::string(fn) + ": null array");
    const int64_t m = u->dim[0], n = u->dim[1];
    // weno.rs:371 assert!(c >= 4); :385 m-5 / :461 n-6 (x) and :525 n-5 / :590 m-6 (y) in usize
    if (u->dim[2] < 4) return shape_err(fn, "u must have at least 4 channels, got " + dims(u));
    if (dir == 0 && (m < 5 || n < 6)) return shape_err(fn, "x reconstruction needs m >= 5 and n >= 6, got " + dims(u));
    if (dir == 1 && (n < 5 || m < 6)) return shape_err(fn, "y reconstruction needs n >= 5 and m >= 6, got " + dims(u));
    const int64_t F1 = dir == 0 ? m - 5 : m - 6, F2 = dir == 0 ? n - 6 : n - 5;
    const int64_t p1 = dir == 0 ? (m - 5) * n * 4 : m * (n - 5) * 4;
    const size_t nu = (size_t)(m * n * 4), nf = (size_t)(F1 * F2 * 4);
    if (F1 * F2 > 0 && (!ql || !fl || !qr || !fr)) return fail(WENOX_E_ARG, std::string(fn) + ": null output");
    std::lock_guard<std::mutex> lk(g_drop.mu);
    int rc = dropin_reserve(nu + 2 * (size_t)p1 + 4 * nf, nu);
    if (rc) return rc;
    double* du = g_drop.dev;
    double* pl = du + nu;
    double* pr = pl + p1;
    double* o[4];
    o[0] = pr + p1;
    for (int k = 1; k < 4; ++k) o[k] = o[k - 1] + nf;
    if ((rc = upload4(u, m, n, du, g_drop.pin))) return rc;
    ops_for(g_variant)->weno(du, lay_aos4(m, n), m, n, dir, cons, gamma, pl, pr, o[0], o[1], o[2],
                             o[3], lay_aos4(F1, F2), g_drop.stream);
    double* outs[4] = {ql, fl, qr, fr};
    for (int k = 0; k < 4; ++k)
        if ((rc = sync_copy_out(outs[k], o[k], nf))) return rc;
    return finish();
}

// riemann.rs:6-486 entry: kind 0 HLLC, 1 HLLE, 2 blend
int riemann_common(const char* fn, const wenox_view3* ql, const wenox_view3* fl,
                   const wenox_view3* qr, const wenox_view3* fr, double gamma, int dir, int kind,
                   double lo, double hi, double* out)
{
    if (bad_view(ql) || bad_view(fl) || bad_view(qr) || bad_view(fr))
        return fail(WENOX_E_ARG, std::string(fn) + ": null array");
    const int64_t F1 = ql->dim[0], F2 = ql->dim[1];
    const wenox_view3* in[4] = {ql, fl, qr, fr};
    if (F1 * F2 > 0) {
        if (!out) return fail(WENOX_E_ARG, std::string(fn) + ": null output");
        for (int k = 0; k < 4; ++k)
            if (in[k]->dim[0] < F1 || in[k]->dim[1] < F2 || in[k]->dim[2] < 4)
                return shape_err(fn, "face arrays must cover q_l's (F1, F2, 4) " + dims(ql) + ", got " + dims(in[k]));
    }
    if (kind == 2) {
        // shock_sensor_grad_p (riemann.rs:360-387) indexes p[1], p[m-2], p[:, 1], p[:, n-2]
        if (!((F1 >= 2 && F2 >= 2) || (F1 == 0 && F2 == 0)))
            return shape_err(fn, "the pressure-gradient sensor needs a face grid of at least 2 x 2, got " + dims(ql));
        // con2primi(q_r) is formed on q_r's own extent and read on q_l's
        if (qr->dim[0] < F1 || qr->dim[1] < F2) return shape_err(fn, "q_r smaller than q_l");
    }
    std::lock_guard<std::mutex> lk(g_drop.mu);
    const size_t nf = (size_t)(F1 * F2 * 4);
    int rc = dropin_reserve(5 * nf + (size_t)(F1 * F2), 4 * nf);
    if (rc) return rc;
    double* d[4];
    for (int k = 0; k < 4; ++k) d[k] = g_drop.dev + k * nf;
    double* dout = g_drop.dev + 4 * nf;
    double* pbar = dout + nf;
    for (int k = 0; k < 4; ++k)
        if ((rc = upload4(in[k], F1, F2, d[k], g_drop.pin + k * nf))) return rc;
    if (F1 * F2 > 0)
        ops_for(g_variant)->riemann(d[0], d[1], d[2], d[3], lay_aos4(F1, F2), F1, F2, dir, kind,
                                    gamma, lo, hi, pbar, dout, lay_aos4(F1, F2), 0, g_drop.stream);
    if ((rc = sync_copy_out(out, dout, nf))) return rc;
    return finish();
}

}  // namespace

namespace wenox {
// error reporting shared with the batched engine's ABI (wenox_batch.cpp)
int abi_fail(int code, const std::string& msg) { return fail(code, msg); }
}  // namespace wenox

// ==========================================================================
// extern "C": misc
// ==========================================================================
extern "C" const char* wenox_last_error(void) { return g_err.c_str(); }

extern "C" const char* wenox_version(void) { return "wenox 0.1.0 (gfx950)"; }

extern "C" int wenox_set_variant(int variant)
{
    if (variant != WENOX_EXACT && variant != WENOX_FAST && variant != WENOX_TWIN)
        return fail(WENOX_E_ARG, "variant must be 0 (exact), 1 (fast) or 2 (twin)");
    g_variant = variant;
    return WENOX_OK;
}

extern "C" int wenox_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

// ==========================================================================
// extern "C": drop-in functions (weno_ext/src/lib.rs:18-153)
// ==========================================================================
extern "C" int wenox_flux_x(const wenox_view3* q, double gamma, double* out)
{
    return flux_common("flux_x_py", q, gamma, 0, out);
}

extern "C" int wenox_flux_y(const wenox_view3* q, double gamma, double* out)
{
    return flux_common("flux_y_py", q, gamma, 1, out);
}

extern "C" int wenox_con2primi(const wenox_view3* q, double gamma, double* p, double* a,
                               double* rho, double* u, double* v, double* h)
{
    const char* fn = "con2primi_py";
    if (bad_view(q)) return fail(WENOX_E_ARG, "con2primi_py: null array");
    const int64_t M = q->dim[0], N = q->dim[1];
    if (M * N > 0 && q->dim[2] < 4) return shape_err(fn, "q needs at least 4 channels, got " + dims(q));
    double* outs[6] = {p, a, rho, u, v, h};
    for (int k = 0; k < 6; ++k)
        if (!outs[k] && M * N > 0) return fail(WENOX_E_ARG, "con2primi_py: null output");
    std::lock_guard<std::mutex> lk(g_drop.mu);
    const size_t n4 = (size_t)(M * N * 4), n1 = (size_t)(M * N);
    int rc = dropin_reserve(n4 + 6 * n1, n4);
    if (rc) return rc;
    double* dq = g_drop.dev;
    double* o = dq + n4;
    if ((rc = upload4(q, M, N, dq, g_drop.pin))) return rc;
    if (M * N > 0)
        ops_for(g_variant)->con2primi(dq, lay_aos4(M, N), M, N, gamma, o, o + n1, o + 2 * n1,
                                      o + 3 * n1, o + 4 * n1, o + 5 * n1, g_drop.stream);
    for (int k = 0; k < 6; ++k)
        if ((rc = sync_copy_out(outs[k], o + k * n1, n1))) return rc;
    return finish();
}

extern "C" int wenox_weno_x(const wenox_view3* u, double gamma, int cons, double* ql, double* fl,
                            double* qr, double* fr)
{
    return weno_common("weno_x_py", u, gamma, cons, 0, ql, fl, qr, fr);
}

extern "C" int wenox_weno_y(const wenox_view3* u, double gamma, int cons, double* ql, double* fl,
                            double* qr, double* fr)
{
    return weno_common("weno_y_py", u, gamma, cons, 1, ql, fl, qr, fr);
}

extern "C" int wenox_hllc_x(const wenox_view3* ql, const wenox_view3* fl, const wenox_view3* qr,
                            const wenox_view3* fr, double gamma, double* out)
{
    return riemann_common("hllc_x_rs_py", ql, fl, qr, fr, gamma, 0, 0, 0, 0, out);
}

extern "C" int wenox_hllc_y(const wenox_view3* ql, const wenox_view3* fl, const wenox_view3* qr,
                            const wenox_view3* fr, double gamma, double* out)
{
    return riemann_common("hllc_y_rs_py", ql, fl, qr, fr, gamma, 1, 0, 0, 0, out);
}

extern "C" int wenox_hlle_x(const wenox_view3* ql, const wenox_view3* fl, const wenox_view3* qr,
                            const wenox_view3* fr, double gamma, double* out)
{
    return riemann_common("hlle_x_rs_py", ql, fl, qr, fr, gamma, 0, 1, 0, 0, out);
}

extern "C" int wenox_hlle_y(const wenox_view3* ql, const wenox_view3* fl, const wenox_view3* qr,
                            const wenox_view3* fr, double gamma, double* out)
{
    return riemann_common("hlle_y_rs_py", ql, fl, qr, fr, gamma, 1, 1, 0, 0, out);
}

extern "C" int wenox_hllc_hlle_x(const wenox_view3* ql, const wenox_view3* fl,
                                 const wenox_view3* qr, const wenox_view3* fr, double gamma,
                                 double jp_lo, double jp_hi, int force_hlle, double* out)
{
    return riemann_common("hllc_hlle_x_rs_py", ql, fl, qr, fr, gamma, 0, force_hlle ? 1 : 2,
                          jp_lo, jp_hi, out);
}

extern "C" int wenox_hllc_hlle_y(const wenox_view3* ql, const wenox_view3* fl,
                                 const wenox_view3* qr, const wenox_view3* fr, double gamma,
                                 double jp_lo, double jp_hi, int force_hlle, double* out)
{
    return riemann_common("hllc_hlle_y_rs_py", ql, fl, qr, fr, gamma, 1, force_hlle ? 1 : 2,
                          jp_lo, jp_hi, out);
}

namespace {

// The reference pass structure of l_local (riemann.rs:489-541) on one padded
// array of m x n cells (element layout lu) on stream s: face arrays in work
// (rhs_unfused_scratch(m - 6, n - 6) doubles), the (m - 6, n - 6, 4) AoS4
// result in dout.  nx_all: the interior rows of the whole grid when the
// array is a band of its rows (the sensor's h follows the whole grid).
void l_local_passes(const Ops* ops, const double* du, Lay lu, int64_t m, int64_t n, int64_t nx_all, double dx,
                    double gamma, int kind, int cons, double jp_lo, double jp_hi, double* work, double* dout,
                    hipStream_t s)
{
    const int64_t nx = m - 6, ny = n - 6;
    const int64_t p1 = std::max((m - 5) * n, m * (n - 5)) * 4;
    const int64_t fxs = 4 * (nx + 1) * ny, fys = 4 * nx * (ny + 1), fm = std::max(fxs, fys);
    double* pl = work;
    double* pr = pl + p1;
    double* f0 = pr + p1;
    double* fx = f0 + 4 * fm;
    double* fy = fx + fxs;
    double* pbar = fy + fys;
    Lay lx = lay_aos4(nx + 1, ny), ly = lay_aos4(nx, ny + 1);
    ops->weno(du, lu, m, n, 0, cons, gamma, pl, pr, f0, f0 + fm, f0 + 2 * fm, f0 + 3 * fm, lx, s);
    ops->riemann(f0, f0 + fm, f0 + 2 * fm, f0 + 3 * fm, lx, nx + 1, ny, 0, kind, gamma, jp_lo, jp_hi, pbar, fx, lx,
                 std::min(nx_all + 1, ny), s);
    ops->weno(du, lu, m, n, 1, cons, gamma, pl, pr, f0, f0 + fm, f0 + 2 * fm, f0 + 3 * fm, ly, s);
    ops->riemann(f0, f0 + fm, f0 + 2 * fm, f0 + 3 * fm, ly, nx, ny + 1, 1, kind, gamma, jp_lo, jp_hi, pbar, fy, ly,
                 std::min(nx_all, ny + 1), s);
    ops->divergence(fx, lx, fy, ly, nx, ny, dx, dout, lay_aos4(nx, ny), s);
}

// Interior rows [i0, i1) of slab k of P over nx rows, and the padded rows
// [a, b) its passes run on: one interior row more on each side that is not
// a grid edge, so that the rows it keeps see the same faces and the same
// central sensor differences as one pass over the whole grid.
struct Slab {
    int64_t i0, i1, a, b;
};
Slab dropin_slab(int k, int P, int64_t nx)
{
    Slab q;
    q.i0 = nx * k / P;
    q.i1 = nx * (k + 1) / P;
    q.a = k == 0 ? 0 : q.i0 - 1;
    q.b = (k == P - 1 ? nx : q.i1 + 1) + 6;
    return q;
}

}  // namespace

// Grids of at least DROPIN_PIPE_MIN output doubles run in DROPIN_SLABS slabs of
// rows: the host packs and uploads slab k + 1 while the device computes slab
// k and copies slab k - 1 back, which the host then copies into the caller's
// array.  A slab's passes cover one row more on each inner side and keep the
// rows in between, so every cell's result is the one a single pass gives, bit
// for bit.  Measured at the double Mach reflection's size (1926 x 486 x 8
// input, fast variant, median of 30 calls): one pass 2.86-2.91 ms a call; 2,
// 3, 4, 6, 8 slabs 2.24, 1.89-1.99, 1.86-2.08, 2.14-2.26, 2.32-2.42 ms (each
// slab costs about 0.1 ms of launches and copies issued from the host).
constexpr size_t DROPIN_PIPE_MIN = (size_t)1 << 20;

extern "C" int wenox_l_local(const wenox_view3* u, double dx, double gamma, int force_hlle,
                             double jp_lo, double jp_hi, int cons, double* out)
{
    const char* fn = "l_local_py";
    if (bad_view(u)) return fail(WENOX_E_ARG, "l_local_py: null array");
    const int64_t m = u->dim[0], n = u->dim[1];
    if (u->dim[2] < 4) return shape_err(fn, "u must have at least 4 channels, got " + dims(u));
    if (m < 6 || n < 6) return shape_err(fn, "u needs m >= 6 and n >= 6 (3 ghost cells per side), got " + dims(u));
    if (!force_hlle && (m < 8 || n < 8))
        return shape_err(fn, "the HLLC/HLLE blend's sensor needs at least 2 x 2 interior cells, got " + dims(u));
    const int64_t nx = m - 6, ny = n - 6;
    const size_t nu = (size_t)(m * n * 4), no = (size_t)(nx * ny * 4);
    if (no > 0 && !out) return fail(WENOX_E_ARG, "l_local_py: null output");
    const Ops* ops = ops_for(g_variant);
    const int kind = force_hlle ? 1 : 2;
    std::lock_guard<std::mutex> lk(g_drop.mu);
    const int P = no >= DROPIN_PIPE_MIN ? (int)std::min<int64_t>(DROPIN_SLABS, nx / 16) : 1;
    if (P <= 1) {
        const size_t scratch = rhs_unfused_scratch(nx, ny);
        int rc = dropin_reserve(nu + scratch + no, nu);
        if (rc) return rc;
        double* du = g_drop.dev;
        double* work = du + nu;
        double* dout = work + scratch;
        if ((rc = upload4(u, m, n, du, g_drop.pin))) return rc;
        l_local_passes(ops, du, lay_aos4(m, n), m, n, nx, dx, gamma, kind, cons, jp_lo, jp_hi, work, dout,
                       g_drop.stream);
        if ((rc = sync_copy_out(out, dout, no))) return rc;
        return finish();
    }
    // device: the input, each slab's result rows (with its extra rows), the
    // passes' face arrays for the largest slab; pinned: input and result
    size_t slab_out = 0, scratch = 0;
    for (int k = 0; k < P; ++k) {
        const Slab q = dropin_slab(k, P, nx);
        slab_out += (size_t)(q.b - 6 - q.a) * ny * 4;
        scratch = std::max(scratch, rhs_unfused_scratch(q.b - 6 - q.a, ny));
    }
    int rc = dropin_reserve(nu + slab_out + scratch, nu + no);
    if (rc) return rc;
    double* du = g_drop.dev;
    double* dslab = du + nu;
    double* work = dslab + slab_out;
    double* pin_in = g_drop.pin;
    double* pin_out = pin_in + nu;
    const int64_t row = n * 4, orow = ny * 4;   // doubles per input / output row
    // the host pool packs the next slab's input rows while this thread issues
    // the current slab's copies and kernels (an early return still waits for
    // the packing before the caller's array may go)
    struct PackJob {
        bool on = false;
        void start(const wenox_view3* v, int64_t r0, int64_t r1, int64_t N, double* dst)
        {
            // two host threads fewer than the pool has: this thread and the
            // runtime's keep running meanwhile
            const int T = std::max(1, std::min(par_threads(r1 - r0, (r1 - r0) * N * 4), host_pool().threads() - 2));
            host_pool().start(r1 - r0, T, pack4_job(v, r0, N, dst));
            on = true;
        }
        void wait()
        {
            if (on) host_pool().wait();
            on = false;
        }
        ~PackJob() { wait(); }
    } pack;
    // every exit, an error's included, leaves no copy in flight on the staging
    // buffers or the arena (declared after `pack`, so it runs first)
    struct Drain {
        ~Drain()
        {
            (void)hipStreamSynchronize(g_drop.h2d);
            (void)hipStreamSynchronize(g_drop.stream);
            (void)hipStreamSynchronize(g_drop.d2h);
        }
    } drain;
    pack.start(u, 0, dropin_slab(0, P, nx).b, n, pin_in);
    size_t off = 0;
    int64_t up = 0;   // input rows uploaded
    for (int k = 0; k < P; ++k) {
        const Slab q = dropin_slab(k, P, nx);
        // rows [up, q.b): the input rows this slab needs beyond those already on their way
        pack.wait();
        if (k + 1 < P) pack.start(u, q.b, dropin_slab(k + 1, P, nx).b, n, pin_in);
        HIPCHK(hipMemcpyAsync(du + up * row, pin_in + up * row, sizeof(double) * (size_t)((q.b - up) * row),
                              hipMemcpyHostToDevice, g_drop.h2d));
        HIPCHK(hipEventRecord(g_drop.ev[2][k], g_drop.h2d));
        HIPCHK(hipStreamWaitEvent(g_drop.stream, g_drop.ev[2][k], 0));
        up = q.b;
        double* dk = dslab + off;
        l_local_passes(ops, du + q.a * row, lay_aos4(q.b - q.a, n), q.b - q.a, n, nx, dx, gamma, kind, cons, jp_lo,
                       jp_hi, work, dk, g_drop.stream);
        HIPCHK(hipEventRecord(g_drop.ev[0][k], g_drop.stream));
        HIPCHK(hipStreamWaitEvent(g_drop.d2h, g_drop.ev[0][k], 0));
        HIPCHK(hipMemcpyAsync(pin_out + q.i0 * orow, dk + (q.i0 - q.a) * orow,
                              sizeof(double) * (size_t)((q.i1 - q.i0) * orow), hipMemcpyDeviceToHost, g_drop.d2h));
        HIPCHK(hipEventRecord(g_drop.ev[1][k], g_drop.d2h));
        off += (size_t)(q.b - 6 - q.a) * orow;
    }
    for (int k = 0; k < P; ++k) {
        const Slab q = dropin_slab(k, P, nx);
        HIPCHK(hipEventSynchronize(g_drop.ev[1][k]));
        par_rows(q.i1 - q.i0, (q.i1 - q.i0) * orow, [=](int64_t r0, int64_t r1) {
            std::memcpy(out + (q.i0 + r0) * orow, pin_out + (q.i0 + r0) * orow, sizeof(double) * (size_t)((r1 - r0) * orow));
        });
    }
    HIPCHK(hipStreamSynchronize(g_drop.d2h));
    return finish();
}

// ==========================================================================
// l_local on device arrays with no allocation
// ==========================================================================
// The face arrays of the reference pass structure live in one file-scope
// device array of fixed size instead of the drop-in arena, so that the call
// allocates nothing (for callers under hipGraph capture or in loops that
// must not allocate).  Every device holds its own copy for the life of the
// process; host code reaches it only through the address hipGetSymbolAddress
// returns for the current device.
namespace {

constexpr size_t NOALLOC_SCRATCH_DOUBLES = ((size_t)64 << 20) / sizeof(double);   // 64 MiB
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// Device address of `scratch` on the current device (looked up once per device).
int scratch_for_current_device(double** out)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(WENOX_E_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        int rc = require_device();
        if (rc) return rc;
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(WENOX_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = (double*)p;
    }
    *out = g_scratch_ptr[dev];
    return WENOX_OK;
}

}  // namespace

extern "C" int wenox_l_local_noalloc(const wenox_view3* u, double dx, double gamma, int force_hlle,
                                     double jp_lo, double jp_hi, int cons, double* out, void* stream)
{
    const char* fn = "l_local_noalloc";
    if (bad_view(u) || !u->data) return fail(WENOX_E_ARG, "l_local_noalloc: null array");
    const int64_t m = u->dim[0], n = u->dim[1];
    if (u->dim[2] < 4) return shape_err(fn, "u must have at least 4 channels, got " + dims(u));
    if (m < 6 || n < 6) return shape_err(fn, "u needs m >= 6 and n >= 6 (3 ghost cells per side), got " + dims(u));
    if (!force_hlle && (m < 8 || n < 8))
        return shape_err(fn, "the HLLC/HLLE blend's sensor needs at least 2 x 2 interior cells, got " + dims(u));
    const int64_t nx = m - 6, ny = n - 6;
    if (nx * ny > 0 && !out) return fail(WENOX_E_ARG, "l_local_noalloc: null output");
    const size_t need = rhs_unfused_scratch(nx, ny);
    if (need > NOALLOC_SCRATCH_DOUBLES)
        return shape_err(fn, "grid " + dims(u) + " needs " + std::to_string(need * 8 >> 20) +
                                 " MiB of scratch; this entry point has " +
                                 std::to_string(NOALLOC_SCRATCH_DOUBLES * 8 >> 20) +
                                 " MiB (use wenox_l_local or the solver for larger grids)");
    double* work = nullptr;
    int rc = scratch_for_current_device(&work);
    if (rc) return rc;
    if (nx * ny == 0) return WENOX_OK;
    const Lay lu{u->stride[0], u->stride[1], u->stride[2]};
    l_local_passes(ops_for(g_variant), u->data, lu, m, n, nx, dx, gamma, force_hlle ? 1 : 2, cons, jp_lo, jp_hi,
                   work, out, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return WENOX_OK;
}

// ==========================================================================
// Device-resident solver
// ==========================================================================
struct wenox_solver {
    wenox_solver_desc d{};
    const Ops* ops = nullptr;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // State buffers: padded SoA4 planes with row pitch ld (U, A, B) or ldT
    // (their transposed copies), each starting STATE_SHIFT elements into its
    // allocation mem[] (see wenox_ops.h).
    double* U = nullptr;     // U^n
    double* A = nullptr;     // stage 1 output
    double* B = nullptr;     // stage 2 output
    double* UT = nullptr;    // transposed copies (fused engine only)
    double* AT = nullptr;
    double* BT = nullptr;
    void* mem[6] = {};
    int64_t ld = 0, ldT = 0;
    double* L = nullptr;     // RHS (unfused, nx*ny SoA4) / Dx (fused, row pitch plan.ldx)
    double* DyT = nullptr;   // fused: y flux differences, transposed interior (row pitch plan.ldyT)
    double* work = nullptr;  // unfused face arrays, RHS output, FP64 copy of a split state
    size_t work_n = 0;
    StepCtl* ctl = nullptr;  // device
    double* cfl_part = nullptr;   // device, CFL_SLOTS per-block maxima
    BcParams bp{};
    hipGraph_t graph = nullptr;      // one step
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph_n = nullptr;    // GRAPH_STEPS steps (fewer graph launches on long runs)
    hipGraphExec_t exec_n = nullptr;
    double last_ms = 0.0;
    SweepPlan plan{};
    const SweepOps* sops = nullptr;
    bool split = false;      // FP32-split state: buffers hold float planes [8][M][N]
    int* flags = nullptr;    // FP32-split: per-workgroup FP64-recompute flags
    double cfl_rho_floor = 1e-14;   // Rie_prob_2d_rs.py:314
    double cfl_cmax_eps = 1e-16;    // Rie_prob_2d_rs.py:315
    double pfloor = 0.0;            // > 0: the twin's pressure floor after every step
    double* sums = nullptr;  // compensated sums: SUM_BLOCKS x 16 block partials, then 8 results
    // fused engine: side stream and events of the slab-pipelined stage (plan.parts)
    hipStream_t side = nullptr;
    hipEvent_t pipe_ev[PIPE_MAX_PARTS + 1] = {};
    // blocked layout of the fast FP64 step (wenox_solver_set_layout): Ub holds
    // the state, Ab and Bb the stage states, Eb the x sweep's part E of each
    // stage, all blocked states (plan.nb, plan.Pb) in one allocation; the
    // row-major U / U^T are brought up to date by sync_rows after steps ran
    bool blk = false;
    bool rows_stale = false;
    double* blk_mem = nullptr;
    double *Ub = nullptr, *Ab = nullptr, *Bb = nullptr, *Eb = nullptr;
};

namespace {

// plane strides of the pitched state and of its transposed copy
int64_t plane(const wenox_solver* s) { return (s->d.nx + 6) * s->ld; }
int64_t plane_t(const wenox_solver* s) { return (s->d.ny + 6) * s->ldT; }

// Entries of the CFL maxima array: CFL_SLOTS row blocks of k_cfl, or one per
// 32 x 32 tile of the fused engine's last RK combine, or (FP64 fused engine)
// one per workgroup of the blocked step's stage-3 y sweep.
size_t cfl_slots(const wenox_solver_desc& d, const SweepPlan* blk_plan)
{
    const size_t tiles = (size_t)((d.nx + 31) / 32) * (size_t)((d.ny + 31) / 32);
    const size_t blk = blk_plan ? (size_t)blk_part_slots(*blk_plan) : 0;
    return std::max<size_t>(std::max<size_t>(CFL_SLOTS, tiles), blk);
}

// Steps per replay of the multi-step graph: consecutive graph launches leave
// a gap of several microseconds between them, consecutive kernels of one
// graph do not.
constexpr int64_t GRAPH_STEPS = 8;

// Forget the captured step graphs (after a change of what a step launches).
void drop_graphs(wenox_solver* s)
{
    if (s->exec) { (void)hipGraphExecDestroy(s->exec); s->exec = nullptr; }
    if (s->graph) { (void)hipGraphDestroy(s->graph); s->graph = nullptr; }
    if (s->exec_n) { (void)hipGraphExecDestroy(s->exec_n); s->exec_n = nullptr; }
    if (s->graph_n) { (void)hipGraphDestroy(s->graph_n); s->graph_n = nullptr; }
}
// FP64 view of work with the state geometry (the split download's staging copy)
double* work_state(const wenox_solver* s) { return s->work + STATE_SHIFT; }

// After blocked steps: the row-major state U and its transposed copy from the
// blocked one (ghost cells included), for the calls that read U.
void sync_rows(wenox_solver* s)
{
    if (!s->blk || !s->rows_stale) return;
    s->sops->from_blk(s->plan, s->Ub, s->U, s->stream);
    s->sops->transpose(s->U, s->UT, s->d.nx + 6, s->d.ny + 6, s->ld, s->ldT, s->stream);
    s->rows_stale = false;
}

// One time step on the solver's stream: CFL -> 3 x (RHS, RK combine, ghost fill) -> advance
// (fused engine: dt from the previous step's tile maxima, the advance applied
// at the next step's start or by a settle launch; the combine writes the
// ghost cells where ghost_on_store allows it).
void enqueue_step(wenox_solver* s)
{
    const auto& d = s->d;
    const Ops* o = s->ops;
    hipStream_t st = s->stream;
    if (s->blk) {
        // blocked: per stage the x sweep (E) and the y sweep (the new state's
        // interior, and its ghost cells where ghost_in_ysweep allows it;
        // stage 3 also the next CFL maxima), then the ghost fill if needed
        double* src[3] = {s->Ub, s->Ab, s->Bb};
        double* dst[3] = {s->Ab, s->Bb, s->Ub};
        const bool ghost_fill = !ghost_in_ysweep(s->plan);
        o->step_dt(s->ctl, st);
        for (int stage = 1; stage <= 3; ++stage) {
            s->sops->stage_blk(s->plan, src[stage - 1], s->Ub, s->Eb, dst[stage - 1], stage, s->ctl, st);
            if (ghost_fill) s->sops->bc_blk(s->plan, dst[stage - 1], s->ctl, stage, st);
        }
        return;
    }
    if (s->split) {
        float* src[3] = {(float*)s->U, (float*)s->A, (float*)s->B};
        float* dst[3] = {(float*)s->A, (float*)s->B, (float*)s->U};
        float* srcT[3] = {(float*)s->UT, (float*)s->AT, (float*)s->BT};
        float* dstT[3] = {(float*)s->AT, (float*)s->BT, (float*)s->UT};
        // dt from the previous step's last combine (or the upload), as FP64
        o->step_dt(s->ctl, st);
        for (int stage = 1; stage <= 3; ++stage) {
            const int i = stage - 1;
            s->sops->stage_split(s->plan, src[i], srcT[i], (const float*)s->U, (float*)s->L,
                                 (float*)s->DyT, dst[i], dstT[i], s->flags, stage, s->ctl, st);
            if (!ghost_on_store(s->plan))
                s->sops->bc_split(dst[i], dstT[i], d.nx, d.ny, s->ld, s->ldT, s->bp, s->ctl, stage, st);
        }
        // the last combine left the next step's CFL maxima and the advance pending
        return;
    }
    const bool fused = d.engine == WENOX_ENGINE_FUSED;
    // fused: the previous step's last combine (or the upload) left the CFL
    // maxima; unfused: reduce the state here
    if (fused) o->step_dt(s->ctl, st);
    else o->cfl(s->U, d.nx, d.ny, s->ld, d.gamma, s->ctl, st);
    double* src[3] = {s->U, s->A, s->B};
    double* dst[3] = {s->A, s->B, s->U};
    double* srcT[3] = {s->UT, s->AT, s->BT};
    double* dstT[3] = {s->AT, s->BT, s->UT};
    for (int stage = 1; stage <= 3; ++stage) {
        const int i = stage - 1;
        if (d.engine == WENOX_ENGINE_FUSED) {
            s->sops->stage(s->plan, src[i], srcT[i], s->U, s->L, s->DyT, dst[i], dstT[i], stage,
                           s->ctl, st);
        } else {
            o->rhs_unfused(src[i], d.nx, d.ny, s->ld, d.dx, d.gamma, d.cons_switch, d.force_hlle,
                           d.jp_lo, d.jp_hi, s->work, s->L, s->ctl, st);
            o->rk(s->U, src[i], s->L, dst[i], d.nx, d.ny, s->ld, stage, s->ctl, st);
        }
        if (!fused || !ghost_on_store(s->plan))
            o->bc(dst[i], dstT[i], d.nx, d.ny, s->ld, s->ldT, s->bp, s->ctl, stage, st);
    }
    if (s->pfloor > 0.0) {
        // WENO_HLLC_2d.py:903 (main): con_var2primi_pfloor after each step, every padded cell
        o->pfloor(s->U, d.nx + 6, d.ny + 6, s->ld, d.gamma, s->pfloor, s->ctl, 1, st);
        if (s->UT) o->pfloor(s->UT, d.ny + 6, d.nx + 6, s->ldT, d.gamma, s->pfloor, s->ctl, 0, st);
    }
    // fused: the last combine marked the advance pending (the next step_dt or
    // a settle applies it)
    if (!fused) o->advance(s->ctl, st);
}

int read_ctl(wenox_solver* s, StepCtl* h)
{
    HIPCHK(hipMemcpyAsync(h, s->ctl, sizeof(StepCtl), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return WENOX_OK;
}

int write_ctl(wenox_solver* s, const StepCtl* h)
{
    StepCtl c = *h;
    c.part = s->cfl_part;   // per-block maxima of the CFL reduction
    HIPCHK(hipMemcpyAsync(s->ctl, &c, sizeof(StepCtl), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return WENOX_OK;
}

// Host copy of the DMR ghost fill (unitest.py:207-227) on an 8-channel array,
// applied after the primitive sync exactly as rk3_refac does (:127-134).
void dmr_ref_fill8(double* q, int64_t nx, int64_t ny, int bound)
{
    static const double post[8] = {8, 33, -57.1576, 563.499, 8, 4.125, -7.1447, 116.5};
    const int64_t n = ny + 6;
    auto at = [&](int64_t i, int64_t j) { return q + (i * n + j) * 8; };
    for (int64_t i = 0; i < nx + 6; ++i)
        for (int64_t j = 0; j < n; ++j) {
            bool ghost = i < 3 || i >= nx + 3 || j < 3 || j >= ny + 3;
            if (!ghost) continue;
            const double* src = nullptr;
            bool flip = false;
            int64_t si = i, sj = j;
            if (i < 3) {
                src = post;
            } else {
                if (i >= nx + 3) si = nx + 2;
                if (j >= ny + 3) sj = ny + 2;
                else if (j < 3) {
                    if (i >= nx + 3) { sj = 5 - j; flip = true; }
                    else if (i <= bound) src = post;
                    else if (i <= bound + 9) sj = 3;
                    else { sj = 5 - j; flip = true; }
                }
            }
            double v[8];
            std::memcpy(v, src ? src : at(si, sj), sizeof v);
            if (flip) { v[2] = -v[2]; v[6] = -v[6]; }
            std::memcpy(at(i, j), v, sizeof v);
        }
}

}  // namespace

extern "C" int wenox_solver_create(const wenox_solver_desc* desc, wenox_solver** out)
{
    if (!desc || !out) return fail(WENOX_E_ARG, "wenox_solver_create: null argument");
    *out = nullptr;
    const auto& d = *desc;
    if (d.nx < 3 || d.ny < 3) return fail(WENOX_E_SHAPE, "solver needs nx >= 3 and ny >= 3");
    if (!d.force_hlle && (d.nx < 2 || d.ny < 2)) return fail(WENOX_E_SHAPE, "blend needs nx, ny >= 2");
    if (d.bc < 0 || d.bc > 4) return fail(WENOX_E_ARG, "unknown boundary kind");
    if (d.bc == WENOX_BC_DMR_STD && !(std::fabs((double)d.ny * d.dx - 1.0) <= 1e-9))
        return fail(WENOX_E_ARG, "standard double Mach reflection: the domain height ny * dx must be 1 "
                                 "(the shock trace on the top boundary is placed for [0, 4] x [0, 1])");
    if (d.bc == WENOX_BC_DMR_REF && d.nx + 2 < d.dmr_bound + 10)
        return fail(WENOX_E_SHAPE, "DMR reference ghost fill needs nx + 2 >= bound + 10");
    if (d.variant != WENOX_EXACT && d.variant != WENOX_FAST && d.variant != WENOX_TWIN)
        return fail(WENOX_E_ARG, "unknown variant");
    if (d.variant == WENOX_TWIN && (d.engine != WENOX_ENGINE_UNFUSED || d.precision != WENOX_PREC_FP64))
        return fail(WENOX_E_ARG, "the twin variant runs on the unfused engine in FP64");
    if (d.variant == WENOX_TWIN && !d.cons_switch)
        return fail(WENOX_E_ARG, "the twin variant reconstructs component-wise (cons_switch = 1)");
    if (d.engine != WENOX_ENGINE_UNFUSED && d.engine != WENOX_ENGINE_FUSED)
        return fail(WENOX_E_ARG, "unknown engine");
    if (d.precision != WENOX_PREC_FP64 && d.precision != WENOX_PREC_FP32SPLIT)
        return fail(WENOX_E_ARG, "unknown precision");
    if (d.precision == WENOX_PREC_FP32SPLIT && d.engine != WENOX_ENGINE_FUSED)
        return fail(WENOX_E_ARG, "the FP32-split precision runs on the fused engine only");
    if (!(d.dx > 0.0) || !(d.cfl > 0.0)) return fail(WENOX_E_ARG, "dx and cfl must be positive");
    int rc = require_device();
    if (rc) return rc;
    auto* s = new wenox_solver();
    s->d = d;
    s->ops = ops_for(d.variant);
    (void)hipGetDevice(&s->device);
    auto cleanup = [&](int code) {
        wenox_solver_destroy(s);
        return code;
    };
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&s->ev0) != hipSuccess || hipEventCreate(&s->ev1) != hipSuccess)
        return cleanup(fail(WENOX_E_DEVICE, "stream/event creation failed"));
    const size_t I = (size_t)(d.nx * d.ny) * 4;
    const bool fused = d.engine == WENOX_ENGINE_FUSED;
    s->ld = row_pitch(d.ny + 6);
    s->ldT = row_pitch(d.nx + 6);
    // doubles per state buffer: four pitched planes plus the shift and its tail
    const size_t SP = (size_t)plane(s) * 4 + ROW_ALIGN, SPT = (size_t)plane_t(s) * 4 + ROW_ALIGN;
    size_t nL = I, nDyT = 4;
    s->work_n = std::max(rhs_unfused_scratch(d.nx, d.ny), I);
    if (fused) {
        s->plan = make_sweep_plan(d.nx, d.ny, d.dx, d.gamma, d.cons_switch, d.force_hlle, d.jp_lo, d.jp_hi);
        s->sops = d.variant == WENOX_FAST ? &sweep_ops_fast : &sweep_ops_exact;
        nL = std::max(I, (size_t)(s->plan.PDx() * 4));
        nDyT = (size_t)(s->plan.PDyT() * 4);
        s->work_n = I;
        if (d.precision == WENOX_PREC_FP32SPLIT) {
            // the FP32 kernels are built in the fast variant only
            s->split = true;
            s->sops = &sweep_ops_fast;
            s->work_n = std::max(I, SP);
        }
    }
    // the shift is STATE_SHIFT elements: doubles, or the floats of a split state
    const size_t shift_bytes = (size_t)STATE_SHIFT * (s->split ? sizeof(float) : sizeof(double));
    auto state_buf = [&](int slot, double** base, size_t n_doubles) {
        if (hipMalloc(&s->mem[slot], n_doubles * 8) != hipSuccess) return false;
        *base = (double*)((char*)s->mem[slot] + shift_bytes);
        return true;
    };
    bool ok = state_buf(0, &s->U, SP) && state_buf(1, &s->A, SP) && state_buf(2, &s->B, SP) &&
              hipMalloc(&s->L, std::max<size_t>(nL, 4) * 8) == hipSuccess &&
              hipMalloc(&s->work, std::max<size_t>(s->work_n, 4) * 8) == hipSuccess &&
              hipMalloc(&s->ctl, sizeof(StepCtl)) == hipSuccess &&
              hipMalloc(&s->cfl_part, cfl_slots(d, fused && !s->split ? &s->plan : nullptr) * sizeof(double)) ==
                  hipSuccess &&
              hipMalloc(&s->sums, ((size_t)SUM_BLOCKS * 16 + 8) * sizeof(double)) == hipSuccess;
    if (ok && fused)
        ok = state_buf(3, &s->UT, SPT) && state_buf(4, &s->AT, SPT) && state_buf(5, &s->BT, SPT) &&
             hipMalloc(&s->DyT, std::max<size_t>(nDyT, 4) * 8) == hipSuccess;
    if (ok && s->split)
        ok = hipMalloc(&s->flags, (size_t)std::max<int64_t>(sweep_flag_count(s->plan), 1) * sizeof(int)) == hipSuccess;
    if (!ok)
        return cleanup(fail(WENOX_E_DEVICE, "device allocation failed (state of " +
                                                std::to_string(((fused ? 3 * (SP + SPT) : 3 * SP) + nL + nDyT + s->work_n) * 8 >> 20) +
                                                " MiB)"));
    s->bp.kind = d.bc;
    s->bp.dmr_bound = d.dmr_bound;
    s->bp.dx = d.dx;
    s->bp.dy = d.dx;
    s->bp.x0 = d.dmr_x0;
    s->bp.gamma = d.gamma;
    if (fused) {
        // the step lets the RK combine write the ghost cells where the rule
        // allows it (ghost_on_store)
        s->plan.bp = s->bp;
        s->plan.ghost = 1;
        // default stream priority (as the solver's own stream)
        bool ev_ok = hipStreamCreateWithFlags(&s->side, hipStreamNonBlocking) == hipSuccess;
        for (auto& e : s->pipe_ev) ev_ok = ev_ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
        if (!ev_ok) return cleanup(fail(WENOX_E_DEVICE, "side stream/event creation failed"));
        s->plan.side = s->side;
        s->plan.ev = s->pipe_ev;
    }
    StepCtl h{};
    h.cfl_dx = d.cfl * d.dx;
    h.rho_floor = s->cfl_rho_floor;
    h.cmax_eps = s->cfl_cmax_eps;
    h.t_final = -1.0;
    h.active = 1;
    if ((rc = write_ctl(s, &h))) return cleanup(rc);
    // the fast FP64 step runs on the blocked state by default
    if (fused && d.variant == WENOX_FAST && !s->split && (rc = wenox_solver_set_layout(s, 1))) return cleanup(rc);
    *out = s;
    return WENOX_OK;
}

extern "C" void wenox_solver_destroy(wenox_solver* s)
{
    if (!s) return;
    drop_graphs(s);
    for (void* p : s->mem)
        if (p) (void)hipFree(p);
    for (double* p : {s->L, s->DyT, s->work})
        if (p) (void)hipFree(p);
    if (s->ctl) (void)hipFree(s->ctl);
    if (s->cfl_part) (void)hipFree(s->cfl_part);
    if (s->sums) (void)hipFree(s->sums);
    if (s->flags) (void)hipFree(s->flags);
    if (s->blk_mem) (void)hipFree(s->blk_mem);
    for (hipEvent_t e : s->pipe_ev)
        if (e) (void)hipEventDestroy(e);
    if (s->side) (void)hipStreamDestroy(s->side);
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

extern "C" int wenox_solver_upload(wenox_solver* s, const wenox_view3* q, double t0)
{
    if (!s || bad_view(q)) return fail(WENOX_E_ARG, "wenox_solver_upload: null argument");
    const auto& d = s->d;
    const int64_t m = d.nx + 6, n = d.ny + 6;
    if (q->dim[0] != m || q->dim[1] != n || q->dim[2] < 4)
        return fail(WENOX_E_SHAPE, "state must be (nx+6, ny+6, c>=4), got " + dims(q));
    HIPCHK(hipSetDevice(s->device));
    // pack to the pitched SoA4 planes in pinned memory
    const int64_t ld = s->ld;
    const size_t P = (size_t)plane(s);
    double* pin = nullptr;
    HIPCHK(hipHostMalloc(&pin, P * 4 * sizeof(double)));
    std::memset(pin, 0, P * 4 * sizeof(double));
    const int64_t s0 = q->stride[0], s1 = q->stride[1], s2 = q->stride[2];
    for (int64_t i = 0; i < m; ++i)
        for (int64_t j = 0; j < n; ++j) {
            const double* src = q->data + i * s0 + j * s1;
            for (int k = 0; k < 4; ++k) pin[k * P + i * ld + j] = src[k * s2];
        }
    StepCtl h{};
    h.cfl_dx = d.cfl * d.dx;
    h.rho_floor = s->cfl_rho_floor;
    h.cmax_eps = s->cfl_cmax_eps;
    h.t_final = -1.0;
    h.T = t0;
    h.active = 1;
    if (q->dim[2] >= 8) {
        // first-step wave speed from the caller's primitives (Rie_prob_2d_rs.py:312-315)
        double cm = 0.0;
        bool nan = false;
        for (int64_t i = 3; i < 3 + d.nx && !nan; ++i)
            for (int64_t j = 3; j < 3 + d.ny; ++j) {
                const double* c = q->data + i * s0 + j * s1;
                double u = c[5 * s2], v = c[6 * s2], p = c[7 * s2], r = c[0];
                double w = std::sqrt(d.gamma * p / std::max(r, s->cfl_rho_floor)) + std::sqrt(u * u + v * v);
                if (w != w) { cm = w; nan = true; break; }
                if (w > cm) cm = w;
            }
        h.cmax0 = cm;
        h.use_cmax0 = 1;
    }
    hipError_t e = hipMemcpy(s->split ? work_state(s) : s->U, pin, P * 4 * sizeof(double), hipMemcpyHostToDevice);
    (void)hipHostFree(pin);
    if (e != hipSuccess) return fail(WENOX_E_DEVICE, std::string("upload: ") + hipGetErrorString(e));
    int rc = write_ctl(s, &h);
    if (rc) return rc;
    if (s->split) {
        s->sops->to_split(work_state(s), (float*)s->U, (int64_t)P, s->stream);
        s->sops->transpose_split((const float*)s->U, (float*)s->UT, m, n, ld, s->ldT, s->stream);
        s->sops->bc_split((float*)s->U, (float*)s->UT, d.nx, d.ny, ld, s->ldT, s->bp, s->ctl, 0, s->stream);
        // the first step reads the CFL maxima of this state
        s->sops->cfl_split((const float*)s->U, d.nx, d.ny, ld, d.gamma, s->ctl, s->stream, 0);
    } else {
        if (s->UT) s->sops->transpose(s->U, s->UT, m, n, ld, s->ldT, s->stream);
        s->ops->bc(s->U, s->UT, d.nx, d.ny, ld, s->ldT, s->bp, s->ctl, 0, s->stream);
        // the fused engine's first step reads the CFL maxima of this state
        if (d.engine == WENOX_ENGINE_FUSED) s->ops->cfl_part(s->U, d.nx, d.ny, ld, d.gamma, s->ctl, s->stream);
        if (s->blk) {
            s->sops->to_blk(s->plan, s->U, s->Ub, s->stream);
            s->rows_stale = false;
        }
    }
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipGetLastError());
    return WENOX_OK;
}

extern "C" int wenox_solver_conserved_sums(wenox_solver* s, double* out, int n_out)
{
    if (!s || !out) return fail(WENOX_E_ARG, "wenox_solver_conserved_sums: null argument");
    if (n_out != 4 && n_out != 8) return fail(WENOX_E_ARG, "wenox_solver_conserved_sums: n_out must be 4 or 8");
    HIPCHK(hipSetDevice(s->device));
    sync_rows(s);
    double* res = s->sums + (size_t)SUM_BLOCKS * 16;
    conserved_sums(s->U, s->split, s->d.nx, s->d.ny, s->ld, plane(s), s->sums, res, s->stream);
    double h[8];
    HIPCHK(hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipGetLastError());
    for (int k = 0; k < n_out; ++k) out[k] = h[k];
    return WENOX_OK;
}

extern "C" int wenox_solver_download(wenox_solver* s, double* q8)
{
    if (!s || !q8) return fail(WENOX_E_ARG, "wenox_solver_download: null argument");
    const auto& d = s->d;
    const int64_t m = d.nx + 6, n = d.ny + 6;
    const size_t P = (size_t)plane(s);
    HIPCHK(hipSetDevice(s->device));
    sync_rows(s);
    const double* state = s->U;
    if (s->split) {
        // FP64 value hi + lo of every cell
        s->sops->from_split((const float*)s->U, work_state(s), (int64_t)P, s->stream);
        state = work_state(s);
    }
    // the (m, n, 8) array is formed on the device in row blocks staged in the
    // stage-1 buffer A (4 P doubles of space, free between steps) and copied
    // straight into the caller's array
    double* stage = (double*)s->A;
    const int64_t rows_per = std::max<int64_t>(1, (int64_t)(4 * P) / (8 * n));
    for (int64_t r0 = 0; r0 < m; r0 += rows_per) {
        const int64_t rows = std::min(rows_per, m - r0);
        s->ops->snapshot8(state, m, n, s->ld, d.gamma, r0, rows, stage, s->stream);
        HIPCHK(hipMemcpyAsync(q8 + r0 * n * 8, stage, (size_t)(rows * n) * 8 * 8, hipMemcpyDeviceToHost,
                              s->stream));
    }
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipGetLastError());
    if (d.bc == WENOX_BC_DMR_REF) dmr_ref_fill8(q8, d.nx, d.ny, d.dmr_bound);
    return WENOX_OK;
}

extern "C" int wenox_solver_rhs(wenox_solver* s, double* out)
{
    if (!s || !out) return fail(WENOX_E_ARG, "wenox_solver_rhs: null argument");
    const auto& d = s->d;
    HIPCHK(hipSetDevice(s->device));
    sync_rows(s);
    StepCtl h;
    int rc = read_ctl(s, &h);
    if (rc) return rc;
    StepCtl act = h;
    act.active = 1;
    act.dt = 0.0;
    if ((rc = write_ctl(s, &act))) return rc;
    if (s->split) {
        s->sops->rhs_split(s->plan, (const float*)s->U, (const float*)s->UT, (float*)s->L,
                           (float*)s->DyT, s->flags, s->work, s->ctl, s->stream);
        HIPCHK(hipMemcpyAsync(s->L, s->work, (size_t)(d.nx * d.ny) * 4 * 8, hipMemcpyDeviceToDevice, s->stream));
    } else if (d.engine == WENOX_ENGINE_FUSED) {
        s->sops->rhs(s->plan, s->U, s->UT, s->L, s->DyT, s->work, s->ctl, s->stream);
        HIPCHK(hipMemcpyAsync(s->L, s->work, (size_t)(d.nx * d.ny) * 4 * 8, hipMemcpyDeviceToDevice, s->stream));
    } else {
        s->ops->rhs_unfused(s->U, d.nx, d.ny, s->ld, d.dx, d.gamma, d.cons_switch, d.force_hlle, d.jp_lo,
                            d.jp_hi, s->work, s->L, s->ctl, s->stream);
    }
    const size_t I = (size_t)(d.nx * d.ny);
    std::vector<double> soa(I * 4);
    HIPCHK(hipMemcpyAsync(soa.data(), s->L, I * 4 * 8, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipGetLastError());
    for (size_t c = 0; c < I; ++c)
        for (int k = 0; k < 4; ++k) out[c * 4 + k] = soa[k * I + c];
    return write_ctl(s, &h);
}

extern "C" int wenox_solver_run(wenox_solver* s, int64_t max_steps, double t_final, double* t_out,
                                int64_t* steps_out)
{
    if (!s) return fail(WENOX_E_ARG, "wenox_solver_run: null solver");
    if (max_steps < 0) return fail(WENOX_E_ARG, "max_steps must be >= 0");
    HIPCHK(hipSetDevice(s->device));
    StepCtl h;
    int rc = read_ctl(s, &h);
    if (rc) return rc;
    const long long step0 = h.step;
    h.max_steps = step0 + max_steps;
    h.t_final = t_final;
    h.active = 1;
    if ((rc = write_ctl(s, &h))) return rc;
    // the steps write the blocked state only
    if (s->blk) s->rows_stale = true;
    if (s->d.use_graph && !s->exec) {
        HIPCHK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        enqueue_step(s);
        HIPCHK(hipStreamEndCapture(s->stream, &s->graph));
        HIPCHK(hipGraphInstantiate(&s->exec, s->graph, nullptr, nullptr, 0));
    }
    // both graphs are captured on the first run() call, whatever its length,
    // so that no later call pays for a capture
    if (s->d.use_graph && !s->exec_n) {
        HIPCHK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        for (int64_t i = 0; i < GRAPH_STEPS; ++i) enqueue_step(s);
        HIPCHK(hipStreamEndCapture(s->stream, &s->graph_n));
        HIPCHK(hipGraphInstantiate(&s->exec_n, s->graph_n, nullptr, nullptr, 0));
    }
    auto launch = [&](int64_t k) -> int {
        if (s->exec_n)
            for (; k >= GRAPH_STEPS; k -= GRAPH_STEPS) HIPCHK(hipGraphLaunch(s->exec_n, s->stream));
        for (int64_t i = 0; i < k; ++i) {
            if (s->exec) {
                HIPCHK(hipGraphLaunch(s->exec, s->stream));
            } else {
                enqueue_step(s);
            }
        }
        return WENOX_OK;
    };
    HIPCHK(hipEventRecord(s->ev0, s->stream));
    int64_t launched = 0;
    if (t_final < 0.0) {
        if ((rc = launch(max_steps))) return rc;
        launched = max_steps;
    } else {
        // Time-controlled: launch chunks sized by the remaining time / dt.
        while (launched < max_steps) {
            s->ops->settle(s->ctl, s->stream);
            if ((rc = read_ctl(s, &h))) return rc;
            if (h.step >= h.max_steps) break;
            if (launched > 0 && !h.active) break;
            if (launched > 0 && !(h.T < t_final - 1e-12)) break;
            int64_t chunk = 1;
            if (h.dt > 0.0 && launched > 0) {
                double est = (t_final - h.T) / h.dt;
                chunk = (int64_t)std::min(64.0, std::max(1.0, std::floor(est)));
            }
            chunk = std::min(chunk, max_steps - launched);
            if ((rc = launch(chunk))) return rc;
            launched += chunk;
        }
    }
    HIPCHK(hipEventRecord(s->ev1, s->stream));
    s->ops->settle(s->ctl, s->stream);   // the last step's pending advance
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipGetLastError());
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    if ((rc = read_ctl(s, &h))) return rc;
    const int64_t taken = h.step - step0;
    s->last_ms = taken > 0 ? ms / (double)taken : 0.0;
    if (t_out) *t_out = h.T;
    if (steps_out) *steps_out = taken;
    return WENOX_OK;
}

extern "C" double wenox_solver_last_ms_per_step(wenox_solver* s) { return s ? s->last_ms : 0.0; }

extern "C" int wenox_solver_profile(wenox_solver* s, int reps, double* out_ms, int n_out)
{
    if (!s || !out_ms || n_out < 6 || reps < 1) return fail(WENOX_E_ARG, "wenox_solver_profile: bad argument");
    const auto& d = s->d;
    if (d.engine != WENOX_ENGINE_FUSED) return fail(WENOX_E_ARG, "per-kernel profile needs the fused engine");
    HIPCHK(hipSetDevice(s->device));
    StepCtl h;
    int rc = read_ctl(s, &h);
    if (rc) return rc;
    StepCtl act = h;
    act.active = 1;
    act.dt = 0.0;
    if ((rc = write_ctl(s, &act))) return rc;
    const int split = s->split ? 1 : 0;
    hipStream_t st = s->stream;
    auto timed = [&](int slot, auto&& launch) -> int {
        launch();   // warm
        HIPCHK(hipEventRecord(s->ev0, st));
        for (int r = 0; r < reps; ++r) launch();
        HIPCHK(hipEventRecord(s->ev1, st));
        HIPCHK(hipEventSynchronize(s->ev1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
        out_ms[slot] = ms / reps;
        return WENOX_OK;
    };
    // scratch targets: the sweeps read the stage-1 input (U, U^T) and write Dx / Dy^T;
    // the combine writes into B / B^T, which the next step overwrites anyway.
    // slot 0: the step's CFL launch, the one-block dt reduction (the maxima
    // come from the previous step's last combine)
    if ((rc = timed(0, [&] {
             s->ops->step_dt(s->ctl, st);
         })))
        return rc;
    if ((rc = write_ctl(s, &act))) return rc;
    if (s->blk) {
        // blocked step from Ub into the scratch Eb / Bb: the x sweep (E), the
        // y sweep (new state), no combine, the ghost fill.  The sweeps are
        // averaged over the three stages, as a step runs them (the x sweep
        // of stage 1 reads no U^n; the y sweep of stage 3 also reduces the
        // next step's wave speeds).
        auto over_stages = [&](int slot, auto&& launch) -> int {
            double sum = 0.0;
            for (int stage = 1; stage <= 3; ++stage) {
                int r = timed(slot, [&] { launch(stage); });
                if (r) return r;
                sum += out_ms[slot];
            }
            out_ms[slot] = sum / 3.0;
            return WENOX_OK;
        };
        auto sweep = [&](int dir, int stage) {
            s->sops->sweep1_blk(s->plan, dir, s->Ub, s->Ub, s->Eb, s->Bb, stage, s->ctl, st);
        };
        if ((rc = over_stages(1, [&](int stage) { sweep(0, stage); }))) return rc;
        if ((rc = over_stages(2, [&](int stage) { sweep(1, stage); }))) return rc;
        out_ms[3] = 0.0;
        // slot 4: the ghost fill, 0 where the y sweep writes the ghost cells
        if (ghost_in_ysweep(s->plan)) out_ms[4] = 0.0;
        else if ((rc = timed(4, [&] { s->sops->bc_blk(s->plan, s->Bb, s->ctl, 2, st); }))) return rc;
        // the stage's two sweeps are the whole stage: one measurement for
        // slots 6 (sweeps) and 7 (stage)
        if (n_out >= 7) {
            if ((rc = over_stages(6, [&](int stage) {
                     s->sops->stage_blk(s->plan, s->Ub, s->Ub, s->Eb, s->Bb, stage, s->ctl, st);
                 })))
                return rc;
            if (n_out >= 8) out_ms[7] = out_ms[6];
        }
        const double st_ms = n_out >= 7 ? out_ms[6] : out_ms[1] + out_ms[2];
        out_ms[5] = 3.0 * (st_ms + out_ms[4]) + out_ms[0];
        return write_ctl(s, &h);
    }
    if ((rc = timed(1, [&] { s->sops->sweep1(s->plan, 0, split, s->U, s->L, s->flags, s->ctl, st); })))
        return rc;
    if ((rc = timed(2, [&] { s->sops->sweep1(s->plan, 1, split, s->UT, s->DyT, s->flags, s->ctl, st); })))
        return rc;
    if ((rc = timed(3, [&] {
             s->sops->combine1(s->plan, split, s->L, s->DyT, s->U, s->U, s->B, s->BT, s->ctl, st);
         })))
        return rc;
    // slot 4: the ghost fill, 0 where the combine writes the ghost cells
    if (ghost_on_store(s->plan)) {
        out_ms[4] = 0.0;
    } else if ((rc = timed(4, [&] {
                    if (split)
                        s->sops->bc_split((float*)s->B, (float*)s->BT, d.nx, d.ny, s->ld, s->ldT, s->bp, s->ctl, 2, st);
                    else s->ops->bc(s->B, s->BT, d.nx, d.ny, s->ld, s->ldT, s->bp, s->ctl, 2, st);
                })))
        return rc;
    double pair = out_ms[1] + out_ms[2];
    if (n_out >= 7) {
        // the two sweeps as a step launches them
        if ((rc = timed(6, [&] {
                 s->sops->sweep2(s->plan, split, s->U, s->UT, s->L, s->DyT, s->flags, s->ctl, st);
             })))
            return rc;
        pair = out_ms[6];
    }
    out_ms[5] = 3.0 * (pair + out_ms[3] + out_ms[4]) + out_ms[0];
    if (n_out >= 8) {
        // slot 7: one whole RK stage as the step launches it (sweeps and
        // combine, the slab pipeline's side stream included), stage 2 into B
        if ((rc = timed(7, [&] {
                 if (split)
                     s->sops->stage_split(s->plan, (const float*)s->U, (const float*)s->UT, (const float*)s->U,
                                          (float*)s->L, (float*)s->DyT, (float*)s->B, (float*)s->BT, s->flags, 2,
                                          s->ctl, st);
                 else
                     s->sops->stage(s->plan, s->U, s->UT, s->U, s->L, s->DyT, s->B, s->BT, 2, s->ctl, st);
             })))
            return rc;
        out_ms[5] = 3.0 * (out_ms[7] + out_ms[4]) + out_ms[0];
    }
    return write_ctl(s, &h);
}

extern "C" int wenox_solver_set_cfl_rule(wenox_solver* s, double rho_floor, double cmax_eps)
{
    if (!s || !(rho_floor >= 0.0) || !(cmax_eps >= 0.0)) return fail(WENOX_E_ARG, "wenox_solver_set_cfl_rule: bad argument");
    s->cfl_rho_floor = rho_floor;
    s->cfl_cmax_eps = cmax_eps;
    StepCtl h;
    int rc = read_ctl(s, &h);
    if (rc) return rc;
    h.rho_floor = rho_floor;
    h.cmax_eps = cmax_eps;
    rc = write_ctl(s, &h);
    if (rc) return rc;
    if (s->d.engine == WENOX_ENGINE_FUSED) {
        sync_rows(s);
        // the stored CFL maxima were formed with the previous floor
        if (s->split)
            s->sops->cfl_split((const float*)s->U, s->d.nx, s->d.ny, s->ld, s->d.gamma, s->ctl, s->stream, 0);
        else s->ops->cfl_part(s->U, s->d.nx, s->d.ny, s->ld, s->d.gamma, s->ctl, s->stream);
        HIPCHK(hipStreamSynchronize(s->stream));
    }
    return WENOX_OK;
}

extern "C" int wenox_solver_layout(wenox_solver* s, int* blocked)
{
    if (!s || !blocked) return fail(WENOX_E_ARG, "wenox_solver_layout: null argument");
    *blocked = s->blk ? 1 : 0;
    return WENOX_OK;
}

extern "C" int wenox_solver_set_layout(wenox_solver* s, int blocked)
{
    if (!s || (blocked != 0 && blocked != 1)) return fail(WENOX_E_ARG, "wenox_solver_set_layout: blocked is 0 or 1");
    if ((blocked != 0) == s->blk) return WENOX_OK;
    HIPCHK(hipSetDevice(s->device));
    if (blocked) {
        if (s->d.engine != WENOX_ENGINE_FUSED || s->split || !s->sops || !s->sops->stage_blk)
            return fail(WENOX_E_ARG, "the blocked layout serves the fused engine's fast FP64 step");
        if (s->pfloor > 0.0) return fail(WENOX_E_ARG, "the pressure floor runs on the row-major layout");
        if (!s->blk_mem) {
            const size_t n = (size_t)s->plan.Pb;
            if (hipMalloc(&s->blk_mem, n * 4 * sizeof(double)) != hipSuccess)
                return fail(WENOX_E_DEVICE, "device allocation failed (blocked state of " +
                                                std::to_string(n * 4 * 8 >> 20) + " MiB)");
            s->Ub = s->blk_mem;
            s->Ab = s->Ub + n;
            s->Bb = s->Ab + n;
            s->Eb = s->Bb + n;
        }
        // the row-major U holds the state (its ghost cells included)
        s->sops->to_blk(s->plan, s->U, s->Ub, s->stream);
        s->blk = true;
        s->rows_stale = false;
    } else {
        sync_rows(s);
        s->blk = false;
    }
    // the captured step graphs launch the other layout's kernels
    drop_graphs(s);
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipGetLastError());
    return WENOX_OK;
}

extern "C" int wenox_solver_set_pfloor(wenox_solver* s, double pmin)
{
    if (!s || !(pmin >= 0.0)) return fail(WENOX_E_ARG, "wenox_solver_set_pfloor: pmin >= 0");
    if (pmin > 0.0 && s->split) return fail(WENOX_E_ARG, "the pressure floor applies to FP64 states");
    if (pmin > 0.0 && s->blk) return fail(WENOX_E_ARG, "the pressure floor runs on the row-major layout");
    s->pfloor = pmin;
    // the captured step graph lacks (or holds) the floor launch
    drop_graphs(s);
    return WENOX_OK;
}

extern "C" int wenox_solver_plan(wenox_solver* s, int* chunk_x, int* chunk_y)
{
    if (!s) return fail(WENOX_E_ARG, "wenox_solver_plan: null solver");
    if (chunk_x) *chunk_x = s->plan.chunk_x;
    if (chunk_y) *chunk_y = s->plan.chunk_y;
    return WENOX_OK;
}

extern "C" int wenox_solver_tiling(wenox_solver* s, int* chunk_x, int* chunk_y, int* waves_x, int* waves_y)
{
    if (!s) return fail(WENOX_E_ARG, "wenox_solver_tiling: null solver");
    if (chunk_x) *chunk_x = s->plan.chunk_x;
    if (chunk_y) *chunk_y = s->plan.chunk_y;
    if (waves_x) *waves_x = s->plan.waves_x;
    if (waves_y) *waves_y = s->plan.waves_y;
    return WENOX_OK;
}

extern "C" int wenox_solver_set_tiling(wenox_solver* s, int chunk_x, int chunk_y, int waves_x, int waves_y)
{
    if (!s) return fail(WENOX_E_ARG, "wenox_solver_set_tiling: null solver");
    auto ok_w = [](int w) { return w == 1 || w == 2 || w == 4; };
    if (chunk_x < 1 || chunk_y < 1 || !ok_w(waves_x) || !ok_w(waves_y) || (waves_x > 1 && chunk_x < 3) ||
        (waves_y > 1 && chunk_y < 3))
        return fail(WENOX_E_ARG, "wenox_solver_set_tiling: chunk >= 1 (>= 3 with several waves), waves 1, 2 or 4");
    if (s->d.engine != WENOX_ENGINE_FUSED) return fail(WENOX_E_ARG, "tiling applies to the fused engine");
    SweepPlan p = s->plan;
    p.chunk_x = chunk_x;
    p.chunk_y = chunk_y;
    p.waves_x = waves_x;
    p.waves_y = waves_y;
    if (s->split && sweep_flag_count(p) > sweep_flag_count(s->plan)) {
        int* f = nullptr;
        HIPCHK(hipMalloc(&f, (size_t)sweep_flag_count(p) * sizeof(int)));
        if (s->flags) (void)hipFree(s->flags);
        s->flags = f;
    }
    s->plan = p;
    // the captured step graph holds the old launch shapes
    drop_graphs(s);
    return WENOX_OK;
}

extern "C" int wenox_solver_combine_cols(wenox_solver* s, int* cols)
{
    if (!s || !cols) return fail(WENOX_E_ARG, "wenox_solver_combine_cols: null argument");
    *cols = s->plan.combine_cols;
    return WENOX_OK;
}

extern "C" int wenox_solver_set_combine_cols(wenox_solver* s, int cols)
{
    if (!s || (cols != 32 && cols != 64)) return fail(WENOX_E_ARG, "wenox_solver_set_combine_cols: cols is 32 or 64");
    if (s->d.engine != WENOX_ENGINE_FUSED) return fail(WENOX_E_ARG, "the combine tile applies to the fused engine");
    s->plan.combine_cols = cols;
    drop_graphs(s);
    return WENOX_OK;
}

extern "C" int wenox_solver_pipeline(wenox_solver* s, int* parts)
{
    if (!s || !parts) return fail(WENOX_E_ARG, "wenox_solver_pipeline: null argument");
    *parts = s->d.engine == WENOX_ENGINE_FUSED ? s->plan.parts : 1;
    return WENOX_OK;
}

extern "C" int wenox_solver_set_pipeline(wenox_solver* s, int parts)
{
    if (!s || parts < 1 || parts > PIPE_MAX_PARTS)
        return fail(WENOX_E_ARG, "wenox_solver_set_pipeline: parts is 1 .. " + std::to_string(PIPE_MAX_PARTS));
    if (s->d.engine != WENOX_ENGINE_FUSED) return fail(WENOX_E_ARG, "the pipeline applies to the fused engine");
    s->plan.parts = parts;
    drop_graphs(s);
    return WENOX_OK;
}

extern "C" int wenox_solver_split_jump(wenox_solver* s, double* tau)
{
    if (!s || !tau) return fail(WENOX_E_ARG, "wenox_solver_split_jump: null argument");
    *tau = s->plan.split_jump;
    return WENOX_OK;
}

extern "C" int wenox_solver_set_split_jump(wenox_solver* s, double tau)
{
    if (!s || !(tau > 0.0)) return fail(WENOX_E_ARG, "wenox_solver_set_split_jump: tau must be > 0 (inf: no jump test)");
    s->plan.split_jump = (float)tau;
    drop_graphs(s);   // the captured sweeps hold the old threshold
    return WENOX_OK;
}

extern "C" int wenox_solver_set_chunks(wenox_solver* s, int chunk_x, int chunk_y)
{
    if (!s) return fail(WENOX_E_ARG, "wenox_solver_set_chunks: null solver");
    return wenox_solver_set_tiling(s, chunk_x, chunk_y, s->plan.waves_x, s->plan.waves_y);
}

// ---------------------------------------------------------------------------
// extern "C": jp_cri tuning objective (metrics.py:1-77)
// ---------------------------------------------------------------------------
namespace {

const double kDefaultWeights[4] = {1.0, 1.0, 50.0, 1.0};   // metrics.py:69-70

int check_view2(const char* fn, const wenox_view2* v, int64_t nx, int64_t ny)
{
    if (!v || !v->data) return fail(WENOX_E_ARG, std::string(fn) + ": null array");
    if (v->dim[0] != nx || v->dim[1] != ny) return fail(WENOX_E_ARG, std::string(fn) + ": arrays differ in shape");
    return WENOX_OK;
}

void pack2(const wenox_view2* v, std::vector<double>& out)
{
    const int64_t nx = v->dim[0], ny = v->dim[1];
    out.resize((size_t)(nx * ny));
    for (int64_t i = 0; i < nx; ++i)
        for (int64_t j = 0; j < ny; ++j) out[(size_t)(i * ny + j)] = v->data[i * v->stride[0] + j * v->stride[1]];
}

int objective_host(const char* fn, const wenox_view2* rho, const wenox_view2* u, const wenox_view2* v, double dx,
                   double dy, const double* weights, double* J, double* metrics, uint8_t* S, uint8_t* C,
                   uint8_t* band, double* median)
{
    if (!rho || !J) return fail(WENOX_E_ARG, std::string(fn) + ": null argument");
    const int64_t nx = rho->dim[0], ny = rho->dim[1];
    if (nx < 1 || ny < 1) return fail(WENOX_E_SHAPE, std::string(fn) + ": empty array");
    int rc;
    if ((rc = check_view2(fn, rho, nx, ny)) || (rc = check_view2(fn, u, nx, ny)) || (rc = check_view2(fn, v, nx, ny)))
        return rc;
    if ((rc = require_device())) return rc;
    std::vector<double> r, uu, vv;
    pack2(rho, r);
    pack2(u, uu);
    pack2(v, vv);
    ObjSource src;
    src.rho = r.data();
    src.u = uu.data();
    src.v = vv.data();
    std::lock_guard<std::mutex> lk(g_drop.mu);
    if (!g_drop.stream) HIPCHK(hipStreamCreateWithFlags(&g_drop.stream, hipStreamNonBlocking));
    return objective_eval(src, nx, ny, dx, dy, weights ? weights : kDefaultWeights, g_drop.stream, J, metrics, S, C,
                          band, median);
}

}  // namespace

extern "C" int wenox_objective_2d(const wenox_view2* rho, const wenox_view2* u, const wenox_view2* v, double dx,
                                  double dy, const double* weights, double* J, double* metrics)
{
    return objective_host("wenox_objective_2d", rho, u, v, dx, dy, weights, J, metrics, nullptr, nullptr, nullptr,
                          nullptr);
}

extern "C" int wenox_objective_fields(const wenox_view2* rho, double dx, double dy, uint8_t* S, uint8_t* C,
                                      uint8_t* band, double* median)
{
    double J = 0.0;
    return objective_host("wenox_objective_fields", rho, rho, rho, dx, dy, nullptr, &J, nullptr, S, C, band, median);
}

extern "C" int wenox_solver_objective(wenox_solver* s, double dx, double dy, const double* weights, double* J,
                                      double* metrics)
{
    if (!s || !J) return fail(WENOX_E_ARG, "wenox_solver_objective: null argument");
    if (s->split) return fail(WENOX_E_ARG, "wenox_solver_objective: FP64 states only");
    HIPCHK(hipSetDevice(s->device));
    sync_rows(s);
    ObjSource src;
    src.U = s->U;
    src.ld = s->ld;
    src.plane = plane(s);
    return objective_eval(src, s->d.nx, s->d.ny, dx, dy, weights ? weights : kDefaultWeights, s->stream, J, metrics,
                          nullptr, nullptr, nullptr, nullptr);
}
