// wenox_objective.hip -- the jp_cri tuning objective on the device:
// metrics.compute_2d_objective (metrics.py:1-77), which opt_jp_cro.py
// (:41-47) evaluates on the final state of every tuning run.
//
// The objective is built from:
//   rho_x, rho_y   central differences with edge replication (:8-9);
//   S              G = |grad rho| > 0.1 (max G + 1e-16) (:10-13);
//   C              S thinned to curves (:15-19): the Zhang-Suen algorithm
//                  (T. Y. Zhang, C. Y. Suen, CACM 27(3), 1984), which
//                  scikit-image's skeletonize implements for 2-D images;
//   band           S dilated by a 5 x 5 square, background outside (:22);
//   I_curve        std of the skeleton's residual about its least-squares line
//                  col = a row + b (:25-34);
//   E_perp         sum over the band of (u, v) . t^2, t the unit tangent of
//                  the density level set (:37-45);
//   O              max over the band of |rho - median7x7(rho)| (scipy's
//                  "reflect" border), over (max rho - min rho + 1e-12) (:48-55);
//   TV             total variation per cell (:58-59);
//   scales         I0 = max(1, sqrt nx), E0 = max(1e-8, median_band(|v|)^2
//                  |band|), TV0 = max(1e-8, mean |rho|) (:62-64).
//
// Integer and selection work (S, band, C, the line-fit sums, medians, maxima)
// is exact; floating sums are reduced per thread, per block (fixed tree) and
// over blocks in block order, so results are deterministic and differ from
// numpy's pairwise sums by rounding only.  Compiled without FMA contraction.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <string>
#include <vector>

#include "../../include/wenox.h"
#include "wenox_objective.h"

namespace wenox {

int abi_fail(int code, const std::string& msg);

namespace {

constexpr int OT = 256;            // threads per block
constexpr int64_t OBJ_BLOCKS = 1024;

__device__ __forceinline__ double nan_max(double a, double b) { return (a > b || a != a) ? a : b; }
__device__ __forceinline__ double nan_min(double a, double b) { return (a < b || a != a) ? a : b; }

// element (i, j) of a row-major nx x ny array
struct Grid {
    int64_t nx, ny;
    __device__ int64_t at(int64_t i, int64_t j) const { return i * ny + j; }
};

// :8-9  (np.pad(..., "edge")[2:] - np.pad(..., "edge")[:-2]) / (2 dx)
__device__ __forceinline__ void grad(const double* R, Grid g, int64_t i, int64_t j, double dx2, double dy2,
                                     double& rx, double& ry)
{
    const int64_t ip = i + 1 < g.nx ? i + 1 : g.nx - 1, im = i > 0 ? i - 1 : 0;
    const int64_t jp = j + 1 < g.ny ? j + 1 : g.ny - 1, jm = j > 0 ? j - 1 : 0;
    rx = (R[g.at(ip, j)] - R[g.at(im, j)]) / dx2;
    ry = (R[g.at(i, jp)] - R[g.at(i, jm)]) / dy2;
}

// block-wide reduction of one value per thread (fixed tree)
template <class Op>
__device__ double block_reduce(double v, double* sh, Op op)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = OT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] = op(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

struct Add { __device__ double operator()(double a, double b) const { return a + b; } };
struct Max { __device__ double operator()(double a, double b) const { return nan_max(a, b); } };
struct Min { __device__ double operator()(double a, double b) const { return nan_min(a, b); } };

// fields of a padded SoA4 state: rho, and u, v over max(rho, 1e-12) (utils.rs:9-35)
__global__ void k_prep(const double* U, int64_t ld, int64_t plane, Grid g, double* R, double* Uv, double* Vv)
{
    for (int64_t t = (int64_t)blockIdx.x * OT + threadIdx.x; t < g.nx * g.ny; t += (int64_t)gridDim.x * OT) {
        const int64_t i = t / g.ny, j = t - i * g.ny, o = (i + 3) * ld + (j + 3);
        const double r = U[o], rs = fmax(r, 1e-12);
        R[t] = r;
        Uv[t] = U[plane + o] / rs;
        Vv[t] = U[2 * plane + o] / rs;
    }
}

// per block: max G, max rho, min rho, sum |rho|, sum |rho[i+1,j] - rho[i,j]|, sum |rho[i,j+1] - rho[i,j]|
__global__ void __launch_bounds__(OT) k_stats(const double* R, Grid g, double dx2, double dy2, double* part)
{
    __shared__ double sh[OT];
    double gm = 0.0, rmax = -INFINITY, rmin = INFINITY, sabs = 0.0, tvx = 0.0, tvy = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * OT + threadIdx.x; t < g.nx * g.ny; t += (int64_t)gridDim.x * OT) {
        const int64_t i = t / g.ny, j = t - i * g.ny;
        double rx, ry;
        grad(R, g, i, j, dx2, dy2, rx, ry);
        const double r = R[t];
        gm = nan_max(gm, sqrt(rx * rx + ry * ry));
        rmax = nan_max(rmax, r);
        rmin = nan_min(rmin, r);
        sabs += fabs(r);
        if (i + 1 < g.nx) tvx += fabs(R[t + g.ny] - r);
        if (j + 1 < g.ny) tvy += fabs(R[t + 1] - r);
    }
    double* p = part + blockIdx.x * 6;
    const double v0 = block_reduce(gm, sh, Max{});
    const double v1 = block_reduce(rmax, sh, Max{});
    const double v2 = block_reduce(rmin, sh, Min{});
    const double v3 = block_reduce(sabs, sh, Add{});
    const double v4 = block_reduce(tvx, sh, Add{});
    const double v5 = block_reduce(tvy, sh, Add{});
    if (threadIdx.x == 0) { p[0] = v0; p[1] = v1; p[2] = v2; p[3] = v3; p[4] = v4; p[5] = v5; }
}

// :13  S = G > alpha Gmax, and the first image of the thinning
__global__ void k_mask(const double* R, Grid g, double dx2, double dy2, double thr, uint8_t* S, uint8_t* C)
{
    for (int64_t t = (int64_t)blockIdx.x * OT + threadIdx.x; t < g.nx * g.ny; t += (int64_t)gridDim.x * OT) {
        const int64_t i = t / g.ny, j = t - i * g.ny;
        double rx, ry;
        grad(R, g, i, j, dx2, dy2, rx, ry);
        const uint8_t s = sqrt(rx * rx + ry * ry) > thr ? 1 : 0;
        S[t] = s;
        C[t] = s;
    }
}

// :22  binary_dilation(S, ones((5, 5))): background outside the grid
__global__ void k_dilate(const uint8_t* S, Grid g, uint8_t* band)
{
    for (int64_t t = (int64_t)blockIdx.x * OT + threadIdx.x; t < g.nx * g.ny; t += (int64_t)gridDim.x * OT) {
        const int64_t i = t / g.ny, j = t - i * g.ny;
        uint8_t b = 0;
        for (int64_t a = i - 2; a <= i + 2; ++a) {
            if (a < 0 || a >= g.nx) continue;
            for (int64_t c = j - 2; c <= j + 2; ++c)
                if (c >= 0 && c < g.ny) b |= S[g.at(a, c)];
        }
        band[t] = b;
    }
}

// One Zhang-Suen subiteration, in -> out (background outside the grid).
// Neighbours P2..P9 clockwise from north; rows (i) run south.
__global__ void k_thin(const uint8_t* in, uint8_t* out, Grid g, int first, int* removed)
{
    for (int64_t t = (int64_t)blockIdx.x * OT + threadIdx.x; t < g.nx * g.ny; t += (int64_t)gridDim.x * OT) {
        if (!in[t]) { out[t] = 0; continue; }
        const int64_t i = t / g.ny, j = t - i * g.ny;
        auto px = [&](int64_t a, int64_t c) -> int {
            return (a >= 0 && a < g.nx && c >= 0 && c < g.ny) ? in[g.at(a, c)] : 0;
        };
        const int p[8] = {px(i - 1, j), px(i - 1, j + 1), px(i, j + 1), px(i + 1, j + 1),
                          px(i + 1, j), px(i + 1, j - 1), px(i, j - 1), px(i - 1, j - 1)};
        int b = 0, a = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            b += p[k];
            a += (p[k] == 0 && p[(k + 1) & 7] == 1) ? 1 : 0;
        }
        const bool c = first ? (p[0] * p[2] * p[4] == 0 && p[2] * p[4] * p[6] == 0)
                             : (p[0] * p[2] * p[6] == 0 && p[0] * p[4] * p[6] == 0);
        const bool kill = b >= 2 && b <= 6 && a == 1 && c;
        out[t] = kill ? 0 : 1;
        if (kill) *removed = 1;
    }
}

// :25-30  line-fit sums over the skeleton: n, sum row, sum col, sum row^2, sum row col
__global__ void __launch_bounds__(OT) k_line(const uint8_t* C, Grid g, long long* part)
{
    __shared__ long long sh[5][OT];
    long long s[5] = {0, 0, 0, 0, 0};
    for (int64_t t = (int64_t)blockIdx.x * OT + threadIdx.x; t < g.nx * g.ny; t += (int64_t)gridDim.x * OT) {
        if (!C[t]) continue;
        const long long i = t / g.ny, j = t - i * g.ny;
        s[0] += 1;
        s[1] += i;
        s[2] += j;
        s[3] += i * i;
        s[4] += i * j;
    }
    for (int k = 0; k < 5; ++k) sh[k][threadIdx.x] = s[k];
    __syncthreads();
    for (int w = OT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 5; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 5; ++k) part[blockIdx.x * 5 + k] = sh[k][0];
}

// :31-32  sum of r = col - (a row + b) (pass 0), of (r - mean)^2 (pass 1)
__global__ void __launch_bounds__(OT) k_resid(const uint8_t* C, Grid g, double a, double b, double mean, int pass,
                                              double* part)
{
    __shared__ double sh[OT];
    double s = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * OT + threadIdx.x; t < g.nx * g.ny; t += (int64_t)gridDim.x * OT) {
        if (!C[t]) continue;
        const int64_t i = t / g.ny, j = t - i * g.ny;
        const double r = (double)j - (a * (double)i + b);
        if (pass == 0) {
            s += r;
        } else {
            const double d = r - mean;
            s += d * d;
        }
    }
    const double v = block_reduce(s, sh, Add{});
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// scipy.ndimage "reflect" (half-sample symmetric) index
__device__ __forceinline__ int64_t reflect(int64_t k, int64_t n)
{
    const int64_t m = 2 * n;
    k %= m;
    if (k < 0) k += m;
    return k < n ? k : m - 1 - k;
}

__device__ __forceinline__ void order2(double& x, double& y)
{
    const bool sw = y < x;
    const double lo = sw ? y : x, hi = sw ? x : y;
    x = lo;
    y = hi;
}

// median_filter(rho, size=7) at (i, j): the 25th smallest of the 7 x 7
// window by forgetful selection -- hold 26 values, drop the smallest and the
// largest (neither can be the median), take in the next value, until three
// remain.
__device__ double median7x7(const double* R, Grid g, int64_t i, int64_t j)
{
    int64_t rows[7], cols[7];
#pragma unroll
    for (int d = 0; d < 7; ++d) {
        rows[d] = reflect(i + d - 3, g.nx) * g.ny;
        cols[d] = reflect(j + d - 3, g.ny);
    }
    double b[26];
#pragma unroll
    for (int k = 0; k < 26; ++k) b[k] = R[rows[k / 7] + cols[k % 7]];
#pragma unroll
    for (int t = 0; t < 23; ++t) {
#pragma unroll
        for (int k = 25; k > t; --k) order2(b[k - 1], b[k]);        // smallest to b[t]
#pragma unroll
        for (int k = t + 1; k < 25; ++k) order2(b[k], b[k + 1]);    // largest to b[25]
        const int e = 26 + t;
        b[25] = R[rows[e / 7] + cols[e % 7]];
    }
    order2(b[23], b[24]);
    order2(b[24], b[25]);
    order2(b[23], b[24]);
    return b[24];
}

// :37-55, :63  over the band: max |rho - median|, sum vperp^2, count, and
// |v| appended to keys (slots from *nkeys, in no particular order: only its
// median is used).  med != nullptr: write the median of every cell to med.
__global__ void __launch_bounds__(OT) k_band(const double* R, const double* Uv, const double* Vv,
                                             const uint8_t* band, Grid g, double dx2, double dy2, double* part,
                                             double* keys, unsigned int* nkeys, double* med)
{
    __shared__ double sh[OT];
    double dmax = 0.0, e = 0.0, cnt = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * OT + threadIdx.x; t < g.nx * g.ny; t += (int64_t)gridDim.x * OT) {
        const int64_t i = t / g.ny, j = t - i * g.ny;
        const bool in = band[t] != 0;
        if (in) keys[atomicAdd(nkeys, 1u)] = fabs(Vv[t]);
        if (!in && !med) continue;
        const double m = median7x7(R, g, i, j);
        if (med) med[t] = m;
        if (!in) continue;
        dmax = nan_max(dmax, fabs(R[t] - m));
        double rx, ry;
        grad(R, g, i, j, dx2, dy2, rx, ry);
        const double mag = sqrt(rx * rx + ry * ry) + 1e-16;
        const double nxn = rx / mag, nyn = ry / mag;
        const double tx = -nyn, ty = nxn;
        const double vp = Uv[t] * tx + Vv[t] * ty;
        e += vp * vp;
        cnt += 1.0;
    }
    const double v0 = block_reduce(dmax, sh, Max{});
    const double v1 = block_reduce(e, sh, Add{});
    const double v2 = block_reduce(cnt, sh, Add{});
    if (threadIdx.x == 0) {
        part[blockIdx.x * 3] = v0;
        part[blockIdx.x * 3 + 1] = v1;
        part[blockIdx.x * 3 + 2] = v2;
    }
}

double h_nan_max(double a, double b) { return (a > b || a != a) ? a : b; }
double h_nan_min(double a, double b) { return (a < b || a != a) ? a : b; }

// device allocations of one evaluation, released on every exit path
struct Buffers {
    std::vector<void*> ptrs;
    ~Buffers()
    {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <class T> hipError_t get(T** p, size_t count)
    {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, count * sizeof(T) > 0 ? count * sizeof(T) : 1);
        if (e == hipSuccess) ptrs.push_back(q);
        *p = (T*)q;
        return e;
    }
};

}  // namespace

#define OCHK(x)                                                                                      \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) return abi_fail(WENOX_E_DEVICE, std::string("objective: ") + #x + ": " + \
                                                                  hipGetErrorString(e_));            \
    } while (0)

int objective_eval(const ObjSource& src, int64_t nx, int64_t ny, double dx, double dy, const double w[4],
                   hipStream_t s, double* J, double* metrics, uint8_t* S_out, uint8_t* C_out, uint8_t* band_out,
                   double* median_out)
{
    const int64_t N = nx * ny;
    const Grid g{nx, ny};
    const unsigned nb = (unsigned)std::max<int64_t>(1, std::min<int64_t>(OBJ_BLOCKS, (N + OT - 1) / OT));
    const double dx2 = 2 * dx, dy2 = 2 * dy;
    Buffers buf;
    double *R, *Uv, *Vv, *keys, *sorted, *dpart, *med = nullptr;
    uint8_t *S, *C, *C2, *band;
    long long* lpart;
    int* removed;
    unsigned int* nkeys;
    OCHK(buf.get(&R, N));
    OCHK(buf.get(&Uv, N));
    OCHK(buf.get(&Vv, N));
    OCHK(buf.get(&keys, N));
    OCHK(buf.get(&sorted, N));
    OCHK(buf.get(&dpart, (size_t)nb * 6));
    OCHK(buf.get(&lpart, (size_t)nb * 5));
    OCHK(buf.get(&S, N));
    OCHK(buf.get(&C, N));
    OCHK(buf.get(&C2, N));
    OCHK(buf.get(&band, N));
    OCHK(buf.get(&removed, 1));
    OCHK(buf.get(&nkeys, 1));
    if (median_out) OCHK(buf.get(&med, N));

    if (src.U) {
        k_prep<<<nb, OT, 0, s>>>(src.U, src.ld, src.plane, g, R, Uv, Vv);
    } else {
        OCHK(hipMemcpyAsync(R, src.rho, N * sizeof(double), hipMemcpyHostToDevice, s));
        OCHK(hipMemcpyAsync(Uv, src.u, N * sizeof(double), hipMemcpyHostToDevice, s));
        OCHK(hipMemcpyAsync(Vv, src.v, N * sizeof(double), hipMemcpyHostToDevice, s));
    }

    // field statistics
    k_stats<<<nb, OT, 0, s>>>(R, g, dx2, dy2, dpart);
    std::vector<double> hp((size_t)nb * 6);
    OCHK(hipMemcpyAsync(hp.data(), dpart, hp.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    OCHK(hipStreamSynchronize(s));
    double gmax = hp[0], rmax = hp[1], rmin = hp[2], sabs = 0.0, tvx = 0.0, tvy = 0.0;
    for (unsigned k = 0; k < nb; ++k) {
        gmax = h_nan_max(gmax, hp[k * 6]);
        rmax = h_nan_max(rmax, hp[k * 6 + 1]);
        rmin = h_nan_min(rmin, hp[k * 6 + 2]);
        sabs += hp[k * 6 + 3];
        tvx += hp[k * 6 + 4];
        tvy += hp[k * 6 + 5];
    }
    const double Gmax = gmax + 1e-16;

    // shock mask, band, skeleton
    k_mask<<<nb, OT, 0, s>>>(R, g, dx2, dy2, 0.10 * Gmax, S, C);
    k_dilate<<<nb, OT, 0, s>>>(S, g, band);
    for (;;) {
        int flag = 0;
        OCHK(hipMemsetAsync(removed, 0, sizeof(int), s));
        k_thin<<<nb, OT, 0, s>>>(C, C2, g, 1, removed);
        k_thin<<<nb, OT, 0, s>>>(C2, C, g, 0, removed);
        OCHK(hipMemcpyAsync(&flag, removed, sizeof(int), hipMemcpyDeviceToHost, s));
        OCHK(hipStreamSynchronize(s));
        if (!flag) break;
    }

    // curve irregularity (:25-34): exact integer normal equations, then the
    // residual's population standard deviation in two passes
    k_line<<<nb, OT, 0, s>>>(C, g, lpart);
    std::vector<long long> lp((size_t)nb * 5);
    OCHK(hipMemcpyAsync(lp.data(), lpart, lp.size() * sizeof(long long), hipMemcpyDeviceToHost, s));
    OCHK(hipStreamSynchronize(s));
    long long ls[5] = {0, 0, 0, 0, 0};
    for (unsigned k = 0; k < nb; ++k)
        for (int q = 0; q < 5; ++q) ls[q] += lp[k * 5 + q];
    double I_curve = 0.0;
    if (ls[0] >= 10) {
        const __int128 n = ls[0], si = ls[1], sj = ls[2], sii = ls[3], sij = ls[4];
        const __int128 den = n * sii - si * si, num = n * sij - si * sj;
        double a = 0.0, b;
        if (den != 0) {
            a = (double)num / (double)den;
            b = ((double)ls[2] - a * (double)ls[1]) / (double)ls[0];
        } else {
            b = (double)ls[2] / (double)ls[0];   // one row: the fit is the mean column
        }
        double mean = 0.0;
        for (int pass = 0; pass < 2; ++pass) {
            k_resid<<<nb, OT, 0, s>>>(C, g, a, b, mean, pass, dpart);
            OCHK(hipMemcpyAsync(hp.data(), dpart, nb * sizeof(double), hipMemcpyDeviceToHost, s));
            OCHK(hipStreamSynchronize(s));
            double sum = 0.0;
            for (unsigned k = 0; k < nb; ++k) sum += hp[k];
            if (pass == 0) mean = sum / (double)ls[0];
            else I_curve = std::sqrt(sum / (double)ls[0]);
        }
    }

    // band terms and the median of |v| over the band
    OCHK(hipMemsetAsync(nkeys, 0, sizeof(unsigned int), s));
    k_band<<<nb, OT, 0, s>>>(R, Uv, Vv, band, g, dx2, dy2, dpart, keys, nkeys, med);
    OCHK(hipMemcpyAsync(hp.data(), dpart, (size_t)nb * 3 * sizeof(double), hipMemcpyDeviceToHost, s));
    OCHK(hipStreamSynchronize(s));
    double dmax = hp[0], E_perp = 0.0, cnt = 0.0;
    for (unsigned k = 0; k < nb; ++k) {
        dmax = h_nan_max(dmax, hp[k * 3]);
        E_perp += hp[k * 3 + 1];
        cnt += hp[k * 3 + 2];
    }
    const int64_t K = (int64_t)cnt;
    double med_v = 0.0;
    if (K > 0) {
        size_t tmp_bytes = 0;
        OCHK(hipcub::DeviceRadixSort::SortKeys(nullptr, tmp_bytes, keys, sorted, (int)K, 0, 64, s));
        void* tmp = nullptr;
        OCHK(buf.get((uint8_t**)&tmp, tmp_bytes));
        OCHK(hipcub::DeviceRadixSort::SortKeys(tmp, tmp_bytes, keys, sorted, (int)K, 0, 64, s));
        double two[2] = {0.0, 0.0};
        const int64_t lo = (K - 1) / 2, take = std::min<int64_t>(2, K - lo);
        OCHK(hipMemcpyAsync(two, sorted + lo, (size_t)take * sizeof(double), hipMemcpyDeviceToHost, s));
        OCHK(hipStreamSynchronize(s));
        med_v = (K % 2) ? two[0] : (two[0] + two[1]) / 2;
    }
    const double O = K > 0 ? dmax / (rmax - rmin + 1e-12) : 0.0;
    const double TV = (tvx + tvy) / (double)N;

    // :62-73
    const double I0 = std::max(1.0, std::sqrt((double)nx));
    const double E0 = K > 0 ? std::max(1e-8, med_v * med_v * (double)K) : 1.0;
    const double TV0 = std::max(1e-8, sabs / (double)N);
    const double cI = I_curve / I0, cE = E_perp / E0, cO = O, cTV = TV / TV0;
    *J = w[0] * cI + w[1] * cE + w[2] * cO + w[3] * cTV;
    const double m[8] = {I_curve, E_perp, O, TV, cI, cE, cO, cTV};
    if (metrics)
        for (int k = 0; k < 8; ++k) metrics[k] = m[k];

    if (S_out) OCHK(hipMemcpyAsync(S_out, S, N, hipMemcpyDeviceToHost, s));
    if (C_out) OCHK(hipMemcpyAsync(C_out, C, N, hipMemcpyDeviceToHost, s));
    if (band_out) OCHK(hipMemcpyAsync(band_out, band, N, hipMemcpyDeviceToHost, s));
    if (median_out) OCHK(hipMemcpyAsync(median_out, med, N * sizeof(double), hipMemcpyDeviceToHost, s));
    OCHK(hipStreamSynchronize(s));
    OCHK(hipGetLastError());
    return WENOX_OK;
}

}  // namespace wenox
