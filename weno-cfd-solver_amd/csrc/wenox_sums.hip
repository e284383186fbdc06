// wenox_sums.hip -- compensated global sums of the conserved variables.
//
// Mass, x- and y-momentum and total energy summed over the interior cells of
// a solver state (the conserved channels 0-3 of the reference's
// (nx+6, ny+6, 8) array, Rie_prob_2d_rs.py:28-32; the reference itself forms
// no such sum).  Every value is added in FP64 with Knuth's TwoSum, whose
// rounding error is carried in a second accumulator (the "Sum2" scheme of
// Ogita, Rump and Oishi, Accurate sum and dot product, SIAM J. Sci. Comput.
// 26, 2005): the result is within about one rounding of the exact sum of the
// stored values plus (n u)^2 times the sum of their magnitudes, for FP64 and
// FP32-split states alike (a split value hi + lo is formed exactly in FP64).
// The order of the additions depends only on the grid size, so the result is
// reproducible.  Built without multiply-add contraction.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "wenox_sums.h"

namespace wenox {
namespace {

constexpr int TB = 256;

// a + b = s + e exactly (Knuth, TAOCP vol. 2, 4.2.2, Theorem B)
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e)
{
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

struct Acc2 {
    double s = 0.0, c = 0.0;
    __device__ __forceinline__ void add(double v)
    {
        double t, e;
        two_sum(s, v, t, e);
        s = t;
        c += e;
    }
    __device__ __forceinline__ void merge(double os, double oc)
    {
        double t, e;
        two_sum(s, os, t, e);
        s = t;
        c += e + oc;
    }
};

// Each block takes rows i = 3 + blockIdx.x + k gridDim.x of the interior, each
// thread the cells j = 3 + threadIdx.x + l TB of them; per block the 64 lanes
// of a wavefront and then the 4 wavefronts are merged in a fixed order.
// out: per block {s, c} of the 4 sums, then {s, c} of the 4 sums of |value|.
template <bool SPLIT>
__global__ void __launch_bounds__(TB) k_sums_part(const void* U, int64_t nx, int64_t ny, int64_t ld, int64_t P,
                                                  double* part)
{
    Acc2 a[4], m[4];
    for (int64_t i = 3 + blockIdx.x; i < nx + 3; i += gridDim.x)
        for (int64_t j = 3 + threadIdx.x; j < ny + 3; j += TB) {
            const int64_t idx = i * ld + j;
            for (int k = 0; k < 4; ++k) {
                double v;
                if (SPLIT) {
                    const float* f = (const float*)U;
                    v = (double)f[k * P + idx] + (double)f[(k + 4) * P + idx];
                } else {
                    v = ((const double*)U)[k * P + idx];
                }
                a[k].add(v);
                m[k].add(fabs(v));
            }
        }
    // butterfly over the wavefront: every lane ends with the same merged value
    for (int off = 32; off >= 1; off >>= 1)
        for (int k = 0; k < 4; ++k) {
            const double as = __shfl_xor(a[k].s, off), ac = __shfl_xor(a[k].c, off);
            const double ms = __shfl_xor(m[k].s, off), mc = __shfl_xor(m[k].c, off);
            // merge in lane order so that both partners compute the same sum
            const bool low = (threadIdx.x & off) == 0;
            Acc2 x, y;
            x.s = low ? a[k].s : as; x.c = low ? a[k].c : ac;
            x.merge(low ? as : a[k].s, low ? ac : a[k].c);
            y.s = low ? m[k].s : ms; y.c = low ? m[k].c : mc;
            y.merge(low ? ms : m[k].s, low ? mc : m[k].c);
            a[k] = x;
            m[k] = y;
        }
    __shared__ double lds[TB / 64][16];
    const int w = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 4; ++k) {
            lds[w][2 * k] = a[k].s;
            lds[w][2 * k + 1] = a[k].c;
            lds[w][8 + 2 * k] = m[k].s;
            lds[w][8 + 2 * k + 1] = m[k].c;
        }
    __syncthreads();
    if (threadIdx.x < 16) {
        Acc2 r;
        for (int ww = 0; ww < TB / 64; ++ww) r.merge(lds[ww][threadIdx.x & ~1], lds[ww][threadIdx.x | 1]);
        if ((threadIdx.x & 1) == 0) {
            part[(int64_t)blockIdx.x * 16 + threadIdx.x] = r.s;
            part[(int64_t)blockIdx.x * 16 + threadIdx.x + 1] = r.c;
        }
    }
}

// One thread per quantity merges the block partials in block order; out[q] =
// s + c rounded once.
__global__ void k_sums_final(const double* part, int blocks, double* out)
{
    const int q = threadIdx.x;   // 0..7: the 4 sums, then the 4 sums of |value|
    if (q >= 8) return;
    Acc2 r;
    for (int b = 0; b < blocks; ++b) r.merge(part[(int64_t)b * 16 + 2 * q], part[(int64_t)b * 16 + 2 * q + 1]);
    out[q] = r.s + r.c;
}

}  // namespace

void conserved_sums(const void* U, bool split, int64_t nx, int64_t ny, int64_t ld, int64_t plane, double* part,
                    double* out, hipStream_t s)
{
    const int blocks = (int)std::min<int64_t>(SUM_BLOCKS, nx);
    if (split)
        k_sums_part<true><<<blocks, TB, 0, s>>>(U, nx, ny, ld, plane, part);
    else
        k_sums_part<false><<<blocks, TB, 0, s>>>(U, nx, ny, ld, plane, part);
    k_sums_final<<<1, 64, 0, s>>>(part, blocks, out);
}

}  // namespace wenox
