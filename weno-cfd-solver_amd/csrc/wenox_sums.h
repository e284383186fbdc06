// wenox_sums.h -- compensated global sums of the conserved variables.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wenox {

// Blocks of the partial-sum launch; part must hold SUM_BLOCKS * 16 doubles.
constexpr int SUM_BLOCKS = 512;

// out[0..3]: sums of the conserved channels 0-3 over the interior cells of a
// padded SoA state (planes of (nx+6) rows, row pitch ld, plane stride plane
// elements; FP64 planes, or for split the eight float planes hi[4], lo[4]);
// out[4..7]: the sums of their magnitudes.  Asynchronous on s.
void conserved_sums(const void* U, bool split, int64_t nx, int64_t ny, int64_t ld, int64_t plane, double* part,
                    double* out, hipStream_t s);

}  // namespace wenox
