// wenox_objective.h -- the jp_cri tuning objective (metrics.py:1-77) on the device.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wenox {

// Where the objective's fields come from: the interior of a padded SoA4 FP64
// state on the device (U != nullptr: rho = plane 0, u and v = planes 1 and 2
// over max(rho, 1e-12)), or packed host arrays rho, u, v of nx * ny values
// (row-major, y fastest).
struct ObjSource {
    const double* U = nullptr;
    int64_t ld = 0, plane = 0;
    const double* rho = nullptr;
    const double* u = nullptr;
    const double* v = nullptr;
};

// J and metrics[8] = {I_curve, E_perp, O, TV, cI, cE, cO, cTV}; w = {I, E, O,
// TV}.  S, C, band (nx * ny bytes) and median (nx * ny doubles) receive the
// intermediates when non-null.  Returns a WENOX_* code (errors are recorded
// as the library's last error).
int objective_eval(const ObjSource& src, int64_t nx, int64_t ny, double dx, double dy, const double w[4],
                   hipStream_t s, double* J, double* metrics, uint8_t* S, uint8_t* C, uint8_t* band,
                   double* median);

}  // namespace wenox
