// Launchers of the device-side frame products (frames.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "sweep_kernels.h"

namespace cfd {

// Schlieren image |grad rho| / max |grad rho| of the state g.cur into out (nx*ny doubles,
// row-major); *mx receives the bit pattern of the maximum. Stream-ordered.
hipError_t launch_schlieren(const GridView& g, int model, double* out, unsigned long long* mx, hipStream_t st);

}  // namespace cfd
