// Launchers of the device-side frame products (frames.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "sweep_kernels.h"

namespace cfd {

// Schlieren image |grad rho| / max |grad rho| of the state g.cur into out (nx*ny doubles,
// row-major); *mx receives the bit pattern of the maximum. Stream-ordered.
hipError_t launch_schlieren(const GridView& g, int model, double* out, unsigned long long* mx, hipStream_t st);

// Plain device copy of n doubles (grid-stride, ordinary loads and stores): the reference kernel
// of known traffic that hardware byte counters are calibrated against. Stream-ordered.
hipError_t launch_copy(const double* src, double* dst, long long n, hipStream_t st);

}  // namespace cfd
