// Launchers of the device-side initial conditions and checksums (ic_device.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "cfdart_host.h"
#include "ic_common.h"

namespace cfd {

struct Ambient4 {
    double U[4];
};
struct BubbleStates {
    double water[6], air[6], post[6];
    double shock_x;
};

// Interior SoA planes of buf (row pitch `pitch`, plane stride `plane`) from the seeded draws;
// d_discs / d_bub are device copies of the draws (draw_discs / draw_bubbles).
hipError_t launch_ic_discs(double* buf, int pitch, long long plane, const cfdart_desc& d, const double* ambient_U,
                           const Disc* d_discs, int n, hipStream_t st);
hipError_t launch_ic_bubbles(double* buf, int pitch, long long plane, const cfdart_desc& d,
                             const cfdart_bubble_params& p, const Bubble* d_bub, int n, hipStream_t st);
// Adds the state checksum of the interior planes to *out (which the caller zeroes).
hipError_t launch_checksum(const double* buf, int pitch, long long plane, int nx, int ny, int nc,
                           unsigned long long* out, hipStream_t st);
// The same checksum of bufs[*sel & 1], the buffer chosen on the device when the kernel runs
// (bufs and sel in device memory: a batch grid whose state buffer changes from step to step).
hipError_t launch_checksum_sel(const double* const* bufs, const int* sel, int pitch, long long plane, int nx, int ny,
                               int nc, unsigned long long* out, hipStream_t st);

// Bytes of device memory generate_state needs for the draws.
constexpr size_t IC_DRAW_BYTES = sizeof(Disc) * IC_MAX_DISCS;

// Fills the interior planes of buf from exactly one of dp (Euler random discs) or bp (Allaire
// shocked bubbles): host draws, copied into d_draws (IC_DRAW_BYTES of device memory), then the
// device generator on stream st. Without h_stage the draws are copied from the stack and the
// stream is synchronised before returning; with h_stage (IC_DRAW_BYTES of pinned host memory)
// the draws are copied from there and nothing waits: the caller keeps h_stage unchanged until the
// stream has passed the copy. 0 or CFDART_E_*.
int generate_state(const cfdart_desc& d, const cfdart_disc_params* dp, const cfdart_bubble_params* bp, double* buf,
                   int pitch, long long plane, void* d_draws, hipStream_t st, std::string& err,
                   void* h_stage = nullptr);

}  // namespace cfd
