// Device data structures shared by the sweep kernels and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>

#include "cfdart.h"
#include "physics.h"

namespace cfd {

// Per-grid device control block: the state of the reference's time loop
// (E/headers/simulation.hpp:48-61) kept on the device so a run of many steps needs no
// host round trip per step.
constexpr int MAX_SLOTS = 64;  // spread of the wave-speed maxima over addresses (no hot atomic)
struct Ctrl {
    double t, dt, T, cfl;
    int n, active;
    unsigned long long cnt[4];      // slope fallbacks, evolve fallbacks, repairs, unphysical
    int error;
    int pad;
    // bit patterns of non-negative doubles, max-reduced by wavefronts into slot (id % MAX_SLOTS):
    // slot i's u and v maxima share a 128-byte line of their own, so the atomics of different
    // slots go to different lines (and memory channels) instead of queueing on a few lines
    alignas(128) unsigned long long slots[MAX_SLOTS][16];
    __host__ __device__ unsigned long long& maxu(int i) { return slots[i][0]; }
    __host__ __device__ unsigned long long& maxv(int i) { return slots[i][1]; }
    __host__ __device__ const unsigned long long& maxu(int i) const { return slots[i][0]; }
    __host__ __device__ const unsigned long long& maxv(int i) const { return slots[i][1]; }
};

// Device view of one grid: SoA planes, interior cells only (the ghost frame is produced on
// the fly by mirrored indexing, which is the reference's pre_sweep fill).
struct GridView {
    const double* cur;
    double* nxt;
    long long plane;  // doubles between component planes
    int pitch;        // doubles between rows
    int nx, ny;
    int bc[4];        // left, right, bottom, top: 0 transmissive, 1 reflective
    Material mat;
    SchemeOpts opt;
    double dx, dy;
    double dx_half_col;  // spacing of the column sweep's half-step (reference: dx)
    Ctrl* ctrl;
};

// Sweep-kernel launch geometry.
constexpr int XTILE = 60;         // output columns per wavefront in the x sweep (64 lanes - 4 halo)
constexpr int YBLOCK = 256;       // threads (columns) per y-sweep block
constexpr int YT_ROWS = 60;       // output rows per transposed y-sweep tile (64 lanes - 4 halo)
// y-sweep tile selectors (the strip height's place): transposed LDS tiles of 16 or 8 columns
constexpr int Y_TRANSPOSED = CFDART_Y_TRANSPOSED;
constexpr int Y_TRANSPOSED_NARROW = CFDART_Y_TRANSPOSED_NARROW;
constexpr bool y_transposed(int yH) { return yH == Y_TRANSPOSED || yH == Y_TRANSPOSED_NARROW; }
constexpr int yt_cols(int yH) { return yH == Y_TRANSPOSED_NARROW ? 8 : 16; }
constexpr int BATCH_YT_COLS = 16;  // the batch's y tile: the 16-column transposed tile

// Host-side launchers (sweep_kernels.hip). A sweep launch given a pair of events records in them
// the start and end of the kernel's own execution (hipExtLaunchKernelGGL); without (the default)
// it is a plain launch.
struct Stamps {
    hipEvent_t start = nullptr, stop = nullptr;
};
hipError_t launch_sweep(const GridView& g, int model, int axis, bool last, double dt_arg,
                        double dt_factor, int use_ctrl, int yH, int xrows, hipStream_t st, Stamps ts = Stamps());
hipError_t launch_reduce_dt(const GridView& g, int model, double* partials, hipStream_t st);
// The fast arithmetic's launchers (sweep_kernels_fast.hip); launch_sweep and launch_reduce_dt
// forward to them for a grid whose opt.fast is set.
hipError_t launch_sweep_fast(const GridView& g, int model, int axis, bool last, double dt_arg, double dt_factor,
                             int use_ctrl, int yH, int xrows, hipStream_t st, Stamps ts);
hipError_t launch_reduce_dt_fast(const GridView& g, int model, double* partials, hipStream_t st);
hipError_t launch_init_dt(Ctrl* c, double dx, double dy, hipStream_t st);
// Mass history entry of the current state (t, mass1, mass2 into hist[0..2]) when the time loop
// is running: the sums compute_dt records, from the same per-block partials.
hipError_t launch_mass_record(const GridView& g, int model, double* partials, double* hist, hipStream_t st);
hipError_t launch_finalize(Ctrl* c, double dx, double dy, hipStream_t st);
hipError_t launch_set_run(Ctrl* c, double T, double cfl, hipStream_t st);
hipError_t launch_set_time(Ctrl* c, double t, int n, hipStream_t st);
int reduce_blocks(const GridView& g);

}  // namespace cfd
