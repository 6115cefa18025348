// HIP kernels for gfx950: the fused x and y sweeps of the MUSCL-Hancock scheme, the CFL
// reduction and the device-side time-loop bookkeeping.
//
// x sweep: one cell per lane along x (coalesced 512-byte rows per component), a wavefront
//   covers 64 consecutive columns of which the middle 60 are written; neighbour values move
//   between lanes with DPP wave shifts, so every per-cell quantity is computed once.
// y sweep, transposed tile (the default for MUSCL-Hancock): a workgroup stages 64 rows x 16 (or
//   8) columns through LDS and each wavefront computes columns as the x sweep computes rows,
//   lanes along y with DPP neighbours; strip tile (the Godunov scheme, or on request): one
//   column per lane, each lane marching down a strip of rows with a three-cell register window.
// Both fuse the ghost fill (mirrored indexing) of E/sourcecode/euler_ppsweep.cpp:7-118, the
// face fluxes, the conservative update, the Euler repair / Allaire z-update and snapping, and,
// in the last sweep of a step, the per-cell wave speeds of compute_dt
// (E/sourcecode/euler_timestep.cpp:14-25) reduced to one atomic max per wavefront.
//
// This file is compiled twice. As itself it builds the exact arithmetic (F = false, no FP
// contraction: bitwise the reference's arithmetic) and everything that has one arithmetic only
// (bookkeeping, mass, the launchers' dispatch). Through sweep_kernels_fast.hip (CFD_FAST_TU = 1)
// it builds the same kernel templates for the fast arithmetic (F = true) with FMA contraction
// allowed, and that arithmetic's launchers.
#include "sweep_device.h"

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#ifndef CFD_FAST_TU
#define CFD_FAST_TU 0
#endif

namespace cfd {

// Occupancy of the MUSCL-Hancock x sweeps, as measured fastest: Allaire at 3 wavefronts per SIMD
// (at most 168 VGPRs), the Euler last sweep at 4 (at most 128 VGPRs instead of its natural 134).
template <int M, bool LAST, bool GOD, bool F>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(M == ALLAIRE && !GOD ? 3 : (LAST && !GOD ? 4 : 1)))) void sweep_x_kernel(GridView g, double dt_arg, double dt_factor, int use_ctrl,
                                                     int xrows)
{
    // tiles in row-major order: neighbours along x (which share halo columns) are consecutive
    const int b = blockIdx.x + blockIdx.y * gridDim.x;
    const int L = xcd_order(b, gridDim.x * gridDim.y);
    const int tx = L % gridDim.x, ty = L / gridDim.x;
    const int row0 = (ty * 4 + wave_index()) * xrows;
    double dt;
    if (!sweep_dt(g, dt_arg, dt_factor, use_ctrl, dt)) {
        const int lane = threadIdx.x & 63, x = tx * XTILE - 2 + lane;
        copy_rows<M>(g, x, lane >= 2 && lane < 62 && x < g.nx, row0, min(row0 + xrows, g.ny));
        return;
    }
    Tally tl;
    if (GOD) x_tile_godunov<M, LAST, F>(g, dt, tx, row0, xrows, tl);
    else x_tile<M, LAST, F>(g, dt, tx, row0, xrows, tl);
    tl.flush(g.ctrl, LAST, blockIdx.x + blockIdx.y * gridDim.x);
}

template <int M, bool LAST, bool GOD, bool F>
__global__ __launch_bounds__(YBLOCK) void sweep_y_kernel(GridView g, double dt_arg, double dt_factor,
                                                        int use_ctrl, int H)
{
    // tiles in column-major order: neighbours along y (which share halo rows) are consecutive
    const int b = blockIdx.x + blockIdx.y * gridDim.x;
    const int L = xcd_order(b, gridDim.x * gridDim.y);
    const int bx = L / gridDim.y, by = L % gridDim.y;
    const int y0 = by * H;
    const int y1 = min(y0 + H, g.ny);
    double dt;
    if (!sweep_dt(g, dt_arg, dt_factor, use_ctrl, dt)) {
        const int x = bx * YBLOCK + threadIdx.x;
        copy_rows<M>(g, x, x < g.nx, y0, y1);
        return;
    }
    Tally tl;
    if (GOD) y_tile_godunov<M, LAST, F>(g, dt, bx, y0, y1, tl);
    else y_tile<M, LAST, F>(g, dt, bx, y0, y1, tl);
    tl.flush(g.ctrl, LAST, blockIdx.x + blockIdx.y * gridDim.x);
}

// Transposed y sweep (MUSCL-Hancock): a workgroup stages 64 rows x COLS columns through LDS and
// its four wavefronts run lanes along y; tiles in column-major order like sweep_y_kernel.
template <int M, bool LAST, int COLS, bool F>
__global__ __launch_bounds__(256) void sweep_yt_kernel(GridView g, double dt_arg, double dt_factor, int use_ctrl)
{
    __shared__ double sU[Model<M>::NC * 64 * (COLS + 1)];
    const int b = blockIdx.x + blockIdx.y * gridDim.x;
    const int L = xcd_order(b, gridDim.x * gridDim.y);
    const int bx = L / gridDim.y, by = L % gridDim.y;
    const int x0 = bx * COLS, y0 = by * YT_ROWS;
    double dt;
    if (!sweep_dt(g, dt_arg, dt_factor, use_ctrl, dt)) {
        if (threadIdx.x < COLS) {
            const int x = x0 + threadIdx.x;
            copy_rows<M>(g, x, x < g.nx, y0, min(y0 + YT_ROWS, g.ny));
        }
        return;
    }
    Tally tl;
    y_tile_t<M, LAST, COLS, F>(g, dt, x0, y0, sU, tl);
    tl.flush(g.ctrl, LAST, b);
}

// ------------------------------------------------------------------ dt reduction / time loop

// compute_dt's per-cell pass over the current state (E/sourcecode/euler_timestep.cpp:12-30):
// maxima into the control block, per-block mass partial sums (fixed-shape tree, deterministic).
// The wave speeds take the grid's arithmetic, as the last sweep of a step computes them, so that a
// run's result does not depend on how it is split into cfdart_run calls.
template <int M, bool F>
__global__ __launch_bounds__(256) void reduce_dt_kernel(GridView g, double* partials)
{
    constexpr int NC = Model<M>::NC;
    __shared__ double sm[2][4];
    const long long n = (long long)g.nx * g.ny;
    double su = 0.0, sv = 0.0, m1 = 0.0, m2 = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        int y = (int)(i / g.nx), x = (int)(i % g.nx);
        double U[NC];
        const double* b = g.cur + (long long)y * g.pitch + x;
#pragma unroll
        for (int k = 0; k < NC; ++k) U[k] = b[k * g.plane];
        double a, c;
        wave_speeds<F>(g.mat, U, a, c, (Model<M>*)nullptr);
        su = F ? fmax(su, a) : smax(su, a);
        sv = F ? fmax(sv, c) : smax(sv, c);
        m1 += U[0];
        if (M == ALLAIRE) m2 += U[1];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) { m1 += __shfl_xor(m1, o); m2 += __shfl_xor(m2, o); }
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        sm[0][w] = m1;
        sm[1][w] = m2;
    }
    Tally tl;
    tl.su = su;
    tl.sv = sv;
    tl.flush(g.ctrl, true, blockIdx.x);
    __syncthreads();
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = ((sm[0][0] + sm[0][1]) + (sm[0][2] + sm[0][3]));
        partials[2 * blockIdx.x + 1] = ((sm[1][0] + sm[1][1]) + (sm[1][2] + sm[1][3]));
    }
}

#if !CFD_FAST_TU
// The mass sums of compute_dt (E/sourcecode/euler_timestep.cpp:26,32-33): the same per-block
// partials as reduce_dt_kernel (same decomposition, same order), reading the density planes only.
template <int M>
__global__ __launch_bounds__(256) void mass_kernel(GridView g, double* partials)
{
    __shared__ double sm[2][4];
    const long long n = (long long)g.nx * g.ny;
    double m1 = 0.0, m2 = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        int y = (int)(i / g.nx), x = (int)(i % g.nx);
        const double* b = g.cur + (long long)y * g.pitch + x;
        m1 += b[0];
        if (M == ALLAIRE) m2 += b[g.plane];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) { m1 += __shfl_xor(m1, o); m2 += __shfl_xor(m2, o); }
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        sm[0][w] = m1;
        sm[1][w] = m2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = ((sm[0][0] + sm[0][1]) + (sm[0][2] + sm[0][3]));
        partials[2 * blockIdx.x + 1] = ((sm[1][0] + sm[1][1]) + (sm[1][2] + sm[1][3]));
    }
}

// One entry of the mass history, as compute_dt records it at the start of a step that runs
// (E/sourcecode/euler_timestep.cpp:32-33): t, then masssum * dx * dy per fluid.
__global__ void mass_record_kernel(const Ctrl* c, const double* partials, int nb, double dx, double dy, double* hist)
{
    if (threadIdx.x != 0 || blockIdx.x != 0 || !c->active) return;
    double m1 = 0.0, m2 = 0.0;
    for (int b = 0; b < nb; ++b) {
        m1 += partials[2 * b];
        m2 += partials[2 * b + 1];
    }
    hist[0] = c->t;
    hist[1] = m1 * dx * dy;
    hist[2] = m2 * dx * dy;
}

// Start of the time loop: dt for (n, t) from a fresh reduction, and the 'while (t < T)' test. A
// grid whose Allaire state has failed the physical-state test (where the reference's assert
// aborts the run) takes no further steps.
__global__ __launch_bounds__(64) void init_dt_kernel(Ctrl* c, double dx, double dy)
{
    double mu, mv;
    fold_maxima_wave(c, mu, mv);  // the whole wavefront
    if (threadIdx.x != 0) return;
    c->active = c->t < c->T && !c->error;
    if (c->active) c->dt = dt_from_maxima(c, mu, mv, c->n, c->t, dx, dy);
}

// End of a step (E/headers/simulation.hpp:68-70): t += dt, n++, snap to T, loop test, and the
// next step's dt from the maxima that the step's last sweep accumulated. The loop ends after a
// step in which an Allaire state failed the physical-state test (the reference's assert aborts
// there), as it does for a grid in a batch.
__global__ __launch_bounds__(64) void finalize_step_kernel(Ctrl* c, double dx, double dy)
{
    // everything the step's end reads is loaded at once (one memory latency rather than a chain):
    // lane i's slot maxima and the control scalars
    const int lane = threadIdx.x & 63;
    const unsigned long long su = c->maxu(lane), sv = c->maxv(lane);
    const int active = c->active, n0 = c->n, error = c->error;
    const double t0 = c->t, dt0 = c->dt, T = c->T, cfl = c->cfl;
    if (!active) return;  // the same value in every lane
    c->maxu(lane) = 0;
    c->maxv(lane) = 0;
    // the slots hold non-negative doubles (or zero), whose bit patterns order as their values
    const double mu = wave_max(__longlong_as_double((long long)su));
    const double mv = wave_max(__longlong_as_double((long long)sv));
    if (lane != 0) return;
    double t = t0 + dt0;
    const int n = n0 + 1;
    if (fabs(t - T) < 1e-12) t = T;
    c->t = t;
    c->n = n;
    const bool go = t < T && !error;
    c->active = go;
    if (go) c->dt = dt_from_values(cfl, T, mu, mv, n, t, dx, dy);
}

__global__ void set_run_kernel(Ctrl* c, double T, double cfl)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) { c->T = T; c->cfl = cfl; }
}
__global__ void set_time_kernel(Ctrl* c, double t, int n)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) { c->t = t; c->n = n; }
}

// ------------------------------------------------------------------ launchers

hipError_t launch_set_run(Ctrl* c, double T, double cfl, hipStream_t st)
{
    hipLaunchKernelGGL(set_run_kernel, dim3(1), dim3(64), 0, st, c, T, cfl);
    return hipGetLastError();
}
hipError_t launch_set_time(Ctrl* c, double t, int n, hipStream_t st)
{
    hipLaunchKernelGGL(set_time_kernel, dim3(1), dim3(64), 0, st, c, t, n);
    return hipGetLastError();
}

#endif  // !CFD_FAST_TU

// A sweep launch on stream st. With a pair of events (stamps.start non-null) the launch goes
// through hipExtLaunchKernelGGL, which stamps the two events with the dispatch's own start and
// end times: their elapsed time is the kernel's execution, as a profiler's kernel trace reports
// it, without the time the stream spends processing separate event records around the launch.
#define CFD_SWEEP_LAUNCH(K, grid, stamps, st, ...)                                                         \
    do {                                                                                                 \
        if ((stamps).start) hipExtLaunchKernelGGL(K, grid, dim3(256), 0, st, (stamps).start, (stamps).stop, 0, __VA_ARGS__); \
        else hipLaunchKernelGGL(K, grid, dim3(256), 0, st, __VA_ARGS__);                                 \
    } while (0)

template <int M, bool LAST, bool GOD, bool F>
static hipError_t launch_x(const GridView& g, double dt_arg, double f, int use_ctrl, int xrows, hipStream_t st,
                           Stamps ts)
{
    dim3 grid((g.nx + XTILE - 1) / XTILE, (g.ny + 4 * xrows - 1) / (4 * xrows));
    CFD_SWEEP_LAUNCH((sweep_x_kernel<M, LAST, GOD, F>), grid, ts, st, g, dt_arg, f, use_ctrl, xrows);
    return hipGetLastError();
}
template <int M, bool LAST, bool GOD, bool F>
static hipError_t launch_y(const GridView& g, double dt_arg, double f, int use_ctrl, int H, hipStream_t st, Stamps ts)
{
    if (y_transposed(H) && !GOD) {
        const int cols = yt_cols(H);
        dim3 grid((g.nx + cols - 1) / cols, (g.ny + YT_ROWS - 1) / YT_ROWS);
        if (cols == 16) CFD_SWEEP_LAUNCH((sweep_yt_kernel<M, LAST, 16, F>), grid, ts, st, g, dt_arg, f, use_ctrl);
        else CFD_SWEEP_LAUNCH((sweep_yt_kernel<M, LAST, 8, F>), grid, ts, st, g, dt_arg, f, use_ctrl);
        return hipGetLastError();
    }
    if (H <= 0) H = 8;  // the first-order scheme has no transposed tile: 8-row strips (its fastest at 2048^2)
    static_assert(YBLOCK == 256, "strip tiles launch 256-thread workgroups");
    dim3 grid((g.nx + YBLOCK - 1) / YBLOCK, (g.ny + H - 1) / H);
    CFD_SWEEP_LAUNCH((sweep_y_kernel<M, LAST, GOD, F>), grid, ts, st, g, dt_arg, f, use_ctrl, H);
    return hipGetLastError();
}
#undef CFD_SWEEP_LAUNCH
template <int M, bool GOD, bool F>
static hipError_t launch_model(const GridView& g, int axis, bool last, double dt_arg, double dt_factor, int use_ctrl,
                               int yH, int xrows, hipStream_t st, Stamps ts)
{
    if (axis == 0) return last ? launch_x<M, true, GOD, F>(g, dt_arg, dt_factor, use_ctrl, xrows, st, ts)
                               : launch_x<M, false, GOD, F>(g, dt_arg, dt_factor, use_ctrl, xrows, st, ts);
    return last ? launch_y<M, true, GOD, F>(g, dt_arg, dt_factor, use_ctrl, yH, st, ts)
                : launch_y<M, false, GOD, F>(g, dt_arg, dt_factor, use_ctrl, yH, st, ts);
}

template <bool F>
static hipError_t launch_sweep_t(const GridView& g, int model, int axis, bool last, double dt_arg, double dt_factor,
                                 int use_ctrl, int yH, int xrows, hipStream_t st, Stamps ts)
{
#define CFD_LAUNCH(M, GOD) launch_model<M, GOD, F>(g, axis, last, dt_arg, dt_factor, use_ctrl, yH, xrows, st, ts)
    if (model == EULER) return g.opt.godunov ? CFD_LAUNCH(EULER, true) : CFD_LAUNCH(EULER, false);
    return g.opt.godunov ? CFD_LAUNCH(ALLAIRE, true) : CFD_LAUNCH(ALLAIRE, false);
#undef CFD_LAUNCH
}

template <bool F>
static hipError_t launch_reduce_dt_t(const GridView& g, int model, double* partials, hipStream_t st)
{
    const int nb = reduce_blocks(g);
    if (model == EULER) hipLaunchKernelGGL((reduce_dt_kernel<EULER, F>), dim3(nb), dim3(256), 0, st, g, partials);
    else hipLaunchKernelGGL((reduce_dt_kernel<ALLAIRE, F>), dim3(nb), dim3(256), 0, st, g, partials);
    return hipGetLastError();
}

#if CFD_FAST_TU
hipError_t launch_sweep_fast(const GridView& g, int model, int axis, bool last, double dt_arg, double dt_factor,
                             int use_ctrl, int yH, int xrows, hipStream_t st, Stamps ts)
{
    return launch_sweep_t<true>(g, model, axis, last, dt_arg, dt_factor, use_ctrl, yH, xrows, st, ts);
}
hipError_t launch_reduce_dt_fast(const GridView& g, int model, double* partials, hipStream_t st)
{
    return launch_reduce_dt_t<true>(g, model, partials, st);
}
#else
hipError_t launch_sweep(const GridView& g, int model, int axis, bool last, double dt_arg,
                        double dt_factor, int use_ctrl, int yH, int xrows, hipStream_t st, Stamps ts)
{
    if (g.opt.fast) return launch_sweep_fast(g, model, axis, last, dt_arg, dt_factor, use_ctrl, yH, xrows, st, ts);
    return launch_sweep_t<false>(g, model, axis, last, dt_arg, dt_factor, use_ctrl, yH, xrows, st, ts);
}
hipError_t launch_reduce_dt(const GridView& g, int model, double* partials, hipStream_t st)
{
    if (g.opt.fast) return launch_reduce_dt_fast(g, model, partials, st);
    return launch_reduce_dt_t<false>(g, model, partials, st);
}

int reduce_blocks(const GridView& g)
{
    long long n = (long long)g.nx * g.ny;
    long long b = (n + 255) / 256;
    return (int)(b < 1024 ? b : 1024);
}

hipError_t launch_mass_record(const GridView& g, int model, double* partials, double* hist, hipStream_t st)
{
    const int nb = reduce_blocks(g);
    if (model == EULER) hipLaunchKernelGGL(mass_kernel<EULER>, dim3(nb), dim3(256), 0, st, g, partials);
    else hipLaunchKernelGGL(mass_kernel<ALLAIRE>, dim3(nb), dim3(256), 0, st, g, partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(mass_record_kernel, dim3(1), dim3(64), 0, st, g.ctrl, partials, nb, g.dx, g.dy, hist);
    return hipGetLastError();
}

hipError_t launch_init_dt(Ctrl* c, double dx, double dy, hipStream_t st)
{
    hipLaunchKernelGGL(init_dt_kernel, dim3(1), dim3(64), 0, st, c, dx, dy);
    return hipGetLastError();
}

hipError_t launch_finalize(Ctrl* c, double dx, double dy, hipStream_t st)
{
    hipLaunchKernelGGL(finalize_step_kernel, dim3(1), dim3(64), 0, st, c, dx, dy);
    return hipGetLastError();
}

#endif  // CFD_FAST_TU

}  // namespace cfd
