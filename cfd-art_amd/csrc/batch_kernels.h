// Device side of the multi-grid sweeps: per-grid state, tile work-lists, and the batch's x and
// y sweep kernels, templated on the arithmetic F. Included by batch.hip (F = false, the exact
// arithmetic) and by batch_fast.hip (F = true, compiled with FMA contraction allowed), each of
// which instantiates only its own arithmetic.
#pragma once

#include <hip/hip_runtime.h>

#include "cfdart.h"
#include "sweep_device.h"

namespace cfd {

struct BatchGrid {
    double* buf[2];
    long long plane;
    int pitch, nx, ny;
    int model, splitting;
    int bc[4];
    Material mat;
    SchemeOpts opt;
    double dx, dy, dx_half_col;
    int yH, xrows;
    int steps_left;
    int live;
    int sw;  // the state at a step boundary is in buf[sw] (a Strang step, three sweeps, flips it)
    Ctrl ctrl;
};

struct Tile {
    int grid, axis, a, b;  // x: (column tile, row group); y: (column block, row block)
};

// Work-list lengths, per model and axis (each list has its own kernel; the launch covers the
// list's capacity and workgroups past its length exit at once).
struct BatchQueue {
    int ntiles[2][2];  // [model][axis]
    int xcd[2][2];     // 1: the list's items are alike, so they may go to the XCDs in contiguous runs
};

// The work-list item of workgroup b of a launch over a list of n items, or -1 (the launch covers
// the list's capacity). Items that are all alike go to the 8 XCDs in contiguous runs (xcd_order), so
// that tiles sharing halo lines meet in one L2; otherwise in list order (heaviest items first).
__device__ __forceinline__ int batch_item(int b, int n, int xcd)
{
    if (b >= n) return -1;
    return xcd ? xcd_order(b, n) : b;
}

// The grid's view for its sweep number si of the step (0, 1 or 2): the sweeps alternate between
// its two buffers from buf[sw].
__device__ __forceinline__ GridView batch_view(BatchGrid* bg, int si)
{
    GridView g;
    const int a = (bg->sw + si) & 1;
    g.cur = bg->buf[a];
    g.nxt = bg->buf[a ^ 1];
    g.plane = bg->plane;
    g.pitch = bg->pitch;
    g.nx = bg->nx;
    g.ny = bg->ny;
    for (int i = 0; i < 4; ++i) g.bc[i] = bg->bc[i];
    g.mat = bg->mat;
    g.opt = bg->opt;
    g.dx = bg->dx;
    g.dy = bg->dy;
    g.dx_half_col = bg->dx_half_col;
    g.ctrl = &bg->ctrl;
    return g;
}

// What a grid does in phase `phase` of a step of nph phases (E/headers/splitting_scheme.hpp:37-122
// per grid): the axis it sweeps (-1: none: dead, finished, or no sweep of its scheme falls in
// this phase), which of its own sweeps of the step it is (si, for its buffers), whether that
// sweep ends its step (the wave speeds of its next dt), and the factor on its dt (Strang's x
// sweeps take dt/2, as the reference passes 0.5 * dt). A two-phase step is x then y (y then x for
// an alternating grid on an odd step); a three-phase step is x, y, x: a Strang grid takes all
// three, a grid going x then y the first two, a grid going y then x the last two.
struct BatchSweep {
    int axis, si;
    bool last;
    double f;
};
__device__ __forceinline__ BatchSweep batch_sweep(const BatchGrid* bg, int phase, int nph)
{
    BatchSweep s{-1, 0, false, 1.0};
    if (!bg->live || !bg->ctrl.active) return s;
    if (bg->splitting == CFDART_STRANG) {
        if (nph != 3) return s;  // a batch holding a Strang grid launches only three-phase steps
        s.axis = phase == 1 ? 1 : 0;
        s.si = phase;
        s.last = phase == 2;
        s.f = phase == 1 ? 1.0 : 0.5;
        return s;
    }
    const bool y_first = bg->splitting == CFDART_ALTERNATING && (bg->ctrl.n & 1);
    if (nph == 2) {
        s.axis = y_first ? 1 - phase : phase;
        s.si = phase;
        s.last = phase == 1;
        return s;
    }
    if (y_first ? phase == 0 : phase == 2) return s;
    s.si = y_first ? phase - 1 : phase;
    s.axis = y_first ? (phase == 1 ? 1 : 0) : phase;
    s.last = s.si == 1;
    return s;
}

// x sweep of one phase over the x tiles of every live grid of model M (one tile per workgroup,
// the tile of sweep_x_kernel). Tiles of grids on the other axis this phase exit at once.
// (Every x sweep of a LAST launch ends its grid's step: x is the last sweep of every scheme
// whose step has an x sweep after the first phase.)
template <int M, bool LAST, bool F>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(M == ALLAIRE ? 3 : (LAST ? 4 : 1)))) void
batch_x_kernel(BatchGrid* grids, const Tile* tiles, const BatchQueue* q, int phase, int nph)
{
    const int t = batch_item(blockIdx.x, q->ntiles[M][0], q->xcd[M][0]);
    if (t < 0) return;
    const Tile tl = tiles[t];
    BatchGrid* bg = grids + tl.grid;
    const BatchSweep sw = batch_sweep(bg, phase, nph);
    if (sw.axis != 0) return;
    const GridView g = batch_view(bg, sw.si);
    const int xrows = bg->xrows;
    Tally ta;
    x_tile<M, LAST, F>(g, sw.f * bg->ctrl.dt, tl.a, (tl.b * 4 + wave_index()) * xrows, xrows, ta);
    ta.flush(&bg->ctrl, LAST, t);
}

// y sweep of one phase over the y tiles of every live grid of model M: the transposed tile of
// sweep_yt_kernel (64 rows x 16 columns through LDS, lanes along y).
// (A LAST launch's y sweep ends the step of a grid going x then y, but not that of a Strang grid
// or of a grid going y then x in a three-phase step: the wave speeds are taken per grid.)
template <int M, bool LAST, bool F>
__global__ __launch_bounds__(256) void batch_y_kernel(BatchGrid* grids, const Tile* tiles, const BatchQueue* q,
                                                      int phase, int nph)
{
    __shared__ double sU[Model<M>::NC * 64 * (BATCH_YT_COLS + 1)];
    const int t = batch_item(blockIdx.x, q->ntiles[M][1], q->xcd[M][1]);
    if (t < 0) return;
    const Tile tl = tiles[t];
    BatchGrid* bg = grids + tl.grid;
    const BatchSweep sw = batch_sweep(bg, phase, nph);
    if (sw.axis != 1) return;
    const GridView g = batch_view(bg, sw.si);
    Tally ta;
    y_tile_t<M, LAST, BATCH_YT_COLS, F>(g, sw.f * bg->ctrl.dt, tl.a * BATCH_YT_COLS, tl.b * YT_ROWS, sU, ta, sw.last);
    ta.flush(&bg->ctrl, LAST && sw.last, t);
}

// Which sweeps a captured step launches: every axis in both phases (grids of the batch sweep
// different axes in a phase), or only x then y, or only y then x (every running grid of the batch
// sweeps that axis in that phase: the other launches would be all exits), or x, y, x in three
// phases (the batch holds a Strang grid).
enum StepShape { STEP_BOTH = 0, STEP_XY = 1, STEP_YX = 2, STEP_XYX = 3, STEP_SHAPES = 4 };

// The x and y kernels of model M for the sweep phases of a step of the given shape, on one stream
// (each kernel's grid is its work-list's capacity; items past the live count exit at once).
template <int M, bool F>
void launch_batch_phases(BatchGrid* grids, Tile* const* tiles, const size_t* cap, const BatchQueue* q, hipStream_t st,
                         int shape)
{
    const dim3 bx((unsigned)cap[0]), by((unsigned)cap[1]);
    if (shape == STEP_XYX) {
        hipLaunchKernelGGL((batch_x_kernel<M, false, F>), bx, dim3(256), 0, st, grids, tiles[0], q, 0, 3);
        hipLaunchKernelGGL((batch_y_kernel<M, true, F>), by, dim3(256), 0, st, grids, tiles[1], q, 1, 3);
        hipLaunchKernelGGL((batch_x_kernel<M, true, F>), bx, dim3(256), 0, st, grids, tiles[0], q, 2, 3);
        return;
    }
    if (shape != STEP_YX) hipLaunchKernelGGL((batch_x_kernel<M, false, F>), bx, dim3(256), 0, st, grids, tiles[0], q, 0, 2);
    if (shape != STEP_XY) hipLaunchKernelGGL((batch_y_kernel<M, false, F>), by, dim3(256), 0, st, grids, tiles[1], q, 0, 2);
    if (shape != STEP_XY) hipLaunchKernelGGL((batch_x_kernel<M, true, F>), bx, dim3(256), 0, st, grids, tiles[0], q, 1, 2);
    if (shape != STEP_YX) hipLaunchKernelGGL((batch_y_kernel<M, true, F>), by, dim3(256), 0, st, grids, tiles[1], q, 1, 2);
}

// The fast arithmetic's phases (batch_fast.hip).
void launch_batch_phases_fast(int model, BatchGrid* grids, Tile* const* tiles, const size_t* cap, const BatchQueue* q,
                              hipStream_t st, int shape);

}  // namespace cfd
