// Device side of the multi-grid sweeps: per-grid state, tile work-lists, and the batch's x and
// y sweep kernels, templated on the arithmetic F. Included by batch.hip (F = false, the exact
// arithmetic) and by batch_fast.hip (F = true, compiled with FMA contraction allowed), each of
// which instantiates only its own arithmetic.
#pragma once

#include <hip/hip_runtime.h>

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

__device__ __forceinline__ GridView batch_view(BatchGrid* bg, int phase)
{
    GridView g;
    g.cur = bg->buf[phase];
    g.nxt = bg->buf[1 - phase];
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

// The sweep axis of a grid in a phase, or -1 when the grid does not run (dead or finished).
__device__ __forceinline__ int batch_axis(const BatchGrid* bg, int phase)
{
    if (!bg->live || !bg->ctrl.active) return -1;
    return (bg->splitting == 2 && (bg->ctrl.n & 1)) ? 1 - phase : phase;
}

// x sweep of one phase over the x tiles of every live grid of model M (one tile per workgroup,
// the tile of sweep_x_kernel). Tiles of grids on the other axis this phase exit at once.
template <int M, bool LAST, bool F>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(M == ALLAIRE ? 3 : (LAST ? 4 : 1)))) void
batch_x_kernel(BatchGrid* grids, const Tile* tiles, const BatchQueue* q, int phase)
{
    const int t = batch_item(blockIdx.x, q->ntiles[M][0], q->xcd[M][0]);
    if (t < 0) return;
    const Tile tl = tiles[t];
    BatchGrid* bg = grids + tl.grid;
    if (batch_axis(bg, phase) != 0) return;
    const GridView g = batch_view(bg, phase);
    const int xrows = bg->xrows;
    Tally ta;
    x_tile<M, LAST, F>(g, bg->ctrl.dt, tl.a, (tl.b * 4 + wave_index()) * xrows, xrows, ta);
    ta.flush(&bg->ctrl, LAST, t);
}

// y sweep of one phase over the y tiles of every live grid of model M: the transposed tile of
// sweep_yt_kernel (64 rows x 16 columns through LDS, lanes along y).
template <int M, bool LAST, bool F>
__global__ __launch_bounds__(256) void batch_y_kernel(BatchGrid* grids, const Tile* tiles, const BatchQueue* q,
                                                      int phase)
{
    __shared__ double sU[Model<M>::NC * 64 * (BATCH_YT_COLS + 1)];
    const int t = batch_item(blockIdx.x, q->ntiles[M][1], q->xcd[M][1]);
    if (t < 0) return;
    const Tile tl = tiles[t];
    BatchGrid* bg = grids + tl.grid;
    if (batch_axis(bg, phase) != 1) return;
    const GridView g = batch_view(bg, phase);
    Tally ta;
    y_tile_t<M, LAST, BATCH_YT_COLS, F>(g, bg->ctrl.dt, tl.a * BATCH_YT_COLS, tl.b * YT_ROWS, sU, ta);
    ta.flush(&bg->ctrl, LAST, t);
}

// Which sweeps a captured step launches: every axis in both phases (grids of the batch sweep
// different axes in a phase), or only x then y, or only y then x (every running grid of the batch
// sweeps that axis in that phase: the other launches would be all exits).
enum StepShape { STEP_BOTH = 0, STEP_XY = 1, STEP_YX = 2, STEP_SHAPES = 3 };

// The x and y kernels of model M for the sweep phases of a step of the given shape, on one stream
// (each kernel's grid is its work-list's capacity; items past the live count exit at once).
template <int M, bool F>
void launch_batch_phases(BatchGrid* grids, Tile* const* tiles, const size_t* cap, const BatchQueue* q, hipStream_t st,
                         int shape)
{
    const dim3 bx((unsigned)cap[0]), by((unsigned)cap[1]);
    if (shape != STEP_YX) hipLaunchKernelGGL((batch_x_kernel<M, false, F>), bx, dim3(256), 0, st, grids, tiles[0], q, 0);
    if (shape != STEP_XY) hipLaunchKernelGGL((batch_y_kernel<M, false, F>), by, dim3(256), 0, st, grids, tiles[1], q, 0);
    if (shape != STEP_XY) hipLaunchKernelGGL((batch_x_kernel<M, true, F>), bx, dim3(256), 0, st, grids, tiles[0], q, 1);
    if (shape != STEP_YX) hipLaunchKernelGGL((batch_y_kernel<M, true, F>), by, dim3(256), 0, st, grids, tiles[1], q, 1);
}

// The fast arithmetic's phases (batch_fast.hip).
void launch_batch_phases_fast(int model, BatchGrid* grids, Tile* const* tiles, const size_t* cap, const BatchQueue* q,
                              hipStream_t st, int shape);

}  // namespace cfd
