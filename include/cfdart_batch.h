/*
 * cfdart batch: many independent grids (Euler and Allaire, uneven sizes) advancing together.
 *
 * No counterpart in the reference, which runs one grid per process
 * (euler_equations/euler_simulation_code/sourcecode/main.cpp:17-49). Each grid in a batch
 * follows exactly the reference's time loop (simulation.hpp:48-61) with its own CFL time step,
 * step budget and end time; its results are bitwise those of the same grid run alone through
 * cfdart.h in the same arithmetic (cfdart_batch_set_arith). Per sweep phase, one x-sweep and
 * one y-sweep launch per model process the tiles of every live grid from device work-lists (one
 * tile per workgroup; a grid's tiles for the axis it does not sweep in that phase exit at once),
 * and the step (two or three phases, loop bookkeeping) is captured once in a HIP graph and
 * replayed. Grids join and leave between steps by editing
 * device-resident tables; the graph is re-captured only when a work-list outgrows its capacity or
 * the first grid of a model joins (or the last one leaves) the batch.
 *
 * Splitting schemes: CFDART_NAIVE, CFDART_ALTERNATING (two sweeps per step) and CFDART_STRANG
 * (x(dt/2), y(dt), x(dt/2)); while the batch holds a Strang grid, its steps launch three sweep
 * phases (x, y, x) and every other grid takes two of them.
 * Admissions, removals and tile changes take effect at the next cfdart_batch_step, which
 * rebuilds the work-list once for all of them.
 * Conventions as in cfdart.h: 0 or a negative CFDART_E_* code; one host thread per batch.
 */
#ifndef CFDART_BATCH_H
#define CFDART_BATCH_H

#include <stddef.h>

#include "cfdart.h"
#include "cfdart_host.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cfdart_batch cfdart_batch;

/* capacity: maximum number of live grids; mem_budget_bytes: admission budget for grid state
 * (two FP64 buffers per grid), 0 = no limit. */
int cfdart_batch_create(int capacity, size_t mem_budget_bytes, cfdart_batch** out);
void cfdart_batch_destroy(cfdart_batch* b);

/* Admits a grid: uploads host_state (format of cfdart.h), computes its first dt and adds its
 * tiles to the work-list. It then runs until it has taken max_steps steps or reached t >= T.
 * y_strip_rows / x_rows_per_wave: tile shape. The y sweep of a batch runs the 16-column
 * transposed tile: y_strip_rows is 0 or CFDART_Y_TRANSPOSED; x_rows_per_wave as cfdart_set_tile
 * (0 = default 8). Returns CFDART_E_NOMEM when the
 * memory budget or the capacity would be exceeded (nothing is changed). slot_out receives the
 * grid's slot id. */
int cfdart_batch_add(cfdart_batch* b, const cfdart_desc* desc, const double* host_state, int splitting,
                     int max_steps, double T, double cfl, int y_strip_rows, int x_rows_per_wave,
                     int* slot_out);
/* As cfdart_batch_add with the initial state generated on the device from exactly one of
 * discs (Euler, cfdart_ic_random_discs) or bubbles (Allaire, cfdart_ic_shocked_bubbles);
 * the other is NULL. Bitwise the state the host generator would upload. Returns once the
 * generation and the first dt are queued on the batch stream, without waiting for the device. */
int cfdart_batch_add_generated(cfdart_batch* b, const cfdart_desc* desc, const cfdart_disc_params* discs,
                               const cfdart_bubble_params* bubbles, int splitting, int max_steps, double T,
                               double cfl, int y_strip_rows, int x_rows_per_wave, int* slot_out);
/* Removes a grid (finished or not) and frees its device memory for reuse, in stream order and
 * without waiting: work queued earlier still sees the grid, work queued later (steps, an
 * admission that reuses its memory) does not. */
int cfdart_batch_remove(cfdart_batch* b, int slot);
/* Advances every live grid by nsteps steps (grids that have finished stay unchanged). */
int cfdart_batch_step(cfdart_batch* b, int nsteps);
/* Copies a grid's current state to the host (format of cfdart.h, ghost frame filled). */
int cfdart_batch_download(cfdart_batch* b, int slot, double* host_state);
/* Loop state of a grid: t, steps taken, and whether it is still running; synchronous. */
int cfdart_batch_status(cfdart_batch* b, int slot, double* t, int* n, int* active);
/* Loop state of every slot after the work queued so far (synchronous, one small copy):
 * steps[slot] = steps taken, flags[slot] = 1 while running | 2 if an Allaire state failed the
 * physical-state test (the grid then stops); 0 for empty slots. Arrays of the batch capacity;
 * either may be NULL. */
int cfdart_batch_poll(cfdart_batch* b, int* steps, int* flags);
/* Checksum of a grid's current state, as cfdart_checksum; synchronous. */
int cfdart_batch_checksum(cfdart_batch* b, int slot, unsigned long long* out);
/* The same checksum queued on the batch stream without waiting: *host_out (pinned memory,
 * cfdart_host_alloc) holds it after the next cfdart_batch_poll or cfdart_batch_sync. The state
 * is the grid's state at that point of the stream (the last step queued before the call). */
int cfdart_batch_checksum_async(cfdart_batch* b, int slot, unsigned long long* host_out);
/* Fallback counters of a grid (see CFDART_CNT_*), 4 values; synchronous. */
int cfdart_batch_stats(cfdart_batch* b, int slot, long long* counters);
int cfdart_batch_sync(cfdart_batch* b);
/* Changes a grid's tile shape (used by the autotuner between steps); y_strip_rows as in
 * cfdart_batch_add. */
int cfdart_batch_set_tile(cfdart_batch* b, int slot, int y_strip_rows, int x_rows_per_wave);
/* Arithmetic of every grid of the batch, as cfdart_set_arith (default CFDART_ARITH_EXACT); takes
 * effect from the next step. No reference counterpart. */
int cfdart_batch_set_arith(cfdart_batch* b, int mode);
/* Live grids, device bytes counted against the memory budget (grid states and the frame
 * staging ring), tiles in the work-list (each may be NULL; asking for the tile count rebuilds a
 * work-list left stale by admissions or removals). */
int cfdart_batch_info(cfdart_batch* b, int* live, size_t* bytes, int* tiles);
/* Frame snapshot: device-to-device copy of a grid's state into a staging ring in device memory
 * on the batch stream, then an asynchronous copy to host_dst (pinned memory, at least the
 * state's interior size: ncomp * nx * ny doubles, SoA planes) on the copy stream, so stepping
 * continues while frames drain. The batch stream waits only when the ring wraps onto a frame
 * still being copied. The ring is allocated on the first call and regrown to four frames of a
 * larger request (at least 128 MiB); its bytes count against the memory budget, and a ring that
 * would exceed the budget returns CFDART_E_NOMEM. Completion: cfdart_batch_sync or
 * cfdart_batch_frames_done. */
int cfdart_batch_frame(cfdart_batch* b, int slot, double* host_dst);
/* Frames of n grids at one point of the batch stream, as n cfdart_batch_frame calls but with
 * one event between the streams for the group (the copies of a group are queued together; a
 * group that would reuse its own ring space is split). status (n entries, may be NULL) receives
 * 0 for each frame taken and CFDART_E_NOMEM for one skipped because the ring could not grow
 * within the memory budget; the call then still takes the others and returns 0. No reference
 * counterpart (the serving loop's frames of one chunk). */
int cfdart_batch_frames(cfdart_batch* b, int n, const int* slots, double* const* host_dsts, int* status);
int cfdart_batch_frames_done(cfdart_batch* b);
void* cfdart_batch_stream(cfdart_batch* b);

/* Pinned host memory for frames (hipHostMalloc), freed with cfdart_host_free. */
void* cfdart_host_alloc(size_t bytes);
void cfdart_host_free(void* p);

/* Autotuner measurement: times one sweep of each axis for a grid of this description with a
 * given tile shape, in the arithmetic 'arith' (CFDART_ARITH_EXACT or CFDART_ARITH_FAST), on
 * host_state (format of cfdart.h; NULL: a uniform state), returning the mean kernel time in
 * microseconds over 'reps' launches per axis. */
int cfdart_time_tile(const cfdart_desc* desc, const double* host_state, int y_strip_rows,
                     int x_rows_per_wave, int reps, int arith, double* us_x, double* us_y);

#ifdef __cplusplus
}
#endif
#endif
