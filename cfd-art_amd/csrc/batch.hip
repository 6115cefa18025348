// Multi-grid execution (include/cfdart_batch.h): per sweep phase, an x-sweep and a y-sweep
// kernel per model over device work-lists of tiles from every live grid, per-grid time-loop state
// on the device, and a HIP graph per step.
//
// Buffer roles: at a step boundary a grid's state is in buf[sw] (sw = 0 unless a Strang step,
// three sweeps, has flipped it); the sweeps of a step alternate between the two buffers from
// there. A grid whose loop has ended is skipped by every sweep, so its state stays put. The sweep
// axis of each phase follows the grid's splitting (alternating: x then y on even steps, y then x
// on odd; Strang: x, y, x). Readers of the state on the stream (frames, checksums) select
// buf[sw] on the device when they run.
#include "cfdart_batch.h"
#include "cfdart_host.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <list>
#include <map>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "batch_kernels.h"
#include "ic_device.h"

namespace cfd {
void set_last_error(const std::string& msg);  // cfdart_api.hip
int check_desc(const cfdart_desc* d);         // cfdart_api.hip: cfdart_create's descriptor checks
}

namespace cfd {

// End of a step for every live grid (E/headers/simulation.hpp:56-58 per grid, plus the
// step budget): one wavefront per grid slot, lane i folding wave-speed slot i (one memory round
// trip for the 64 slots instead of a serial fold in one thread).
__global__ __launch_bounds__(64) void batch_finalize_kernel(BatchGrid* grids, int* status)
{
    const int i = blockIdx.x;
    BatchGrid* bg = grids + i;
    Ctrl* c = &bg->ctrl;
    if (!bg->live || !c->active) return;  // the same in every lane
    double mu, mv;
    fold_maxima_wave(c, mu, mv);  // the whole wavefront
    if (threadIdx.x != 0) return;
    double t = c->t + c->dt;
    int n = c->n + 1;
    if (fabs(t - c->T) < 1e-12) t = c->T;
    c->t = t;
    c->n = n;
    if (bg->splitting == CFDART_STRANG) bg->sw ^= 1;  // three sweeps: the state moved to the other buffer
    bg->steps_left -= 1;
    c->active = (t < c->T) && bg->steps_left > 0 && !c->error;
    if (c->active) c->dt = dt_from_maxima(c, mu, mv, n, t, bg->dx, bg->dy);
    status[2 * i] = n;
    status[2 * i + 1] = (c->active ? 1 : 0) | (c->error ? 2 : 0);
}

// Removal of a grid from the device table (stream-ordered: sweeps queued before it still see the
// grid, whose finished loop makes its tiles exit at once).
__global__ void batch_mark_dead_kernel(BatchGrid* bg)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) bg->live = 0;
}

__global__ void batch_zero_kernel(unsigned long long* p)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0;
}

// Frame staging: a grid's interior planes from buf[sw] (read when the kernel runs) into dst,
// packed [component][y][x]; each workgroup copies whole rows, grid-stride over the rows.
__global__ __launch_bounds__(256) void batch_stage_frame_kernel(const BatchGrid* bg, double* dst)
{
    const int nx = bg->nx, ny = bg->ny, nc = bg->model == ALLAIRE ? 6 : 4;
    const long long plane = bg->plane, pitch = bg->pitch;
    const double* src = bg->buf[bg->sw & 1];
    const long long rows = (long long)ny * nc;
    for (long long r = blockIdx.x; r < rows; r += gridDim.x) {
        const long long k = r / ny, y = r - k * ny;
        const double* a = src + k * plane + y * pitch;
        double* d = dst + r * nx;
        for (int x = threadIdx.x; x < nx; x += blockDim.x) d[x] = a[x];
    }
}

__global__ void batch_set_lengths_kernel(BatchQueue* q, BatchQueue lengths)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *q = lengths;
}

// After the reduction of a newly admitted grid: its first dt and loop test.
__global__ void batch_admit_kernel(BatchGrid* bg, int* status)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    Ctrl* c = &bg->ctrl;
    double mu, mv;
    fold_maxima(c, mu, mv);
    c->active = (c->t < c->T) && bg->steps_left > 0;
    if (c->active) c->dt = dt_from_maxima(c, mu, mv, c->n, c->t, bg->dx, bg->dy);
    status[0] = c->n;
    status[1] = c->active ? 1 : 0;
}

// A work-list segment: the tiles of one grid on one axis, in the order the sweeps take them
// (x: row band by row band, column tiles inner; y: column block by column block, row blocks
// inner), written at [offset, offset + n_outer * n_inner) of its model's list for that axis.
struct TileSeg {
    int grid, model, axis, n_outer, n_inner, offset;
};
struct TileLists {
    Tile* l[2][2];  // [model][axis]
};

// Writes the tiles of every segment, one workgroup per segment: the host sends one record per
// grid and axis rather than the tiles themselves.
__global__ __launch_bounds__(256) void batch_expand_tiles_kernel(const TileSeg* segs, TileLists lists)
{
    const TileSeg sg = segs[blockIdx.x];
    Tile* const base = sg.model == 0 ? (sg.axis == 0 ? lists.l[0][0] : lists.l[0][1])
                                     : (sg.axis == 0 ? lists.l[1][0] : lists.l[1][1]);
    Tile* out = base + sg.offset;
    const int n = sg.n_outer * sg.n_inner;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int o = i / sg.n_inner, in = i - o * sg.n_inner;
        // x: (column tile, row band) = (inner, outer); y: (column block, row block) = (outer, inner)
        out[i] = sg.axis == 0 ? Tile{sg.grid, 0, in, o} : Tile{sg.grid, 1, o, in};
    }
}

}  // namespace cfd

using cfd::BatchGrid;
using cfd::BatchQueue;
using cfd::Tile;

struct AdmitStage {
    BatchGrid* rec;   // the grid record copied into the device table
    void* draws;      // the initial condition's draws (IC_DRAW_BYTES)
    hipEvent_t done;  // recorded after the entry's copies were queued
    bool pending;
};
constexpr int ADMIT_STAGES = 32;

// An event shared by the frames of one group (destroyed with its last holder).
using EventRef = std::shared_ptr<std::remove_pointer<hipEvent_t>::type>;
inline EventRef event_ref(hipEvent_t e)
{
    return EventRef(e, [](hipEvent_t x) { (void)hipEventDestroy(x); });
}

struct FrameRegion {
    size_t off, len;  // doubles
    EventRef done;    // the device-to-host copies out of the region's group have finished
};

struct cfdart_batch {
    int cap;
    size_t budget, used;
    BatchGrid* d_grids;
    std::vector<BatchGrid> h_grids;
    std::vector<char> live;
    std::vector<size_t> bytes;
    Tile* d_tiles[2][2];  // per model and axis
    size_t tile_cap[2][2];
    int ntile[2][2];
    BatchQueue* d_q;
    double* d_partials;
    int nred;
    hipStream_t st, copy_st;
    hipGraphExec_t exec[cfd::STEP_SHAPES];  // the captured step, per shape
    bool have_graph;
    int graph_models;           // bit m: the captured steps launch model m's sweeps
    long long gstep;            // batch steps launched since creation
    std::vector<int> cls;       // per slot: 0 / 1 alternating grid admitted at an even / odd step, 2 naive, 3 Strang
    int nclass[4];              // live grids per class
    int arith;  // CFDART_ARITH_EXACT or CFDART_ARITH_FAST, for every grid of the batch
    int ntiles;
    std::multimap<size_t, double*> pool;  // freed state buffers for reuse, by size
    int* d_status;           // per slot: steps taken, flags (1 running, 2 error)
    int* h_status;           // its pinned host copy (cfdart_batch_poll)
    bool tiles_dirty;        // the work-list is rebuilt before the next step
    void* d_draws;           // initial-condition draws of the admission in progress
    unsigned long long* d_sum;
    // pinned host staging of generated admissions (grid record and draws), reused cyclically once
    // the stream has passed an entry's copies: the admission returns without waiting
    std::vector<AdmitStage> stages;
    size_t stage_next;
    // the work-lists' segments (one per live grid and axis): pinned host staging, its device
    // copy, their capacity, and the event after their last upload and expansion
    cfd::TileSeg* h_segs;
    cfd::TileSeg* d_segs;
    size_t seg_cap;
    hipEvent_t tiles_ev;
    bool tiles_ev_pending;
    // frame staging ring in device memory: regions still being copied to the host
    double* d_ring;
    size_t ring_cap, ring_head;  // in doubles
    std::list<FrameRegion> inflight;
};

namespace {

int bfail(int code, const std::string& m)
{
    cfd::set_last_error(m);
    return code;
}

#define BCHK(x)                                                                                 \
    do {                                                                                        \
        hipError_t e_ = (x);                                                                    \
        if (e_ != hipSuccess) return bfail(CFDART_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

size_t round_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

double* pool_get(cfdart_batch* b, size_t bytes)
{
    auto it = b->pool.find(bytes);
    if (it != b->pool.end()) {
        double* p = it->second;
        b->pool.erase(it);
        return p;
    }
    double* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        // release cached buffers and retry once
        for (auto& kv : b->pool) (void)hipFree(kv.second);
        b->pool.clear();
        if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    }
    return p;
}

// The captured steps (they hold the work-lists' addresses and launch sizes).
void destroy_graphs(cfdart_batch* b)
{
    if (!b->have_graph) return;
    for (int k = 0; k < cfd::STEP_SHAPES; ++k) (void)hipGraphExecDestroy(b->exec[k]);
    b->have_graph = false;
}

// The work-lists of the live grids, per model and axis. x tiles go largest first (longest
// processing time first: decreasing rows per wavefront, grids in slot order within a shape), which
// shortens the tail of the launch, each grid's row bands in the outer loop so that tiles sharing
// halo columns are consecutive; y tiles are all alike and go in slot order, each grid's column
// blocks in the outer loop so that tiles sharing halo rows are consecutive. A list of alike items
// goes to the XCDs in contiguous runs (batch_item). The host lays out one segment per grid and
// axis (no comparison sort: a batch rebuilds its lists at every admission and removal) and a
// kernel writes the tiles, so a rebuild costs the host O(grids), not O(tiles).
int rebuild_tiles(cfdart_batch* b)
{
    std::vector<int> shapes[2];  // distinct x rows per wavefront, per model
    size_t nseg = 0;
    for (int s = 0; s < b->cap; ++s) {
        if (!b->live[s]) continue;
        const BatchGrid& g = b->h_grids[s];
        const int m = g.model == cfd::ALLAIRE ? 1 : 0;
        if (std::find(shapes[m].begin(), shapes[m].end(), g.xrows) == shapes[m].end()) shapes[m].push_back(g.xrows);
        nseg += 2;
    }
    // the pinned staging of the previous upload is rewritten only once that upload (and the
    // expansion that read its device copy) has finished
    if (b->tiles_ev_pending) {
        BCHK(hipEventSynchronize(b->tiles_ev));
        b->tiles_ev_pending = false;
    }
    if (nseg > b->seg_cap) {
        if (b->h_segs) BCHK(hipHostFree(b->h_segs));
        if (b->d_segs) BCHK(hipFree(b->d_segs));
        b->h_segs = nullptr;
        b->d_segs = nullptr;
        b->seg_cap = 0;
        const size_t cap = std::max<size_t>(nseg, 256);
        BCHK(hipHostMalloc((void**)&b->h_segs, sizeof(cfd::TileSeg) * cap, hipHostMallocDefault));
        BCHK(hipMalloc(&b->d_segs, sizeof(cfd::TileSeg) * cap));
        b->seg_cap = cap;
    }
    size_t len[2][2] = {{0, 0}, {0, 0}};
    size_t k = 0;
    for (int s = 0; s < b->cap; ++s) {
        if (!b->live[s]) continue;
        const BatchGrid& g = b->h_grids[s];
        const int m = g.model == cfd::ALLAIRE ? 1 : 0;
        const int nbx = (g.nx + cfd::BATCH_YT_COLS - 1) / cfd::BATCH_YT_COLS, nby = (g.ny + cfd::YT_ROWS - 1) / cfd::YT_ROWS;
        b->h_segs[k++] = cfd::TileSeg{s, m, 1, nbx, nby, (int)len[m][1]};
        len[m][1] += (size_t)nbx * nby;
    }
    for (int m = 0; m < 2; ++m) {
        std::sort(shapes[m].begin(), shapes[m].end(), [](int x, int y) { return x > y; });
        for (int xr : shapes[m])
            for (int s = 0; s < b->cap; ++s) {
                if (!b->live[s]) continue;
                const BatchGrid& g = b->h_grids[s];
                if ((g.model == cfd::ALLAIRE ? 1 : 0) != m || g.xrows != xr) continue;
                const int ntx = (g.nx + cfd::XTILE - 1) / cfd::XTILE, nry = (g.ny + 4 * g.xrows - 1) / (4 * g.xrows);
                b->h_segs[k++] = cfd::TileSeg{s, m, 0, nry, ntx, (int)len[m][0]};
                len[m][0] += (size_t)nry * ntx;
            }
    }
    if (k != nseg) return bfail(CFDART_E_HIP, "work-list segments miscounted");
    for (int m = 0; m < 2; ++m)
        for (int ax = 0; ax < 2; ++ax)
            if (len[m][ax] > (size_t)INT32_MAX) return bfail(CFDART_E_NOMEM, "work-list longer than 2^31 tiles");
    b->ntiles = 0;
    BatchQueue lengths{};
    cfd::TileLists lists;
    for (int m = 0; m < 2; ++m)
        for (int ax = 0; ax < 2; ++ax) {
            if (len[m][ax] > b->tile_cap[m][ax]) {
                BCHK(hipStreamSynchronize(b->st));
                if (b->d_tiles[m][ax]) BCHK(hipFree(b->d_tiles[m][ax]));
                b->d_tiles[m][ax] = nullptr;
                b->tile_cap[m][ax] = std::max(len[m][ax], 2 * b->tile_cap[m][ax]);
                BCHK(hipMalloc(&b->d_tiles[m][ax], sizeof(Tile) * b->tile_cap[m][ax]));
                destroy_graphs(b);  // re-captured with the new list and launch size
            }
            lists.l[m][ax] = b->d_tiles[m][ax];
            b->ntile[m][ax] = (int)len[m][ax];
            lengths.ntiles[m][ax] = (int)len[m][ax];
            lengths.xcd[m][ax] = ax == 1 || shapes[m].size() <= 1;  // y tiles are all alike
            b->ntiles += (int)len[m][ax];
        }
    if (nseg > 0) {
        BCHK(hipMemcpyAsync(b->d_segs, b->h_segs, sizeof(cfd::TileSeg) * nseg, hipMemcpyHostToDevice, b->st));
        hipLaunchKernelGGL(cfd::batch_expand_tiles_kernel, dim3((unsigned)nseg), dim3(256), 0, b->st, b->d_segs, lists);
        BCHK(hipGetLastError());
    }
    hipLaunchKernelGGL(cfd::batch_set_lengths_kernel, dim3(1), dim3(64), 0, b->st, b->d_q, lengths);
    BCHK(hipGetLastError());
    BCHK(hipEventRecord(b->tiles_ev, b->st));
    b->tiles_ev_pending = true;
    return 0;
}

// One step = both sweep phases of the Euler kernels, then of the Allaire kernels, then the
// per-grid loop bookkeeping, on the batch stream; captured once per step shape (cfd::StepShape).
// The two models' launches are not run concurrently on two streams: sharing the CUs, the batched
// workload's step took 3 % longer than with the launches one after another.
int models_present(const cfdart_batch* b)
{
    return (b->ntile[0][0] + b->ntile[0][1] > 0 ? 1 : 0) | (b->ntile[1][0] + b->ntile[1][1] > 0 ? 2 : 0);
}

int capture_graph(cfdart_batch* b)
{
    destroy_graphs(b);
    b->graph_models = models_present(b);
    for (int m = 0; m < 2; ++m)
        for (int ax = 0; ax < 2; ++ax)
            if (!b->d_tiles[m][ax]) {
                b->tile_cap[m][ax] = 1;
                BCHK(hipMalloc(&b->d_tiles[m][ax], sizeof(Tile)));
            }
    using cfd::ALLAIRE;
    using cfd::EULER;
    for (int k = 0; k < cfd::STEP_SHAPES; ++k) {
        hipGraph_t graph;
        BCHK(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
        // a model with no grid in the batch gets no launches (a grid of it joining re-captures)
        const bool euler = (b->graph_models & 1) != 0, allaire = (b->graph_models & 2) != 0;
        if (b->arith == CFDART_ARITH_FAST) {
            if (euler) cfd::launch_batch_phases_fast(EULER, b->d_grids, b->d_tiles[EULER], b->tile_cap[EULER], b->d_q, b->st, k);
            if (allaire)
                cfd::launch_batch_phases_fast(ALLAIRE, b->d_grids, b->d_tiles[ALLAIRE], b->tile_cap[ALLAIRE], b->d_q, b->st, k);
        } else {
            if (euler) cfd::launch_batch_phases<EULER, false>(b->d_grids, b->d_tiles[EULER], b->tile_cap[EULER], b->d_q, b->st, k);
            if (allaire)
                cfd::launch_batch_phases<ALLAIRE, false>(b->d_grids, b->d_tiles[ALLAIRE], b->tile_cap[ALLAIRE], b->d_q, b->st, k);
        }
        hipLaunchKernelGGL(cfd::batch_finalize_kernel, dim3(b->cap), dim3(64), 0, b->st, b->d_grids, b->d_status);
        BCHK(hipStreamEndCapture(b->st, &graph));
        hipError_t e = hipGraphInstantiate(&b->exec[k], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            for (int j = 0; j < k; ++j) (void)hipGraphExecDestroy(b->exec[j]);
            return bfail(CFDART_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
    }
    b->have_graph = true;
    return 0;
}

// The shape of batch step number s (counted from the batch's creation): an alternating grid
// admitted at step g has taken s - g steps, so it sweeps x first when s + g is even; a naive grid
// always sweeps x first. When every admitted grid sweeps the same axis first, the step launches
// only the sweeps that have work (a grid that has finished exits at once whatever the shape).
int step_shape(const cfdart_batch* b, long long s)
{
    if (b->nclass[3] > 0) return cfd::STEP_XYX;  // a Strang grid: x, y, x (the others take two of the three)
    const bool x_first = b->nclass[2] > 0 || (b->nclass[0] > 0 && (s & 1) == 0) || (b->nclass[1] > 0 && (s & 1) == 1);
    const bool y_first = (b->nclass[0] > 0 && (s & 1) == 1) || (b->nclass[1] > 0 && (s & 1) == 0);
    if (x_first && y_first) return cfd::STEP_BOTH;
    return y_first ? cfd::STEP_YX : cfd::STEP_XY;
}

int check_slot(cfdart_batch* b, int slot)
{
    if (!b || slot < 0 || slot >= b->cap || !b->live[slot]) return bfail(CFDART_E_ARG, "bad batch slot");
    return 0;
}

}  // namespace

extern "C" {

int cfdart_batch_create(int capacity, size_t mem_budget_bytes, cfdart_batch** out)
{
    if (!out || capacity < 1 || capacity > (1 << 20)) return bfail(CFDART_E_ARG, "bad capacity");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return bfail(CFDART_E_NODEV, "no HIP device");
    cfdart_batch* b = new cfdart_batch();
    b->cap = capacity;
    b->budget = mem_budget_bytes;
    b->used = 0;
    b->h_grids.assign(capacity, BatchGrid{});
    b->live.assign(capacity, 0);
    b->bytes.assign(capacity, 0);
    for (int m = 0; m < 2; ++m)
        for (int ax = 0; ax < 2; ++ax) {
            b->d_tiles[m][ax] = nullptr;
            b->tile_cap[m][ax] = 0;
            b->ntile[m][ax] = 0;
        }
    b->have_graph = false;
    b->graph_models = 0;
    b->gstep = 0;
    b->cls.assign(capacity, 0);
    for (int k = 0; k < 4; ++k) b->nclass[k] = 0;
    b->arith = CFDART_ARITH_EXACT;
    b->ntiles = 0;
    b->d_status = nullptr;
    b->h_status = nullptr;
    b->tiles_dirty = false;
    b->d_draws = nullptr;
    b->d_sum = nullptr;
    b->d_ring = nullptr;
    b->ring_cap = b->ring_head = 0;
    b->stage_next = 0;
    b->h_segs = nullptr;
    b->d_segs = nullptr;
    b->seg_cap = 0;
    b->tiles_ev_pending = false;
    b->tiles_ev = nullptr;
    if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&b->copy_st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->tiles_ev, hipEventDisableTiming) != hipSuccess) {
        delete b;
        return bfail(CFDART_E_HIP, "stream creation failed");
    }
    for (int i = 0; i < ADMIT_STAGES; ++i) {
        AdmitStage a{nullptr, nullptr, nullptr, false};
        void* p = nullptr;
        if (hipHostMalloc(&p, sizeof(BatchGrid) + cfd::IC_DRAW_BYTES, hipHostMallocDefault) != hipSuccess) break;
        if (hipEventCreateWithFlags(&a.done, hipEventDisableTiming) != hipSuccess) {
            (void)hipHostFree(p);
            break;
        }
        a.rec = (BatchGrid*)p;
        a.draws = (char*)p + sizeof(BatchGrid);
        b->stages.push_back(a);
    }
    if (hipHostMalloc((void**)&b->h_status, sizeof(int) * 2 * (size_t)capacity, hipHostMallocDefault) != hipSuccess)
        b->h_status = nullptr;
    if (b->stages.size() != (size_t)ADMIT_STAGES || !b->h_status) {
        cfdart_batch_destroy(b);
        return bfail(CFDART_E_NOMEM, "pinned host staging");
    }
    if (hipMalloc(&b->d_grids, sizeof(BatchGrid) * capacity) != hipSuccess ||
        hipMalloc(&b->d_q, sizeof(BatchQueue)) != hipSuccess ||
        hipMalloc(&b->d_partials, sizeof(double) * 2 * 1024) != hipSuccess ||
        hipMalloc(&b->d_status, sizeof(int) * 2 * capacity) != hipSuccess ||
        hipMalloc(&b->d_draws, cfd::IC_DRAW_BYTES) != hipSuccess ||
        hipMalloc(&b->d_sum, sizeof(unsigned long long)) != hipSuccess) {
        cfdart_batch_destroy(b);
        return bfail(CFDART_E_NOMEM, "batch tables");
    }
    (void)hipMemset(b->d_grids, 0, sizeof(BatchGrid) * capacity);
    (void)hipMemset(b->d_status, 0, sizeof(int) * 2 * capacity);
    (void)hipMemset(b->d_q, 0, sizeof(BatchQueue));
    if (hipDeviceSynchronize() != hipSuccess) {
        cfdart_batch_destroy(b);
        return bfail(CFDART_E_HIP, "device initialisation failed");
    }
    *out = b;
    return 0;
}

void cfdart_batch_destroy(cfdart_batch* b)
{
    if (!b) return;
    (void)hipStreamSynchronize(b->st);
    (void)hipStreamSynchronize(b->copy_st);
    for (int s = 0; s < b->cap; ++s)
        if (b->live[s]) {
            (void)hipFree(b->h_grids[s].buf[0]);
            (void)hipFree(b->h_grids[s].buf[1]);
        }
    for (auto& kv : b->pool) (void)hipFree(kv.second);
    destroy_graphs(b);
    for (int m = 0; m < 2; ++m)
        for (int ax = 0; ax < 2; ++ax)
            if (b->d_tiles[m][ax]) (void)hipFree(b->d_tiles[m][ax]);
    if (b->d_grids) (void)hipFree(b->d_grids);
    if (b->d_q) (void)hipFree(b->d_q);
    if (b->d_partials) (void)hipFree(b->d_partials);
    if (b->d_status) (void)hipFree(b->d_status);
    if (b->d_draws) (void)hipFree(b->d_draws);
    if (b->d_sum) (void)hipFree(b->d_sum);
    if (b->d_ring) (void)hipFree(b->d_ring);
    b->inflight.clear();
    for (AdmitStage& a : b->stages) {
        (void)hipEventDestroy(a.done);
        (void)hipHostFree(a.rec);
    }
    if (b->h_status) (void)hipHostFree(b->h_status);
    if (b->h_segs) (void)hipHostFree(b->h_segs);
    if (b->d_segs) (void)hipFree(b->d_segs);
    if (b->tiles_ev) (void)hipEventDestroy(b->tiles_ev);
    (void)hipStreamDestroy(b->st);
    (void)hipStreamDestroy(b->copy_st);
    delete b;
}

}  // extern "C"

namespace {

// Admission shared by the host-state and the generated-state entry points: budget and slot
// checks, buffers, the initial state (from host_state, or generated on the device from dp / bp),
// then compute_dt's reduction and the grid's first dt. The work-list is rebuilt lazily.
int admit(cfdart_batch* b, const cfdart_desc* d, const double* host, const cfdart_disc_params* dp,
          const cfdart_bubble_params* bp, int splitting, int max_steps, double T, double cfl, int yH, int xrows,
          int* slot_out)
{
    if (!b || !d || !slot_out) return bfail(CFDART_E_ARG, "null argument");
    if (splitting != CFDART_NAIVE && splitting != CFDART_ALTERNATING && splitting != CFDART_STRANG)
        return bfail(CFDART_E_ARG, "bad splitting scheme");
    if (max_steps < 1 || !(cfl > 0.0)) return bfail(CFDART_E_ARG, "need max_steps >= 1 and cfl > 0");
    // every check of cfdart_create, then the batch's own limits (MUSCL-Hancock, ghost width 2)
    const int rc = cfd::check_desc(d);
    if (rc) return rc;
    if (d->flux_solver != CFDART_MUSCL_HANCOCK) return bfail(CFDART_E_ARG, "batched grids use MUSCL-Hancock");
    // the y sweep of a batch runs the 16-column transposed tile
    if ((yH != 0 && yH != CFDART_Y_TRANSPOSED) || xrows < 0 || xrows > 256)
        return bfail(CFDART_E_ARG, "tile shape out of range (batched grids: y_strip_rows 0 or CFDART_Y_TRANSPOSED, "
                                   "x_rows_per_wave in [0, 256])");
    yH = CFDART_Y_TRANSPOSED;
    if (xrows == 0) xrows = 8;
    const int nc = d->model == CFDART_ALLAIRE ? 6 : 4;
    const int pitch = (d->nx + 31) / 32 * 32;
    const long long plane = (long long)pitch * d->ny;
    const size_t one = round_up(sizeof(double) * (size_t)plane * nc, 2u << 20);
    const size_t need = 2 * one;
    if (b->budget && b->used + need > b->budget) return bfail(CFDART_E_NOMEM, "admission refused: memory budget");
    int slot = -1;
    for (int s = 0; s < b->cap; ++s)
        if (!b->live[s]) { slot = s; break; }
    if (slot < 0) return bfail(CFDART_E_NOMEM, "admission refused: batch is full");

    BatchGrid g{};
    g.buf[0] = pool_get(b, one);
    g.buf[1] = g.buf[0] ? pool_get(b, one) : nullptr;
    if (!g.buf[0] || !g.buf[1]) {
        if (g.buf[0]) b->pool.insert({one, g.buf[0]});
        return bfail(CFDART_E_NOMEM, "hipMalloc of the grid state failed");
    }
    g.plane = plane;
    g.pitch = pitch;
    g.nx = d->nx;
    g.ny = d->ny;
    g.model = d->model;
    g.splitting = splitting;
    for (int i = 0; i < 4; ++i) g.bc[i] = d->bc[i];
    g.mat = cfd::make_material(d->gamma1, d->pinf1, d->gamma2, d->pinf2);
    g.opt.limiter = d->limiter;
    g.opt.wave_speeds = d->wave_speeds;
    g.opt.z_lo = d->z_lo;
    g.opt.z_hi = d->z_hi;
    g.opt.godunov = 0;
    g.opt.fast = 0;
    g.dx = d->dx;
    g.dy = d->dy;
    g.dx_half_col = d->colsweep_uses_dx ? d->dx : d->dy;
    g.yH = yH;
    g.xrows = xrows;
    g.steps_left = max_steps;
    g.live = 1;
    g.sw = 0;
    std::memset(&g.ctrl, 0, sizeof(g.ctrl));
    g.ctrl.T = T;
    g.ctrl.cfl = cfl;

    auto release = [&]() {
        (void)hipStreamSynchronize(b->st);
        b->pool.insert({one, g.buf[0]});
        b->pool.insert({one, g.buf[1]});
    };
    // a generated state is admitted without waiting: its draws and grid record go through a
    // pinned staging entry, reused once the stream has passed that entry's copies
    AdmitStage* stage = nullptr;
    if (!host) {
        stage = &b->stages[b->stage_next];
        b->stage_next = (b->stage_next + 1) % b->stages.size();
        if (stage->pending) {
            hipError_t e = hipEventSynchronize(stage->done);
            if (e != hipSuccess) {
                release();
                return bfail(CFDART_E_HIP, std::string("admission staging: ") + hipGetErrorString(e));
            }
            stage->pending = false;
        }
    }
    std::vector<double> soa;
    if (host) {
        // interior of the host grid -> SoA buf[0]
        soa.assign((size_t)plane * nc, 0.0);
        const int W = d->nx + 2 * d->ng;
        for (int k = 0; k < nc; ++k)
            for (int y = 0; y < d->ny; ++y)
                for (int x = 0; x < d->nx; ++x)
                    soa[(size_t)k * plane + (size_t)y * pitch + x] = host[((size_t)(y + d->ng) * W + (x + d->ng)) * nc + k];
        hipError_t e = hipMemcpyAsync(g.buf[0], soa.data(), sizeof(double) * soa.size(), hipMemcpyHostToDevice, b->st);
        if (e != hipSuccess) {
            release();
            return bfail(CFDART_E_HIP, std::string("state upload: ") + hipGetErrorString(e));
        }
    } else {
        std::string err;
        int rc = cfd::generate_state(*d, dp, bp, g.buf[0], pitch, plane, b->d_draws, b->st, err, stage->draws);
        if (rc) {
            release();
            return bfail(rc, err);
        }
    }
    const BatchGrid* rec = &g;
    if (stage) {
        *stage->rec = g;
        rec = stage->rec;
    }
    hipError_t e = hipMemcpyAsync(b->d_grids + slot, rec, sizeof(BatchGrid), hipMemcpyHostToDevice, b->st);
    if (e != hipSuccess) {
        release();
        return bfail(CFDART_E_HIP, std::string("grid table upload: ") + hipGetErrorString(e));
    }
    // first dt: compute_dt's reduction on the new state into the grid's control block
    cfd::GridView v;
    v.cur = g.buf[0];
    v.nxt = g.buf[1];
    v.plane = plane;
    v.pitch = pitch;
    v.nx = g.nx;
    v.ny = g.ny;
    for (int i = 0; i < 4; ++i) v.bc[i] = g.bc[i];
    v.mat = g.mat;
    v.opt = g.opt;
    v.opt.fast = b->arith == CFDART_ARITH_FAST ? 1 : 0;  // the first dt in the batch's arithmetic
    v.dx = g.dx;
    v.dy = g.dy;
    v.dx_half_col = g.dx_half_col;
    v.ctrl = &(b->d_grids + slot)->ctrl;
    e = cfd::launch_reduce_dt(v, g.model, b->d_partials, b->st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(cfd::batch_admit_kernel, dim3(1), dim3(64), 0, b->st, b->d_grids + slot,
                           b->d_status + 2 * slot);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        if (stage) {
            e = hipEventRecord(stage->done, b->st);
            stage->pending = e == hipSuccess;
        } else {
            e = hipStreamSynchronize(b->st);  // the staging vector and g go out of scope
        }
    }
    if (e != hipSuccess) {
        release();
        return bfail(CFDART_E_HIP, std::string("admission kernels: ") + hipGetErrorString(e));
    }
    b->h_grids[slot] = g;
    b->live[slot] = 1;
    b->cls[slot] = splitting == CFDART_ALTERNATING ? (int)(b->gstep & 1) : (splitting == CFDART_STRANG ? 3 : 2);
    b->nclass[b->cls[slot]] += 1;
    b->bytes[slot] = need;
    b->used += need;
    b->tiles_dirty = true;
    *slot_out = slot;
    return 0;
}

}  // namespace

extern "C" {

int cfdart_batch_add(cfdart_batch* b, const cfdart_desc* d, const double* host, int splitting, int max_steps,
                     double T, double cfl, int yH, int xrows, int* slot_out)
{
    if (!host) return bfail(CFDART_E_ARG, "null host state");
    return admit(b, d, host, nullptr, nullptr, splitting, max_steps, T, cfl, yH, xrows, slot_out);
}

int cfdart_batch_add_generated(cfdart_batch* b, const cfdart_desc* d, const cfdart_disc_params* discs,
                               const cfdart_bubble_params* bubbles, int splitting, int max_steps, double T,
                               double cfl, int yH, int xrows, int* slot_out)
{
    return admit(b, d, nullptr, discs, bubbles, splitting, max_steps, T, cfl, yH, xrows, slot_out);
}

int cfdart_batch_remove(cfdart_batch* b, int slot)
{
    int rc = check_slot(b, slot);
    if (rc) return rc;
    BatchGrid& g = b->h_grids[slot];
    g.live = 0;
    // marked dead on the device in stream order: work queued later (an admission reusing the
    // buffers, the next steps) sees it dead, work queued earlier still sees its state
    hipLaunchKernelGGL(cfd::batch_mark_dead_kernel, dim3(1), dim3(64), 0, b->st, b->d_grids + slot);
    BCHK(hipGetLastError());
    b->live[slot] = 0;
    b->nclass[b->cls[slot]] -= 1;
    b->tiles_dirty = true;
    size_t one = b->bytes[slot] / 2;
    b->pool.insert({one, g.buf[0]});
    b->pool.insert({one, g.buf[1]});
    b->used -= b->bytes[slot];
    b->bytes[slot] = 0;
    return 0;
}

int cfdart_batch_step(cfdart_batch* b, int nsteps)
{
    if (!b || nsteps < 0) return bfail(CFDART_E_ARG, "bad argument");
    if (b->tiles_dirty) {
        int rc = rebuild_tiles(b);
        if (rc) return rc;
        b->tiles_dirty = false;
    }
    if (b->ntiles == 0) return 0;
    if (b->have_graph && b->graph_models != models_present(b)) {  // a model joined or left the batch
        BCHK(hipStreamSynchronize(b->st));
        destroy_graphs(b);
    }
    if (!b->have_graph) {
        int rc = capture_graph(b);
        if (rc) return rc;
    }
    for (int i = 0; i < nsteps; ++i) {
        BCHK(hipGraphLaunch(b->exec[step_shape(b, b->gstep)], b->st));
        ++b->gstep;
    }
    return 0;
}

int cfdart_batch_sync(cfdart_batch* b)
{
    if (!b) return bfail(CFDART_E_ARG, "null batch");
    BCHK(hipStreamSynchronize(b->st));
    BCHK(hipStreamSynchronize(b->copy_st));
    return 0;
}

int cfdart_batch_status(cfdart_batch* b, int slot, double* t, int* n, int* active)
{
    int rc = check_slot(b, slot);
    if (rc) return rc;
    cfd::Ctrl c;
    BCHK(hipMemcpyAsync(&c, &(b->d_grids + slot)->ctrl, sizeof(c), hipMemcpyDeviceToHost, b->st));
    BCHK(hipStreamSynchronize(b->st));
    if (t) *t = c.t;
    if (n) *n = c.n;
    if (active) *active = c.active;
    if (c.error) return bfail(CFDART_E_UNPHYS, "an Allaire state failed the physical-state test");
    return 0;
}

int cfdart_batch_stats(cfdart_batch* b, int slot, long long* counters)
{
    int rc = check_slot(b, slot);
    if (rc) return rc;
    if (!counters) return bfail(CFDART_E_ARG, "null counters");
    cfd::Ctrl c;
    BCHK(hipMemcpyAsync(&c, &(b->d_grids + slot)->ctrl, sizeof(c), hipMemcpyDeviceToHost, b->st));
    BCHK(hipStreamSynchronize(b->st));
    for (int i = 0; i < 4; ++i) counters[i] = (long long)c.cnt[i];
    return 0;
}

int cfdart_batch_download(cfdart_batch* b, int slot, double* host)
{
    int rc = check_slot(b, slot);
    if (rc) return rc;
    if (!host) return bfail(CFDART_E_ARG, "null host buffer");
    const BatchGrid& g = b->h_grids[slot];
    const int nc = g.model == cfd::ALLAIRE ? 6 : 4, ng = 2, nx = g.nx, ny = g.ny, W = nx + 2 * ng;
    std::vector<double> soa((size_t)g.plane * nc);
    BCHK(hipStreamSynchronize(b->st));
    int sw = 0;  // the buffer holding the state (flipped by Strang steps)
    BCHK(hipMemcpy(&sw, &b->d_grids[slot].sw, sizeof(int), hipMemcpyDeviceToHost));
    BCHK(hipMemcpy(soa.data(), g.buf[sw & 1], sizeof(double) * soa.size(), hipMemcpyDeviceToHost));
    auto at = [&](int i, int j, int k) -> double& { return host[((size_t)i * W + j) * nc + k]; };
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x)
            for (int k = 0; k < nc; ++k) at(y + ng, x + ng, k) = soa[(size_t)k * g.plane + (size_t)y * g.pitch + x];
    const int mx = g.model == cfd::ALLAIRE ? 2 : 1, my = mx + 1;
    for (int i = ng; i < ny + ng; ++i)
        for (int j = 0; j < ng; ++j)
            for (int k = 0; k < nc; ++k) {
                at(i, j, k) = at(i, 2 * ng - j - 1, k) * ((k == mx && g.bc[0]) ? -1.0 : 1.0);
                at(i, j + nx + ng, k) = at(i, nx + ng - 1 - j, k) * ((k == mx && g.bc[1]) ? -1.0 : 1.0);
            }
    for (int j = ng; j < nx + ng; ++j)
        for (int i = 0; i < ng; ++i)
            for (int k = 0; k < nc; ++k) {
                at(i, j, k) = at(2 * ng - i - 1, j, k) * ((k == my && g.bc[2]) ? -1.0 : 1.0);
                at(i + ny + ng, j, k) = at(ny + ng - 1 - i, j, k) * ((k == my && g.bc[3]) ? -1.0 : 1.0);
            }
    return 0;
}

int cfdart_batch_set_tile(cfdart_batch* b, int slot, int yH, int xrows)
{
    int rc = check_slot(b, slot);
    if (rc) return rc;
    if ((yH != 0 && yH != CFDART_Y_TRANSPOSED) || xrows < 1 || xrows > 256)
        return bfail(CFDART_E_ARG, "tile shape out of range (batched grids: y_strip_rows 0 or CFDART_Y_TRANSPOSED)");
    yH = CFDART_Y_TRANSPOSED;
    BatchGrid& g = b->h_grids[slot];
    g.yH = yH;
    g.xrows = xrows;
    // only the two tile-shape fields change on the device
    BCHK(hipMemcpyAsync(&(b->d_grids + slot)->yH, &g.yH, 2 * sizeof(int), hipMemcpyHostToDevice, b->st));
    BCHK(hipStreamSynchronize(b->st));
    b->tiles_dirty = true;
    return 0;
}

int cfdart_batch_set_arith(cfdart_batch* b, int mode)
{
    if (!b || (mode != CFDART_ARITH_EXACT && mode != CFDART_ARITH_FAST)) return bfail(CFDART_E_ARG, "bad batch or arithmetic mode");
    if (mode == b->arith) return 0;
    b->arith = mode;
    if (b->have_graph) {  // the captured steps hold the other mode's kernels
        BCHK(hipStreamSynchronize(b->st));
        destroy_graphs(b);
    }
    return 0;
}

int cfdart_batch_info(cfdart_batch* b, int* live, size_t* bytes, int* tiles)
{
    if (!b) return bfail(CFDART_E_ARG, "null batch");
    if (tiles && b->tiles_dirty) {  // the tile count needs the work-list of the live grids
        int rc = rebuild_tiles(b);
        if (rc) return rc;
        b->tiles_dirty = false;
    }
    int n = 0;
    for (int s = 0; s < b->cap; ++s) n += b->live[s];
    if (live) *live = n;
    if (bytes) *bytes = b->used;
    if (tiles) *tiles = b->ntiles;
    return 0;
}

int cfdart_batch_frames(cfdart_batch* b, int n, const int* slots, double* const* host_dsts, int* status)
{
    if (!b || n < 0 || (n > 0 && (!slots || !host_dsts))) return bfail(CFDART_E_ARG, "bad frame request");
    for (int i = 0; i < n; ++i) {
        int rc = check_slot(b, slots[i]);
        if (rc) return rc;
        if (!host_dsts[i]) return bfail(CFDART_E_ARG, "null frame buffer");
    }
    // frames staged in the ring whose copies to the host are not queued yet
    struct Pending {
        size_t off, len;
        double* dst;
    };
    std::vector<Pending> pend;
    // queues the copies of the staged frames to the host behind one event on the batch stream,
    // and one event after them on the copy stream, shared by their ring regions
    auto flush = [&]() -> int {
        if (pend.empty()) return 0;
        hipEvent_t ready, done;
        BCHK(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
        EventRef hold_ready = event_ref(ready);
        BCHK(hipEventRecord(ready, b->st));
        BCHK(hipStreamWaitEvent(b->copy_st, ready, 0));
        for (const Pending& p : pend)
            BCHK(hipMemcpyAsync(p.dst, b->d_ring + p.off, sizeof(double) * p.len, hipMemcpyDeviceToHost, b->copy_st));
        BCHK(hipEventCreateWithFlags(&done, hipEventDisableTiming));
        EventRef hold_done = event_ref(done);
        BCHK(hipEventRecord(done, b->copy_st));
        for (const Pending& p : pend) b->inflight.push_back(FrameRegion{p.off, p.len, hold_done});
        pend.clear();
        return 0;
    };
    for (int i = 0; i < n; ++i) {
        if (status) status[i] = 0;
        const BatchGrid& g = b->h_grids[slots[i]];
        const int nc = g.model == cfd::ALLAIRE ? 6 : 4;
        const size_t len = (size_t)g.nx * g.ny * nc;
        if (len > b->ring_cap / 2) {
            // grow the ring to hold four frames of this size (at least 16 Mi doubles, 128 MiB); its
            // bytes count against the batch's memory budget like the grids' states
            const size_t cap = std::max<size_t>(4 * len, size_t(1) << 24);
            const size_t used_without = b->used - sizeof(double) * b->ring_cap;
            if (b->budget && used_without + sizeof(double) * cap > b->budget) {
                if (status) status[i] = CFDART_E_NOMEM;
                if (n == 1) return bfail(CFDART_E_NOMEM, "frame staging ring exceeds the memory budget");
                continue;  // this frame is skipped; the others are taken
            }
            int rc = flush();
            if (rc) return rc;
            BCHK(hipStreamSynchronize(b->st));
            BCHK(hipStreamSynchronize(b->copy_st));
            b->inflight.clear();
            if (b->d_ring) BCHK(hipFree(b->d_ring));
            b->d_ring = nullptr;
            b->used = used_without;
            b->ring_cap = 0;
            b->ring_head = 0;
            if (hipMalloc(&b->d_ring, sizeof(double) * cap) != hipSuccess) return bfail(CFDART_E_NOMEM, "frame staging ring");
            b->ring_cap = cap;
            b->used += sizeof(double) * cap;
        }
        size_t off = b->ring_head;
        if (off + len > b->ring_cap) off = 0;
        // a region staged earlier in this call has its copy queued before the region is reused
        for (const Pending& p : pend)
            if (p.off < off + len && off < p.off + p.len) {
                int rc = flush();
                if (rc) return rc;
                break;
            }
        // the region must not hold a frame whose copy to the host is still running
        for (auto it = b->inflight.begin(); it != b->inflight.end();) {
            if (it->off < off + len && off < it->off + it->len) {
                BCHK(hipStreamWaitEvent(b->st, it->done.get(), 0));
                it = b->inflight.erase(it);
            } else {
                ++it;
            }
        }
        // stage the interior planes (contiguous, pitch removed) on the batch stream, from the
        // buffer that holds the grid's state when the copy runs
        const long long rows = (long long)g.ny * nc;
        hipLaunchKernelGGL(cfd::batch_stage_frame_kernel, dim3((unsigned)std::min<long long>(rows, 4096)), dim3(256), 0,
                           b->st, b->d_grids + slots[i], b->d_ring + off);
        BCHK(hipGetLastError());
        pend.push_back(Pending{off, len, host_dsts[i]});
        b->ring_head = off + len;
    }
    return flush();
}

int cfdart_batch_frame(cfdart_batch* b, int slot, double* host_dst)
{
    return cfdart_batch_frames(b, 1, &slot, &host_dst, nullptr);
}

int cfdart_batch_frames_done(cfdart_batch* b)
{
    if (!b) return bfail(CFDART_E_ARG, "null batch");
    BCHK(hipStreamSynchronize(b->copy_st));
    b->inflight.clear();
    return 0;
}

int cfdart_batch_poll(cfdart_batch* b, int* steps, int* flags)
{
    if (!b) return bfail(CFDART_E_ARG, "null batch");
    // only the slots up to the last live one are copied (into pinned memory)
    int hi = 0;
    for (int s = 0; s < b->cap; ++s)
        if (b->live[s]) hi = s + 1;
    if (hi > 0) BCHK(hipMemcpyAsync(b->h_status, b->d_status, sizeof(int) * 2 * hi, hipMemcpyDeviceToHost, b->st));
    BCHK(hipStreamSynchronize(b->st));
    for (int s = 0; s < b->cap; ++s) {
        const bool on = s < hi && b->live[s];
        if (steps) steps[s] = on ? b->h_status[2 * s] : 0;
        if (flags) flags[s] = on ? b->h_status[2 * s + 1] : 0;
    }
    return 0;
}

int cfdart_batch_checksum(cfdart_batch* b, int slot, unsigned long long* out)
{
    int rc = check_slot(b, slot);
    if (rc) return rc;
    if (!out) return bfail(CFDART_E_ARG, "null output");
    const BatchGrid& g = b->h_grids[slot];
    BCHK(hipMemsetAsync(b->d_sum, 0, sizeof(unsigned long long), b->st));
    BCHK(cfd::launch_checksum_sel(b->d_grids[slot].buf, &b->d_grids[slot].sw, g.pitch, g.plane, g.nx, g.ny,
                                  g.model == cfd::ALLAIRE ? 6 : 4, b->d_sum, b->st));
    BCHK(hipMemcpyAsync(out, b->d_sum, sizeof(unsigned long long), hipMemcpyDeviceToHost, b->st));
    BCHK(hipStreamSynchronize(b->st));
    return 0;
}

int cfdart_batch_checksum_async(cfdart_batch* b, int slot, unsigned long long* host_out)
{
    int rc = check_slot(b, slot);
    if (rc) return rc;
    if (!host_out) return bfail(CFDART_E_ARG, "null output");
    const BatchGrid& g = b->h_grids[slot];
    hipLaunchKernelGGL(cfd::batch_zero_kernel, dim3(1), dim3(64), 0, b->st, b->d_sum);
    BCHK(hipGetLastError());
    BCHK(cfd::launch_checksum_sel(b->d_grids[slot].buf, &b->d_grids[slot].sw, g.pitch, g.plane, g.nx, g.ny,
                                  g.model == cfd::ALLAIRE ? 6 : 4, b->d_sum, b->st));
    BCHK(hipMemcpyAsync(host_out, b->d_sum, sizeof(unsigned long long), hipMemcpyDeviceToHost, b->st));
    return 0;
}

void* cfdart_batch_stream(cfdart_batch* b) { return b ? (void*)b->st : nullptr; }

void* cfdart_host_alloc(size_t bytes)
{
    void* p = nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) return nullptr;
    return p;
}

void cfdart_host_free(void* p)
{
    if (p) (void)hipHostFree(p);
}

int cfdart_time_tile(const cfdart_desc* d, const double* host_state, int yH, int xrows, int reps, int arith,
                     double* us_x, double* us_y)
{
    if (!d || !us_x || !us_y || reps < 1) return bfail(CFDART_E_ARG, "bad argument");
    cfdart_grid* g = nullptr;
    int rc = cfdart_create(d, &g);
    if (rc) return rc;
    const int nc = d->model == CFDART_ALLAIRE ? 6 : 4;
    std::vector<double> h((size_t)(d->nx + 4) * (d->ny + 4) * nc);
    double W[6], U[6];
    if (d->model == CFDART_ALLAIRE) {
        double z = 0.5 * (d->z_lo + d->z_hi);
        double w[6] = {z * 1000.0, (1 - z) * 1.0, 10.0, 5.0, 1e5, z};
        std::memcpy(W, w, sizeof(w));
    } else {
        double w[4] = {1.0, 0.3, 0.2, 1.0};
        std::memcpy(W, w, sizeof(w));
    }
    cfdart_prim_to_cons(d, W, U);
    for (size_t c = 0; c < h.size() / nc; ++c)
        for (int k = 0; k < nc; ++k) h[c * nc + k] = U[k];
    rc = cfdart_upload(g, host_state ? host_state : h.data());
    if (!rc) rc = cfdart_set_tile(g, yH, xrows);
    if (!rc) rc = cfdart_set_arith(g, arith);
    if (!rc) rc = cfdart_set_profiling(g, 1);
    double dt = 0.0;
    if (!rc) rc = cfdart_compute_dt(g, 10, 0.0, INFINITY, 0.5, &dt, nullptr);
    for (int i = 0; i < reps + 1 && !rc; ++i) {
        if (i == 1) rc = cfdart_set_profiling(g, 1);  // first launches are warm-up
        if (!rc) rc = cfdart_sweep(g, 0, dt);
        if (!rc) rc = cfdart_sweep(g, 1, dt);
    }
    double ms[2];
    long long nl[2];
    if (!rc) rc = cfdart_profile(g, ms, nl);
    if (!rc) {
        *us_x = 1e3 * ms[0] / (double)std::max(1LL, nl[0]);
        *us_y = 1e3 * ms[1] / (double)std::max(1LL, nl[1]);
    }
    if (rc) {
        std::string m = cfdart_last_error();
        cfdart_destroy(g);
        return bfail(rc, m);
    }
    cfdart_destroy(g);
    return 0;
}

}  // extern "C"
