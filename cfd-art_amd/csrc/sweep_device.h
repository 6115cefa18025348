// Device functions of the sweeps, shared by the single-grid kernels (sweep_kernels.hip) and
// the persistent multi-grid kernel (batch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "sweep_kernels.h"

namespace cfd {

// ------------------------------------------------------------------ lane exchange

__device__ __forceinline__ double lane_from_left(double x)  // lane i receives lane i-1
{
    int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), 0x138, 0xf, 0xf, false);
    int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), 0x138, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane_from_right(double x)  // lane i receives lane i+1
{
    int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), 0x130, 0xf, 0xf, false);
    int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), 0x130, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}
// As above, but a lane without a source lane (lane 0 shifting right, lane 63 shifting left)
// keeps its own value instead of 0 (bound_ctrl off: the DPP move leaves such lanes unwritten).
__device__ __forceinline__ double lane_from_left_or_self(double x)
{
    int lo = __builtin_amdgcn_update_dpp(__double2loint(x), __double2loint(x), 0x138, 0xf, 0xf, false);
    int hi = __builtin_amdgcn_update_dpp(__double2hiint(x), __double2hiint(x), 0x138, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane_from_right_or_self(double x)
{
    int lo = __builtin_amdgcn_update_dpp(__double2loint(x), __double2loint(x), 0x130, 0xf, 0xf, false);
    int hi = __builtin_amdgcn_update_dpp(__double2hiint(x), __double2hiint(x), 0x130, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ int lane_from_left_or_self_i(int x)
{
    return __builtin_amdgcn_update_dpp(x, x, 0x138, 0xf, 0xf, false);
}

// The right-hand side of the left neighbour's cell stage; lane 0 takes its own.
template <int M> __device__ __forceinline__ void side_from_left(const Side<M>& s, Side<M>& o)
{
#pragma unroll
    for (int k = 0; k < Model<M>::NC; ++k) o.U[k] = lane_from_left_or_self(s.U[k]);
    o.rho = lane_from_left_or_self(s.rho);
    o.u = lane_from_left_or_self(s.u);
    o.p = lane_from_left_or_self(s.p);
    o.c = lane_from_left_or_self(s.c);
    if (M == ALLAIRE) {
        o.eps = lane_from_left_or_self(s.eps);
        o.k = lane_from_left_or_self(s.k);
    }
    o.ok = lane_from_left_or_self_i(s.ok ? 1 : 0) != 0;
}

// Workgroups go to the 8 XCDs (each with its own L2) round-robin in order of linear id.
// xcd_order maps the linear id b of n workgroups to a tile order in which each XCD receives
// one contiguous run of tiles (the last n % 8 workgroups keep their id), so tiles that
// share halo cache lines are fetched into the same L2.
__device__ __forceinline__ int xcd_order(int b, int n)
{
    constexpr int NXCD = 8;
    const int q = n / NXCD;
    if (b >= q * NXCD) return b;
    return (b % NXCD) * q + b / NXCD;
}

// Ghost index of the reference's pre_sweep: interior index i of an axis of n cells, with
// i in [-2, n+2) mapped to its mirror. Indices further out (dead lanes) are clamped in range.
__device__ __forceinline__ int mirror(int i, int n)
{
    int r = i < 0 ? -1 - i : (i >= n ? 2 * n - 1 - i : i);
    return r < 0 ? 0 : (r >= n ? n - 1 : r);
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = smax(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ int wave_sum(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

struct Tally {
    double su = 0.0, sv = 0.0;  // running max of |u|+c, |v|+c (NaN never enters: smax keeps the old value)
    int slope = 0, evolve = 0, repair = 0, bad = 0;
    // Wavefront-wide reduction, then lane 0 issues the atomics into a slot of the control block
    // (slot = the workgroup's number times its wavefront count plus the wavefront's). Every lane
    // of the wavefront must call it. There is no workgroup barrier: a barrier's release fence
    // would hold the whole workgroup, and its LDS and wavefront slots, until every global store
    // of its tile had completed (1.6 us per transposed y tile at 2048^2).
    __device__ void flush(Ctrl* c, bool last, int slot)
    {
        double mu = 0.0, mv = 0.0;
        if (last) { mu = wave_max(su); mv = wave_max(sv); }
        const int a = wave_sum(slope), b = wave_sum(evolve), d = wave_sum(repair), e = wave_sum(bad);
        if ((threadIdx.x & 63) == 0) {
            const int sl = (slot * (int)(blockDim.x >> 6) + (int)(threadIdx.x >> 6)) & (MAX_SLOTS - 1);
            if (last) {
                atomicMax(&c->maxu[sl], (unsigned long long)__double_as_longlong(mu));
                atomicMax(&c->maxv[sl], (unsigned long long)__double_as_longlong(mv));
            }
            if (a) atomicAdd(&c->cnt[0], (unsigned long long)a);
            if (b) atomicAdd(&c->cnt[1], (unsigned long long)b);
            if (d) atomicAdd(&c->cnt[2], (unsigned long long)d);
            if (e) {
                atomicAdd(&c->cnt[3], (unsigned long long)e);
                atomicOr(&c->error, 1);
            }
        }
    }
};

__device__ __forceinline__ bool sweep_dt(const GridView& g, double dt_arg, double dt_factor,
                                         int use_ctrl, double& dt)
{
    if (!use_ctrl) { dt = dt_arg; return true; }
    if (!g.ctrl->active) return false;
    dt = dt_factor * g.ctrl->dt;
    return true;
}

// Conservative update + model-specific post-processing of one cell (row-sweep frame or
// rotated column frame): E/sourcecode/euler_rowsweep.cpp:36-39,
// A/sourcecode/allaire_diffuse_rowsweep.cpp:36-50.
template <int M, bool F = false>
__device__ __forceinline__ void update_cell(const GridView& g, double r, const double* U,
                                           const double* FL, const double* FR, double usL, double usR,
                                           double zsL, double zsR, double* Un, bool& bad, bool checked = true)
{
    constexpr int NC = Model<M>::NC;
#pragma unroll
    for (int k = 0; k < NC; ++k) Un[k] = U[k] + r * (FL[k] - FR[k]);
    bad = false;
    if (M == ALLAIRE) {
        Un[5] = zupdate(r, U[5], usL, usR, zsL, zsR, g.opt.z_lo, g.opt.z_hi);
        snap(Un);
        bad = checked && !allaire_physical_screened<F>(Un);
    }
}

// ------------------------------------------------------------------ x sweep

// One line of an x-sweep tile: the calling wavefront's lanes hold columns [tx*60-2, tx*60+62)
// of row y (U: their conserved state as loaded, before the wall's sign flip); it writes the
// middle 60 columns.
template <int M, bool LAST, bool F = false>
__device__ __forceinline__ void x_line(const GridView& g, double s, double r, int tx, int y, double* U, Tally& tl)
{
    constexpr int NC = Model<M>::NC;
    constexpr int MN = Model<M>::MN;
    const int lane = threadIdx.x & 63;
    const int x = tx * XTILE - 2 + lane;
    const bool store = lane >= 2 && lane < 62 && x < g.nx;
    const bool neg = (x < 0 && g.bc[0]) || (x >= g.nx && g.bc[1]);
    const bool face_counted = (lane >= 2 && lane < 62 && x <= g.nx) || (lane == 62 && x == g.nx);
    if (neg) U[MN] = -U[MN];

    Prim<M> P;
    P.template from_cons<F>(g.mat, U);
    double Wm[NC], Wp[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) { Wm[k] = lane_from_left_or_self(P.W[k]); Wp[k] = lane_from_right_or_self(P.W[k]); }
    CellStage<M> cs;
    cell_stage<M, F>(g.mat, g.opt.limiter, s, P, Wm, Wp, cs);

    // left neighbour's data for this lane's left face (lane 0, which has no left neighbour
    // and whose face is not used, takes its own cell stage: a physical state, so it sends
    // the wavefront into no fallback branch)
    Prim<M> Pa;
#pragma unroll
    for (int k = 0; k < NC; ++k) Pa.W[k] = Wm[k];
    Pa.c = lane_from_left(P.c);
    double Ua[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) Ua[k] = lane_from_left(U[k]);
    CellStage<M> a;
    a.fR = lane_from_left_or_self_i(cs.fR ? 1 : 0) != 0;
    side_from_left<M>(cs.R, a.R);

    FaceOut<M> fo;
    int fl = face_stage<M, F>(g.mat, g.opt, s, Pa, Ua, a, P, U, cs, fo);

    double Fr[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) Fr[k] = lane_from_right(fo.F[k]);
    double usr = lane_from_right(fo.ustar);
    double zsr = M == ALLAIRE ? lane_from_right(fo.zstar) : 0.0;

    // lanes 0 and 63 (halo lanes without a neighbour on one side) update from incomplete
    // fluxes and are not stored: the repair and the physical-state test skip them
    double Un[NC];
    bool bad;
    update_cell<M, F>(g, r, U, fo.F, Fr, fo.ustar, usr, fo.zstar, zsr, Un, bad, store);
    bool rep = false;
    if (M == EULER && store) rep = euler_repair<F>(g.mat, U, Un);

    if (face_counted) { tl.slope += fl & 1; tl.evolve += (fl >> 1) & 1; }
    if (store) {
        double* out = g.nxt + (long long)y * g.pitch + x;
#pragma unroll
        for (int k = 0; k < NC; ++k) __builtin_nontemporal_store(Un[k], out + k * g.plane);
        tl.repair += rep;
        tl.bad += bad;
        if (LAST) {
            double su, sv;
            wave_speeds<F>(g.mat, Un, su, sv, (Model<M>*)nullptr);
            tl.su = smax(tl.su, su);
            tl.sv = smax(tl.sv, sv);
        }
    }
}

// The state of the lanes' columns of row y of tile tx.
template <int M>
__device__ __forceinline__ void x_load(const GridView& g, int tx, int y, double* U)
{
    const int xs = mirror(tx * XTILE - 2 + (int)(threadIdx.x & 63), g.nx);
    const double* base = g.cur + (long long)y * g.pitch + xs;
#pragma unroll
    for (int k = 0; k < Model<M>::NC; ++k) U[k] = base[k * g.plane];
}

// One x-sweep tile: the calling wavefront covers columns [tx*60-2, tx*60+62) and rows
// [row0, row0+nrows); it writes the middle 60 columns.
template <int M, bool LAST, bool F = false>
__device__ __forceinline__ void x_tile(const GridView& g, double dt, int tx, int row0, int nrows, Tally& tl)
{
    constexpr int NC = Model<M>::NC;
    const double s = 0.5 * (dt / g.dx);
    const double r = dt / g.dx;
    const int yend = min(row0 + nrows, g.ny);

    // rows are independent: for Euler the next row's state is loaded while this one is computed;
    // Allaire, whose registers are the tighter limit, leaves the load latency to the other
    // wavefronts (the prefetch registers would spill at its three wavefronts per SIMD)
    constexpr bool PREFETCH = M == EULER;
    double Unext[NC];
    if (PREFETCH && row0 < yend) x_load<M>(g, tx, row0, Unext);
    for (int y = row0; y < yend; ++y) {
        double U[NC];
        if constexpr (PREFETCH) {
#pragma unroll
            for (int k = 0; k < NC; ++k) U[k] = Unext[k];
            if (y + 1 < yend) x_load<M>(g, tx, y + 1, Unext);
        } else {
            x_load<M>(g, tx, y, U);
        }
        x_line<M, LAST, F>(g, s, r, tx, y, U, tl);
    }
}

// Godunov update of one cell: conservative update, then for Allaire the upwind z update and
// the snapping physical-state test (A/sourcecode/allaire_diffuse_rowsweep.cpp:36-50 with
// zupdate_upwind).
template <int M, bool F = false>
__device__ __forceinline__ void update_cell_godunov(const GridView& g, double r, const double* U, double zL,
                                                    double zR, const double* FL, const double* FR, double usL,
                                                    double usR, double* Un, bool& bad, bool checked)
{
    constexpr int NC = Model<M>::NC;
#pragma unroll
    for (int k = 0; k < NC; ++k) Un[k] = U[k] + r * (FL[k] - FR[k]);
    bad = false;
    if (M == ALLAIRE) {
        Un[5] = zupdate_upwind(r, zL, U[5], zR, usL, usR, g.opt.z_lo, g.opt.z_hi);
        snap(Un);
        bad = checked && !allaire_physical_screened<F>(Un);
    }
}

// One x-sweep tile of the first-order Godunov scheme (same tile geometry as x_tile): the face
// left of each lane's cell is the Riemann problem of the left neighbour's and the cell's stored
// states.
template <int M, bool LAST, bool F = false>
__device__ __forceinline__ void x_tile_godunov(const GridView& g, double dt, int tx, int row0, int nrows, Tally& tl)
{
    constexpr int NC = Model<M>::NC;
    constexpr int MN = Model<M>::MN;
    const int lane = threadIdx.x & 63;
    const int x = tx * XTILE - 2 + lane;
    const bool store = lane >= 2 && lane < 62 && x < g.nx;
    const double r = dt / g.dx;
    const int xs = mirror(x, g.nx);
    const bool neg = (x < 0 && g.bc[0]) || (x >= g.nx && g.bc[1]);
    const int yend = min(row0 + nrows, g.ny);

    double Unext[NC];
    if (row0 < yend) {
        const double* base = g.cur + (long long)row0 * g.pitch + xs;
#pragma unroll
        for (int k = 0; k < NC; ++k) Unext[k] = base[k * g.plane];
    }
    for (int y = row0; y < yend; ++y) {
        double U[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) U[k] = Unext[k];
        if (y + 1 < yend) {
            const double* base = g.cur + (long long)(y + 1) * g.pitch + xs;
#pragma unroll
            for (int k = 0; k < NC; ++k) Unext[k] = base[k * g.plane];
        }
        if (neg) U[MN] = -U[MN];

        double Ua[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) Ua[k] = lane_from_left_or_self(U[k]);
        FaceOut<M> fo;
        godunov_face<M, F>(g.mat, g.opt.wave_speeds, Ua, U, fo);
        double Fr[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) Fr[k] = lane_from_right(fo.F[k]);
        const double usr = lane_from_right(fo.ustar);
        const double zL = M == ALLAIRE ? Ua[NC - 1] : 0.0;
        const double zR = M == ALLAIRE ? lane_from_right_or_self(U[NC - 1]) : 0.0;

        double Un[NC];
        bool bad;
        update_cell_godunov<M, F>(g, r, U, zL, zR, fo.F, Fr, fo.ustar, usr, Un, bad, store);
        bool rep = false;
        if (M == EULER && store) rep = euler_repair<F>(g.mat, U, Un);
        if (store) {
            double* out = g.nxt + (long long)y * g.pitch + x;
#pragma unroll
            for (int k = 0; k < NC; ++k) __builtin_nontemporal_store(Un[k], out + k * g.plane);
            tl.repair += rep;
            tl.bad += bad;
            if (LAST) {
                double su, sv;
                wave_speeds<F>(g.mat, Un, su, sv, (Model<M>*)nullptr);
                tl.su = smax(tl.su, su);
                tl.sv = smax(tl.sv, sv);
            }
        }
    }
}

// Identity copy of a tile (a grid whose time loop has ended keeps its state through the swap).
template <int M>
__device__ __forceinline__ void copy_rows(const GridView& g, int x, bool valid, int y0, int y1)
{
    if (!valid) return;
    for (int y = y0; y < y1; ++y)
        for (int k = 0; k < Model<M>::NC; ++k)
            g.nxt[k * g.plane + (long long)y * g.pitch + x] = g.cur[k * g.plane + (long long)y * g.pitch + x];
}

// ------------------------------------------------------------------ y sweep

template <int M>
__device__ __forceinline__ void load_rot(const GridView& g, int y, int xs, double* U)
{
    constexpr int NC = Model<M>::NC;
    constexpr int MX = Model<M>::MN, MY = Model<M>::MT;
    const int ys = mirror(y, g.ny);
    const double* b = g.cur + (long long)ys * g.pitch + xs;
    double raw[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) raw[k] = b[k * g.plane];
    // ghost rows of a reflective bottom/top wall negate rho*v (E/sourcecode/euler_ppsweep.cpp:79,105)
    if ((y < 0 && g.bc[2]) || (y >= g.ny && g.bc[3])) raw[MY] = -raw[MY];
    // rotation (m_x, m_y) -> (m_y, -m_x) of E/sourcecode/euler_colsweep.cpp:118-124
#pragma unroll
    for (int k = 0; k < NC; ++k) U[k] = raw[k];
    U[MX] = raw[MY];
    U[MY] = -raw[MX];
}

// One y-sweep tile: each lane owns column x = bx*256 + lane and marches rows [y0, y1).
template <int M, bool LAST, bool F = false>
__device__ __forceinline__ void y_tile(const GridView& g, double dt, int bx, int y0, int y1, Tally& tl)
{
    constexpr int NC = Model<M>::NC;
    constexpr int MX = Model<M>::MN, MY = Model<M>::MT;
    const int x = bx * YBLOCK + threadIdx.x;
    const bool valid = x < g.nx;
    const double s = 0.5 * (dt / g.dx_half_col);
    const double r = dt / g.dy;
    const int xs = valid ? x : g.nx - 1;

    double Um[NC], U0[NC], Up[NC];
    Prim<M> Pm, P0, Pp;
    CellStage<M> a;   // cell c-1 (its R side and fR are used)
    FaceOut<M> Fm;    // face c-1

    load_rot<M>(g, y0 - 2, xs, Um);
    Pm.template from_cons<F>(g.mat, Um);
    load_rot<M>(g, y0 - 1, xs, U0);
    P0.template from_cons<F>(g.mat, U0);
    load_rot<M>(g, y0, xs, Up);
    Pp.template from_cons<F>(g.mat, Up);
    cell_stage<M, F>(g.mat, g.opt.limiter, s, P0, Pm.W, Pp.W, a);  // cell y0-1
    // shift: c-1 <- y0-1, c <- y0
#pragma unroll
    for (int k = 0; k < NC; ++k) { Um[k] = U0[k]; U0[k] = Up[k]; }
    Pm = P0;
    P0 = Pp;

    double Unext[NC];
    load_rot<M>(g, y0 + 1, xs, Unext);
    for (int c = y0; c <= y1; ++c) {
        // window: Um = row c-1, U0 = row c, Unext = row c+1
#pragma unroll
        for (int k = 0; k < NC; ++k) Up[k] = Unext[k];
        if (c + 2 <= y1 + 1) load_rot<M>(g, c + 2, xs, Unext);
        Pp.template from_cons<F>(g.mat, Up);
        CellStage<M> cs;
        cell_stage<M, F>(g.mat, g.opt.limiter, s, P0, Pm.W, Pp.W, cs);
        FaceOut<M> fo;
        int fl = face_stage<M, F>(g.mat, g.opt, s, Pm, Um, a, P0, U0, cs, fo);
        if (valid && (c < y1 || c == g.ny)) { tl.slope += fl & 1; tl.evolve += (fl >> 1) & 1; }
        if (c > y0) {
            // update cell c-1 from faces c-1 (Fm) and c (fo)
            double Un[NC];
            bool bad;
            update_cell<M, F>(g, r, Um, Fm.F, fo.F, Fm.ustar, fo.ustar, Fm.zstar, fo.zstar, Un, bad);
            // rotate back (m_x, m_y) -> (-m_y, m_x): E/sourcecode/euler_colsweep.cpp:163-168
            double t = Un[MX];
            Un[MX] = -Un[MY];
            Un[MY] = t;
            bool rep = false;
            if (M == EULER) {
                double old[NC];
#pragma unroll
                for (int k = 0; k < NC; ++k) old[k] = Um[k];
                old[MX] = -Um[MY];
                old[MY] = Um[MX];
                rep = euler_repair<F>(g.mat, old, Un);
            }
            if (valid) {
                double* out = g.nxt + (long long)(c - 1) * g.pitch + x;
#pragma unroll
                for (int k = 0; k < NC; ++k) __builtin_nontemporal_store(Un[k], out + k * g.plane);
                tl.repair += rep;
                tl.bad += bad;
                if (LAST) {
                    double su, sv;
                    wave_speeds<F>(g.mat, Un, su, sv, (Model<M>*)nullptr);
                    tl.su = smax(tl.su, su);
                    tl.sv = smax(tl.sv, sv);
                }
            }
        }
        // shift the window
#pragma unroll
        for (int k = 0; k < NC; ++k) { Um[k] = U0[k]; U0[k] = Up[k]; }
        Pm = P0;
        P0 = Pp;
        a.fR = cs.fR;
        a.R = cs.R;
        Fm = fo;
    }
}

// Transposed y-sweep tile: rows [y0-2, y0+62) x columns [x0, x0+COLS) staged in the LDS array
// sU (NC * 64 * (COLS+1) doubles; one 8*COLS-byte segment per plane row, loaded and stored
// coalesced). Each wavefront runs the x sweep's lane-parallel line computation down COLS/4 of
// the columns -- lanes along y, neighbours by DPP -- and overwrites its columns in LDS with the
// results; the workgroup (256 threads) then stores the middle 60 rows. Same values as y_tile
// (same cell and face stages, rotated frame).
template <int M, bool LAST, int COLS, bool F = false>
__device__ __forceinline__ void y_tile_t(const GridView& g, double dt, int x0, int y0, double* sU, Tally& tl)
{
    constexpr int NC = Model<M>::NC;
    constexpr int MX = Model<M>::MN, MY = Model<M>::MT;
    constexpr int PAD = COLS + 1;  // LDS row pitch in doubles: conflict-free column reads
    static_assert(256 % COLS == 0 && COLS % 4 == 0, "a 256-thread tile load covers whole rows; 4 waves share the columns");
    const int tid = threadIdx.x;
    {
        const int col = tid % COLS, rsub = tid / COLS;
        const int xs = min(x0 + col, g.nx - 1);
#pragma unroll
        for (int pass = 0; pass < 64 * COLS / 256; ++pass) {
            const int r = pass * (256 / COLS) + rsub;
            const int ys = mirror(y0 - 2 + r, g.ny);
            const double* b = g.cur + (long long)ys * g.pitch + xs;
#pragma unroll
            for (int k = 0; k < NC; ++k) sU[(k * 64 + r) * PAD + col] = b[k * g.plane];
        }
    }
    __syncthreads();
    const int lane = tid & 63, w = tid >> 6;
    const int y = y0 - 2 + lane;
    const bool in_rows = lane >= 2 && lane < 62 && y < g.ny;
    const bool face_row = (lane >= 2 && lane < 62 && y <= g.ny) || (lane == 62 && y == g.ny);
    const bool neg = (y < 0 && g.bc[2]) || (y >= g.ny && g.bc[3]);
    const double s = 0.5 * (dt / g.dx_half_col);
    const double r = dt / g.dy;
    for (int cc = 0; cc < COLS / 4; ++cc) {
        const int c = w + 4 * cc;
        const bool col_ok = x0 + c < g.nx;
        const bool store = in_rows && col_ok;
        double raw[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) raw[k] = sU[(k * 64 + lane) * PAD + c];
        // ghost rows of a reflective bottom/top wall negate rho*v (E/sourcecode/euler_ppsweep.cpp:79,105);
        // rotation (m_x, m_y) -> (m_y, -m_x) of E/sourcecode/euler_colsweep.cpp:118-124
        if (neg) raw[MY] = -raw[MY];
        double U[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) U[k] = raw[k];
        U[MX] = raw[MY];
        U[MY] = -raw[MX];

        Prim<M> P;
        P.template from_cons<F>(g.mat, U);
        double Wm[NC], Wp[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) { Wm[k] = lane_from_left_or_self(P.W[k]); Wp[k] = lane_from_right_or_self(P.W[k]); }
        CellStage<M> cs;
        cell_stage<M, F>(g.mat, g.opt.limiter, s, P, Wm, Wp, cs);
        Prim<M> Pa;
#pragma unroll
        for (int k = 0; k < NC; ++k) Pa.W[k] = Wm[k];
        Pa.c = lane_from_left(P.c);
        double Ua[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) Ua[k] = lane_from_left(U[k]);
        CellStage<M> a;
        a.fR = lane_from_left_or_self_i(cs.fR ? 1 : 0) != 0;
        side_from_left<M>(cs.R, a.R);
        FaceOut<M> fo;
        int fl = face_stage<M, F>(g.mat, g.opt, s, Pa, Ua, a, P, U, cs, fo);
        double Fr[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) Fr[k] = lane_from_right(fo.F[k]);
        double usr = lane_from_right(fo.ustar);
        double zsr = M == ALLAIRE ? lane_from_right(fo.zstar) : 0.0;
        double Un[NC];
        bool bad;
        update_cell<M, F>(g, r, U, fo.F, Fr, fo.ustar, usr, fo.zstar, zsr, Un, bad, store);
        // rotate back (m_x, m_y) -> (-m_y, m_x): E/sourcecode/euler_colsweep.cpp:163-168
        double t = Un[MX];
        Un[MX] = -Un[MY];
        Un[MY] = t;
        bool rep = false;
        if (M == EULER && store) rep = euler_repair<F>(g.mat, raw, Un);
        if (face_row && col_ok) { tl.slope += fl & 1; tl.evolve += (fl >> 1) & 1; }
        if (store) {
            tl.repair += rep;
            tl.bad += bad;
            if (LAST) {
                double su, sv;
                wave_speeds<F>(g.mat, Un, su, sv, (Model<M>*)nullptr);
                tl.su = smax(tl.su, su);
                tl.sv = smax(tl.sv, sv);
            }
        }
        // the column is this wavefront's alone: its results replace it
#pragma unroll
        for (int k = 0; k < NC; ++k) sU[(k * 64 + lane) * PAD + c] = Un[k];
    }
    __syncthreads();
    {
        const int col = tid % COLS, rsub = tid / COLS;
        const int x = x0 + col;
        if (x < g.nx) {
#pragma unroll
            for (int pass = 0; pass < 64 * COLS / 256; ++pass) {
                const int rr = pass * (256 / COLS) + rsub;
                const int yy = y0 - 2 + rr;
                if (rr >= 2 && rr < 62 && yy < g.ny) {
                    double* out = g.nxt + (long long)yy * g.pitch + x;
#pragma unroll
                    for (int k = 0; k < NC; ++k) __builtin_nontemporal_store(sU[(k * 64 + rr) * PAD + col], out + k * g.plane);
                }
            }
        }
    }
}

// One y-sweep tile of the first-order Godunov scheme: each lane marches rows [y0, y1) of its
// column with a two-row window of stored states (faces y0 .. y1 need rows y0-1 .. y1).
template <int M, bool LAST, bool F = false>
__device__ __forceinline__ void y_tile_godunov(const GridView& g, double dt, int bx, int y0, int y1, Tally& tl)
{
    constexpr int NC = Model<M>::NC;
    constexpr int MX = Model<M>::MN, MY = Model<M>::MT;
    const int x = bx * YBLOCK + threadIdx.x;
    const bool valid = x < g.nx;
    const double r = dt / g.dy;
    const int xs = valid ? x : g.nx - 1;

    double Um[NC], U0[NC], Up[NC], Unext[NC];
    load_rot<M>(g, y0 - 1, xs, Um);
    load_rot<M>(g, y0, xs, U0);
    FaceOut<M> Fm;
    godunov_face<M, F>(g.mat, g.opt.wave_speeds, Um, U0, Fm);  // face y0
    load_rot<M>(g, y0 + 1, xs, Unext);
    for (int c = y0 + 1; c <= y1; ++c) {
        // window: Um = row c-2, U0 = row c-1, Up = row c
#pragma unroll
        for (int k = 0; k < NC; ++k) Up[k] = Unext[k];
        if (c + 1 <= y1) load_rot<M>(g, c + 1, xs, Unext);
        FaceOut<M> fo;
        godunov_face<M, F>(g.mat, g.opt.wave_speeds, U0, Up, fo);  // face c
        // update row c-1 from faces c-1 (Fm) and c (fo)
        double Un[NC];
        bool bad;
        update_cell_godunov<M, F>(g, r, U0, Um[NC - 1], Up[NC - 1], Fm.F, fo.F, Fm.ustar, fo.ustar, Un, bad, valid);
        double t = Un[MX];
        Un[MX] = -Un[MY];
        Un[MY] = t;
        bool rep = false;
        if (M == EULER && valid) {
            double old[NC];
#pragma unroll
            for (int k = 0; k < NC; ++k) old[k] = U0[k];
            old[MX] = -U0[MY];
            old[MY] = U0[MX];
            rep = euler_repair<F>(g.mat, old, Un);
        }
        if (valid) {
            double* out = g.nxt + (long long)(c - 1) * g.pitch + x;
#pragma unroll
            for (int k = 0; k < NC; ++k) __builtin_nontemporal_store(Un[k], out + k * g.plane);
            tl.repair += rep;
            tl.bad += bad;
            if (LAST) {
                double su, sv;
                wave_speeds<F>(g.mat, Un, su, sv, (Model<M>*)nullptr);
                tl.su = smax(tl.su, su);
                tl.sv = smax(tl.sv, sv);
            }
        }
#pragma unroll
        for (int k = 0; k < NC; ++k) { Um[k] = U0[k]; U0[k] = Up[k]; }
        Fm = fo;
    }
}

// dt from the reduced maxima (E/sourcecode/euler_timestep.cpp:35-43).
// Folds the slot maxima of the control block (and clears them for the next reduction).
__device__ __forceinline__ void fold_maxima(Ctrl* c, double& mu, double& mv)
{
    unsigned long long a = 0, b = 0;
    for (int i = 0; i < MAX_SLOTS; ++i) {
        a = c->maxu[i] > a ? c->maxu[i] : a;
        b = c->maxv[i] > b ? c->maxv[i] : b;
        c->maxu[i] = 0;
        c->maxv[i] = 0;
    }
    mu = __longlong_as_double((long long)a);
    mv = __longlong_as_double((long long)b);
}

// The same fold by one whole wavefront, lane i taking slot i: every lane of the calling
// wavefront receives the maxima. The slots hold non-negative doubles, whose bit patterns order
// as their values, so the floating-point maximum equals the serial fold's.
__device__ __forceinline__ void fold_maxima_wave(Ctrl* c, double& mu, double& mv)
{
    static_assert(MAX_SLOTS == 64, "one slot per lane");
    const int lane = threadIdx.x & 63;
    const double a = __longlong_as_double((long long)c->maxu[lane]);
    const double b = __longlong_as_double((long long)c->maxv[lane]);
    c->maxu[lane] = 0;
    c->maxv[lane] = 0;
    mu = wave_max(a);
    mv = wave_max(b);
}

// dt from the folded maxima (E/sourcecode/euler_timestep.cpp:35-43).
__device__ __forceinline__ double dt_from_maxima(const Ctrl* c, double mu, double mv, int n, double t, double dx,
                                                 double dy)
{
    double CFL = c->cfl;
    if (n < 5) CFL = smin(CFL, 0.2);
    double dt = CFL * smin(dx / mu, dy / mv);
    if (t + dt > c->T) dt = c->T - t;
    return dt;
}

}  // namespace cfd
