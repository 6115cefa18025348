// Device functions of the sweeps, shared by the single-grid kernels (sweep_kernels.hip) and
// the persistent multi-grid kernel (batch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "sweep_kernels.h"

namespace cfd {

// ------------------------------------------------------------------ lane exchange

// Wavefront rotations by one lane (DPP wave_ror:1 and wave_rol:1): lane i receives lane i-1
// (from_left) or lane i+1 (from_right), and lanes 0 and 63 receive lanes 63 and 0. Every lane
// has a source lane, so each 32-bit half is one v_mov_b32_dpp that writes its destination
// outright (a shift leaves a lane without a source unwritten, which costs a copy of a fallback
// value into the destination first). Every lane of the wavefront must be active. What lanes 0
// and 63 receive is no neighbour's: the line computations below never store those lanes' cells
// and never use a value derived from them in a stored cell.
__device__ __forceinline__ int rot_from_left_i(int x) { return __builtin_amdgcn_mov_dpp(x, 0x13C, 0xf, 0xf, false); }
__device__ __forceinline__ int rot_from_right_i(int x) { return __builtin_amdgcn_mov_dpp(x, 0x134, 0xf, 0xf, false); }
__device__ __forceinline__ double rot_from_left(double x)
{
    return __hiloint2double(rot_from_left_i(__double2hiint(x)), rot_from_left_i(__double2loint(x)));
}
__device__ __forceinline__ double rot_from_right(double x)
{
    return __hiloint2double(rot_from_right_i(__double2hiint(x)), rot_from_right_i(__double2loint(x)));
}

// True on every lane when the predicate holds on any lane (a wavefront-uniform branch condition:
// code under it runs with every lane active, so it may move values between lanes).
__device__ __forceinline__ bool wave_any(bool p) { return __builtin_amdgcn_ballot_w64(p) != 0; }

// Lane i receives lane i-1's predicate (lane 0 receives false): the wavefront's lane mask shifted
// by one lane, in scalar registers. Every lane of the wavefront must be active.
__device__ __forceinline__ bool from_left_lane(bool p)
{
    return __builtin_amdgcn_inverse_ballot_w64(__builtin_amdgcn_ballot_w64(p) << 1);
}

// The right-hand evolved side of the left neighbour's cell stage (its 'ok' flag travels
// separately). Only what cannot be rebuilt is moved: rho is U[0] (Euler) or U[0] + U[1]
// (Allaire) on every side, and a fast-arithmetic Euler side is its rho, u, v, p (and c), from
// which form_cons builds U where a face needs it.
template <int M, bool F>
__device__ __forceinline__ void side_from_left(const Side<M>& s, Side<M>& o)
{
    if constexpr (F && M == EULER) {
        o.rho = rot_from_left(s.rho);
        o.u = rot_from_left(s.u);
        o.v = rot_from_left(s.v);
        o.p = rot_from_left(s.p);
        o.c = rot_from_left(s.c);
    } else {
#pragma unroll
        for (int k = 0; k < Model<M>::NC; ++k) o.U[k] = rot_from_left(s.U[k]);
        o.rho = M == EULER ? o.U[0] : o.U[0] + o.U[1];
        o.u = rot_from_left(s.u);
        o.p = rot_from_left(s.p);
        o.c = rot_from_left(s.c);
        if (M == ALLAIRE) {
            o.eps = rot_from_left(s.eps);
            o.k = rot_from_left(s.k);
        }
    }
}

// The wavefront's index in its workgroup, as a wavefront-uniform value: rows derived from it
// are uniform too, so a row's address is formed once in scalar registers.
__device__ __forceinline__ int wave_index() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

// A wavefront-uniform row address kept in scalar registers (the loop optimiser would otherwise
// carry per-lane 64-bit copies of each row pointer across the row loop).
__device__ __forceinline__ unsigned long long uniform_addr(const void* p)
{
    const unsigned long long a = (unsigned long long)p;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)a);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(a >> 32));
    return ((unsigned long long)hi << 32) | lo;
}

// A wavefront-uniform row of n doubles as a buffer resource in scalar registers: loads and
// stores through it take the lane's 32-bit byte offset and need no per-lane address arithmetic.
// (An offset at or past 8n bytes would read zero and drop the write; callers' columns are in
// [0, n).) The descriptor's data format is 32-bit, as untyped buffer accesses require on CDNA.
typedef unsigned int u32x2 __attribute__((__vector_size__(8)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t row_rsrc(const double* row, int n)
{
    return __builtin_amdgcn_make_buffer_rsrc((void*)uniform_addr(row), 0, n * 8, 0x00020000);
}
__device__ __forceinline__ double row_load(__amdgpu_buffer_rsrc_t r, int x)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, x * 8, 0, 0));
}
__device__ __forceinline__ void row_store(__amdgpu_buffer_rsrc_t r, int x, double v)
{
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, x * 8, 0, 0);
}

// New states are written with ordinary stores, which leave them in the memory-side cache (256 MB
// on MI355X) from which the next sweep reads them back. Nontemporal (streaming) stores bypass it:
// each sweep then read its input from HBM, and the memory traffic of a 2048^2 Euler x sweep alone
// took 56 instead of 44 us.
__device__ __forceinline__ void store_state(double v, double* p) { *p = v; }

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

// A DPP move of a double within the wavefront (both 32-bit halves); lanes of rows outside
// ROW_MASK keep their own value.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ double dpp_move(double v)
{
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(v), __double2hiint(v), CTRL, ROW_MASK, 0xf, false);
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(v), __double2loint(v), CTRL, ROW_MASK, 0xf, false);
    return __hiloint2double(hi, lo);
}

// Maximum over the wavefront, received by every lane: pairs, quads (quad_perm), half-rows and
// rows (mirrors), then rows 0+1 and 2+3 and all four (row_bcast:15 and :31 into the rows of
// row mask 0xA and 0xC) leave it in row 3, whose lane 63 is read into scalar registers. Every
// lane must be active. The inputs here are never NaN and never below +0, so v_max_f64 gives the
// maximum a serial smax fold gives, in any order.
__device__ __forceinline__ double wave_max(double v)
{
    v = fmax(v, dpp_move<0xB1>(v));        // quad_perm [1,0,3,2]
    v = fmax(v, dpp_move<0x4E>(v));        // quad_perm [2,3,0,1]
    v = fmax(v, dpp_move<0x141>(v));       // row_half_mirror
    v = fmax(v, dpp_move<0x140>(v));       // row_mirror
    v = fmax(v, dpp_move<0x142, 0xA>(v));  // row_bcast:15
    v = fmax(v, dpp_move<0x143, 0xC>(v));  // row_bcast:31
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 63),
                            __builtin_amdgcn_readlane(__double2loint(v), 63));
}

struct Tally {
    double su = 0.0, sv = 0.0;  // running max of |u|+c, |v|+c (NaN never enters: smax keeps the old value)
    // Wavefront-wide counts, the same in every lane: each increment is the population count of a
    // ballot, taken where every lane of the wavefront is active (the counting costs scalar
    // instructions only).
    int slope = 0, evolve = 0, repair = 0, bad = 0;
    __device__ static void count(int& n, bool p) { n += __builtin_popcountll(__builtin_amdgcn_ballot_w64(p)); }
    __device__ static void count(int& n, unsigned long long lanes) { n += __builtin_popcountll(lanes); }
    // running maxima of a final state's speeds: the fast arithmetic takes v_max_f64, which agrees
    // with smax here (speeds are +0 or more, and a NaN speed keeps the old maximum in both)
    template <bool F> __device__ void speeds(double a, double b)
    {
        su = F ? fmax(su, a) : smax(su, a);
        sv = F ? fmax(sv, b) : smax(sv, b);
    }
    // Wavefront-wide reduction of the maxima, then lane 0 issues the atomics into a slot of the control block
    // (slot = the workgroup's number times its wavefront count plus the wavefront's). Every lane
    // of the wavefront must call it. There is no workgroup barrier: a barrier's release fence
    // would hold the whole workgroup, and its LDS and wavefront slots, until every global store
    // of its tile had completed (1.6 us per transposed y tile at 2048^2).
    __device__ void flush(Ctrl* c, bool last, int slot)
    {
        double mu = 0.0, mv = 0.0;
        if (last) { mu = wave_max(su); mv = wave_max(sv); }
        const int a = slope, b = evolve, d = repair, e = bad;
        if ((threadIdx.x & 63) == 0) {
            const int sl = (slot * (int)(blockDim.x >> 6) + (int)(threadIdx.x >> 6)) & (MAX_SLOTS - 1);
            if (last) {
                atomicMax(&c->maxu(sl), (unsigned long long)__double_as_longlong(mu));
                atomicMax(&c->maxv(sl), (unsigned long long)__double_as_longlong(mv));
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

// ------------------------------------------------------------------ line step

// One step of a line of 64 consecutive cells, one per lane, shared by the x sweep (a row) and
// the transposed y sweep (a column in the rotated frame). U is the lane's conserved state in the
// sweep's frame, with a reflective wall's ghost sign flip applied. Each lane forms its cell stage
// (primitives, limited slopes from its neighbours' primitives, M3 flags, evolved sides) and the
// face to its left from its left neighbour's right-hand side and its own left-hand side (M3 and M5
// fallbacks, HLLC), receives the face to its right from the next lane, and updates its cell into
// Un (bad: the Allaire physical-state test failed, tested where 'checked'). slope_fb and
// evolve_fb: the lanes whose face took a fallback (wavefront-uniform lane masks). A stored cell's result needs its lane's neighbours
// two lanes either side, so lanes 0, 1, 62 and 63 are halo lanes; lanes 0 and 63 receive values
// from no neighbour and are never stored, and neither is anything derived from them.
template <int M, bool F>
__device__ __forceinline__ void line_step(const GridView& g, double s, double r, const double* U, bool checked,
                                          double* Un, bool& bad, unsigned long long& slope_lanes,
                                          unsigned long long& evolve_lanes)
{
    constexpr int NC = Model<M>::NC;
    Prim<M> P;
    P.template from_cons<F>(g.mat, U);
    double Wm[NC], Wp[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) { Wm[k] = rot_from_left(P.W[k]); Wp[k] = rot_from_right(P.W[k]); }
    CellStage<M> cs;
    cell_stage<M, F>(g.mat, g.opt.limiter, s, P, Wm, Wp, cs);

    // the face's left input: the left neighbour's right-hand side (its M5 and M3 flags arrive as
    // shifted lane masks); its right input: this cell's left-hand side
    Side<M> L, R = cs.L;
    side_from_left<M, F>(cs.R, L);
    L.ok = from_left_lane(cs.R.ok);
    // M3: both slopes of the face are zeroed when either extrapolation is unphysical. The left
    // neighbour's primitives and sound speed are moved (again) only when some lane of the
    // wavefront needs them, so they need not stay live across the cell stage. Sides that carry
    // only primitives get their U before any fallback replaces a side.
    // Only the faces of lanes 2..62 reach a stored cell (lanes 2..61 store; lane i updates from
    // the faces of lanes i and i+1), so only their fallbacks make the wavefront take the
    // fallback branches (lane 0, whose left neighbour is no lane, always reports one).
    constexpr unsigned long long FACES = 0x7FFFFFFFFFFFFFFCull;
    const bool slope_fb = !(from_left_lane(cs.fR) && cs.fL);
    slope_lanes = __builtin_amdgcn_ballot_w64(slope_fb) & FACES;
    const bool any_slope = slope_lanes != 0;
    if (any_slope) {
        form_cons<M, F>(g.mat, L);
        form_cons<M, F>(g.mat, R);
        Prim<M> Pa;
#pragma unroll
        for (int k = 0; k < NC; ++k) Pa.W[k] = rot_from_left(P.W[k]);
        Pa.c = rot_from_left(P.c);
        if (slope_fb) {
            unsloped_side<M, F>(g.mat, s, Pa, true, L);
            unsloped_side<M, F>(g.mat, s, P, false, R);
        }
    }
    // M5: an unphysical evolved side sends the face to the raw stencil cells; the left
    // neighbour's raw state is moved only when some lane needs it
    const bool evolve_fb = !(L.ok && R.ok);
    evolve_lanes = __builtin_amdgcn_ballot_w64(evolve_fb) & FACES;
    const bool any_evolve = evolve_lanes != 0;
    if (any_evolve) {
        if (!any_slope) {
            form_cons<M, F>(g.mat, L);
            form_cons<M, F>(g.mat, R);
        }
        double Ua[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) Ua[k] = rot_from_left(U[k]);
        if (evolve_fb) {
            side_from_raw<M, F>(g.mat, Ua, L);
            side_from_raw<M, F>(g.mat, U, R);
        }
    }
    FaceOut<M> fo;
    // without a fallback in the wavefront, a fast Euler face forms only its upwind side's U
    if (F && M == EULER && !any_slope && !any_evolve) hllc<M, F, true>(g.mat, g.opt.wave_speeds, L, R, fo);
    else hllc<M, F>(g.mat, g.opt.wave_speeds, L, R, fo);

    // the face to the right, from the next lane (Allaire's sixth flux component is zero, and the
    // z update replaces that component of the new state: it is not moved)
    constexpr int NF = M == ALLAIRE ? 5 : NC;
    double Fr[NC];
#pragma unroll
    for (int k = 0; k < NF; ++k) Fr[k] = rot_from_right(fo.F[k]);
    if (M == ALLAIRE) Fr[5] = 0.0;
    const double usr = rot_from_right(fo.ustar);
    const double zsr = M == ALLAIRE ? rot_from_right(fo.zstar) : 0.0;
    update_cell<M, F>(g, r, U, fo.F, Fr, fo.ustar, usr, fo.zstar, zsr, Un, bad, checked);
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

    // the repair and the physical-state test skip the halo lanes, which are not stored
    double Un[NC];
    bool bad;
    unsigned long long sfb, efb;
    line_step<M, F>(g, s, r, U, store, Un, bad, sfb, efb);
    bool rep = false;
    if (M == EULER && store) rep = euler_repair<F>(g.mat, U, Un);

    // rep and bad hold only on stored lanes
    const unsigned long long counted = __builtin_amdgcn_ballot_w64(face_counted);
    Tally::count(tl.slope, sfb & counted);
    Tally::count(tl.evolve, efb & counted);
    Tally::count(tl.repair, rep);
    Tally::count(tl.bad, bad);
    if (store) {
        // y is wavefront-uniform: each component's row address is a scalar
#pragma unroll
        for (int k = 0; k < NC; ++k)
            row_store(row_rsrc(g.nxt + k * g.plane + (long long)y * g.pitch, g.pitch), x, Un[k]);
        if (LAST) {
            double su, sv;
            wave_speeds<F>(g.mat, Un, su, sv, (Model<M>*)nullptr);
            tl.speeds<F>(su, sv);
        }
    }
}

// The state of the lanes' columns of row y (wavefront-uniform) of tile tx.
template <int M>
__device__ __forceinline__ void x_load(const GridView& g, int tx, int y, double* U)
{
    const int xs = mirror(tx * XTILE - 2 + (int)(threadIdx.x & 63), g.nx);
#pragma unroll
    for (int k = 0; k < Model<M>::NC; ++k) U[k] = row_load(row_rsrc(g.cur + k * g.plane + (long long)y * g.pitch, g.pitch), xs);
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

        // lane 0 receives lane 63's state and lane 63 lane 0's flux and z: neither is stored
        double Ua[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) Ua[k] = rot_from_left(U[k]);
        FaceOut<M> fo;
        godunov_face<M, F>(g.mat, g.opt.wave_speeds, Ua, U, fo);
        constexpr int NF = M == ALLAIRE ? 5 : NC;  // Allaire's sixth flux component is zero
        double Fr[NC];
#pragma unroll
        for (int k = 0; k < NF; ++k) Fr[k] = rot_from_right(fo.F[k]);
        if (M == ALLAIRE) Fr[5] = 0.0;
        const double usr = rot_from_right(fo.ustar);
        const double zL = M == ALLAIRE ? Ua[NC - 1] : 0.0;
        const double zR = M == ALLAIRE ? rot_from_right(U[NC - 1]) : 0.0;

        double Un[NC];
        bool bad;
        update_cell_godunov<M, F>(g, r, U, zL, zR, fo.F, Fr, fo.ustar, usr, Un, bad, store);
        bool rep = false;
        if (M == EULER && store) rep = euler_repair<F>(g.mat, U, Un);
        Tally::count(tl.repair, store && rep);
        Tally::count(tl.bad, store && bad);
        if (store) {
            double* out = g.nxt + (long long)y * g.pitch + x;
#pragma unroll
            for (int k = 0; k < NC; ++k) store_state(Un[k], out + k * g.plane);
            if (LAST) {
                double su, sv;
                wave_speeds<F>(g.mat, Un, su, sv, (Model<M>*)nullptr);
                tl.speeds<F>(su, sv);
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
        const bool counted = valid && (c < y1 || c == g.ny);
        Tally::count(tl.slope, counted && (fl & 1) != 0);
        Tally::count(tl.evolve, counted && (fl & 2) != 0);
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
            Tally::count(tl.repair, valid && rep);
            Tally::count(tl.bad, valid && bad);
            if (valid) {
                double* out = g.nxt + (long long)(c - 1) * g.pitch + x;
#pragma unroll
                for (int k = 0; k < NC; ++k) store_state(Un[k], out + k * g.plane);
                if (LAST) {
                    double su, sv;
                    wave_speeds<F>(g.mat, Un, su, sv, (Model<M>*)nullptr);
                    tl.speeds<F>(su, sv);
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
// last_here (LAST only): this sweep ends the grid's step, so the wave speeds are taken; a batch
// launch of LAST sweeps can hold grids whose step ends in a later sweep.
template <int M, bool LAST, int COLS, bool F = false>
__device__ __forceinline__ void y_tile_t(const GridView& g, double dt, int x0, int y0, double* sU, Tally& tl,
                                         bool last_here = true)
{
    constexpr int NC = Model<M>::NC;
    constexpr int MX = Model<M>::MN, MY = Model<M>::MT;
    constexpr int PAD = COLS + 1;  // LDS row pitch in doubles: conflict-free column reads
    static_assert(256 % COLS == 0 && COLS % 4 == 0, "a 256-thread tile load covers whole rows; 4 waves share the columns");
    const int tid = threadIdx.x;
    {
        // every load of the tile is issued before the first LDS write (one memory latency per
        // tile rather than one per pass)
        constexpr int PASSES = 64 * COLS / 256;
        const int col = tid % COLS, rsub = tid / COLS;
        const int xs = min(x0 + col, g.nx - 1);
        double v[PASSES][NC];
#pragma unroll
        for (int pass = 0; pass < PASSES; ++pass) {
            const int ys = mirror(y0 - 2 + pass * (256 / COLS) + rsub, g.ny);
            const double* b = g.cur + (long long)ys * g.pitch + xs;
#pragma unroll
            for (int k = 0; k < NC; ++k) v[pass][k] = b[k * g.plane];
        }
#pragma unroll
        for (int pass = 0; pass < PASSES; ++pass) {
            const int r = pass * (256 / COLS) + rsub;
#pragma unroll
            for (int k = 0; k < NC; ++k) sU[(k * 64 + r) * PAD + col] = v[pass][k];
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

        double Un[NC];
        bool bad;
        unsigned long long sfb, efb;
        line_step<M, F>(g, s, r, U, store, Un, bad, sfb, efb);
        // rotate back (m_x, m_y) -> (-m_y, m_x): E/sourcecode/euler_colsweep.cpp:163-168
        double t = Un[MX];
        Un[MX] = -Un[MY];
        Un[MY] = t;
        bool rep = false;
        if (M == EULER && store) rep = euler_repair<F>(g.mat, raw, Un);
        // rep and bad hold only on stored lanes
        const unsigned long long counted = __builtin_amdgcn_ballot_w64(face_row && col_ok);
        Tally::count(tl.slope, sfb & counted);
        Tally::count(tl.evolve, efb & counted);
        Tally::count(tl.repair, rep);
        Tally::count(tl.bad, bad);
        // the column is this wavefront's alone: its results replace it
#pragma unroll
        for (int k = 0; k < NC; ++k) sU[(k * 64 + lane) * PAD + c] = Un[k];
    }
    __syncthreads();
    {
        // The last sweep of a step takes each new cell's wave speeds here, from the state it
        // stores: the same cells (rows 2..61 of the tile inside the grid) and values as in the
        // line loop above, but with few registers live, so the line loop keeps the occupancy of
        // a sweep that does not end a step.
        const int col = tid % COLS, rsub = tid / COLS;
        const int x = x0 + col;
        if (x < g.nx) {
#pragma unroll
            for (int pass = 0; pass < 64 * COLS / 256; ++pass) {
                const int rr = pass * (256 / COLS) + rsub;
                const int yy = y0 - 2 + rr;
                if (rr >= 2 && rr < 62 && yy < g.ny) {
                    double* out = g.nxt + (long long)yy * g.pitch + x;
                    double Un[NC];
#pragma unroll
                    for (int k = 0; k < NC; ++k) {
                        Un[k] = sU[(k * 64 + rr) * PAD + col];
                        store_state(Un[k], out + k * g.plane);
                    }
                    if (LAST && last_here) {
                        double su, sv;
                        wave_speeds<F>(g.mat, Un, su, sv, (Model<M>*)nullptr);
                        tl.speeds<F>(su, sv);
                    }
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
        Tally::count(tl.repair, valid && rep);
        Tally::count(tl.bad, valid && bad);
        if (valid) {
            double* out = g.nxt + (long long)(c - 1) * g.pitch + x;
#pragma unroll
            for (int k = 0; k < NC; ++k) store_state(Un[k], out + k * g.plane);
            if (LAST) {
                double su, sv;
                wave_speeds<F>(g.mat, Un, su, sv, (Model<M>*)nullptr);
                tl.speeds<F>(su, sv);
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
        a = c->maxu(i) > a ? c->maxu(i) : a;
        b = c->maxv(i) > b ? c->maxv(i) : b;
        c->maxu(i) = 0;
        c->maxv(i) = 0;
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
    const double a = __longlong_as_double((long long)c->maxu(lane));
    const double b = __longlong_as_double((long long)c->maxv(lane));
    c->maxu(lane) = 0;
    c->maxv(lane) = 0;
    mu = wave_max(a);
    mv = wave_max(b);
}

// dt from the folded maxima (E/sourcecode/euler_timestep.cpp:35-43).
__device__ __forceinline__ double dt_from_values(double cfl, double T, double mu, double mv, int n, double t, double dx,
                                                 double dy)
{
    double CFL = cfl;
    if (n < 5) CFL = smin(CFL, 0.2);
    double dt = CFL * smin(dx / mu, dy / mv);
    if (t + dt > T) dt = T - t;
    return dt;
}
__device__ __forceinline__ double dt_from_maxima(const Ctrl* c, double mu, double mv, int n, double t, double dx,
                                                 double dy)
{
    return dt_from_values(c->cfl, c->T, mu, mv, n, t, dx, dy);
}

}  // namespace cfd
