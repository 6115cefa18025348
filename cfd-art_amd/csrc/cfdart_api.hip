// C-ABI of the sweep library (declarations and conventions in include/cfdart.h).
#include "cfdart.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "cfdart_host.h"
#include "frames.h"
#include "ic_device.h"
#include "sweep_kernels.h"

struct cfdart_grid {
    cfdart_desc d;
    int nc;
    int pitch;
    long long plane;
    double* buf[2];
    int cur;
    cfd::Ctrl* ctrl;
    double* partials;
    int nred;
    hipStream_t st;
    int yH, xrows;
    int arith;  // CFDART_ARITH_EXACT or CFDART_ARITH_FAST
    // optional per-kernel timing: an event before and after every sweep launch
    int profiling;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, int>> ev_marks;  // (kind, index of start event); end = start + 1
    size_t ev_used;
    double prof_ms[3];
    long long prof_n[3];
    unsigned long long* d_sum;  // checksum accumulator
};

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(x)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) return fail(CFDART_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

int validate_desc(const cfdart_desc* d)
{
    if (!d) return fail(CFDART_E_ARG, "null descriptor");
    if (d->model != CFDART_EULER && d->model != CFDART_ALLAIRE) return fail(CFDART_E_ARG, "model must be 0 (Euler) or 1 (Allaire)");
    if (d->nx < 2 || d->ny < 2) return fail(CFDART_E_ARG, "nx and ny must be >= 2");
    if (d->flux_solver != CFDART_MUSCL_HANCOCK && d->flux_solver != CFDART_GODUNOV)
        return fail(CFDART_E_ARG, "flux_solver must be 0 (MUSCL-Hancock) or 1 (Godunov)");
    if (d->flux_solver == CFDART_MUSCL_HANCOCK && d->ng != 2)
        return fail(CFDART_E_ARG, "ng must be 2 for MUSCL-Hancock (its stencil)");
    if (d->flux_solver == CFDART_GODUNOV && d->ng != 1 && d->ng != 2)
        return fail(CFDART_E_ARG, "ng must be 1 or 2 for Godunov");
    if (!(d->dx > 0.0) || !(d->dy > 0.0)) return fail(CFDART_E_ARG, "dx and dy must be positive");
    if (!(d->gamma1 > 1.0) || (d->model == CFDART_ALLAIRE && !(d->gamma2 > 1.0)))
        return fail(CFDART_E_ARG, "gamma must be > 1");
    for (int i = 0; i < 4; ++i)
        if (d->bc[i] != 0 && d->bc[i] != 1) return fail(CFDART_E_ARG, "bc must be 0 (transmissive) or 1 (reflective)");
    if (d->limiter != 0 && d->limiter != 1) return fail(CFDART_E_ARG, "limiter must be 0 (minmod) or 1 (van Leer)");
    if (d->wave_speeds != 0 && d->wave_speeds != 1) return fail(CFDART_E_ARG, "wave_speeds must be 0 or 1");
    if (d->wave_speeds == 1 && d->model != CFDART_ALLAIRE) return fail(CFDART_E_ARG, "PVRS wave speeds exist for the Allaire model only");
    if (d->model == CFDART_ALLAIRE && !(d->z_lo >= 0.0 && d->z_lo <= d->z_hi && d->z_hi <= 1.0))
        return fail(CFDART_E_ARG, "need 0 <= z_lo <= z_hi <= 1");
    if ((long long)d->nx * d->ny > (1LL << 31)) return fail(CFDART_E_ARG, "grid too large");
    if (d->nx > (1 << 28)) return fail(CFDART_E_ARG, "nx must be at most 2^28 (a row's byte offsets are 32-bit)");
    return 0;
}

cfd::GridView make_view(const cfdart_desc& d, const double* cur, double* nxt, int pitch, long long plane, int arith,
                        cfd::Ctrl* ctrl)
{
    cfd::GridView v;
    v.cur = cur;
    v.nxt = nxt;
    v.plane = plane;
    v.pitch = pitch;
    v.nx = d.nx;
    v.ny = d.ny;
    for (int i = 0; i < 4; ++i) v.bc[i] = d.bc[i];
    v.mat = cfd::make_material(d.gamma1, d.pinf1, d.gamma2, d.pinf2);
    v.opt.limiter = d.limiter;
    v.opt.wave_speeds = d.wave_speeds;
    v.opt.z_lo = d.z_lo;
    v.opt.z_hi = d.z_hi;
    v.opt.godunov = d.flux_solver == CFDART_GODUNOV ? 1 : 0;
    v.opt.fast = arith == CFDART_ARITH_FAST ? 1 : 0;
    v.dx = d.dx;
    v.dy = d.dy;
    v.dx_half_col = d.colsweep_uses_dx ? d.dx : d.dy;
    v.ctrl = ctrl;
    return v;
}

cfd::GridView view(const cfdart_grid* g)
{
    return make_view(g->d, g->buf[g->cur], g->buf[1 - g->cur], g->pitch, g->plane, g->arith, g->ctrl);
}

hipEvent_t next_event(cfdart_grid* g)
{
    if (g->ev_used == g->ev_pool.size()) {
        hipEvent_t e;
        // timing only: the elapsed times are read after a stream synchronisation, so the events
        // need no system-scope release (which writes back the L2 and costs about 3 us per event)
        if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
        g->ev_pool.push_back(e);
    }
    return g->ev_pool[g->ev_used++];
}

// Accumulates the recorded kernel times (needs the stream to be idle).
void harvest_profile(cfdart_grid* g)
{
    for (auto& m : g->ev_marks) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, g->ev_pool[m.second], g->ev_pool[m.second + 1]) == hipSuccess) {
            g->prof_ms[m.first] += ms;
            g->prof_n[m.first] += 1;
        }
    }
    g->ev_marks.clear();
    g->ev_used = 0;
}

int sweep_once(cfdart_grid* g, int axis, bool last, double dt_arg, double factor, int use_ctrl)
{
    cfd::GridView v = view(g);
    // profiling: the launch itself stamps a pair of events with the kernel's start and end
    cfd::Stamps ts;
    if (g->profiling) {
        if (g->ev_used + 2 > 4096) {
            if (hipStreamSynchronize(g->st) != hipSuccess) return fail(CFDART_E_HIP, "sync failed");
            harvest_profile(g);
        }
        size_t i0 = g->ev_used;
        ts.start = next_event(g);
        ts.stop = next_event(g);
        if (!ts.start || !ts.stop) return fail(CFDART_E_HIP, "event creation failed");
        g->ev_marks.push_back({axis, (int)i0});
    }
    hipError_t e = cfd::launch_sweep(v, g->d.model, axis, last, dt_arg, factor, use_ctrl, g->yH, g->xrows, g->st, ts);
    if (e != hipSuccess) return fail(CFDART_E_HIP, std::string("sweep launch: ") + hipGetErrorString(e));
    g->cur = 1 - g->cur;  // post_sweep: grid.swap(future_grid)
    return 0;
}

// Sweep sequence of one step of the splitting scheme: axes and dt factors.
int step_plan(int splitting, int n, int* axes, double* factors)
{
    if (splitting == CFDART_STRANG) {
        axes[0] = 0; factors[0] = 0.5;
        axes[1] = 1; factors[1] = 1.0;
        axes[2] = 0; factors[2] = 0.5;
        return 3;
    }
    if (splitting == CFDART_ALTERNATING && (n % 2) == 1) {
        axes[0] = 1; factors[0] = 1.0;
        axes[1] = 0; factors[1] = 1.0;
        return 2;
    }
    axes[0] = 0; factors[0] = 1.0;
    axes[1] = 1; factors[1] = 1.0;
    return 2;
}

}  // namespace

namespace cfd {
void set_last_error(const std::string& msg) { g_err = msg; }
// The descriptor checks of cfdart_create, shared with batch admission.
int check_desc(const cfdart_desc* d) { return validate_desc(d); }
}  // namespace cfd

extern "C" {

const char* cfdart_last_error(void) { return g_err.c_str(); }

int cfdart_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int cfdart_create(const cfdart_desc* desc, cfdart_grid** out)
{
    if (!out) return fail(CFDART_E_ARG, "null output pointer");
    *out = nullptr;
    int rc = validate_desc(desc);
    if (rc) return rc;
    if (cfdart_device_count() < 1) return fail(CFDART_E_NODEV, "no HIP device");
    cfdart_grid* g = new cfdart_grid();
    g->d = *desc;
    g->nc = desc->model == CFDART_ALLAIRE ? 6 : 4;
    g->pitch = (desc->nx + 31) / 32 * 32;  // 256-byte aligned rows
    g->plane = (long long)g->pitch * desc->ny;
    g->cur = 0;
    g->yH = CFDART_Y_TRANSPOSED;
    g->xrows = 4;
    g->arith = CFDART_ARITH_EXACT;
    g->buf[0] = g->buf[1] = nullptr;
    g->ctrl = nullptr;
    g->partials = nullptr;
    g->d_sum = nullptr;
    g->profiling = 0;
    g->ev_used = 0;
    for (int i = 0; i < 3; ++i) { g->prof_ms[i] = 0.0; g->prof_n[i] = 0; }
    size_t bytes = sizeof(double) * (size_t)g->plane * g->nc;
    if (hipStreamCreateWithFlags(&g->st, hipStreamNonBlocking) != hipSuccess) {
        delete g;
        return fail(CFDART_E_HIP, "hipStreamCreate failed");
    }
    if (hipMalloc(&g->buf[0], bytes) != hipSuccess || hipMalloc(&g->buf[1], bytes) != hipSuccess ||
        hipMalloc(&g->ctrl, sizeof(cfd::Ctrl)) != hipSuccess ||
        hipMalloc(&g->d_sum, sizeof(unsigned long long)) != hipSuccess) {
        cfdart_destroy(g);
        return fail(CFDART_E_NOMEM, "hipMalloc of the grid state failed");
    }
    cfd::GridView v = view(g);
    g->nred = cfd::reduce_blocks(v);
    if (hipMalloc(&g->partials, sizeof(double) * 2 * g->nred) != hipSuccess) {
        cfdart_destroy(g);
        return fail(CFDART_E_NOMEM, "hipMalloc of the reduction buffer failed");
    }
    (void)hipMemset(g->buf[0], 0, bytes);
    (void)hipMemset(g->buf[1], 0, bytes);
    (void)hipMemset(g->ctrl, 0, sizeof(cfd::Ctrl));
    if (hipDeviceSynchronize() != hipSuccess) {
        cfdart_destroy(g);
        return fail(CFDART_E_HIP, "device initialisation failed");
    }
    *out = g;
    return 0;
}

void cfdart_destroy(cfdart_grid* g)
{
    if (!g) return;
    if (g->st) (void)hipStreamSynchronize(g->st);
    if (g->buf[0]) (void)hipFree(g->buf[0]);
    if (g->buf[1]) (void)hipFree(g->buf[1]);
    if (g->ctrl) (void)hipFree(g->ctrl);
    if (g->partials) (void)hipFree(g->partials);
    if (g->d_sum) (void)hipFree(g->d_sum);
    for (hipEvent_t e : g->ev_pool) (void)hipEventDestroy(e);
    if (g->st) (void)hipStreamDestroy(g->st);
    delete g;
}

int cfdart_upload(cfdart_grid* g, const double* host)
{
    if (!g || !host) return fail(CFDART_E_ARG, "null argument");
    const int nx = g->d.nx, ny = g->d.ny, ng = g->d.ng, nc = g->nc;
    const int W = nx + 2 * ng;
    std::vector<double> soa((size_t)g->plane * nc, 0.0);
    for (int k = 0; k < nc; ++k)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x)
                soa[(size_t)k * g->plane + (size_t)y * g->pitch + x] =
                    host[((size_t)(y + ng) * W + (x + ng)) * nc + k];
    HIPCHK(hipStreamSynchronize(g->st));
    g->cur = 0;
    HIPCHK(hipMemcpy(g->buf[0], soa.data(), sizeof(double) * soa.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(g->ctrl, 0, sizeof(cfd::Ctrl)));
    HIPCHK(hipDeviceSynchronize());
    return 0;
}

namespace {

int init_generated(cfdart_grid* g, const cfdart_disc_params* dp, const cfdart_bubble_params* bp)
{
    if (!g) return fail(CFDART_E_ARG, "null grid");
    void* d_draws = nullptr;
    HIPCHK(hipStreamSynchronize(g->st));
    HIPCHK(hipMalloc(&d_draws, cfd::IC_DRAW_BYTES));
    g->cur = 0;
    std::string err;
    int rc = cfd::generate_state(g->d, dp, bp, g->buf[0], g->pitch, g->plane, d_draws, g->st, err);
    (void)hipFree(d_draws);
    if (rc) return fail(rc, err);
    HIPCHK(hipMemsetAsync(g->ctrl, 0, sizeof(cfd::Ctrl), g->st));
    HIPCHK(hipStreamSynchronize(g->st));
    return 0;
}

}  // namespace

int cfdart_init_random_discs(cfdart_grid* g, const cfdart_disc_params* p)
{
    if (!p) return fail(CFDART_E_ARG, "null parameters");
    return init_generated(g, p, nullptr);
}

int cfdart_init_shocked_bubbles(cfdart_grid* g, const cfdart_bubble_params* p)
{
    if (!p) return fail(CFDART_E_ARG, "null parameters");
    return init_generated(g, nullptr, p);
}

int cfdart_checksum(cfdart_grid* g, unsigned long long* out)
{
    if (!g || !out) return fail(CFDART_E_ARG, "null argument");
    HIPCHK(hipMemsetAsync(g->d_sum, 0, sizeof(unsigned long long), g->st));
    HIPCHK(cfd::launch_checksum(g->buf[g->cur], g->pitch, g->plane, g->d.nx, g->d.ny, g->nc, g->d_sum, g->st));
    HIPCHK(hipMemcpyAsync(out, g->d_sum, sizeof(unsigned long long), hipMemcpyDeviceToHost, g->st));
    HIPCHK(hipStreamSynchronize(g->st));
    return 0;
}

int cfdart_schlieren(cfdart_grid* g, double* host_dst, double* max_grad)
{
    if (!g || !host_dst) return fail(CFDART_E_ARG, "null argument");
    // the image is built in the grid's other buffer, which holds nothing between sweeps
    double* img = g->buf[1 - g->cur];
    HIPCHK(cfd::launch_schlieren(view(g), g->d.model, img, g->d_sum, g->st));
    HIPCHK(hipMemcpyAsync(host_dst, img, sizeof(double) * (size_t)g->d.nx * g->d.ny, hipMemcpyDeviceToHost, g->st));
    unsigned long long m = 0;
    HIPCHK(hipMemcpyAsync(&m, g->d_sum, sizeof(m), hipMemcpyDeviceToHost, g->st));
    HIPCHK(hipStreamSynchronize(g->st));
    if (max_grad) std::memcpy(max_grad, &m, sizeof(double));
    return 0;
}

int cfdart_copy_state(cfdart_grid* g, int reps)
{
    if (!g || reps < 1) return fail(CFDART_E_ARG, "bad handle or repetition count");
    for (int i = 0; i < reps; ++i)
        HIPCHK(cfd::launch_copy(g->buf[g->cur], g->buf[1 - g->cur], g->plane * g->nc, g->st));
    HIPCHK(hipStreamSynchronize(g->st));
    return 0;
}

int cfdart_frame(cfdart_grid* g, double* host_dst)
{
    if (!g || !host_dst) return fail(CFDART_E_ARG, "null argument");
    HIPCHK(hipMemcpy2DAsync(host_dst, sizeof(double) * g->d.nx, g->buf[g->cur], sizeof(double) * g->pitch,
                            sizeof(double) * g->d.nx, (size_t)g->d.ny * g->nc, hipMemcpyDeviceToHost, g->st));
    return 0;
}

int cfdart_download(cfdart_grid* g, double* host)
{
    if (!g || !host) return fail(CFDART_E_ARG, "null argument");
    const int nx = g->d.nx, ny = g->d.ny, ng = g->d.ng, nc = g->nc;
    const int W = nx + 2 * ng;
    std::vector<double> soa((size_t)g->plane * nc);
    HIPCHK(hipStreamSynchronize(g->st));
    HIPCHK(hipMemcpy(soa.data(), g->buf[g->cur], sizeof(double) * soa.size(), hipMemcpyDeviceToHost));
    auto at = [&](int i, int j, int k) -> double& { return host[((size_t)i * W + j) * nc + k]; };
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x)
            for (int k = 0; k < nc; ++k) at(y + ng, x + ng, k) = soa[(size_t)k * g->plane + (size_t)y * g->pitch + x];
    // ghost frame as pre_sweep fills it (E/sourcecode/euler_ppsweep.cpp:7-113)
    const int mx = g->d.model == CFDART_ALLAIRE ? 2 : 1, my = mx + 1;
    for (int i = ng; i < ny + ng; ++i)
        for (int j = 0; j < ng; ++j)
            for (int k = 0; k < nc; ++k) {
                at(i, j, k) = at(i, 2 * ng - j - 1, k);
                at(i, j + nx + ng, k) = at(i, nx + ng - 1 - j, k);
            }
    for (int j = ng; j < nx + ng; ++j)
        for (int i = 0; i < ng; ++i)
            for (int k = 0; k < nc; ++k) {
                at(i, j, k) = at(2 * ng - i - 1, j, k);
                at(i + ny + ng, j, k) = at(ny + ng - 1 - i, j, k);
            }
    for (int i = ng; i < ny + ng; ++i)
        for (int j = 0; j < ng; ++j) {
            if (g->d.bc[0]) at(i, j, mx) = -at(i, j, mx);
            if (g->d.bc[1]) at(i, j + nx + ng, mx) = -at(i, j + nx + ng, mx);
        }
    for (int j = ng; j < nx + ng; ++j)
        for (int i = 0; i < ng; ++i) {
            if (g->d.bc[2]) at(i, j, my) = -at(i, j, my);
            if (g->d.bc[3]) at(i + ny + ng, j, my) = -at(i + ny + ng, j, my);
        }
    return 0;
}

int cfdart_sweep(cfdart_grid* g, int axis, double dt)
{
    if (!g || (axis != 0 && axis != 1)) return fail(CFDART_E_ARG, "bad handle or axis");
    return sweep_once(g, axis, false, dt, 1.0, 0);
}

int cfdart_step(cfdart_grid* g, int splitting, int n, double dt)
{
    if (!g || splitting < 0 || splitting > 2) return fail(CFDART_E_ARG, "bad handle or splitting");
    int axes[3];
    double f[3];
    int ns = step_plan(splitting, n, axes, f);
    for (int s = 0; s < ns; ++s) {
        // the reference passes 0.5 * dt to the Strang x sweeps (splitting_scheme.hpp:90,116)
        double dts = f[s] == 1.0 ? dt : f[s] * dt;
        int rc = sweep_once(g, axes[s], false, dts, 1.0, 0);
        if (rc) return rc;
    }
    return 0;
}

int cfdart_compute_dt(cfdart_grid* g, int n, double t, double T, double cfl, double* dt_out, double* mass_out)
{
    if (!g || !dt_out) return fail(CFDART_E_ARG, "null argument");
    cfd::GridView v = view(g);
    hipError_t e = cfd::launch_reduce_dt(v, g->d.model, g->partials, g->st);
    if (e != hipSuccess) return fail(CFDART_E_HIP, hipGetErrorString(e));
    cfd::Ctrl c;
    std::vector<double> part(2 * (size_t)g->nred);
    HIPCHK(hipMemcpyAsync(&c, g->ctrl, sizeof(c), hipMemcpyDeviceToHost, g->st));
    HIPCHK(hipMemcpyAsync(part.data(), g->partials, sizeof(double) * part.size(), hipMemcpyDeviceToHost, g->st));
    HIPCHK(hipStreamSynchronize(g->st));
    HIPCHK(hipMemsetAsync(&g->ctrl->slots[0][0], 0, sizeof(g->ctrl->slots), g->st));
    unsigned long long mu = 0, mv = 0;
    for (int i = 0; i < cfd::MAX_SLOTS; ++i) {
        mu = c.maxu(i) > mu ? c.maxu(i) : mu;
        mv = c.maxv(i) > mv ? c.maxv(i) : mv;
    }
    double maxu, maxv;
    std::memcpy(&maxu, &mu, sizeof(double));
    std::memcpy(&maxv, &mv, sizeof(double));
    // E/sourcecode/euler_timestep.cpp:35-43
    double CFL = cfl;
    if (n < 5) CFL = cfd::smin(CFL, 0.2);
    double dt = CFL * cfd::smin(g->d.dx / maxu, g->d.dy / maxv);
    if (t + dt > T) dt = T - t;
    *dt_out = dt;
    if (mass_out) {
        double m1 = 0.0, m2 = 0.0;
        for (int b = 0; b < g->nred; ++b) { m1 += part[2 * b]; m2 += part[2 * b + 1]; }
        mass_out[0] = m1 * g->d.dx * g->d.dy;
        if (g->d.model == CFDART_ALLAIRE) mass_out[1] = m2 * g->d.dx * g->d.dy;
    }
    return 0;
}

int cfdart_set_time(cfdart_grid* g, double t, int n)
{
    if (!g) return fail(CFDART_E_ARG, "null handle");
    hipError_t e = cfd::launch_set_time(g->ctrl, t, n, g->st);
    if (e != hipSuccess) return fail(CFDART_E_HIP, hipGetErrorString(e));
    return 0;
}

int cfdart_run_mass(cfdart_grid* g, int splitting, int max_steps, double T, double cfl, double* t_out, int* n_out,
                    double* mass_hist)
{
    if (!g || splitting < 0 || splitting > 2 || max_steps < 0) return fail(CFDART_E_ARG, "bad argument");
    if (!(cfl > 0.0)) return fail(CFDART_E_ARG, "cfl must be positive");
    cfd::GridView v = view(g);
    hipError_t e = cfd::launch_set_run(g->ctrl, T, cfl, g->st);
    if (e == hipSuccess) e = cfd::launch_reduce_dt(v, g->d.model, g->partials, g->st);
    if (e == hipSuccess) e = cfd::launch_init_dt(g->ctrl, g->d.dx, g->d.dy, g->st);
    if (e != hipSuccess) return fail(CFDART_E_HIP, hipGetErrorString(e));
    // The alternating splitting needs n on the host to pick the sweep order; fixed-order
    // schemes do not. For alternating, step parity is tracked from the handle's n.
    int n0 = 0;
    if (splitting == CFDART_ALTERNATING) {
        cfd::Ctrl c;
        HIPCHK(hipMemcpyAsync(&c, g->ctrl, sizeof(c), hipMemcpyDeviceToHost, g->st));
        HIPCHK(hipStreamSynchronize(g->st));
        n0 = c.n;
    }
    double* d_hist = nullptr;
    if (mass_hist && max_steps > 0) {
        HIPCHK(hipMalloc(&d_hist, sizeof(double) * 3 * (size_t)max_steps));
        e = hipMemsetAsync(d_hist, 0, sizeof(double) * 3 * (size_t)max_steps, g->st);
        if (e != hipSuccess) {
            (void)hipFree(d_hist);
            return fail(CFDART_E_HIP, std::string("clearing the mass history: ") + hipGetErrorString(e));
        }
    }
    auto release = [&]() {
        if (d_hist) (void)hipFree(d_hist);
    };
    for (int k = 0; k < max_steps; ++k) {
        if (d_hist) {
            e = cfd::launch_mass_record(view(g), g->d.model, g->partials, d_hist + 3 * (size_t)k, g->st);
            if (e != hipSuccess) {
                release();
                return fail(CFDART_E_HIP, hipGetErrorString(e));
            }
        }
        int axes[3];
        double f[3];
        int ns = step_plan(splitting, n0 + k, axes, f);
        for (int s = 0; s < ns; ++s) {
            int rc = sweep_once(g, axes[s], s == ns - 1, 0.0, f[s], 1);
            if (rc) {
                release();
                return rc;
            }
        }
        e = cfd::launch_finalize(g->ctrl, g->d.dx, g->d.dy, g->st);
        if (e != hipSuccess) {
            release();
            return fail(CFDART_E_HIP, hipGetErrorString(e));
        }
    }
    if (d_hist) {
        hipError_t ce = hipMemcpyAsync(mass_hist, d_hist, sizeof(double) * 3 * (size_t)max_steps,
                                       hipMemcpyDeviceToHost, g->st);
        if (ce == hipSuccess) ce = hipStreamSynchronize(g->st);
        release();
        if (ce != hipSuccess) return fail(CFDART_E_HIP, hipGetErrorString(ce));
    }
    if (t_out || n_out) {
        cfd::Ctrl c;
        HIPCHK(hipMemcpyAsync(&c, g->ctrl, sizeof(c), hipMemcpyDeviceToHost, g->st));
        HIPCHK(hipStreamSynchronize(g->st));
        if (t_out) *t_out = c.t;
        if (n_out) *n_out = c.n;
        if (c.error) return fail(CFDART_E_UNPHYS, "an Allaire state failed the physical-state test");
    }
    return 0;
}

int cfdart_run(cfdart_grid* g, int splitting, int max_steps, double T, double cfl, double* t_out, int* n_out)
{
    return cfdart_run_mass(g, splitting, max_steps, T, cfl, t_out, n_out, nullptr);
}

int cfdart_stats(cfdart_grid* g, long long* counters)
{
    if (!g || !counters) return fail(CFDART_E_ARG, "null argument");
    cfd::Ctrl c;
    HIPCHK(hipMemcpyAsync(&c, g->ctrl, sizeof(c), hipMemcpyDeviceToHost, g->st));
    HIPCHK(hipStreamSynchronize(g->st));
    for (int i = 0; i < 4; ++i) counters[i] = (long long)c.cnt[i];
    return 0;
}

int cfdart_sync(cfdart_grid* g)
{
    if (!g) return fail(CFDART_E_ARG, "null handle");
    int err = 0;
    HIPCHK(hipMemcpyAsync(&err, &g->ctrl->error, sizeof(err), hipMemcpyDeviceToHost, g->st));
    HIPCHK(hipStreamSynchronize(g->st));
    if (err) return fail(CFDART_E_UNPHYS, "an Allaire state failed the physical-state test");
    return 0;
}

void* cfdart_stream(cfdart_grid* g) { return g ? (void*)g->st : nullptr; }

int cfdart_set_profiling(cfdart_grid* g, int enable)
{
    if (!g) return fail(CFDART_E_ARG, "null handle");
    HIPCHK(hipStreamSynchronize(g->st));
    harvest_profile(g);
    for (int i = 0; i < 3; ++i) { g->prof_ms[i] = 0.0; g->prof_n[i] = 0; }
    g->profiling = enable ? 1 : 0;
    return 0;
}

int cfdart_profile(cfdart_grid* g, double* ms_total, long long* launches)
{
    if (!g || !ms_total || !launches) return fail(CFDART_E_ARG, "null argument");
    HIPCHK(hipStreamSynchronize(g->st));
    harvest_profile(g);
    for (int i = 0; i < 2; ++i) { ms_total[i] = g->prof_ms[i]; launches[i] = g->prof_n[i]; }
    return 0;
}

int cfdart_set_strip_rows(cfdart_grid* g, int rows)
{
    if (!g || rows < 1 || rows > 4096) return fail(CFDART_E_ARG, "rows must be in [1, 4096]");
    g->yH = rows;
    return 0;
}

int cfdart_set_tile(cfdart_grid* g, int y_strip_rows, int x_rows_per_wave)
{
    if (!g || (y_strip_rows < 1 && !cfd::y_transposed(y_strip_rows)) || y_strip_rows > 4096 ||
        x_rows_per_wave < 1 || x_rows_per_wave > 256)
        return fail(CFDART_E_ARG, "tile shape out of range");
    g->yH = y_strip_rows;
    g->xrows = x_rows_per_wave;
    return 0;
}

int cfdart_set_arith(cfdart_grid* g, int mode)
{
    if (!g || (mode != CFDART_ARITH_EXACT && mode != CFDART_ARITH_FAST)) return fail(CFDART_E_ARG, "bad handle or arithmetic mode");
    g->arith = mode;
    return 0;
}

}  // extern "C"

// ------------------------------------------------------------------ the time loop without allocation

namespace {

// The loop's control block and first-step reduction partials for cfdart_run_noalloc: a file-scope
// device array, one per GPU, held until the process ends. Every call on a GPU uses that GPU's
// array and re-initialises it first, so calls that overlap on one GPU race.
constexpr size_t NOALLOC_SCRATCH_BYTES = 32 * 1024;
constexpr size_t NOALLOC_CTRL_BYTES = (sizeof(cfd::Ctrl) + 255) / 256 * 256;
static_assert(sizeof(cfd::Ctrl) % 8 == 0, "the control block is cleared as 64-bit words");
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

std::mutex g_scratch_mu;
std::vector<void*> g_scratch_addr;  // per device: the array's device address, looked up once

// The array's address on the current device (host code may not use the array's name as an
// address: each device holds its own copy at its own address).
int scratch_address(void** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(CFDART_E_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((size_t)dev >= g_scratch_addr.size()) g_scratch_addr.resize((size_t)dev + 1, nullptr);
    if (!g_scratch_addr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(CFDART_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_addr[dev] = p;
    }
    *out = g_scratch_addr[dev];
    return 0;
}

// A fresh loop state: every word of the control block zero (counters, wave-speed slots, error
// flag), then t, n, T and CFL. A kernel rather than a memset, so that a captured graph replays it.
__global__ void noalloc_init_kernel(cfd::Ctrl* c, double t, int n, double T, double cfl)
{
    unsigned long long* w = reinterpret_cast<unsigned long long*>(c);
    for (int i = threadIdx.x; i < (int)(sizeof(cfd::Ctrl) / 8); i += blockDim.x) w[i] = 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        c->t = t;
        c->n = n;
        c->T = T;
        c->cfl = cfl;
    }
}

}  // namespace

extern "C" int cfdart_run_noalloc(const cfdart_desc* d, double* state, double* work, int pitch, int arith,
                                  int splitting, int max_steps, double T, double cfl, double t0, int n0, void* stream,
                                  double* t_out, int* n_out)
{
    int rc = validate_desc(d);
    if (rc) return rc;
    if (!state || !work || state == work) return fail(CFDART_E_ARG, "state and work must be two distinct device buffers");
    if (pitch < d->nx) return fail(CFDART_E_ARG, "pitch must be at least nx");
    if (arith != CFDART_ARITH_EXACT && arith != CFDART_ARITH_FAST) return fail(CFDART_E_ARG, "bad arithmetic mode");
    if (splitting < 0 || splitting > 2 || max_steps < 0 || !(cfl > 0.0) || n0 < 0) return fail(CFDART_E_ARG, "bad argument");
    const long long plane = (long long)pitch * d->ny;
    const int nc = d->model == CFDART_ALLAIRE ? 6 : 4;
    const int nred = cfd::reduce_blocks(make_view(*d, state, work, pitch, plane, arith, nullptr));
    const size_t need = NOALLOC_CTRL_BYTES + sizeof(double) * 2 * (size_t)nred;
    if (need > NOALLOC_SCRATCH_BYTES)
        return fail(CFDART_E_ARG, "the loop needs " + std::to_string(need) + " bytes of scratch, beyond the " +
                                      std::to_string(NOALLOC_SCRATCH_BYTES) + " the library holds");
    void* base = nullptr;
    if ((rc = scratch_address(&base)) != 0) return rc;
    cfd::Ctrl* ctrl = static_cast<cfd::Ctrl*>(base);
    double* partials = reinterpret_cast<double*>(static_cast<char*>(base) + NOALLOC_CTRL_BYTES);
    hipStream_t st = static_cast<hipStream_t>(stream);

    hipLaunchKernelGGL(noalloc_init_kernel, dim3(1), dim3(64), 0, st, ctrl, t0, n0, T, cfl);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = cfd::launch_reduce_dt(make_view(*d, state, work, pitch, plane, arith, ctrl), d->model, partials, st);
    if (e == hipSuccess) e = cfd::launch_init_dt(ctrl, d->dx, d->dy, st);
    double* cur = state;
    double* nxt = work;
    for (int k = 0; k < max_steps && e == hipSuccess; ++k) {
        int axes[3];
        double f[3];
        const int ns = step_plan(splitting, n0 + k, axes, f);
        for (int s = 0; s < ns && e == hipSuccess; ++s) {
            e = cfd::launch_sweep(make_view(*d, cur, nxt, pitch, plane, arith, ctrl), d->model, axes[s], s == ns - 1,
                                  0.0, f[s], 1, CFDART_Y_TRANSPOSED, 4, st);
            std::swap(cur, nxt);
        }
        if (e == hipSuccess) e = cfd::launch_finalize(ctrl, d->dx, d->dy, st);
    }
    // an odd number of sweeps (Strang) leaves the state in the work buffer
    if (e == hipSuccess && cur != state)
        e = hipMemcpyAsync(state, cur, sizeof(double) * (size_t)plane * nc, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return fail(CFDART_E_HIP, std::string("cfdart_run_noalloc: ") + hipGetErrorString(e));
    if (t_out || n_out) {
        cfd::Ctrl c;
        HIPCHK(hipMemcpyAsync(&c, ctrl, sizeof(c), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (t_out) *t_out = c.t;
        if (n_out) *n_out = c.n;
        if (c.error) return fail(CFDART_E_UNPHYS, "an Allaire state failed the physical-state test");
    }
    return 0;
}
