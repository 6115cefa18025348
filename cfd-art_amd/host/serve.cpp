// Serving: request traces, request setup, the continuous-batching loop and the one-at-a-time
// comparison (include/cfdart_serve.h). Host code only; every GPU action goes through the C-ABI
// of cfdart.h and cfdart_batch.h.
#include "cfdart_serve.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "cfdart_batch.h"

namespace {

using Clock = std::chrono::steady_clock;

double seconds_since(Clock::time_point t0) { return std::chrono::duration<double>(Clock::now() - t0).count(); }

struct Unit {
    std::mt19937_64 g;
    explicit Unit(unsigned long long s) : g(s) {}
    double operator()() { return double(g() >> 11) * 0x1.0p-53; }
};

int ncomp(int model) { return model == CFDART_ALLAIRE ? 6 : 4; }

size_t frame_bytes(const cfdart_desc& d) { return sizeof(double) * (size_t)d.nx * d.ny * ncomp(d.model); }

// Pinned host memory that frames land in, reused cyclically (the frames are not read back:
// a writer would take them from here).
struct FrameRing {
    char* base = nullptr;
    size_t cap = 0, head = 0;
    explicit FrameRing(size_t bytes)
    {
        if (bytes) {
            base = (char*)cfdart_host_alloc(bytes);
            cap = base ? bytes : 0;
        }
    }
    ~FrameRing() { cfdart_host_free(base); }
    double* take(size_t bytes)
    {
        if (bytes > cap) return nullptr;
        if (head + bytes > cap) head = 0;
        double* p = (double*)(base + head);
        head += (bytes + 255) / 256 * 256;
        return p;
    }
};

// Pinned host array of checksums that the device writes (cfdart_batch_checksum_async).
struct PinnedSums {
    unsigned long long* p = nullptr;
    explicit PinnedSums(size_t n)
    {
        if (n) p = (unsigned long long*)cfdart_host_alloc(sizeof(unsigned long long) * n);
        if (p) std::memset(p, 0, sizeof(unsigned long long) * n);
    }
    ~PinnedSums() { cfdart_host_free(p); }
    unsigned long long& operator[](size_t i) { return p[i]; }
};

}  // namespace

extern "C" {

int cfdart_serve_trace(unsigned long long seed, double rate, double duration, double euler_frac, int n_lo, int n_hi,
                       int steps_lo, int steps_hi, cfdart_request* out, int max_out, int* n_out)
{
    if (!out || !n_out || max_out < 0 || !(rate > 0) || n_lo < 2 || n_hi < n_lo || steps_lo < 1 || steps_hi < steps_lo)
        return CFDART_E_ARG;
    Unit u(seed);
    double t = 0.0;
    int k = 0;
    const double a = std::log((double)n_lo), b = std::log((double)n_hi);
    while (k < max_out) {
        t += -std::log(1.0 - u()) / rate;
        if (t > duration) break;
        cfdart_request& r = out[k];
        r.arrival = t;
        r.model = u() < euler_frac ? CFDART_EULER : CFDART_ALLAIRE;
        r.n = (int)std::lround(std::exp(a + u() * (b - a)));
        r.steps = steps_lo + (int)std::floor(u() * (steps_hi - steps_lo + 1));
        if (r.steps > steps_hi) r.steps = steps_hi;
        r.seed = seed * 1000003ULL + (unsigned long long)k;
        ++k;
    }
    *n_out = k;
    return 0;
}

int cfdart_request_setup(const cfdart_request* r, cfdart_desc* d, cfdart_disc_params* dp, cfdart_bubble_params* bp,
                         double* cfl)
{
    if (!r || !d || !dp || !bp || !cfl || r->n < 2 || r->steps < 1) return CFDART_E_ARG;
    std::memset(d, 0, sizeof(*d));
    d->model = r->model;
    d->nx = d->ny = r->n;
    d->ng = 2;
    d->x0 = d->y0 = 0.0;
    d->dx = d->dy = 1.0 / r->n;
    d->colsweep_uses_dx = 1;
    for (int i = 0; i < 4; ++i) d->bc[i] = CFDART_TRANSMISSIVE;
    if (r->model == CFDART_EULER) {
        d->gamma1 = 1.4;
        d->gamma2 = 1.4;
        d->limiter = CFDART_VANLEER;
        d->wave_speeds = CFDART_GARRICK;
        d->z_lo = 0.0;
        d->z_hi = 1.0;
        std::memset(dp, 0, sizeof(*dp));
        dp->seed = r->seed;
        dp->n = 8;
        dp->cx_lo = dp->cy_lo = 0.1;
        dp->cx_hi = dp->cy_hi = 0.9;
        dp->r_lo = 0.03;
        dp->r_hi = 0.10;
        dp->p_lo = 5.0;
        dp->p_hi = 20.0;
        dp->rho_lo = 0.5;
        dp->rho_hi = 2.0;
        const double amb[4] = {1.0, 0.0, 0.0, 1.0};
        for (int i = 0; i < 4; ++i) dp->ambient[i] = amb[i];
        *cfl = 0.8;
    } else if (r->model == CFDART_ALLAIRE) {
        const double eps = 1e-6, zw = 1.0 - eps, za = eps;
        d->gamma1 = 4.4;
        d->pinf1 = 6e8;
        d->gamma2 = 1.4;
        d->pinf2 = 0.0;
        d->limiter = CFDART_MINMOD;
        d->wave_speeds = CFDART_PVRS;
        d->z_lo = eps;
        d->z_hi = 1.0 - eps;
        std::memset(bp, 0, sizeof(*bp));
        bp->seed = r->seed;
        bp->n = 16;
        bp->cx_lo = 0.15;
        bp->cx_hi = 0.95;
        bp->cy_lo = 0.05;
        bp->cy_hi = 0.95;
        bp->r_lo = 0.02;
        bp->r_hi = 0.06;
        const double water[6] = {zw * 1000.0, (1 - zw) * 1.0, 0.0, 0.0, 1e5, zw};
        const double air[6] = {za * 1000.0, (1 - za) * 1.0, 0.0, 0.0, 1e5, za};
        const double post[6] = {zw * 1323.65, (1 - zw) * 1.0, 681.58, 0.0, 1e9, zw};
        for (int i = 0; i < 6; ++i) {
            bp->water[i] = water[i];
            bp->air[i] = air[i];
            bp->postshock[i] = post[i];
        }
        bp->shock_x = 0.1;
        *cfl = 0.5;
    } else {
        return CFDART_E_ARG;
    }
    return 0;
}

int cfdart_serve_batched(const cfdart_request* reqs, int n, const cfdart_serve_opts* o, cfdart_request_result* out,
                         cfdart_serve_summary* sum)
{
    if (!reqs || n < 0 || !o || !out || !sum || o->capacity < 1 || o->chunk_steps < 1) return CFDART_E_ARG;
    if (o->frame_every > 0 && o->frame_every % o->chunk_steps != 0) return CFDART_E_ARG;
    for (int i = 1; i < n; ++i)
        if (reqs[i].arrival < reqs[i - 1].arrival) return CFDART_E_ARG;
    std::memset(sum, 0, sizeof(*sum));
    std::memset(out, 0, sizeof(cfdart_request_result) * (size_t)n);
    cfdart_batch* b = nullptr;
    int rc = cfdart_batch_create(o->capacity, o->mem_budget, &b);
    if (rc) return rc;
    if ((rc = cfdart_batch_set_arith(b, o->arith)) != 0) {
        cfdart_batch_destroy(b);
        return rc;
    }
    FrameRing ring(o->frame_every > 0 ? o->frame_host_bytes : 0);
    std::vector<int> owner(o->capacity, -1), steps(o->capacity), flags(o->capacity);
    std::vector<int> fslots, fstatus;  // the frames of a chunk
    std::vector<double*> fdsts;
    std::vector<size_t> fbytes;
    PinnedSums csum((size_t)n);  // the finished requests' checksums, written by the device
    std::vector<char> summed((size_t)n, 0);
    if (n > 0 && !csum.p) {
        cfdart_batch_destroy(b);
        return CFDART_E_NOMEM;
    }
    std::deque<int> waiting;
    int next = 0, done = 0, live = 0;
    long long live_cells = 0;
    size_t held = 0;  // device bytes the batch counts against the budget (states and frame ring)
    const Clock::time_point t0 = Clock::now();

    auto fail_all = [&](int code) {
        cfdart_batch_destroy(b);
        return code;
    };
    double reported = 0.0;
    while (done < n) {
        const double now = seconds_since(t0);
        if (o->progress_every > 0 && now - reported >= o->progress_every) {
            std::fprintf(stderr, "serve batched: %.1f s, %d/%d requests finished, %d live, %zu waiting\n", now, done, n,
                         live, waiting.size() + (size_t)(n - next));
            std::fflush(stderr);
            reported = now;
        }
        while (next < n && (!o->realtime || reqs[next].arrival <= now)) waiting.push_back(next++);
        // admission, first come first served: stop at the first request that does not fit
        while (!waiting.empty()) {
            const int r = waiting.front();
            cfdart_desc d;
            cfdart_disc_params dp;
            cfdart_bubble_params bp;
            double cfl;
            int slot = -1;
            int e = cfdart_request_setup(&reqs[r], &d, &dp, &bp, &cfl);
            const long long cells = (long long)reqs[r].n * reqs[r].n;
            if (!e && o->live_cells > 0 && live > 0 && live_cells + cells > o->live_cells) break;
            if (!e) {
                const int m = d.model;
                e = cfdart_batch_add_generated(b, &d, m == CFDART_EULER ? &dp : nullptr,
                                               m == CFDART_ALLAIRE ? &bp : nullptr, CFDART_ALTERNATING,
                                               reqs[r].steps, INFINITY, cfl, o->y_strip_rows[m],
                                               o->x_rows_per_wave[m], &slot);
            }
            if (e == CFDART_E_NOMEM && live > 0) break;  // wait for memory or a slot to free up
            if (e) {
                // cannot run even on an empty batch, or a bad request
                out[r].status = e;
                out[r].admitted = out[r].finished = seconds_since(t0);
                waiting.pop_front();
                ++done;
                continue;
            }
            waiting.pop_front();
            owner[slot] = r;
            out[r].admitted = seconds_since(t0);
            cfdart_batch_info(b, nullptr, &held, nullptr);
            ++live;
            live_cells += cells;
        }
        sum->max_live = std::max(sum->max_live, live);
        sum->peak_state_bytes = std::max(sum->peak_state_bytes, (double)held);
        if (live == 0) {
            if (next < n && o->realtime) {
                const double wait = reqs[next].arrival - seconds_since(t0);
                if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
            }
            continue;
        }
        if ((rc = cfdart_batch_step(b, o->chunk_steps)) != 0) return fail_all(rc);
        if ((rc = cfdart_batch_poll(b, steps.data(), flags.data())) != 0) return fail_all(rc);
        sum->chunks += 1;
        // a frame each time a grid's step count reaches a multiple of frame_every: the chunk's
        // frames go to the batch as one group, before any grid of the chunk is retired
        fslots.clear();
        fdsts.clear();
        fbytes.clear();
        for (int s = 0; s < o->capacity && o->frame_every > 0; ++s) {
            const int r = owner[s];
            if (r < 0) continue;
            cfdart_request_result& res = out[r];
            if (steps[s] / o->frame_every > res.frames && steps[s] % o->frame_every == 0) {
                cfdart_desc d;
                cfdart_disc_params dp;
                cfdart_bubble_params bp;
                double cfl;
                cfdart_request_setup(&reqs[r], &d, &dp, &bp, &cfl);
                const size_t fb = frame_bytes(d);
                // a frame that finds no room (host ring smaller than the frame, or a device staging
                // ring the memory budget cannot hold) is skipped; the grid keeps running
                double* dst = ring.take(fb);
                if (dst) {
                    fslots.push_back(s);
                    fdsts.push_back(dst);
                    fbytes.push_back(fb);
                }
                res.frames = steps[s] / o->frame_every;
            }
        }
        if (!fslots.empty()) {
            fstatus.assign(fslots.size(), 0);
            rc = cfdart_batch_frames(b, (int)fslots.size(), fslots.data(), fdsts.data(), fstatus.data());
            if (rc) return fail_all(rc);
            for (size_t i = 0; i < fslots.size(); ++i)
                if (fstatus[i] == 0) {
                    sum->frames += 1;
                    sum->frame_bytes += (double)fbytes[i];
                }
            cfdart_batch_info(b, nullptr, &held, nullptr);
            sum->peak_state_bytes = std::max(sum->peak_state_bytes, (double)held);
        }
        for (int s = 0; s < o->capacity; ++s) {
            const int r = owner[s];
            if (r < 0) continue;
            cfdart_request_result& res = out[r];
            if (!(flags[s] & 1)) {
                res.steps_done = steps[s];
                res.status = (flags[s] & 2) ? CFDART_E_UNPHYS : 0;
                // the checksum lands in pinned memory once the stream has passed it (read at the end)
                if ((rc = cfdart_batch_checksum_async(b, s, &csum[r])) != 0) return fail_all(rc);
                summed[r] = 1;
                res.finished = seconds_since(t0);
                res.service = res.finished - res.admitted;
                sum->cell_updates += (double)reqs[r].n * reqs[r].n * steps[s];
                if ((rc = cfdart_batch_remove(b, s)) != 0) return fail_all(rc);
                cfdart_batch_info(b, nullptr, &held, nullptr);
                owner[s] = -1;
                --live;
                live_cells -= (long long)reqs[r].n * reqs[r].n;
                ++done;
            }
        }
    }
    rc = cfdart_batch_sync(b);
    sum->makespan = seconds_since(t0);
    for (int r = 0; r < n && rc == 0; ++r)
        if (summed[r]) out[r].checksum = csum[r];
    cfdart_batch_destroy(b);
    return rc;
}

int cfdart_serve_one_at_a_time(const cfdart_request* reqs, int n, const cfdart_serve_opts* o,
                               cfdart_request_result* out, cfdart_serve_summary* sum)
{
    if (!reqs || n < 0 || !o || !out || !sum) return CFDART_E_ARG;
    std::memset(sum, 0, sizeof(*sum));
    std::memset(out, 0, sizeof(cfdart_request_result) * (size_t)n);
    FrameRing ring(o->frame_every > 0 ? o->frame_host_bytes : 0);
    double clock = 0.0;  // the schedule a one-at-a-time server would follow
    const Clock::time_point start = Clock::now();
    double reported = 0.0;
    for (int r = 0; r < n; ++r) {
        if (o->progress_every > 0 && seconds_since(start) - reported >= o->progress_every) {
            reported = seconds_since(start);
            std::fprintf(stderr, "serve one at a time: %.1f s, %d/%d requests, schedule at %.1f s\n", reported, r, n, clock);
            std::fflush(stderr);
        }
        cfdart_desc d;
        cfdart_disc_params dp;
        cfdart_bubble_params bp;
        double cfl;
        cfdart_request_result& res = out[r];
        int rc = cfdart_request_setup(&reqs[r], &d, &dp, &bp, &cfl);
        if (rc) {
            res.status = rc;
            continue;
        }
        // the grid's buffers are allocated before and freed after the timed service, as a server
        // that pools its buffers would (the batch likewise keeps its pool across requests)
        cfdart_grid* g = nullptr;
        if ((rc = cfdart_create(&d, &g)) != 0) {
            res.status = rc;
            continue;
        }
        const Clock::time_point t0 = Clock::now();
        rc = d.model == CFDART_EULER ? cfdart_init_random_discs(g, &dp) : cfdart_init_shocked_bubbles(g, &bp);
        if (!rc) rc = cfdart_set_arith(g, o->arith);
        const int m = d.model;
        if (!rc && (o->plain_y_strip_rows[m] != 0 || o->plain_x_rows_per_wave[m] != 0))
            rc = cfdart_set_tile(g, o->plain_y_strip_rows[m] != 0 ? o->plain_y_strip_rows[m] : CFDART_Y_TRANSPOSED,
                                 o->plain_x_rows_per_wave[m] != 0 ? o->plain_x_rows_per_wave[m] : 4);
        int taken = 0;
        const int every = o->frame_every > 0 ? o->frame_every : reqs[r].steps;
        while (!rc && taken < reqs[r].steps) {
            const int k = std::min(every, reqs[r].steps - taken);
            double t;
            int nn = taken;
            rc = cfdart_run(g, CFDART_ALTERNATING, k, INFINITY, cfl, &t, &nn);
            taken = nn;  // also when the run stopped on an unphysical state, as the batch records it
            if (rc) break;
            if (o->frame_every > 0 && taken % o->frame_every == 0) {
                double* dst = ring.take(frame_bytes(d));
                if (dst && (rc = cfdart_frame(g, dst)) == 0) {
                    res.frames += 1;
                    sum->frames += 1;
                    sum->frame_bytes += (double)frame_bytes(d);
                }
            }
        }
        // a grid stopped by an unphysical state still reports its checksum, as the batch does
        if (!rc || rc == CFDART_E_UNPHYS) {
            const int c = cfdart_checksum(g, &res.checksum);
            if (c) rc = c;
        }
        if (!rc) rc = cfdart_sync(g);
        res.service = seconds_since(t0);
        cfdart_destroy(g);
        res.status = rc;
        res.steps_done = taken;
        res.admitted = std::max(clock, reqs[r].arrival);
        res.finished = res.admitted + res.service;
        clock = res.finished;
        sum->cell_updates += (double)reqs[r].n * reqs[r].n * taken;
    }
    sum->makespan = clock;
    return 0;
}

}  // extern "C"
