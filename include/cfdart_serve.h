/*
 * cfdart serving: a request loop with continuous batching over cfdart_batch.h.
 *
 * No counterpart in the reference, which runs one simulation per process to completion
 * (euler_equations/euler_simulation_code/sourcecode/main.cpp:17-49). A request is one
 * simulation: a square grid, a model, a seeded initial condition and a step count. The
 * batched server keeps every admitted simulation's state resident in HBM, admits waiting
 * requests (first come, first served) whenever the memory budget and the slot capacity allow,
 * advances all live grids together in chunks of whole time steps, writes a frame of each grid
 * every frame_every of its steps to pinned host memory while stepping continues, and retires
 * grids as they finish. Each request's final state is bitwise the state the plain single-grid
 * solver (cfdart.h) produces for the same request alone, in the same arithmetic (opts.arith); cfdart_serve_one_at_a_time runs that
 * plain solver request by request for comparison.
 *
 * Conventions as in cfdart.h: 0 or a negative CFDART_E_* code.
 */
#ifndef CFDART_SERVE_H
#define CFDART_SERVE_H

#include <stddef.h>

#include "cfdart.h"
#include "cfdart_host.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cfdart_request {
    double arrival;          /* seconds after the start of the replay */
    int model;               /* CFDART_EULER or CFDART_ALLAIRE */
    int n;                   /* grid side: n x n cells on [0,1]^2 */
    int steps;               /* time steps to take */
    unsigned long long seed; /* initial-condition seed */
} cfdart_request;

/* Poisson request trace from std::mt19937_64(seed), draws u = (x >> 11) * 2^-53, per request in
 * this order: inter-arrival time -ln(1-u)/rate, model (Euler if u < euler_frac), side
 * round(exp(ln n_lo + u (ln n_hi - ln n_lo))), steps steps_lo + floor(u (steps_hi - steps_lo + 1));
 * request i gets the seed seed * 1000003 + i. Arrivals run until `duration` seconds or max_out
 * requests; *n_out receives the count. */
int cfdart_serve_trace(unsigned long long seed, double rate, double duration, double euler_frac, int n_lo,
                       int n_hi, int steps_lo, int steps_hi, cfdart_request* out, int max_out, int* n_out);

/* The grid, initial condition and CFL number of a request. Euler: gamma 1.4, transmissive
 * sides, van Leer + HLLC, 8 discs (centres U[0.1,0.9]^2, radii U[0.03,0.10], p U[5,20],
 * rho U[0.5,2]) on (rho,u,v,p) = (1,0,0,1), CFL 0.8; *bp is unused. Allaire: water (gamma 4.4,
 * p_inf 6e8, rho 1000) with 16 air bubbles (gamma 1.4, rho 1; centres U[0.15,0.95]x[0.05,0.95],
 * radii U[0.02,0.06]) at p 1e5 behind a planar shock at x = 0.1 (post-shock water p 1e9,
 * rho 1323.65, u 681.58), z clipped to [1e-6, 1-1e-6], minmod + PVRS-HLLC, CFL 0.5; *dp is
 * unused. Both use alternating x/y splitting. */
int cfdart_request_setup(const cfdart_request* r, cfdart_desc* d, cfdart_disc_params* dp,
                         cfdart_bubble_params* bp, double* cfl);

typedef struct cfdart_serve_opts {
    size_t mem_budget;       /* bytes of grid state admitted at once (two buffers per grid) */
    int capacity;            /* maximum number of live grids */
    int chunk_steps;         /* batch steps between scheduling decisions; must divide frame_every.
                                An even count admits every grid at an even batch step, so all live
                                grids sweep the same axis first and each step launches only the
                                sweeps with work (odd counts give the same results, more slowly) */
    int frame_every;         /* steps between frames of one grid; 0 = no frames */
    size_t frame_host_bytes; /* pinned host frame ring; frames overwrite it cyclically */
    int realtime;            /* 1: a request is visible from its arrival time on; 0: all at once */
    int y_strip_rows[2];     /* batched tile shape per model (Euler, Allaire), as cfdart_batch_add;
                                0 = default */
    int x_rows_per_wave[2];
    int plain_y_strip_rows[2]; /* one-at-a-time tile shape per model, as cfdart_set_tile
                                  (transposed y tiles included); 0 = default */
    int plain_x_rows_per_wave[2];
    long long live_cells;    /* batched: admit while the live grids hold at most this many cells
                                (enough to fill the GPU; the rest wait first come first served,
                                which shortens each admitted request's latency); 0 = no limit */
    int arith;               /* CFDART_ARITH_EXACT (0) or CFDART_ARITH_FAST, batched and one at a time */
    double progress_every;   /* > 0: a progress line on stderr at most this many seconds apart
                                (requests finished, live, waiting); 0 = silent */
} cfdart_serve_opts;

typedef struct cfdart_request_result {
    double admitted;              /* seconds from the start: admission (batched) or start (one at a time) */
    double finished;              /* seconds from the start when the result was available */
    double service;               /* one at a time: the request's own solve time; batched: finished - admitted */
    int steps_done;
    int frames;
    int status;                   /* 0, or CFDART_E_* (E_NOMEM: larger than the whole budget) */
    unsigned long long checksum;  /* cfdart_checksum of the final state */
} cfdart_request_result;

typedef struct cfdart_serve_summary {
    double makespan;       /* seconds from the start to the last result */
    double cell_updates;   /* sum over requests of n*n*steps_done */
    double frame_bytes;    /* bytes of frames copied to the host */
    long long frames;
    long long chunks;      /* batched: scheduling rounds */
    int max_live;          /* batched: most grids live at once */
    double peak_state_bytes; /* batched: most device bytes held at once (grid states and the
                                frame staging ring, as cfdart_batch_info counts them) */
} cfdart_serve_summary;

/* Continuous batching over reqs (sorted by arrival). out[i] belongs to reqs[i]. */
int cfdart_serve_batched(const cfdart_request* reqs, int n, const cfdart_serve_opts* o, cfdart_request_result* out,
                         cfdart_serve_summary* sum);
/* The same requests one at a time with the plain single-grid solver (cfdart.h): create, generate
 * the state on the device, run with the same frames, checksum, destroy. Requests are solved
 * back to back, each timed on its own (service: state generation to checksum; the grid's
 * allocation and release fall outside it, as for a server that pools its buffers);
 * admitted/finished are the first-come-first-served schedule of those service times against the
 * arrival times, i.e. what a server that solves one request at a time would observe. */
int cfdart_serve_one_at_a_time(const cfdart_request* reqs, int n, const cfdart_serve_opts* o,
                               cfdart_request_result* out, cfdart_serve_summary* sum);

#ifdef __cplusplus
}
#endif
#endif
