/*
 * cfdart: C-ABI of the MI355X (gfx950) MUSCL-Hancock sweep library.
 *
 * The reference (cfd-art) drives its hot path through the problem_base / splitting_scheme
 * C++ interfaces, one grid row or column per call:
 *   problem_base::{pre_sweep, update_row, update_col, post_sweep, compute_dt}
 *     euler_equations/euler_simulation_code/headers/problem_base.hpp:22-39
 *   splitting_scheme_base::advance_timestep
 *     euler_equations/euler_simulation_code/headers/splitting_scheme.hpp:26
 * A row is far too small a unit for a GPU, so this boundary sits one level up: whole sweeps,
 * whole steps and whole runs over an opaque device-resident grid. Plain C types only.
 *
 * Conventions: every call returns 0 or a negative CFDART_E_* code and never aborts the
 * process; cfdart_last_error() describes the last failure of the calling thread. A handle owns
 * its device memory; the caller owns host buffers. One host thread per handle at a time; each
 * handle has its own HIP stream; calls are asynchronous on it except those that return host
 * data (download, compute_dt, stats, run when t_out is non-null, sync).
 *
 * Host state format (upload/download): the reference's grid[i][j](k) including its ghost frame
 * (euler_equations/euler_simulation_code/headers/typedefs.hpp:19-21), row-major AoS:
 *   host[((i * (nx + 2*ng)) + j) * ncomp + k],  i in [0, ny+2ng), j in [0, nx+2ng).
 * Euler: k = rho, rho*u, rho*v, E (ncomp 4). Allaire: z*rho1, (1-z)*rho2, rho*u, rho*v, E, z (6).
 */
#ifndef CFDART_H
#define CFDART_H

#ifdef __cplusplus
extern "C" {
#endif

#define CFDART_E_ARG     (-1) /* invalid argument or descriptor */
#define CFDART_E_NOMEM   (-2) /* device or host allocation failed */
#define CFDART_E_HIP     (-3) /* HIP runtime error */
#define CFDART_E_UNPHYS  (-4) /* an Allaire state failed the physical-state test (reference: assert abort) */
#define CFDART_E_NODEV   (-5) /* no usable gfx950 device */
#define CFDART_E_STATE   (-6) /* call not valid in the handle's current state */

enum { CFDART_EULER = 0, CFDART_ALLAIRE = 1 };
enum { CFDART_TRANSMISSIVE = 0, CFDART_REFLECTIVE = 1 };
enum { CFDART_MINMOD = 0, CFDART_VANLEER = 1 };
enum { CFDART_GARRICK = 0, CFDART_PVRS = 1 };
enum { CFDART_NAIVE = 0, CFDART_STRANG = 1, CFDART_ALTERNATING = 2 };
enum { CFDART_MUSCL_HANCOCK = 0, CFDART_GODUNOV = 1 };

/* Grid and scheme description; mirrors the reference's sim_info + SGparams/binarySGparams +
 * the flux-solver choices (euler_equations/euler_simulation_code/headers/sim_info.hpp:16-63,
 * headers/stiffened_gas_eos.hpp:20-50, sourcecode/euler_io.cpp:13-234,436-491). */
typedef struct cfdart_desc {
    int model;            /* CFDART_EULER or CFDART_ALLAIRE */
    int nx, ny, ng;       /* interior cells (>= 2 each) and ghost width of the host format:
                             2 for MUSCL-Hancock, 1 (the reference's) or 2 for Godunov */
    double x0, y0, dx, dy;
    double gamma1, pinf1; /* Euler material / Allaire fluid 1 */
    double gamma2, pinf2; /* Allaire fluid 2 (ignored for Euler) */
    int bc[4];            /* left, right, bottom, top: CFDART_TRANSMISSIVE / CFDART_REFLECTIVE */
    int limiter;          /* CFDART_MINMOD (reference) or CFDART_VANLEER */
    int wave_speeds;      /* CFDART_GARRICK (reference) or CFDART_PVRS (Allaire only) */
    int colsweep_uses_dx; /* 1 = reference: the column sweep's half-step uses dx */
    double z_lo, z_hi;    /* Allaire volume-fraction clamp (reference 0 and 1) */
    int flux_solver;      /* CFDART_MUSCL_HANCOCK or CFDART_GODUNOV: first order, the Riemann
                             problem of the two cells' own states and, for Allaire, the upwind
                             z update (flux_solver_godunov.hpp:18-41, A/headers/zupdate.hpp:40-55) */
} cfdart_desc;

typedef struct cfdart_grid cfdart_grid;

/* Counters (cfdart_stats): reference diagnostics that it prints instead of counting. */
enum {
    CFDART_CNT_SLOPE_FALLBACK = 0,  /* flux_solver_MUSCLhancock.hpp:184-189 */
    CFDART_CNT_EVOLVE_FALLBACK = 1, /* flux_solver_MUSCLhancock.hpp:216-223 */
    CFDART_CNT_REPAIR = 2,          /* euler_rowsweep.cpp:49-81 */
    CFDART_CNT_UNPHYSICAL = 3       /* allaire_diffuse_rowsweep.cpp:49 (assert) */
};

const char* cfdart_last_error(void);
int cfdart_device_count(void);

/* Allocates two device state buffers and the control block. Replaces the grid allocation of
 * set_ICs (euler_equations/euler_simulation_code/sourcecode/euler_io.cpp:494). */
int cfdart_create(const cfdart_desc* desc, cfdart_grid** out);
void cfdart_destroy(cfdart_grid* g);

/* Interior of a host grid (ghost frame ignored) -> device; resets t, n and the counters. */
int cfdart_upload(cfdart_grid* g, const double* host_state);
/* Device -> host grid; writes the interior and fills the ghost frame as pre_sweep would
 * (euler_equations/euler_simulation_code/sourcecode/euler_ppsweep.cpp:7-113; corners untouched). */
int cfdart_download(cfdart_grid* g, double* host_state);

/* One sweep: pre_sweep + update_row over all rows (axis 0) or update_col over all columns
 * (axis 1) + post_sweep (splitting_scheme.hpp:39-50,53-66). */
int cfdart_sweep(cfdart_grid* g, int axis, double dt);
/* compute_dt on the current state (euler_timestep.cpp:7-44, race-free max): synchronous.
 * mass_out (nullable) receives sum(rho)*dx*dy (Euler) or the two phase masses (Allaire). */
int cfdart_compute_dt(cfdart_grid* g, int n, double t, double T, double cfl, double* dt_out,
                      double* mass_out);
/* One time step with a host-given dt: splitting_scheme::advance_timestep
 * (splitting_scheme.hpp:37-67 naive, :78-122 Strang; alternating = XY on even n, YX on odd n). */
int cfdart_step(cfdart_grid* g, int splitting, int n, double dt);
/* The time loop of simulation::run_simulation (simulation.hpp:48-61) on the device: from the
 * handle's (t, n) take at most max_steps steps while t < T, each with dt = compute_dt(...).
 * The loop state stays on the device; with t_out/n_out non-null the call synchronises and
 * reports the final t and n. Pass T = INFINITY for a fixed number of steps. A step in which an
 * Allaire state fails the physical-state test (the reference's assert aborts the run there)
 * is the loop's last: the call returns CFDART_E_UNPHYS when it synchronises (t_out or n_out
 * non-null), and otherwise the handle's next cfdart_sync does. A batched grid stops at the
 * same step (cfdart_batch.h, status flag 2). */
int cfdart_run(cfdart_grid* g, int splitting, int max_steps, double T, double cfl,
               double* t_out, int* n_out);
/* As cfdart_run, also recording the mass history of compute_dt (E/sourcecode/euler_timestep.cpp:
 * 26,32-33, the reference's *-masschange.dat data) on the device: for each step taken,
 * mass_hist[3k..3k+2] = t at the start of the step and the masses (sum of U0, and of U1 for
 * Allaire, times dx*dy) of the state it starts from, summed in the same fixed order as
 * cfdart_compute_dt's mass_out. mass_hist holds 3*max_steps doubles; entries of steps not taken
 * are 0. Synchronous. */
int cfdart_run_mass(cfdart_grid* g, int splitting, int max_steps, double T, double cfl,
                    double* t_out, int* n_out, double* mass_hist);
/* Sets the loop state used by cfdart_run (defaults after upload: t = 0, n = 0). */
int cfdart_set_time(cfdart_grid* g, double t, int n);
/* Counters accumulated since upload (see CFDART_CNT_*), 4 values; synchronous. */
int cfdart_stats(cfdart_grid* g, long long* counters);
/* Waits for the handle's stream; returns CFDART_E_UNPHYS once an Allaire state of the grid has
 * failed the physical-state test since its last upload or initialisation. */
int cfdart_sync(cfdart_grid* g);
/* The HIP stream (hipStream_t) the handle's kernels run on. */
void* cfdart_stream(cfdart_grid* g);
/* Kernel timing: with profiling enabled every sweep launch is bracketed by HIP events on the
 * handle's stream; cfdart_profile returns, per axis (x, y), the summed kernel time in ms and
 * the number of launches since profiling was (re-)enabled. Synchronous. */
int cfdart_set_profiling(cfdart_grid* g, int enable);
int cfdart_profile(cfdart_grid* g, double* ms_total /* [2] */, long long* launches /* [2] */);
/* Tuning: rows per y-sweep strip (selects the strip tile). */
int cfdart_set_strip_rows(cfdart_grid* g, int rows);
/* Tile shape: the y-sweep tile and rows per x-sweep wavefront (default 4). y_strip_rows > 0
 * selects strips of that many rows, 256 columns wide, one column per lane marching down the
 * strip; y_strip_rows = CFDART_Y_TRANSPOSED (the default) selects the transposed tile: 64 rows x
 * 16 columns staged in LDS, lanes along y, 60 rows written; CFDART_Y_TRANSPOSED_NARROW the same
 * with 8 columns. The Godunov scheme has no transposed tile and runs 8-row strips in its place.
 * Every shape gives the same bits. */
#define CFDART_Y_TRANSPOSED (-1)
#define CFDART_Y_TRANSPOSED_NARROW (-2)
int cfdart_set_tile(cfdart_grid* g, int y_strip_rows, int x_rows_per_wave);

/* Arithmetic of the sweeps. CFDART_ARITH_EXACT (the default): every division and square root is
 * correctly rounded, so the results are bitwise those of the reference's arithmetic.
 * CFDART_ARITH_FAST: a division is the numerator times the divisor's reciprocal refined by one
 * Newton step, a square root takes one Goldschmidt step (each within a few ulp), a product
 * feeding a sum may fuse into one FMA, and work that only reproduces the exact arithmetic's
 * roundings is skipped (the HLLC star state of the side whose weight is zero, conserved-to-
 * primitive round trips of the half-step states, products with zero Jacobian entries); 110 steps
 * of the 2048^2 Euler and 1024^2 Allaire workloads stay within 5.6e-14 relative L1 and 8.2e-13
 * relative Linf of the exact results per field (the accuracy bar is 1e-10), and their sweeps
 * take 19-30 % less time. The
 * per-cell wave speeds of the time step follow the mode (so a run's result does not depend on how
 * it is split into cfdart_run calls); the fold of the maxima into dt is exact. Batches
 * (cfdart_batch_set_arith) and the serving loop (cfdart_serve_opts.arith) take the mode too.
 * No reference counterpart (the reference has one arithmetic). */
/* cfdart_run without a handle and without allocation: no counterpart in the reference. For
 * callers that cannot allocate device memory (code under HIP graph capture, serving loops and
 * harnesses that own all memory). The caller passes the state in the library's device layout
 * (ncomp planes of pitch*ny doubles, interior cells only, rows of pitch >= nx doubles) and a
 * second buffer of the same size; the loop runs from (t0, n0) on 'stream' (NULL: the null stream)
 * with the default tile shapes, and the final state ends in 'state'. The loop's control block and
 * reduction partials live in a 32 KiB device array of the library, one per GPU and held until the
 * process ends; a grid whose partials would not fit is refused with CFDART_E_ARG (none does: the
 * need is at most 24.25 KiB). Calls that overlap on one GPU share that array and race. Asynchronous
 * unless t_out or n_out is non-null (then as cfdart_run, including CFDART_E_UNPHYS). Results are
 * bitwise those of cfdart_run on a handle in the same arithmetic. */
int cfdart_run_noalloc(const cfdart_desc* d, double* state, double* work, int pitch, int arith, int splitting,
                       int max_steps, double T, double cfl, double t0, int n0, void* stream, double* t_out,
                       int* n_out);

#define CFDART_ARITH_EXACT 0
#define CFDART_ARITH_FAST 1
int cfdart_set_arith(cfdart_grid* g, int mode);

/* Checksum of the current interior state: the sum mod 2^64, over every value, of a 64-bit mix
 * of its bit pattern and its position in [component][y][x] order. Equal states give equal
 * checksums whatever their device layout; used to compare a grid run alone with the same grid
 * in a batch. Synchronous. No reference counterpart. */
int cfdart_checksum(cfdart_grid* g, unsigned long long* out);
/* Schlieren image of the current state, computed on the device: per interior cell
 * |grad rho| / max |grad rho| with central differences, rho = U0 (Euler) or U0 + U1 (Allaire),
 * as the reference's gnuplot_schlieren (euler_io.cpp:1142-1185, allaire_diffuse_io.cpp:1109-1150).
 * Cells next to the boundary difference against the boundary condition's mirror of the current
 * state (the reference reads whatever its ghost cells hold since their last fill). host_dst:
 * nx*ny doubles, row-major (y outer); max_grad (nullable) receives max |grad rho|. Synchronous. */
int cfdart_schlieren(cfdart_grid* g, double* host_dst, double* max_grad);
/* Frame: asynchronous copy of the current interior state to host_dst (pinned memory, ncomp SoA
 * planes of nx*ny doubles) on the grid's stream; complete after cfdart_sync. The state the
 * reference's output (euler_io.cpp:1142-1261) would write, without its ghost frame. */
int cfdart_frame(cfdart_grid* g, double* host_dst);

/* Measurement: copies the grid's state buffer (ncomp * pitch * ny doubles) into its other buffer
 * 'reps' times with a plain copy kernel, then synchronises: the kernel of known traffic against
 * which hardware byte counters (rocprofv3 FETCH_SIZE, WRITE_SIZE) are calibrated. The state is
 * unchanged. No reference counterpart. */
int cfdart_copy_state(cfdart_grid* g, int reps);

#ifdef __cplusplus
}
#endif
#endif
