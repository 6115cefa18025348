/*
 * CPU oracle for the dimensionally split MUSCL-Hancock sweep of cfd-art
 * (ideal/stiffened-gas Euler and the Allaire five-equation model).
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain, deliberately unoptimised FP64
 * restatement of the reference arithmetic, in the reference's operation order,
 * used as the checker for the HIP path and as the timed CPU baseline. Nothing in
 * the product (the cfd-art_amd package, its C-ABI library or its driver) links,
 * loads or calls it; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg do.
 *
 * Grid layout: the reference's grid[i][j](k) (E/headers/typedefs.hpp:19-21) as one
 * contiguous AoS array with the ghost frame stored:
 *     grid[((i * (nx + 2*ng)) + j) * ncomp + k],  i in [0, ny+2ng) (y), j in [0, nx+2ng) (x).
 * Euler components: rho, rho*u, rho*v, E.  Allaire: z*rho1, (1-z)*rho2, rho*u, rho*v, E, z.
 *
 * Pinning status of each part is listed in the header of cfd_oracle.c.
 */
#ifndef CFD_ORACLE_H
#define CFD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_EULER = 0, ORC_ALLAIRE = 1 };
enum { ORC_TRANSMISSIVE = 0, ORC_REFLECTIVE = 1 };
enum { ORC_MINMOD = 0, ORC_VANLEER = 1 };
enum { ORC_GARRICK = 0, ORC_PVRS = 1 };
enum { ORC_NAIVE = 0, ORC_STRANG = 1, ORC_ALTERNATING = 2 };
enum { ORC_MUSCL_HANCOCK = 0, ORC_GODUNOV = 1 };

/* Counters of the reference's diagnostic prints / aborts. */
enum {
    ORC_CNT_SLOPE_FALLBACK = 0,   /* E/headers/flux_solver_MUSCLhancock.hpp:184-189 */
    ORC_CNT_EVOLVE_FALLBACK = 1,  /* E/headers/flux_solver_MUSCLhancock.hpp:216-223 */
    ORC_CNT_REPAIR = 2,           /* E/sourcecode/euler_rowsweep.cpp:49-81 */
    ORC_CNT_UNPHYSICAL = 3,       /* A/sourcecode/allaire_diffuse_rowsweep.cpp:49 assert would abort */
    ORC_NCNT = 4
};

typedef struct orc_desc {
    int model;            /* ORC_EULER (4 comps) or ORC_ALLAIRE (6 comps) */
    int nx, ny, ng;       /* interior cells and ghost width (2 for MUSCL, 1 for Godunov) */
    double x0, y0, dx, dy;
    double gamma1, pinf1; /* Euler material, or Allaire fluid 1 */
    double gamma2, pinf2; /* Allaire fluid 2 */
    int bc[4];            /* left, right, bottom, top */
    int limiter;          /* ORC_MINMOD (reference) or ORC_VANLEER */
    int wave_speeds;      /* ORC_GARRICK (reference) or ORC_PVRS (Allaire only) */
    int colsweep_uses_dx; /* 1 = reference: column half-step uses dx */
    double z_lo, z_hi;    /* Allaire volume-fraction clamp (reference 0, 1) */
    int flux_solver;      /* ORC_MUSCL_HANCOCK or ORC_GODUNOV */
} orc_desc;

int  orc_ncomp(int model);
/* pre_sweep: ghost fill (E/sourcecode/euler_ppsweep.cpp:7-118). */
void orc_fill_ghosts(const orc_desc* d, double* grid);
/* One sweep over all rows (axis 0, update_row) or columns (axis 1, update_col);
 * writes the interior of future; counters may be NULL. */
void orc_sweep(const orc_desc* d, int axis, const double* grid, double* future, double dt,
               long long* counters);
/* compute_dt (E/sourcecode/euler_timestep.cpp:7-44) with the max race fixed. */
double orc_compute_dt(const orc_desc* d, const double* grid, int n, double t, double T,
                      double cfl, double* mass_out);
/* One step of the splitting scheme (E/headers/splitting_scheme.hpp:37-122); the new
 * state ends in grid (contents swapped as the reference's grid.swap does). */
void orc_advance(const orc_desc* d, int splitting, double* grid, double* future, double dt,
                 int n, long long* counters);
/* Time loop (E/headers/simulation.hpp:48-61) for at most max_steps steps or until t >= T.
 * dt_hist (nullable) receives each step's dt. Returns the number of steps taken. */
int  orc_run(const orc_desc* d, int splitting, double* grid, double* future, int max_steps,
             double T, double cfl, double* t_inout, int n0, double* dt_hist,
             long long* counters);

/* Building blocks exposed for unit tests (single cell / single face). */
void orc_c2p(const orc_desc* d, const double* U, double* W);
void orc_p2c(const orc_desc* d, const double* W, double* U);
int  orc_is_physical(const orc_desc* d, const double* U);          /* const overload */
int  orc_is_physical_snap(const orc_desc* d, double* U);           /* Allaire snapping overload */
void orc_flux_phys(const orc_desc* d, const double* U, double* F);
void orc_hllc(const orc_desc* d, const double* UL, const double* UR, double* F, double* ustar);
/* First-order upwind z update (A/headers/zupdate.hpp:44-50), clamp bounds parameterised. */
double orc_zupdate_upwind(const orc_desc* d, double dx, double dt, double zL, double z, double zR,
                          double usL, double usR);
void orc_face(const orc_desc* d, const double* s0, const double* s1, const double* s2,
              const double* s3, double dt, double dx, double* F, double* ustar,
              double* zstar, long long* counters);
double orc_limit(int limiter, double a, double b);
/* Euler positivity repair of one updated cell (E/sourcecode/euler_rowsweep.cpp:47-82): fut is
 * replaced in place when unphysical, from old's density or internal energy; returns 1 if so. */
int orc_euler_repair(const orc_desc* d, const double* old, double* fut);
void orc_jacobian(const orc_desc* d, const double* W, double* A);  /* row-major n x n */
double orc_zupdate(const orc_desc* d, double dx, double dt, double zL, double z, double zR,
                   double usL, double usR, double zsL, double zsR);
void orc_mixture_eos(const orc_desc* d, double z, double* eps, double* k);
double orc_pressure(const orc_desc* d, double rho, double e, double z);
double orc_specific_ie(const orc_desc* d, double rho, double p, double z);
double orc_soundspeed(const orc_desc* d, double rho, double p, double z);
double orc_shock_speed_jump(const orc_desc* d, double p, double rho, double pstar, double z);

/* OpenMP threads of the sweeps: set (n > 0) and return the previous maximum; query. */
int orc_set_threads(int n);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
