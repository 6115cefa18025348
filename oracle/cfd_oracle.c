/*
 * CPU oracle: FP64 restatement of cfd-art's dimensionally split MUSCL-Hancock solver.
 *
 * TEST INFRASTRUCTURE ONLY (see cfd_oracle.h): the checker for the HIP path and the timed CPU
 * baseline. It is never linked into, loaded by or called from the product.
 *
 * Reference files followed (paths relative to the reference repository root):
 *   E/ = euler_equations/euler_simulation_code/, A/ = allaire_2d/
 *   EOS                 E/headers/stiffened_gas_eos.hpp:57-71 ; A/headers/mixturemodel.hpp:18-52
 *   state algebra       E/headers/misc_euler.hpp:20-39,63-131 ; A/headers/misc.hpp:31-52,130-163,189-266
 *   limiter             E/headers/flux_solver_MUSCLhancock.hpp:67-89 (minmod; van Leer is new)
 *   MUSCL-Hancock face  E/headers/flux_solver_MUSCLhancock.hpp:128-227 ; A/headers/flux_solver_MUSCLhancock.hpp:272-376
 *   HLLC                E/headers/riemann_solver_HLLC.hpp:50-145 ; A/headers/riemann_solver_HLLC.hpp:49-150
 *                       (PVRS wave speeds: the commented block A/headers/riemann_solver_HLLC.hpp:84-107)
 *   z update            A/headers/zupdate.hpp:58-73 (second order), :40-55 (upwind, with Godunov)
 *   Godunov face        E/headers/flux_solver_godunov.hpp:18-41 ; A/headers/flux_solver_godunov.hpp:17-41
 *   sweeps              E/sourcecode/euler_rowsweep.cpp:10-95, euler_colsweep.cpp:10-114,
 *                       A/sourcecode/allaire_diffuse_rowsweep.cpp:10-52, allaire_diffuse_colsweep.cpp:9-73
 *   ghost fill / swap   E/sourcecode/euler_ppsweep.cpp:7-123 ; A/sourcecode/allaire_diffuse_ppsweep.cpp:7-123
 *   dt                  E/sourcecode/euler_timestep.cpp:7-44 ; A/sourcecode/allaire_diffuse_timestep.cpp:8-48
 *   splitting           E/headers/splitting_scheme.hpp:37-122
 *   time loop           E/headers/simulation.hpp:48-61
 *
 * Arithmetic order: every expression is written in the reference's left-to-right C++ order.
 * std::min(a,b) is (b<a)?b:a and std::max(a,b) is (a<b)?b:a (their definitions), including
 * for signed zeros and NaN. Eigen's small products (0.5*(dt/dx)*A*del, A a fixed 4x4/6x6
 * matrix) are evaluated coefficient-wise as sum_k (s*A(i,k))*del(k), accumulated from k=0
 * upwards -- the order of Eigen's unrolled coefficient-based (lazy) product. Build with
 * -ffp-contract=off so no FMA is formed (the reference targets x86-64 SSE2, which has none).
 *
 * Pinning (tests/test_oracle_*.py): EOS and mixture closed forms and reductions; state
 * round trips; Jacobian vs finite differences; limiter tables; HLLC consistency and
 * contact preservation; free-stream preservation; x/y transposition invariance; conservation
 * budgets; Toro's exact Riemann solutions (L1 convergence); ghost-fill index tables by
 * brute force; dt by brute force; the Godunov scheme by first-order convergence to the exact
 * Riemann solution and the upwind z update by its closed form on a step; the Euler repair at its
 * thresholds and by the source of each recovered quantity; the face's M3 and M5 fallbacks as
 * HLLC of the zero-slope or raw states (Allaire: snapped where the reference snaps); the
 * splitting schemes composed from single sweeps, on a state where the y sweep is the identity
 * and against the transposed problem; shock_speed_jump as the exact Rankine-Hugoniot speed.
 * tests/test_oracle_mutants.py plants a mistake in each of these and checks that its pin fails.
 * Parity unpinned: the exact Eigen evaluation order of the
 * half-step product (reassociation-level only; no reference binary can be built, Eigen is
 * absent) and the reference's racy multithreaded max in compute_dt.
 */
#include "cfd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NCMAX 6

static inline double std_min(double a, double b) { return (b < a) ? b : a; }
static inline double std_max(double a, double b) { return (a < b) ? b : a; }

int orc_ncomp(int model) { return model == ORC_ALLAIRE ? 6 : 4; }

/* ------------------------------------------------------------------ EOS */

/* A/headers/mixturemodel.hpp:20-22 (eps and k, recomputed in every mixture function) */
void orc_mixture_eos(const orc_desc* d, double z, double* eps, double* k)
{
    double e = z / (d->gamma1 - 1.0) + (1.0 - z) / (d->gamma2 - 1.0);
    *eps = e;
    *k = (z * d->gamma1 * d->pinf1 / (d->gamma1 - 1.0)
          + (1.0 - z) * d->gamma2 * d->pinf2 / (d->gamma2 - 1.0)) / e;
}

/* E/headers/stiffened_gas_eos.hpp:57-60 ; A/headers/mixturemodel.hpp:18-24 */
double orc_pressure(const orc_desc* d, double rho, double e, double z)
{
    if (d->model == ORC_EULER)
        return (d->gamma1 - 1.0) * rho * e - d->gamma1 * d->pinf1;
    double eps, k;
    orc_mixture_eos(d, z, &eps, &k);
    return rho * e / eps - k;
}

/* E/headers/stiffened_gas_eos.hpp:62-65 ; A/headers/mixturemodel.hpp:26-32 */
double orc_specific_ie(const orc_desc* d, double rho, double p, double z)
{
    if (d->model == ORC_EULER)
        return (p + d->gamma1 * d->pinf1) / ((d->gamma1 - 1.0) * rho);
    double eps, k;
    orc_mixture_eos(d, z, &eps, &k);
    return (p + k) * eps / rho;
}

/* E/headers/stiffened_gas_eos.hpp:67-70 ; A/headers/mixturemodel.hpp:34-40 */
double orc_soundspeed(const orc_desc* d, double rho, double p, double z)
{
    if (d->model == ORC_EULER)
        return sqrt((d->gamma1 * (p + d->pinf1)) / rho);
    double eps, k;
    orc_mixture_eos(d, z, &eps, &k);
    return sqrt((k + p * ((1.0 / eps) + 1.0)) / rho);
}

/* A/headers/mixturemodel.hpp:42-52 */
double orc_shock_speed_jump(const orc_desc* d, double p, double rho, double pstar, double z)
{
    double eps, k;
    orc_mixture_eos(d, z, &eps, &k);
    double p_avg = 0.5 * (pstar + p);
    return sqrt((pstar + k + p_avg / eps) / rho);
}

/* ------------------------------------------------------------------ state algebra */

/* E/headers/misc_euler.hpp:63-78 ; A/headers/misc.hpp:130-147 */
void orc_c2p(const orc_desc* d, const double* U, double* W)
{
    if (d->model == ORC_EULER) {
        double rho = U[0];
        double u = U[1] / rho;
        double v = U[2] / rho;
        double e = U[3] / rho - 0.5 * (u * u + v * v);
        double p = orc_pressure(d, rho, e, 0.0);
        W[0] = rho; W[1] = u; W[2] = v; W[3] = p;
    } else {
        double rho = U[0] + U[1];
        double u = U[2] / rho;
        double v = U[3] / rho;
        double e = U[4] / rho - 0.5 * (u * u + v * v);
        double p = orc_pressure(d, rho, e, U[5]);
        W[0] = U[0]; W[1] = U[1]; W[2] = u; W[3] = v; W[4] = p; W[5] = U[5];
    }
}

/* E/headers/misc_euler.hpp:82-97 ; A/headers/misc.hpp:149-163 */
void orc_p2c(const orc_desc* d, const double* W, double* U)
{
    if (d->model == ORC_EULER) {
        double rho = W[0], u = W[1], v = W[2], p = W[3];
        double e = orc_specific_ie(d, rho, p, 0.0);
        U[0] = rho;
        U[1] = rho * u;
        U[2] = rho * v;
        U[3] = rho * e + 0.5 * rho * (u * u + v * v);
    } else {
        double rho = W[0] + W[1];
        double e = orc_specific_ie(d, rho, W[4], W[5]);
        U[0] = W[0];
        U[1] = W[1];
        U[2] = rho * W[2];
        U[3] = rho * W[3];
        U[4] = rho * e + 0.5 * rho * (W[2] * W[2] + W[3] * W[3]);
        U[5] = W[5];
    }
}

/* E/headers/misc_euler.hpp:123-131 ; A/headers/misc.hpp:228-237 (const overload) */
int orc_is_physical(const orc_desc* d, const double* U)
{
    if (d->model == ORC_EULER) {
        double rho = U[0];
        double u = U[1] / rho;
        double v = U[2] / rho;
        double e = U[3] / rho - 0.5 * (u * u + v * v);
        return rho >= 0.0 && e >= 0.0;
    }
    double rho = U[0] + U[1];
    double u = U[2] / rho;
    double v = U[3] / rho;
    double e = U[4] / rho - 0.5 * (u * u + v * v);
    double z = U[5];
    return U[0] >= 0.0 && U[1] >= 0.0 && e >= 0.0 && z >= 0.0 && z <= 1.0;
}

/* A/headers/misc.hpp:251-266: the non-const overload snaps the state before testing it. */
int orc_is_physical_snap(const orc_desc* d, double* U)
{
    if (d->model == ORC_EULER) return orc_is_physical(d, U);
    if (fabs(U[0]) < 1e-12) U[0] = 0.0;
    if (fabs(U[1]) < 1e-12) U[1] = 0.0;
    if (fabs(U[5]) < 1e-12) U[5] = 0.0;
    if (fabs(U[5] - 1.0) < 1e-12) U[5] = 1.0;
    return orc_is_physical(d, U);
}

/* E/headers/misc_euler.hpp:20-39 ; A/headers/misc.hpp:31-52 */
void orc_flux_phys(const orc_desc* d, const double* U, double* F)
{
    if (d->model == ORC_EULER) {
        double rho = U[0];
        double u = U[1] / rho;
        double v = U[2] / rho;
        double e = U[3] / rho - 0.5 * (u * u + v * v);
        double p = orc_pressure(d, rho, e, 0.0);
        F[0] = U[0] * u;
        F[1] = U[1] * u + p;
        F[2] = U[2] * u;
        F[3] = (U[3] + p) * u;
    } else {
        double rho = U[0] + U[1];
        double u = U[2] / rho;
        double v = U[3] / rho;
        double e = U[4] / rho - 0.5 * (u * u + v * v);
        double p = orc_pressure(d, rho, e, U[5]);
        F[0] = U[0] * u;
        F[1] = U[1] * u;
        F[2] = U[2] * u + p;
        F[3] = U[3] * u;
        F[4] = (U[4] + p) * u;
        F[5] = 0.0;
    }
}

/* Primitive-variable Jacobian: E/headers/misc_euler.hpp:102-119 ; A/headers/misc.hpp:189-205.
 * Row-major A[i*n+k]. */
void orc_jacobian(const orc_desc* d, const double* W, double* A)
{
    int n = orc_ncomp(d->model);
    if (d->model == ORC_EULER) {
        double rho = W[0], u = W[1], p = W[3];
        double c = orc_soundspeed(d, rho, p, 0.0);
        for (int i = 0; i < n * n; ++i) A[i] = 0.0;
        for (int i = 0; i < n; ++i) A[i * n + i] = u;
        A[0 * n + 1] = rho;
        A[1 * n + 3] = 1.0 / rho;
        A[3 * n + 1] = rho * c * c;
    } else {
        double rho = W[0] + W[1];
        double c = orc_soundspeed(d, rho, W[4], W[5]);
        for (int i = 0; i < n * n; ++i) A[i] = 0.0;
        for (int i = 0; i < n; ++i) A[i * n + i] = W[2];
        A[0 * n + 2] = W[0];
        A[1 * n + 2] = W[1];
        A[2 * n + 4] = 1.0 / rho;
        A[4 * n + 2] = rho * c * c;
    }
}

/* ------------------------------------------------------------------ limiter */

/* minmod: E/headers/flux_solver_MUSCLhancock.hpp:67-89 ; van Leer (not in the reference):
 * ab > 0 ? 2ab/(a+b) : 0, written as ((2*a)*b)/(a+b). */
double orc_limit(int limiter, double a, double b)
{
    if (limiter == ORC_VANLEER)
        return (a * b > 0.0) ? (2.0 * a * b) / (a + b) : 0.0;
    if (b > 0.0)
        return std_max(0.0, std_min(a, b));
    return std_min(0.0, std_max(a, b));
}

/* ------------------------------------------------------------------ HLLC */

/* E/headers/riemann_solver_HLLC.hpp:50-145 ; A/headers/riemann_solver_HLLC.hpp:49-150 */
void orc_hllc(const orc_desc* d, const double* UL, const double* UR, double* F, double* ustar)
{
    int n = orc_ncomp(d->model);
    int all = d->model == ORC_ALLAIRE;
    int m = all ? 2 : 1;  /* normal-momentum component */
    double rhoL, uL, vL, eL, pL, cL, rhoR, uR, vR, eR, pR, cR, zL = 0.0, zR = 0.0;
    if (!all) {
        rhoL = UL[0];
        uL = UL[1] / rhoL;
        vL = UL[2] / rhoL;
        eL = UL[3] / rhoL - 0.5 * (uL * uL + vL * vL);
        pL = orc_pressure(d, rhoL, eL, 0.0);
        cL = orc_soundspeed(d, rhoL, pL, 0.0);
        rhoR = UR[0];
        uR = UR[1] / rhoR;
        vR = UR[2] / rhoR;
        eR = UR[3] / rhoR - 0.5 * (uR * uR + vR * vR);
        pR = orc_pressure(d, rhoR, eR, 0.0);
        cR = orc_soundspeed(d, rhoR, pR, 0.0);
    } else {
        zL = UL[5];
        rhoL = UL[0] + UL[1];
        uL = UL[2] / rhoL;
        vL = UL[3] / rhoL;
        eL = UL[4] / rhoL - 0.5 * (uL * uL + vL * vL);
        pL = orc_pressure(d, rhoL, eL, zL);
        cL = orc_soundspeed(d, rhoL, pL, zL);
        zR = UR[5];
        rhoR = UR[0] + UR[1];
        uR = UR[2] / rhoR;
        vR = UR[3] / rhoR;
        eR = UR[4] / rhoR - 0.5 * (uR * uR + vR * vR);
        pR = orc_pressure(d, rhoR, eR, zR);
        cR = orc_soundspeed(d, rhoR, pR, zR);
    }
    (void)vL; (void)vR;

    double SL, SR;
    if (all && d->wave_speeds == ORC_PVRS) {
        /* A/headers/riemann_solver_HLLC.hpp:86-105 (commented in the reference) */
        double rho_avg = 0.5 * (rhoL + rhoR);
        double c_avg = 0.5 * (cL + cR);
        double p_pvrs = std_max(0.0, 0.5 * (pL + pR) + 0.5 * (uL - uR) * rho_avg * c_avg);
        if (p_pvrs > pL) SL = uL - orc_shock_speed_jump(d, pL, rhoL, p_pvrs, zL);
        else SL = uL - cL;
        if (p_pvrs > pR) SR = uR + orc_shock_speed_jump(d, pR, rhoR, p_pvrs, zR);
        else SR = uR + cR;
    } else {
        /* Garrick, Owkes, Regele: E/headers/riemann_solver_HLLC.hpp:74-77 */
        double u_avg = 0.5 * (uL + uR);
        double c_avg = 0.5 * (cL + cR);
        SL = std_min(u_avg - c_avg, uL - cL);
        SR = std_max(u_avg + c_avg, uR + cR);
    }

    double u_star = (pR - pL + rhoL * uL * (SL - uL) - rhoR * uR * (SR - uR))
                    / (rhoL * (SL - uL) - rhoR * (SR - uR));
    double factorL = (SL - uL) / (SL - u_star);
    double factorR = (SR - uR) / (SR - u_star);
    if (ustar) *ustar = u_star;

    double UsL[NCMAX], UsR[NCMAX], FL[NCMAX], FR[NCMAX];
    if (!all) {
        UsL[0] = factorL * UL[0];
        UsL[1] = factorL * rhoL * u_star;
        UsL[2] = factorL * UL[2];
        UsL[3] = factorL * (UL[3] + (u_star - uL) * (rhoL * u_star + (pL / (SL - uL))));
        UsR[0] = factorR * UR[0];
        UsR[1] = factorR * rhoR * u_star;
        UsR[2] = factorR * UR[2];
        UsR[3] = factorR * (UR[3] + (u_star - uR) * (rhoR * u_star + (pR / (SR - uR))));
    } else {
        UsL[0] = factorL * UL[0];
        UsL[1] = factorL * UL[1];
        UsL[2] = factorL * rhoL * u_star;
        UsL[3] = factorL * UL[3];
        UsL[4] = factorL * (UL[4] + (u_star - uL) * (rhoL * u_star + (pL / (SL - uL))));
        UsL[5] = 0.0;
        UsR[0] = factorR * UR[0];
        UsR[1] = factorR * UR[1];
        UsR[2] = factorR * rhoR * u_star;
        UsR[3] = factorR * UR[3];
        UsR[4] = factorR * (UR[4] + (u_star - uR) * (rhoR * u_star + (pR / (SR - uR))));
        UsR[5] = 0.0;
    }
    (void)m;
    orc_flux_phys(d, UL, FL);
    orc_flux_phys(d, UR, FR);
    double s = copysign(1.0, u_star);
    double wL = (1.0 + s) / 2.0;
    double wR = (1.0 - s) / 2.0;
    double mSL = std_min(0.0, SL);
    double MSR = std_max(0.0, SR);
    for (int k = 0; k < n; ++k)
        F[k] = wL * (FL[k] + mSL * (UsL[k] - UL[k])) + wR * (FR[k] + MSR * (UsR[k] - UR[k]));
    if (all) F[5] = 0.0;
}

/* ------------------------------------------------------------------ MUSCL-Hancock face */

/* E/headers/flux_solver_MUSCLhancock.hpp:128-227 ; A/headers/flux_solver_MUSCLhancock.hpp:272-376 */
void orc_face(const orc_desc* d, const double* s0, const double* s1, const double* s2,
              const double* s3, double dt, double dx, double* F, double* ustar,
              double* zstar, long long* counters)
{
    int n = orc_ncomp(d->model);
    int all = d->model == ORC_ALLAIRE;
    double WL[NCMAX], WR[NCMAX], W0[NCMAX], W3[NCMAX];
    double dL[NCMAX], dC[NCMAX], dR[NCMAX], delL[NCMAX], delR[NCMAX];
    double UL[NCMAX], UR[NCMAX], T[NCMAX], A[NCMAX * NCMAX], tmp[NCMAX];

    orc_c2p(d, s1, WL);
    orc_c2p(d, s2, WR);
    orc_c2p(d, s0, W0);
    orc_c2p(d, s3, W3);
    for (int k = 0; k < n; ++k) {
        dL[k] = WL[k] - W0[k];
        dC[k] = WR[k] - WL[k];
        dR[k] = W3[k] - WR[k];
    }
    for (int k = 0; k < n; ++k) {
        delL[k] = orc_limit(d->limiter, dL[k], dC[k]);
        delR[k] = orc_limit(d->limiter, dC[k], dR[k]);
    }

    /* M3: unphysical slope extrapolation -> zero slopes (:178-189). Allaire applies the
     * snapping overload to these temporaries (A/...MUSCLhancock.hpp:324-325). */
    for (int k = 0; k < n; ++k) T[k] = WL[k] + 0.5 * delL[k];
    orc_p2c(d, T, UL);
    for (int k = 0; k < n; ++k) T[k] = WR[k] - 0.5 * delR[k];
    orc_p2c(d, T, UR);
    int ok = all ? (orc_is_physical_snap(d, UL) && orc_is_physical_snap(d, UR))
                 : (orc_is_physical(d, UL) && orc_is_physical(d, UR));
    if (!ok) {
        for (int k = 0; k < n; ++k) { delL[k] = 0.0; delR[k] = 0.0; }
        if (counters) {
#pragma omp atomic
            counters[ORC_CNT_SLOPE_FALLBACK] += 1;
        }
    }

    /* M4: half-step evolution UL = W + 0.5 del - (0.5 (dt/dx) A) del (:195-201) */
    double s = 0.5 * (dt / dx);
    orc_jacobian(d, WL, A);
    for (int i = 0; i < n; ++i) {
        double acc = (s * A[i * n + 0]) * delL[0];
        for (int k = 1; k < n; ++k) acc = (s * A[i * n + k]) * delL[k] + acc;
        tmp[i] = acc;
    }
    for (int k = 0; k < n; ++k) T[k] = WL[k] + 0.5 * delL[k] - tmp[k];
    orc_p2c(d, T, UL);
    orc_jacobian(d, WR, A);
    for (int i = 0; i < n; ++i) {
        double acc = (s * A[i * n + 0]) * delR[0];
        for (int k = 1; k < n; ++k) acc = (s * A[i * n + k]) * delR[k] + acc;
        tmp[i] = acc;
    }
    for (int k = 0; k < n; ++k) T[k] = WR[k] - 0.5 * delR[k] - tmp[k];
    orc_p2c(d, T, UR);

    /* M5: unphysical evolved states -> first-order states (:212-223). Allaire's snapping
     * overload modifies UL/UR in place (A/...MUSCLhancock.hpp:356-357). */
    ok = all ? (orc_is_physical_snap(d, UL) && orc_is_physical_snap(d, UR))
             : (orc_is_physical(d, UL) && orc_is_physical(d, UR));
    if (!ok) {
        for (int k = 0; k < n; ++k) { UL[k] = s1[k]; UR[k] = s2[k]; }
        if (counters) {
#pragma omp atomic
            counters[ORC_CNT_EVOLVE_FALLBACK] += 1;
        }
    }

    double us;
    orc_hllc(d, UL, UR, F, &us);
    if (ustar) *ustar = us;
    if (all && zstar) {
        double sg = copysign(1.0, us);
        *zstar = 0.5 * (1.0 + sg) * UL[5] + 0.5 * (1.0 - sg) * UR[5];
    }
}

/* A/headers/zupdate.hpp:62-67, clamp bounds parameterised (reference: 0, 1) */
double orc_zupdate(const orc_desc* d, double dx, double dt, double zL, double z, double zR,
                   double usL, double usR, double zsL, double zsR)
{
    (void)zL; (void)zR;
    double newz = z - (dt / dx) * (usR * zsR - usL * zsL - z * (usR - usL));
    return std_max(std_min(d->z_hi, newz), d->z_lo);
}

/* A/headers/zupdate.hpp:44-50 (upwind; paired with the Godunov flux solver,
 * A/sourcecode/allaire_diffuse_io.cpp:446-450), clamp bounds parameterised */
double orc_zupdate_upwind(const orc_desc* d, double dx, double dt, double zL, double z, double zR,
                          double usL, double usR)
{
    double newz = z - (dt / dx) * ((zR - z) * std_min(0.0, usR) + (z - zL) * std_max(0.0, usL));
    return std_max(std_min(d->z_hi, newz), d->z_lo);
}

/* Godunov face (E/headers/flux_solver_godunov.hpp:30-36, A/headers/flux_solver_godunov.hpp:30-39):
 * the Riemann problem of the two cells' stored states, z* as the MUSCL face does. */
static void godunov_face(const orc_desc* d, const double* s0, const double* s1, double* F, double* ustar,
                         double* zstar)
{
    orc_hllc(d, s0, s1, F, ustar);
    if (d->model == ORC_ALLAIRE) {
        double sg = copysign(1.0, *ustar);
        *zstar = ((1.0 + sg) / 2.0) * s0[5] + ((1.0 - sg) / 2.0) * s1[5];
    }
}

/* ------------------------------------------------------------------ ghost fill */

/* E/sourcecode/euler_ppsweep.cpp:7-113 ; A/sourcecode/allaire_diffuse_ppsweep.cpp:7-113.
 * Corners are never written. */
void orc_fill_ghosts(const orc_desc* d, double* grid)
{
    int n = orc_ncomp(d->model), g = d->ng, nx = d->nx, ny = d->ny;
    int W = nx + 2 * g;
    int mx = d->model == ORC_ALLAIRE ? 2 : 1, my = mx + 1;
#define CELL(i, j) (grid + ((size_t)(i) * W + (j)) * n)
    for (int i = g; i < ny + g; ++i)
        for (int j = 0; j < g; ++j) {
            memcpy(CELL(i, j), CELL(i, 2 * g - j - 1), n * sizeof(double));
            if (d->bc[0] == ORC_REFLECTIVE) CELL(i, j)[mx] = -CELL(i, j)[mx];
        }
    for (int i = g; i < ny + g; ++i)
        for (int j = 0; j < g; ++j) {
            memcpy(CELL(i, j + nx + g), CELL(i, nx + g - 1 - j), n * sizeof(double));
            if (d->bc[1] == ORC_REFLECTIVE) CELL(i, j + nx + g)[mx] = -CELL(i, j + nx + g)[mx];
        }
    for (int j = g; j < nx + g; ++j)
        for (int i = 0; i < g; ++i) {
            memcpy(CELL(i, j), CELL(2 * g - i - 1, j), n * sizeof(double));
            if (d->bc[2] == ORC_REFLECTIVE) CELL(i, j)[my] = -CELL(i, j)[my];
        }
    for (int j = g; j < nx + g; ++j)
        for (int i = 0; i < g; ++i) {
            memcpy(CELL(i + ny + g, j), CELL(ny + g - 1 - i, j), n * sizeof(double));
            if (d->bc[3] == ORC_REFLECTIVE) CELL(i + ny + g, j)[my] = -CELL(i + ny + g, j)[my];
        }
#undef CELL
}

/* ------------------------------------------------------------------ sweeps */

/* Euler positivity repair: E/sourcecode/euler_rowsweep.cpp:47-82 (column variant
 * E/sourcecode/euler_colsweep.cpp:78-113 uses the unrotated old state). */
static int euler_repair(const orc_desc* d, const double* old, double* fut, long long* counters)
{
    if (orc_is_physical(d, fut)) return 0;
    if (counters) {
#pragma omp atomic
        counters[ORC_CNT_REPAIR] += 1;
    }
    double rho = fut[0];
    double u = fut[1] / rho;
    double v = fut[2] / rho;
    double e = fut[3] / rho - 0.5 * (u * u + v * v);
    double rho_old = old[0];
    double u_old = old[1] / rho_old;
    double v_old = old[2] / rho_old;
    double e_old = old[3] / rho_old - 0.5 * (u_old * u_old + v_old * v_old);
    if (rho <= 1e-12) rho = rho_old;
    if (e <= 1e-12) e = e_old;
    double p = orc_pressure(d, rho, e, 0.0);
    double W[4] = {rho, u, v, p};
    orc_p2c(d, W, fut);
    return 1;
}

int orc_euler_repair(const orc_desc* d, const double* old, double* fut)
{
    return euler_repair(d, old, fut, NULL);
}

/* One line (row or rotated column): line[l], l in [0, len+2g), the flux stencil runs over it.
 * out_*: write back of interior cell l (rotation handled by the caller). */
static void sweep_line(const orc_desc* d, const double* line, int len, double dt, double dx_half,
                       double dx_update, double* flux, double* ustar, double* zstar,
                       double* out, long long* counters)
{
    int n = orc_ncomp(d->model), g = d->ng;
    int god = d->flux_solver == ORC_GODUNOV;
    /* E/sourcecode/euler_rowsweep.cpp:23-32: faces j in [g, len+g], stencil j-stclsize .. j+stclsize-1 */
    for (int j = g; j < len + g + 1; ++j) {
        if (god)
            godunov_face(d, line + (size_t)(j - 1) * n, line + (size_t)j * n, flux + (size_t)j * n, ustar + j,
                         zstar + j);
        else
            orc_face(d, line + (size_t)(j - 2) * n, line + (size_t)(j - 1) * n, line + (size_t)j * n,
                     line + (size_t)(j + 1) * n, dt, dx_half, flux + (size_t)j * n, ustar + j,
                     zstar + j, counters);
    }
    double r = dt / dx_update;
    for (int j = g; j < len + g; ++j)
        for (int k = 0; k < n; ++k)
            out[(size_t)(j - g) * n + k] =
                line[(size_t)j * n + k] + r * (flux[(size_t)j * n + k] - flux[(size_t)(j + 1) * n + k]);
    if (d->model == ORC_ALLAIRE) {
        /* A/sourcecode/allaire_diffuse_rowsweep.cpp:44-50 */
        for (int j = g; j < len + g; ++j) {
            double* o = out + (size_t)(j - g) * n;
            if (god)
                o[5] = orc_zupdate_upwind(d, dx_update, dt, line[(size_t)(j - 1) * n + 5], line[(size_t)j * n + 5],
                                          line[(size_t)(j + 1) * n + 5], ustar[j], ustar[j + 1]);
            else
                o[5] = orc_zupdate(d, dx_update, dt, line[(size_t)(j - 1) * n + 5], line[(size_t)j * n + 5],
                                   line[(size_t)(j + 1) * n + 5], ustar[j], ustar[j + 1], zstar[j],
                                   zstar[j + 1]);
            if (!orc_is_physical_snap(d, o) && counters) {
#pragma omp atomic
                counters[ORC_CNT_UNPHYSICAL] += 1;
            }
        }
    }
}

void orc_sweep(const orc_desc* d, int axis, const double* grid, double* future, double dt,
               long long* counters)
{
    int n = orc_ncomp(d->model), g = d->ng, nx = d->nx, ny = d->ny;
    int W = nx + 2 * g;
    int all = d->model == ORC_ALLAIRE;
    int mx = all ? 2 : 1, my = mx + 1;
    if (axis == 0) {
        /* E/headers/splitting_scheme.hpp:41-50 -> update_row for each interior row */
#pragma omp parallel
        {
            double* flux = (double*)malloc(sizeof(double) * (size_t)(W + 1) * n);
            double* ust = (double*)malloc(sizeof(double) * (size_t)(W + 1));
            double* zst = (double*)malloc(sizeof(double) * (size_t)(W + 1));
            double* out = (double*)malloc(sizeof(double) * (size_t)nx * n);
#pragma omp for schedule(dynamic)
            for (int i = g; i < ny + g; ++i) {
                const double* row = grid + (size_t)i * W * n;
                sweep_line(d, row, nx, dt, d->dx, d->dx, flux, ust, zst, out, counters);
                for (int j = 0; j < nx; ++j) {
                    double* f = future + ((size_t)i * W + j + g) * n;
                    memcpy(f, out + (size_t)j * n, n * sizeof(double));
                    if (!all) euler_repair(d, row + (size_t)(j + g) * n, f, counters);
                }
            }
            free(flux); free(ust); free(zst); free(out);
        }
    } else {
        /* update_col: rotate (m_x, m_y) -> (m_y, -m_x), sweep, rotate back, repair */
        int H = ny + 2 * g;
        double dxh = d->colsweep_uses_dx ? d->dx : d->dy;
#pragma omp parallel
        {
            double* line = (double*)malloc(sizeof(double) * (size_t)H * n);
            double* flux = (double*)malloc(sizeof(double) * (size_t)(H + 1) * n);
            double* ust = (double*)malloc(sizeof(double) * (size_t)(H + 1));
            double* zst = (double*)malloc(sizeof(double) * (size_t)(H + 1));
            double* out = (double*)malloc(sizeof(double) * (size_t)ny * n);
#pragma omp for schedule(dynamic)
            for (int j = g; j < nx + g; ++j) {
                for (int i = 0; i < H; ++i) {
                    const double* c = grid + ((size_t)i * W + j) * n;
                    double* l = line + (size_t)i * n;
                    memcpy(l, c, n * sizeof(double));
                    double tmp = l[mx];
                    l[mx] = l[my];
                    l[my] = -tmp;
                }
                sweep_line(d, line, ny, dt, dxh, d->dy, flux, ust, zst, out, counters);
                for (int i = 0; i < ny; ++i) {
                    double* f = future + ((size_t)(i + g) * W + j) * n;
                    memcpy(f, out + (size_t)i * n, n * sizeof(double));
                    double tmp = f[mx];
                    f[mx] = -f[my];
                    f[my] = tmp;
                    if (!all) euler_repair(d, grid + ((size_t)(i + g) * W + j) * n, f, counters);
                }
            }
            free(line); free(flux); free(ust); free(zst); free(out);
        }
    }
}

/* ------------------------------------------------------------------ dt */

/* E/sourcecode/euler_timestep.cpp:7-44 ; A/sourcecode/allaire_diffuse_timestep.cpp:8-48.
 * The reference's unprotected max over threads is replaced by a proper max reduction; the
 * mass sums are accumulated row by row in a fixed order. mass_out[0] (and [1] for Allaire). */
double orc_compute_dt(const orc_desc* d, const double* grid, int n, double t, double T,
                      double cfl, double* mass_out)
{
    int nc = orc_ncomp(d->model), g = d->ng, nx = d->nx, ny = d->ny;
    int W = nx + 2 * g;
    double* rowm = (double*)calloc((size_t)ny * 2, sizeof(double));
    double maxu = 0.0, maxv = 0.0;
#pragma omp parallel
    {
        double lu = 0.0, lv = 0.0;
#pragma omp for schedule(dynamic)
        for (int i = g; i < ny + g; ++i) {
            double m1 = 0.0, m2 = 0.0;
            for (int j = g; j < nx + g; ++j) {
                const double* U = grid + ((size_t)i * W + j) * nc;
                double rho, u, v, e, p, c;
                if (d->model == ORC_EULER) {
                    rho = U[0];
                    u = U[1] / rho;
                    v = U[2] / rho;
                    e = U[3] / rho - 0.5 * (u * u + v * v);
                    p = orc_pressure(d, rho, e, 0.0);
                    c = orc_soundspeed(d, rho, p, 0.0);
                    m1 += rho;
                } else {
                    rho = U[0] + U[1];
                    u = U[2] / rho;
                    v = U[3] / rho;
                    e = U[4] / rho - 0.5 * (u * u + v * v);
                    double z = U[5];
                    p = orc_pressure(d, rho, e, z);
                    c = orc_soundspeed(d, rho, p, z);
                    m1 += U[0];
                    m2 += U[1];
                }
                lu = std_max(lu, fabs(u) + c);
                lv = std_max(lv, fabs(v) + c);
            }
            rowm[(size_t)(i - g) * 2] = m1;
            rowm[(size_t)(i - g) * 2 + 1] = m2;
        }
#pragma omp critical
        {
            maxu = std_max(maxu, lu);
            maxv = std_max(maxv, lv);
        }
    }
    if (mass_out) {
        double m1 = 0.0, m2 = 0.0;
        for (int i = 0; i < ny; ++i) { m1 += rowm[2 * i]; m2 += rowm[2 * i + 1]; }
        mass_out[0] = m1 * d->dx * d->dy;
        if (d->model == ORC_ALLAIRE) mass_out[1] = m2 * d->dx * d->dy;
    }
    free(rowm);
    double CFL = cfl;
    if (n < 5) CFL = std_min(CFL, 0.2);
    double dt = CFL * std_min(d->dx / maxu, d->dy / maxv);
    if (t + dt > T) dt = T - t;
    return dt;
}

/* ------------------------------------------------------------------ splitting / time loop */

static void swap_contents(double** a, double** b) { double* t = *a; *a = *b; *b = t; }

void orc_advance(const orc_desc* d, int splitting, double* grid, double* future, double dt,
                 int n, long long* counters)
{
    double* g0 = grid;
    double* a = grid;
    double* b = future;
    int axes[3], nsw = 0;
    double dts[3];
    if (splitting == ORC_STRANG) {
        /* E/headers/splitting_scheme.hpp:78-122: X(dt/2) Y(dt) X(dt/2) */
        axes[0] = 0; dts[0] = 0.5 * dt;
        axes[1] = 1; dts[1] = dt;
        axes[2] = 0; dts[2] = 0.5 * dt;
        nsw = 3;
    } else if (splitting == ORC_ALTERNATING && (n % 2) == 1) {
        axes[0] = 1; dts[0] = dt;
        axes[1] = 0; dts[1] = dt;
        nsw = 2;
    } else {
        /* E/headers/splitting_scheme.hpp:37-67: X(dt) Y(dt) */
        axes[0] = 0; dts[0] = dt;
        axes[1] = 1; dts[1] = dt;
        nsw = 2;
    }
    for (int s = 0; s < nsw; ++s) {
        orc_fill_ghosts(d, a);                       /* pre_sweep */
        orc_sweep(d, axes[s], a, b, dts[s], counters);
        swap_contents(&a, &b);                       /* post_sweep */
    }
    if (a != g0) {
        /* the reference swaps vector contents, so its 'grid' always names the new state */
        size_t sz = (size_t)(d->nx + 2 * d->ng) * (d->ny + 2 * d->ng) * orc_ncomp(d->model);
        memcpy(grid, a, sz * sizeof(double));
    }
}

int orc_run(const orc_desc* d, int splitting, double* grid, double* future, int max_steps,
            double T, double cfl, double* t_inout, int n0, double* dt_hist, long long* counters)
{
    double t = *t_inout;
    int n = n0, k = 0;
    while (t < T && k < max_steps) {
        double dt = orc_compute_dt(d, grid, n, t, T, cfl, NULL);
        if (dt_hist) dt_hist[k] = dt;
        orc_advance(d, splitting, grid, future, dt, n, counters);
        t += dt;
        n++;
        k++;
        if (fabs(t - T) < 1e-12) t = T;
    }
    *t_inout = t;
    return k;
}

int orc_set_threads(int n)
{
    const int prev = omp_get_max_threads();
    if (n > 0) omp_set_num_threads(n);
    return prev;
}

int orc_get_threads(void) { return omp_get_max_threads(); }
