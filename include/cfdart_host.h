/*
 * cfdart initial conditions: the reference's problem-class initial states, restated in C++.
 * The host generators produce the host state format of cfdart.h (the reference's grid[i][j](k)
 * with its ghost frame); the device generators (cfdart_init_*) write the same interior state,
 * bitwise, straight into a grid's device memory.
 */
#ifndef CFDART_HOST_H
#define CFDART_HOST_H

#include <stddef.h>

#include "cfdart.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Seeded random discs on an ambient Euler state. Not in the reference, whose discs are fixed
 * (euler_equations/euler_simulation_code/sourcecode/euler_io.cpp:528-536) or unseeded
 * (:705-724). Per disc, in this order, from std::mt19937_64(seed) mapped to [0,1) as
 * (x >> 11) * 2^-53: centre x, centre y, radius, pressure, density (velocity zero). The
 * cell fraction is the reference's 10x10 midpoint sub-sampling with its far-field cut-off
 * (euler_io.cpp:344-383); a disc replaces what lies under it: U <- (1-f) U + f U_disc. */
typedef struct cfdart_disc_params {
    unsigned long long seed;
    int n;
    double cx_lo, cx_hi, cy_lo, cy_hi;
    double r_lo, r_hi, p_lo, p_hi, rho_lo, rho_hi;
    double ambient[4]; /* rho, u, v, p */
} cfdart_disc_params;
int cfdart_ic_random_discs(const cfdart_desc* d, const cfdart_disc_params* p, double* host_state);

/* Seeded air bubbles in water hit by a planar shock (Allaire). Bubble draws per bubble:
 * centre x, centre y, radius. Each bubble blends primitives as set_circular_IC does
 * (allaire_2d/sourcecode/allaire_diffuse_io.cpp:365-399): W <- f W_air + (1-f) W; cells whose
 * centre has x <= shock_x then take the post-shock state (set_halfspace_IC, :312-330).
 * States are primitive 6-vectors (z rho1, (1-z) rho2, u, v, p, z). */
typedef struct cfdart_bubble_params {
    unsigned long long seed;
    int n;
    double cx_lo, cx_hi, cy_lo, cy_hi;
    double r_lo, r_hi;
    double water[6], air[6], postshock[6];
    double shock_x;
} cfdart_bubble_params;
int cfdart_ic_shocked_bubbles(const cfdart_desc* d, const cfdart_bubble_params* p, double* host_state);

/* At most this many discs / bubbles per generator call. */
#define CFDART_IC_MAX_DISCS 256

/* Planar initial condition: cells with a*x + b*y <= c take W_under, the rest W_over
 * (allaire_2d/sourcecode/allaire_diffuse_io.cpp:332-360); primitive vectors of the model. */
int cfdart_ic_planar(const cfdart_desc* d, const double* W_under, const double* W_over, double a,
                     double b, double c, double* host_state);

/* Device-side generation: the state cfdart_ic_random_discs / cfdart_ic_shocked_bubbles would
 * produce, written straight into the grid's device buffer (no host loop over cells, no
 * host-to-device copy of the state; the reference builds ICs on the host,
 * euler_equations/euler_simulation_code/sourcecode/euler_io.cpp:282-433). Resets the grid's
 * time-loop state as cfdart_upload does; synchronous. */
int cfdart_init_random_discs(cfdart_grid* g, const cfdart_disc_params* p);
int cfdart_init_shocked_bubbles(cfdart_grid* g, const cfdart_bubble_params* p);

/* Primitive -> conserved of the model (euler_equations/.../headers/misc_euler.hpp:82-97,
 * allaire_2d/headers/misc.hpp:149-163). */
int cfdart_prim_to_cons(const cfdart_desc* d, const double* W, double* U);

/* A named test case of the settings-file driver (set_parameters + set_ICs,
 * E/sourcecode/euler_io.cpp:13-234,436-760 and A/sourcecode/allaire_diffuse_io.cpp:13-310,
 * 400-990; problem "euler" or "allaire", flux_solver "MUSCL" or "Godunov"): its grid descriptor
 * with the case's defaults (limiter minmod, Garrick wave speeds, z clip [0, 1], the case's walls),
 * its end time T, and, when host_state is non-null, its initial state in the format of cfdart.h
 * ((nx+2ng)*(ny+2ng)*ncomp doubles, ghost frame included; host_len: the buffer's length in
 * doubles). Host only: no device is needed. CFDART_E_ARG names an unknown case in
 * cfdart_last_error. */
int cfdart_test_case(const char* problem, const char* test_case, int nx, int ny, const char* flux_solver,
                     cfdart_desc* d_out, double* host_state, size_t host_len, double* T_out);

#ifdef __cplusplus
}
#endif
#endif
