// Initial-condition arithmetic shared by the host generator (host/ic.cpp) and the device
// generator (batch.hip), so both produce bitwise the same state.
#pragma once

#include <math.h>

#include "cfdart_host.h"
#include "physics.h"

namespace cfd {

// sim_info::cellcentre_coord (euler_equations/euler_simulation_code/headers/sim_info.hpp:40-46),
// with the interior index (ghost-frame index minus ng).
CFD_HD double cell_centre(double x0, double dx, int interior_index)
{
    return x0 + double(interior_index) * dx + 0.5 * dx;
}

// Fraction of a cell inside a circle by N x N midpoint samples (euler_io.cpp:344-383); with
// far_cut, the reference's shortcut |centre distance| > R + 2 dx -> 0 (euler_io.cpp:351).
CFD_HD double circle_fraction(double x, double y, double dx, double dy, double cx, double cy, double R, int N,
                              bool far_cut)
{
    if (far_cut) {
        double ex = x - cx, ey = y - cy;
        if (sqrt(ex * ex + ey * ey) > R + 2.0 * dx) return 0.0;
    }
    int inside = 0;
    double delx = dx / N, dely = dy / N;
    double blx = x - 0.5 * dx, bly = y - 0.5 * dy;
    for (int a = 0; a < N; ++a)
        for (int b = 0; b < N; ++b) {
            double sx = blx + (a + 0.5) * delx;
            double sy = bly + (b + 0.5) * dely;
            sx -= cx;
            sy -= cy;
            if (sqrt(sx * sx + sy * sy) <= R) inside++;
        }
    return double(inside) / (N * N);
}

// One drawn disc (Euler) or bubble (Allaire): centre, radius and, for discs, the conserved
// state inside.
struct Disc {
    double cx, cy, R;
    double U[4];
};
struct Bubble {
    double cx, cy, R;
};

constexpr int IC_MAX_DISCS = CFDART_IC_MAX_DISCS;

// The seeded draws of the generators (host/ic.cpp): discs with their conserved inner state,
// bubbles with centre and radius, in the generators' order.
int draw_discs(const cfdart_desc* d, const cfdart_disc_params* p, Disc* out, int max_out);
int draw_bubbles(const cfdart_bubble_params* p, Bubble* out, int max_out);

}  // namespace cfd
