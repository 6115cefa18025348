// Initial conditions (host side of the reference's set_ICs): see include/cfdart_host.h.
#include "cfdart_host.h"

#include <cmath>
#include <random>

#include "physics.h"

namespace {

int ncomp(const cfdart_desc* d) { return d->model == CFDART_ALLAIRE ? 6 : 4; }

// sim_info::cellcentre_coord (euler_equations/euler_simulation_code/headers/sim_info.hpp:40-46)
double ccx(const cfdart_desc* d, int j) { return d->x0 + double(j - d->ng) * d->dx + 0.5 * d->dx; }
double ccy(const cfdart_desc* d, int i) { return d->y0 + double(i - d->ng) * d->dy + 0.5 * d->dy; }

// Fraction of cell (i, j) inside a circle by N x N midpoint samples
// (euler_io.cpp:344-383; the far-field cut-off of :351 when far_cut).
double circle_fraction(const cfdart_desc* d, int i, int j, double cx, double cy, double R, int N, bool far_cut)
{
    double x = ccx(d, j), y = ccy(d, i);
    if (far_cut) {
        double ex = x - cx, ey = y - cy;
        if (std::sqrt(ex * ex + ey * ey) > R + 2.0 * d->dx) return 0.0;
    }
    int inside = 0;
    double delx = d->dx / N, dely = d->dy / N;
    double blx = x - 0.5 * d->dx, bly = y - 0.5 * d->dy;
    for (int a = 0; a < N; ++a)
        for (int b = 0; b < N; ++b) {
            double sx = blx + (a + 0.5) * delx;
            double sy = bly + (b + 0.5) * dely;
            sx -= cx;
            sy -= cy;
            if (std::sqrt(sx * sx + sy * sy) <= R) inside++;
        }
    return double(inside) / (N * N);
}

struct Uniform {
    std::mt19937_64 g;
    explicit Uniform(unsigned long long s) : g(s) {}
    double operator()(double lo, double hi) { return lo + (hi - lo) * (double(g() >> 11) * 0x1.0p-53); }
};

void p2c(const cfdart_desc* d, const double* W, double* U)
{
    cfd::Material m = cfd::make_material(d->gamma1, d->pinf1, d->gamma2, d->pinf2);
    if (d->model == CFDART_ALLAIRE) cfd::p2c<cfd::ALLAIRE>(m, W, U);
    else cfd::p2c<cfd::EULER>(m, W, U);
}

}  // namespace

extern "C" {

int cfdart_prim_to_cons(const cfdart_desc* d, const double* W, double* U)
{
    if (!d || !W || !U) return CFDART_E_ARG;
    p2c(d, W, U);
    return 0;
}

int cfdart_ic_random_discs(const cfdart_desc* d, const cfdart_disc_params* p, double* host)
{
    if (!d || !p || !host || d->model != CFDART_EULER || p->n < 0) return CFDART_E_ARG;
    const int nc = 4, W = d->nx + 2 * d->ng, H = d->ny + 2 * d->ng;
    double Ua[4];
    p2c(d, p->ambient, Ua);
    for (long long c = 0; c < (long long)W * H; ++c)
        for (int k = 0; k < nc; ++k) host[c * nc + k] = Ua[k];
    Uniform U(p->seed);
    for (int n = 0; n < p->n; ++n) {
        double cx = U(p->cx_lo, p->cx_hi);
        double cy = U(p->cy_lo, p->cy_hi);
        double R = U(p->r_lo, p->r_hi);
        double pr = U(p->p_lo, p->p_hi);
        double rho = U(p->rho_lo, p->rho_hi);
        double Wd[4] = {rho, 0.0, 0.0, pr}, Ud[4];
        p2c(d, Wd, Ud);
        // only cells near the disc can have a non-zero fraction
        int j0 = (int)std::floor((cx - R - d->x0) / d->dx) + d->ng - 3, j1 = (int)std::ceil((cx + R - d->x0) / d->dx) + d->ng + 3;
        int i0 = (int)std::floor((cy - R - d->y0) / d->dy) + d->ng - 3, i1 = (int)std::ceil((cy + R - d->y0) / d->dy) + d->ng + 3;
        if (j0 < 0) j0 = 0;
        if (i0 < 0) i0 = 0;
        if (j1 > W) j1 = W;
        if (i1 > H) i1 = H;
        for (int i = i0; i < i1; ++i)
            for (int j = j0; j < j1; ++j) {
                double f = circle_fraction(d, i, j, cx, cy, R, 10, true);
                if (f > 0.0) {
                    double* c = host + ((long long)i * W + j) * nc;
                    for (int k = 0; k < nc; ++k) c[k] = (1.0 - f) * c[k] + f * Ud[k];
                }
            }
    }
    return 0;
}

int cfdart_ic_shocked_bubbles(const cfdart_desc* d, const cfdart_bubble_params* p, double* host)
{
    if (!d || !p || !host || d->model != CFDART_ALLAIRE || p->n < 0) return CFDART_E_ARG;
    const int nc = 6, W = d->nx + 2 * d->ng, H = d->ny + 2 * d->ng;
    // primitive working field
    const long long ncell = (long long)W * H;
    double* prim = new double[ncell * nc];
    for (long long c = 0; c < ncell; ++c)
        for (int k = 0; k < nc; ++k) prim[c * nc + k] = p->water[k];
    Uniform U(p->seed);
    for (int n = 0; n < p->n; ++n) {
        double cx = U(p->cx_lo, p->cx_hi);
        double cy = U(p->cy_lo, p->cy_hi);
        double R = U(p->r_lo, p->r_hi);
        int j0 = (int)std::floor((cx - R - d->x0) / d->dx) + d->ng - 3, j1 = (int)std::ceil((cx + R - d->x0) / d->dx) + d->ng + 3;
        int i0 = (int)std::floor((cy - R - d->y0) / d->dy) + d->ng - 3, i1 = (int)std::ceil((cy + R - d->y0) / d->dy) + d->ng + 3;
        if (j0 < 0) j0 = 0;
        if (i0 < 0) i0 = 0;
        if (j1 > W) j1 = W;
        if (i1 > H) i1 = H;
        for (int i = i0; i < i1; ++i)
            for (int j = j0; j < j1; ++j) {
                double f = circle_fraction(d, i, j, cx, cy, R, 10, false);
                if (f > 0.0) {
                    double* w = prim + ((long long)i * W + j) * nc;
                    for (int k = 0; k < nc; ++k) w[k] = f * p->air[k] + (1.0 - f) * w[k];
                }
            }
    }
    for (int i = 0; i < H; ++i)
        for (int j = 0; j < W; ++j) {
            long long c = (long long)i * W + j;
            if (ccx(d, j) <= p->shock_x) for (int k = 0; k < nc; ++k) prim[c * nc + k] = p->postshock[k];
            p2c(d, prim + c * nc, host + c * nc);
        }
    delete[] prim;
    return 0;
}

int cfdart_ic_planar(const cfdart_desc* d, const double* Wu, const double* Wo, double a, double b, double c,
                     double* host)
{
    if (!d || !Wu || !Wo || !host) return CFDART_E_ARG;
    const int nc = ncomp(d), W = d->nx + 2 * d->ng, H = d->ny + 2 * d->ng;
    double Uu[6], Uo[6];
    p2c(d, Wu, Uu);
    p2c(d, Wo, Uo);
    for (int i = 0; i < H; ++i)
        for (int j = 0; j < W; ++j) {
            const double* s = (ccx(d, j) * a + ccy(d, i) * b <= c) ? Uu : Uo;
            for (int k = 0; k < nc; ++k) host[((long long)i * W + j) * nc + k] = s[k];
        }
    return 0;
}

}  // extern "C"
