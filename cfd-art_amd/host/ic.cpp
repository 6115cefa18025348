// Initial conditions (host side of the reference's set_ICs): see include/cfdart_host.h.
#include "cfdart_host.h"

#include <algorithm>
#include <cmath>
#include <random>
#include <vector>

#include "ic_common.h"
#include "physics.h"

namespace {

int ncomp(const cfdart_desc* d) { return d->model == CFDART_ALLAIRE ? 6 : 4; }

// sim_info::cellcentre_coord (euler_equations/euler_simulation_code/headers/sim_info.hpp:40-46)
double ccx(const cfdart_desc* d, int j) { return cfd::cell_centre(d->x0, d->dx, j - d->ng); }
double ccy(const cfdart_desc* d, int i) { return cfd::cell_centre(d->y0, d->dy, i - d->ng); }

struct Uniform {
    std::mt19937_64 g;
    explicit Uniform(unsigned long long s) : g(s) {}
    double operator()(double lo, double hi) { return lo + (hi - lo) * (double(g() >> 11) * 0x1.0p-53); }
};

void host_p2c(const cfdart_desc* d, const double* W, double* U)
{
    cfd::Material m = cfd::make_material(d->gamma1, d->pinf1, d->gamma2, d->pinf2);
    if (d->model == CFDART_ALLAIRE) cfd::p2c<cfd::ALLAIRE>(m, W, U);
    else cfd::p2c<cfd::EULER>(m, W, U);
}

// Cells of the ghost-framed grid that can intersect a circle (others have fraction 0).
void circle_box(const cfdart_desc* d, double cx, double cy, double R, int& i0, int& i1, int& j0, int& j1)
{
    const int W = d->nx + 2 * d->ng, H = d->ny + 2 * d->ng;
    j0 = std::max(0, (int)std::floor((cx - R - d->x0) / d->dx) + d->ng - 3);
    j1 = std::min(W, (int)std::ceil((cx + R - d->x0) / d->dx) + d->ng + 3);
    i0 = std::max(0, (int)std::floor((cy - R - d->y0) / d->dy) + d->ng - 3);
    i1 = std::min(H, (int)std::ceil((cy + R - d->y0) / d->dy) + d->ng + 3);
}

}  // namespace

namespace cfd {

int draw_discs(const cfdart_desc* d, const cfdart_disc_params* p, Disc* out, int max_out)
{
    if (!d || !p || p->n < 0 || p->n > max_out) return CFDART_E_ARG;
    Uniform U(p->seed);
    for (int n = 0; n < p->n; ++n) {
        Disc& c = out[n];
        c.cx = U(p->cx_lo, p->cx_hi);
        c.cy = U(p->cy_lo, p->cy_hi);
        c.R = U(p->r_lo, p->r_hi);
        double pr = U(p->p_lo, p->p_hi);
        double rho = U(p->rho_lo, p->rho_hi);
        double Wd[4] = {rho, 0.0, 0.0, pr};
        host_p2c(d, Wd, c.U);
    }
    return 0;
}

int draw_bubbles(const cfdart_bubble_params* p, Bubble* out, int max_out)
{
    if (!p || p->n < 0 || p->n > max_out) return CFDART_E_ARG;
    Uniform U(p->seed);
    for (int n = 0; n < p->n; ++n) {
        out[n].cx = U(p->cx_lo, p->cx_hi);
        out[n].cy = U(p->cy_lo, p->cy_hi);
        out[n].R = U(p->r_lo, p->r_hi);
    }
    return 0;
}

}  // namespace cfd

extern "C" {

int cfdart_prim_to_cons(const cfdart_desc* d, const double* W, double* U)
{
    if (!d || !W || !U) return CFDART_E_ARG;
    host_p2c(d, W, U);
    return 0;
}

int cfdart_ic_random_discs(const cfdart_desc* d, const cfdart_disc_params* p, double* host)
{
    if (!d || !p || !host || d->model != CFDART_EULER || p->n < 0 || p->n > cfd::IC_MAX_DISCS) return CFDART_E_ARG;
    const int nc = 4, W = d->nx + 2 * d->ng, H = d->ny + 2 * d->ng;
    double Ua[4];
    host_p2c(d, p->ambient, Ua);
    for (long long c = 0; c < (long long)W * H; ++c)
        for (int k = 0; k < nc; ++k) host[c * nc + k] = Ua[k];
    std::vector<cfd::Disc> discs(p->n);
    if (p->n) cfd::draw_discs(d, p, discs.data(), p->n);
    for (const cfd::Disc& c : discs) {
        int i0, i1, j0, j1;
        circle_box(d, c.cx, c.cy, c.R, i0, i1, j0, j1);
        for (int i = i0; i < i1; ++i)
            for (int j = j0; j < j1; ++j) {
                double f = cfd::circle_fraction(ccx(d, j), ccy(d, i), d->dx, d->dy, c.cx, c.cy, c.R, 10, true);
                if (f > 0.0) {
                    double* u = host + ((long long)i * W + j) * nc;
                    for (int k = 0; k < nc; ++k) u[k] = (1.0 - f) * u[k] + f * c.U[k];
                }
            }
    }
    return 0;
}

int cfdart_ic_shocked_bubbles(const cfdart_desc* d, const cfdart_bubble_params* p, double* host)
{
    if (!d || !p || !host || d->model != CFDART_ALLAIRE || p->n < 0 || p->n > cfd::IC_MAX_DISCS) return CFDART_E_ARG;
    const int nc = 6, W = d->nx + 2 * d->ng, H = d->ny + 2 * d->ng;
    const long long ncell = (long long)W * H;
    std::vector<double> prim((size_t)ncell * nc);
    for (long long c = 0; c < ncell; ++c)
        for (int k = 0; k < nc; ++k) prim[c * nc + k] = p->water[k];
    std::vector<cfd::Bubble> bub(p->n);
    if (p->n) cfd::draw_bubbles(p, bub.data(), p->n);
    for (const cfd::Bubble& c : bub) {
        int i0, i1, j0, j1;
        circle_box(d, c.cx, c.cy, c.R, i0, i1, j0, j1);
        for (int i = i0; i < i1; ++i)
            for (int j = j0; j < j1; ++j) {
                double f = cfd::circle_fraction(ccx(d, j), ccy(d, i), d->dx, d->dy, c.cx, c.cy, c.R, 10, false);
                if (f > 0.0) {
                    double* w = &prim[((size_t)i * W + j) * nc];
                    for (int k = 0; k < nc; ++k) w[k] = f * p->air[k] + (1.0 - f) * w[k];
                }
            }
    }
    for (int i = 0; i < H; ++i)
        for (int j = 0; j < W; ++j) {
            size_t c = (size_t)i * W + j;
            if (ccx(d, j) <= p->shock_x) for (int k = 0; k < nc; ++k) prim[c * nc + k] = p->postshock[k];
            host_p2c(d, &prim[c * nc], host + c * nc);
        }
    return 0;
}

int cfdart_ic_planar(const cfdart_desc* d, const double* Wu, const double* Wo, double a, double b, double c,
                     double* host)
{
    if (!d || !Wu || !Wo || !host) return CFDART_E_ARG;
    const int nc = ncomp(d), W = d->nx + 2 * d->ng, H = d->ny + 2 * d->ng;
    double Uu[6], Uo[6];
    host_p2c(d, Wu, Uu);
    host_p2c(d, Wo, Uo);
    for (int i = 0; i < H; ++i)
        for (int j = 0; j < W; ++j) {
            const double* s = (ccx(d, j) * a + ccy(d, i) * b <= c) ? Uu : Uo;
            for (int k = 0; k < nc; ++k) host[((long long)i * W + j) * nc + k] = s[k];
        }
    return 0;
}

}  // extern "C"
