// Settings file, initial conditions, frame output and the time loop of the driver; see
// problem.hpp for the reference files each part mirrors.
#include "problem.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <functional>
#include <iostream>
#include <random>
#include <sstream>

namespace cfdhost {

// ------------------------------------------------------------------ settings file

std::string Settings::read(const std::string& filename)
{
    std::ifstream in(filename);
    if (!in) return "cannot open settings file " + filename;
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream iss(line);
        std::string key;
        iss >> key;
        if (key == "Nx") iss >> Nx;
        else if (key == "Ny") iss >> Ny;
        else if (key == "CFL") iss >> CFL;
        else if (key == "splitting_scheme") iss >> splitting_scheme;
        else if (key == "test_case") iss >> test_case;
        else if (key == "output_freq") iss >> output_freq;
        else if (key == "outputpath") iss >> outputpath;
        else if (key == "flux_solver") iss >> flux_solver;
        else if (key == "riemann_solver") iss >> riemann_solver;
        else if (key == "problem") iss >> problem;
        else if (key == "limiter") iss >> limiter;
        else if (key == "wave_speeds") iss >> wave_speeds;
        else if (key == "num_steps") iss >> num_steps;
        else if (key == "T") iss >> T;
        else if (key == "seed") iss >> seed;
        else if (key == "n_discs") iss >> n_discs;
        else if (key == "bc_left") iss >> bc[0];
        else if (key == "bc_right") iss >> bc[1];
        else if (key == "bc_bottom") iss >> bc[2];
        else if (key == "bc_top") iss >> bc[3];
        else if (key == "z_clip") iss >> z_clip;
        else if (key == "frame_every_steps") iss >> frame_every_steps;
        else if (key == "colsweep_uses_dx") iss >> colsweep_uses_dx;
    }
    // settings_file.hpp:69
    basename = outputpath + problem + "-" + test_case + "-" + flux_solver + "-" + std::to_string(Nx) + "-" +
               std::to_string(Ny);
    if (Nx < 2 || Ny < 2) return "Nx and Ny must be >= 2";
    if (!(CFL > 0.0)) return "CFL must be positive";
    return "";
}

// ------------------------------------------------------------------ initial conditions

namespace {

int ncomp(const cfdart_desc& d) { return d.model == CFDART_ALLAIRE ? 6 : 4; }
double ccx(const cfdart_desc& d, int j) { return d.x0 + double(j - d.ng) * d.dx + 0.5 * d.dx; }
double ccy(const cfdart_desc& d, int i) { return d.y0 + double(i - d.ng) * d.dy + 0.5 * d.dy; }

// eos::specific_ie of one stiffened gas (stiffened_gas_eos.hpp:73-76)
double sg_specific_ie(double gamma, double pinf, double p, double rho) { return (p + gamma * pinf) / ((gamma - 1.0) * rho); }

std::vector<double> p2c(const cfdart_desc& d, std::vector<double> W)
{
    std::vector<double> U(W.size());
    cfdart_prim_to_cons(&d, W.data(), U.data());
    return U;
}

struct Field {
    const cfdart_desc& d;
    std::vector<double>& h;
    int W, H, nc;
    Field(const cfdart_desc& dd, std::vector<double>& hh)
        : d(dd), h(hh), W(dd.nx + 2 * dd.ng), H(dd.ny + 2 * dd.ng), nc(ncomp(dd)) {}
    double* at(int i, int j) { return &h[((size_t)i * W + j) * nc]; }
    void fill(const std::vector<double>& U)
    {
        for (int i = 0; i < H; ++i)
            for (int j = 0; j < W; ++j)
                for (int k = 0; k < nc; ++k) at(i, j)[k] = U[k];
    }
    // add_circle_to_IC (euler_io.cpp:335-389): additive, far-field cut-off, N x N samples
    void add_circle(const std::vector<double>& Win, double cx, double cy, double R, int N = 10)
    {
        std::vector<double> Uin = p2c(d, Win);
        for (int i = 0; i < H; ++i)
            for (int j = 0; j < W; ++j) {
                double x = ccx(d, j), y = ccy(d, i), frac;
                double ex = x - cx, ey = y - cy;
                if (std::sqrt(ex * ex + ey * ey) > R + 2.0 * d.dx) {
                    frac = 0.0;
                } else {
                    int inside = 0;
                    double delx = d.dx / N, dely = d.dy / N, blx = x - 0.5 * d.dx, bly = y - 0.5 * d.dy;
                    for (int a = 0; a < N; ++a)
                        for (int b = 0; b < N; ++b) {
                            double sx = blx + (a + 0.5) * delx - cx, sy = bly + (b + 0.5) * dely - cy;
                            if (std::sqrt(sx * sx + sy * sy) <= R) inside++;
                        }
                    frac = double(inside) / (N * N);
                }
                for (int k = 0; k < nc; ++k) at(i, j)[k] += frac * Uin[k];
            }
    }
    // add_rectangle_to_IC (euler_io.cpp:282-330): additive
    void add_rect(const std::vector<double>& Win, double blx0, double bly0, double w, double hgt, int N = 10)
    {
        std::vector<double> Uin = p2c(d, Win);
        for (int i = 0; i < H; ++i)
            for (int j = 0; j < W; ++j) {
                int inside = 0;
                double delx = d.dx / N, dely = d.dy / N;
                double blx = ccx(d, j) - 0.5 * d.dx, bly = ccy(d, i) - 0.5 * d.dy;
                for (int a = 0; a < N; ++a)
                    for (int b = 0; b < N; ++b) {
                        double sx = blx + (a + 0.5) * delx - blx0, sy = bly + (b + 0.5) * dely - bly0;
                        if ((0.0 <= sx) & (sx <= w) & (0.0 <= sy) & (sy <= hgt)) inside++;
                    }
                double frac = double(inside) / (N * N);
                for (int k = 0; k < nc; ++k) at(i, j)[k] += frac * Uin[k];
            }
    }
    // set_halfspace_IC (allaire_diffuse_io.cpp:312-330): U where a x + b y <= c at the cell centre
    void set_halfspace(const std::vector<double>& U, double a, double b, double c)
    {
        for (int i = 0; i < H; ++i)
            for (int j = 0; j < W; ++j)
                if (ccx(d, j) * a + ccy(d, i) * b <= c)
                    for (int k = 0; k < nc; ++k) at(i, j)[k] = U[k];
    }
    // Fraction of a cell's N x N midpoint samples (the reference's sampling) where inside(x, y)
    template <class In> double sample_frac(int i, int j, int N, const In& inside)
    {
        int count = 0;
        const double delx = d.dx / N, dely = d.dy / N;
        const double blx = ccx(d, j) - 0.5 * d.dx, bly = ccy(d, i) - 0.5 * d.dy;
        for (int a = 0; a < N; ++a)
            for (int b = 0; b < N; ++b)
                if (inside(blx + (a + 0.5) * delx, bly + (b + 0.5) * dely)) count++;
        return double(count) / (N * N);
    }
    // A two-fluid cell from its volume fraction and the fluids' densities and specific internal
    // energies (allaire_diffuse_io.cpp:697-702)
    void set_mixture(int i, int j, double z, double rho1, double rho2, double u, double v, double e1, double e2)
    {
        double* U = at(i, j);
        U[0] = z * rho1;
        U[1] = (1.0 - z) * rho2;
        U[2] = u * (z * rho1 + (1.0 - z) * rho2);
        U[3] = v * (z * rho1 + (1.0 - z) * rho2);
        U[4] = z * rho1 * e1 + (1.0 - z) * rho2 * e2 + 0.5 * (z * rho1 + (1.0 - z) * rho2) * (u * u + v * v);
        U[5] = z;
    }
    // set_circular_IC (euler_io.cpp:392-433): primitive blend of inside/outside states
    void set_circular(const std::vector<double>& Win, const std::vector<double>& Wout, double cx, double cy,
                      double R, int N = 10)
    {
        for (int i = 0; i < H; ++i)
            for (int j = 0; j < W; ++j) {
                int inside = 0;
                double delx = d.dx / N, dely = d.dy / N;
                double blx = ccx(d, j) - 0.5 * d.dx, bly = ccy(d, i) - 0.5 * d.dy;
                for (int a = 0; a < N; ++a)
                    for (int b = 0; b < N; ++b) {
                        double sx = blx + (a + 0.5) * delx - cx, sy = bly + (b + 0.5) * dely - cy;
                        if (std::sqrt(sx * sx + sy * sy) <= R) inside++;
                    }
                double frac = double(inside) / (N * N);
                std::vector<double> Wc(nc);
                for (int k = 0; k < nc; ++k) Wc[k] = frac * Win[k] + (1.0 - frac) * Wout[k];
                std::vector<double> U = p2c(d, Wc);
                for (int k = 0; k < nc; ++k) at(i, j)[k] = U[k];
            }
    }
};

int bc_code(const std::string& s) { return s == "reflective" ? CFDART_REFLECTIVE : CFDART_TRANSMISSIVE; }

// The grid, initial state (the host format of cfdart.h, ghost frame included) and end time of
// a test case: set_parameters and set_ICs of euler_io.cpp:13-234,436-760 and
// allaire_diffuse_io.cpp:13-310,400-990.
std::string build_case(const Settings& sf, cfdart_desc& d, std::vector<double>& host, double& T)
{
    d = cfdart_desc{};
    d.nx = sf.Nx;
    d.ny = sf.Ny;
    d.flux_solver = sf.flux_solver == "Godunov" ? CFDART_GODUNOV : CFDART_MUSCL_HANCOCK;
    d.ng = sf.flux_solver == "Godunov" ? 1 : 2;
    d.limiter = sf.limiter == "vanleer" ? CFDART_VANLEER : CFDART_MINMOD;
    d.wave_speeds = sf.wave_speeds == "pvrs" ? CFDART_PVRS : CFDART_GARRICK;
    d.colsweep_uses_dx = sf.colsweep_uses_dx;
    d.z_lo = 0.0;
    d.z_hi = 1.0;
    std::string bc[4] = {"transmissive", "transmissive", "transmissive", "transmissive"};  // L R B T
    auto domain = [&](double x0, double y0, double lx, double ly, double t_end) {
        d.x0 = x0; d.y0 = y0; d.dx = lx / d.nx; d.dy = ly / d.ny; T = t_end;
    };
    const std::string& tc = sf.test_case;
    host.assign((size_t)(d.nx + 2 * d.ng) * (d.ny + 2 * d.ng) * (sf.problem == "allaire" ? 6 : 4), 0.0);

    if (sf.problem == "euler") {
        // set_parameters (euler_io.cpp:13-234)
        d.model = CFDART_EULER;
        d.gamma1 = 1.4;
        d.pinf1 = 0.0;
        if (tc == "circular_explosion") { domain(0, 0, 2, 2, 1.0); bc[3] = bc[2] = "reflective"; }
        else if (tc == "multiple_circles_1" || tc == "multiple_circles_2" || tc == "cross" || tc == "internet" ||
                 tc == "totem_pole_2") domain(0, 0, 1, 1, 1.0);
        else if (tc == "totem_pole_1") { domain(0, 0, 1, 1, 1.0); bc[3] = "reflective"; }
        else if (tc == "totem_pole_3") { domain(0, 0, 0.5, 1, 1.0); bc[0] = "reflective"; }
        else if (tc == "random_discs") domain(0, 0, 1, double(d.ny) / d.nx, 0.2);
        else return "invalid Euler test_case " + tc;
    } else if (sf.problem == "allaire") {
        // set_parameters (allaire_diffuse_io.cpp:13-310)
        d.model = CFDART_ALLAIRE;
        d.gamma1 = d.gamma2 = 1.4;
        d.pinf1 = d.pinf2 = 0.0;
        if (tc == "ST1_x" || tc == "ST1_y") { d.gamma2 = 2.4; domain(0, 0, 1, 1, 0.14); }
        else if (tc == "ST2_x" || tc == "ST2_y") { d.gamma1 = 4.4; d.pinf1 = 6e8; domain(-2, -2, 4, 4, 0.0009); }
        else if (tc.rfind("TTC1_", 0) == 0) domain(0, 0, 1, 1, 0.25);
        else if (tc.rfind("TTC2_", 0) == 0) domain(0, 0, 1, 1, 0.15);
        else if (tc.rfind("TTC3_", 0) == 0) domain(0, 0, 1, 1, 0.012);
        else if (tc.rfind("TTC4_", 0) == 0 || tc.rfind("TTC5_", 0) == 0) domain(0, 0, 1, 1, 0.035);
        else if (tc == "shocked_bubbles") {
            d.gamma1 = 4.4; d.pinf1 = 6e8;
            domain(0, 0, 1, double(d.ny) / d.nx, 2e-4);
        }
        else if (tc == "circular_explosion") domain(0, 0, 2, 2, 0.25);
        else if (tc == "underwater_shocked_bubble") { d.gamma1 = 7.15; d.pinf1 = 3309.0; domain(0, 0, 12, 12, 0.05); }
        else if (tc == "underwater_explosion") { d.gamma2 = 7.15; d.pinf2 = 3.309e8; domain(-5, -5, 10, 10, 0.005); }
        else if (tc == "underwater_explosion_modified") {
            d.gamma2 = 7.15; d.pinf2 = 3.309e8;
            domain(-5, -5, 10, 10, 0.010);
            bc[2] = bc[3] = "reflective";
        }
        else if (tc == "shocked_helium_bubble") { d.gamma2 = 1.667; domain(0, 0, 325, 89, 280.0); bc[2] = bc[3] = "reflective"; }
        else if (tc == "shocked_R22_bubble") { d.gamma2 = 1.249; domain(0, 0, 0.445, 0.089, 0.001080); bc[2] = bc[3] = "reflective"; }
        else if (tc == "shocked_SF6") { d.gamma2 = 1.076; domain(0, 0, 0.45, 0.2, 0.25); bc[1] = bc[2] = bc[3] = "reflective"; }
        else if (tc == "RMI_SF6") { d.gamma2 = 1.093; domain(0, 0, 4.0, 0.5, 10.0); bc[2] = bc[3] = "reflective"; }
        else if (tc == "tin_air_implosion") { d.gamma1 = 3.27; d.pinf1 = 149500.0; domain(0, -25, 25, 50, 0.08); bc[0] = "reflective"; }
        else if (tc == "TSTM") { d.gamma1 = 1.5; domain(0, 0, 7, 3, 8.0); }
        else return "invalid Allaire test_case " + tc;
        if (sf.z_clip >= 0.0) { d.z_lo = sf.z_clip; d.z_hi = 1.0 - sf.z_clip; }
    } else {
        return "problem must be euler or allaire";
    }
    for (int s = 0; s < 4; ++s)
        if (!sf.bc[s].empty()) bc[s] = sf.bc[s];
    for (int s = 0; s < 4; ++s) d.bc[s] = bc_code(bc[s]);
    if (sf.T > 0.0) T = sf.T;
    if (sf.num_steps >= 0) T = INFINITY;  // fixed step count

    Field f(d, host);
    using V = std::vector<double>;
    if (d.model == CFDART_EULER) {
        V bg = {0.125, 0.0, 0.0, 0.1};
        if (tc == "circular_explosion") {
            f.set_circular({1.0, 0.0, 0.0, 1.0}, {0.125, 0.0, 0.0, 0.1}, 1.0, 1.0, 0.4);
        } else if (tc == "multiple_circles_1") {
            f.fill(p2c(d, bg));
            const double c[9][2] = {{0.5, 0.4}, {0.3, 0.7}, {0.2, 0.7}, {0.6, 0.5}, {0.1, 0.3},
                                    {0.3, 0.5}, {0.4, 0.6}, {0.5, 0.4}, {0.5, 0.3}};
            for (auto& p : c) f.add_circle({1.0, 0.0, 0.0, 1.0}, p[0], p[1], 0.1);
        } else if (tc == "multiple_circles_2") {
            f.fill(p2c(d, bg));
            for (double cx : {0.0, 1.0})
                for (double cy : {0.0, 1.0}) f.add_circle({1.0, 0.0, 0.0, 1.0}, cx, cy, 0.2);
            f.add_circle({0.4, 0.0, 0.0, 0.0}, 0.8, 0.5, 0.1);
            f.add_circle({0.3, 0.0, 0.0, 0.0}, 0.7, 0.5, 0.2);
            f.add_circle({0.2, 0.0, 0.0, 0.0}, 0.6, 0.5, 0.3);
            f.add_circle({0.1, 0.0, 0.0, 0.0}, 0.5, 0.5, 0.4);
        } else if (tc == "cross" || tc == "totem_pole_1" || tc == "totem_pole_2") {
            f.fill(p2c(d, bg));
            for (double cx : {0.0, 1.0})
                for (double cy : {0.0, 1.0}) f.add_circle({1.0, 0.0, 0.0, 5.0}, cx, cy, 0.2);
            if (tc == "cross") {
                const double c[4][2] = {{0.5, 0.1}, {0.5, 0.9}, {0.3, 0.7}, {0.7, 0.7}};
                for (auto& p : c) f.add_circle({0.5, 0.0, 0.0, 0.0}, p[0], p[1], 0.11);
                f.add_rect({0.5, 0.0, 0.0, 0.0}, 0.45, 0.1, 0.1, 0.8);
                f.add_rect({0.5, 0.0, 0.0, 0.0}, 0.3, 0.65, 0.4, 0.1);
            } else if (tc == "totem_pole_1") {
                f.add_rect({0.5, 0, 0, 0}, 0.45, 0.1, 0.1, 0.8);
                f.add_rect({0.5, 0, 0, 0}, 0.3, 0.65, 0.4, 0.1);
                f.add_rect({0.2, 0, 0, 0}, 0.4, 0.5, 0.2, 0.05);
                f.add_rect({0.2, 0, 0, 0}, 0.4, 0.375, 0.2, 0.075);
                f.add_rect({0.2, 0, 0, 0}, 0.4, 0.25, 0.2, 0.1);
                f.add_rect({0.5, 0, 0, 0}, 0.45, 0.1, 0.1, 0.1);
            } else {
                f.add_rect({0.7, 0, 0, 0}, 0.45, 0.1, 0.1, 0.8);
                f.add_rect({0.7, 0, 0, 0}, 0.3, 0.65, 0.4, 0.1);
                f.add_rect({0.4, 0, 0, 0}, 0.45, 0.575, 0.2, 0.05);
                f.add_rect({0.4, 0, 0, 0}, 0.35, 0.5, 0.2, 0.05);
                f.add_rect({0.4, 0, 0, 0}, 0.4, 0.375, 0.2, 0.075);
                f.add_rect({0.4, 0, 0, 0}, 0.35, 0.25, 0.3, 0.1);
                f.add_rect({0.7, 0, 0, 0}, 0.45, 0.1, 0.1, 0.1);
            }
        } else if (tc == "totem_pole_3") {
            f.fill(p2c(d, bg));
            f.add_circle({1.0, 0.0, 0.0, 5.0}, 0.5, 0.0, 0.2);
            f.add_circle({1.0, 0.0, 0.0, 5.0}, 0.5, 1.0, 0.2);
            f.add_rect({0.5, 0, 0, 0}, -0.05, 0.0, 0.1, 0.9);
            f.add_rect({0.5, 0, 0, 0}, -0.2, 0.65, 0.4, 0.1);
            f.add_rect({0.2, 0, 0, 0}, -0.1, 0.5, 0.2, 0.1);
            f.add_rect({0.4, 0, 0, 0}, -0.15, 0.375, 0.3, 0.05);
            f.add_rect({0.2, 0, 0, 0}, -0.1, 0.2, 0.2, 0.1);
            f.add_rect({0.5, 0, 0, 0}, -0.15, 0.0, 0.3, 0.15);
        } else if (tc == "internet") {
            // euler_io.cpp:688-724 draws from an unseeded std::random_device; here the draws are
            // seeded (std::mt19937_64(seed), 53-bit mapping) and keep their order: two
            // discarded centre jitters, then the density of each of the 16 x 16 circles.
            f.fill(p2c(d, bg));
            std::mt19937_64 gen(sf.seed);
            auto dis = [&]() { return double(gen() >> 11) * 0x1.0p-53; };
            const int np = 16;
            const double dxp = 1.0 / np;
            for (int i = 0; i < np; ++i)
                for (int j = 0; j < np; ++j) {
                    double cx = (i + 0.5) * dxp + dis() * 0.0, cy = (j + 0.5) * dxp + dis() * 0.0;
                    double rho = 5.0 * dis();
                    f.add_circle({rho, 0.0, 0.0, rho}, cx, cy, 0.02);
                }
        } else if (tc == "random_discs") {
            cfdart_disc_params p{};
            p.seed = sf.seed;
            p.n = sf.n_discs >= 0 ? sf.n_discs : 8;
            double ly = double(d.ny) / d.nx;
            p.cx_lo = 0.1; p.cx_hi = 0.9; p.cy_lo = 0.1 * ly; p.cy_hi = 0.9 * ly;
            p.r_lo = 0.03; p.r_hi = 0.10; p.p_lo = 5.0; p.p_hi = 20.0; p.rho_lo = 0.5; p.rho_hi = 2.0;
            p.ambient[0] = 1.0; p.ambient[1] = 0.0; p.ambient[2] = 0.0; p.ambient[3] = 1.0;
            if (cfdart_ic_random_discs(&d, &p, host.data())) return "random_discs failed";
        }
    } else {
        // allaire_diffuse_io.cpp:467-564 (planar states; the _y variants keep the x velocity)
        V L, R;
        double cut = 0.5;
        if (tc.rfind("ST1_", 0) == 0) { L = {1.0, 0.0, 0.0, 0.0, 1.0, 1.0}; R = {0.0, 0.125, 0.0, 0.0, 0.1, 0.0}; }
        else if (tc.rfind("ST2_", 0) == 0) { L = {1000.0, 0.0, 0.0, 0.0, 1e9, 1.0}; R = {0.0, 50.0, 0.0, 0.0, 1e5, 0.0}; cut = 0.7; }
        else if (tc.rfind("TTC1_", 0) == 0) { L = {1.0, 0.0, 0.0, 0.0, 1.0, 1.0}; R = {0.125, 0.0, 0.0, 0.0, 0.1, 1.0}; }
        else if (tc.rfind("TTC2_", 0) == 0) { L = {1.0, 0.0, -2.0, 0.0, 0.4, 1.0}; R = {1.0, 0.0, 2.0, 0.0, 0.4, 1.0}; }
        else if (tc.rfind("TTC3_", 0) == 0) { L = {1.0, 0.0, 0.0, 0.0, 1000.0, 1.0}; R = {1.0, 0.0, 0.0, 0.0, 0.01, 1.0}; }
        else if (tc.rfind("TTC4_", 0) == 0) { L = {1.0, 0.0, 0.0, 0.0, 0.1, 1.0}; R = {1.0, 0.0, 0.0, 0.0, 100.0, 1.0}; }
        else if (tc.rfind("TTC5_", 0) == 0) { L = {5.99924, 0.0, 19.5975, 0.0, 460.894, 1.0}; R = {5.99242, 0.0, -6.19633, 0.0, 46.0950, 1.0}; }
        if (!L.empty()) {
            bool ydir = tc.find("_y") != std::string::npos;
            if (cfdart_ic_planar(&d, L.data(), R.data(), ydir ? 0.0 : 1.0, ydir ? 1.0 : 0.0, cut, host.data()))
                return "planar IC failed";
        } else if (tc == "circular_explosion") {  // allaire_diffuse_io.cpp:565-574
            f.set_circular({1.0, 0.0, 0.0, 0.0, 1.0, 1.0}, {0.0, 0.125, 0.0, 0.0, 0.1, 0.0}, 1.0, 1.0, 0.4);
        } else if (tc == "shocked_helium_bubble") {  // :575-588
            f.set_circular({0.0, 0.138, 0.0, 0.0, 1.0, 0.0}, {1.0, 0.0, 0.0, 0.0, 1.0, 1.0}, 175.0, 44.5, 25.0);
            f.set_halfspace(p2c(d, {1.3764, 0.0, -0.394, 0.0, 1.5698, 1.0}), -1.0, 0.0, -225.0);
        } else if (tc == "shocked_R22_bubble") {  // :589-602
            f.set_circular({0.0, 3.863, 0.0, 0.0, 1.01325e5, 0.0}, {1.225, 0.0, 0.0, 0.0, 1.01325e5, 1.0}, 0.225, 0.0445,
                           0.025, 25);
            f.set_halfspace(p2c(d, {1.686, 0.0, -113.5, 0.0, 1.59e5, 1.0}), -1.0, 0.0, -0.275);
        } else if (tc == "underwater_shocked_bubble") {  // :603-616
            f.set_circular({0.0, 0.0012, 0.0, 0.0, 1.0, 0.0}, {1.0, 0.0, 0.0, 0.0, 1.0, 1.0}, 6.0, 6.0, 3.0);
            f.set_halfspace(p2c(d, {1.31, 0.0, 67.32, 0.0, 19000.0, 1.0}), 1.0, 0.0, 2.4);
        } else if (tc == "underwater_explosion" || tc == "underwater_explosion_modified") {  // :617-630
            f.set_circular({1270.0, 0.0, 0.0, 0.0, 8.29e8, 1.0}, {0.0, 1000.0, 0.0, 0.0, 1.0e5, 0.0}, 0.0, 0.0, 1.0);
            f.set_halfspace(p2c(d, {1.0, 0.0, 0.0, 0.0, 1.0e5, 1.0}), 0.0, -1.0, -2.5);
        } else if (tc == "shocked_SF6" || tc == "TSTM" || tc == "RMI_SF6") {
            // :631-704, :705-803, :804-876: fluid 1 pre- or post-shock by the cell centre, fluid 2
            // at rest, z = 1 - the fraction of 10 x 10 samples inside fluid 2's region
            struct Gas { double rho, p, u; };
            Gas pre, post, g2;
            std::function<bool(double, double)> in2;
            std::function<bool(double)> shocked;
            if (tc == "shocked_SF6") {
                pre = {1.153, 9.6856, 0.0}; post = {1.6672, 16.3256, 1.33273}; g2 = {5.805, 9.6856, 0.0};
                in2 = [](double x, double y) { return x >= 0.1 && x <= 0.25 && y >= 0.0 && y <= 0.1; };
                shocked = [](double cx) { return !(cx > 0.05); };
            } else if (tc == "TSTM") {
                pre = {0.125, 0.1, 0.0}; post = {1.0, 1.0, 0.0}; g2 = {1.0, 0.1, 0.0};
                in2 = [](double x, double y) { return x >= 1.0 && x <= 8.0 && y >= -1.0 && y <= 1.5; };
                shocked = [](double cx) { return !(cx > 1.0); };
            } else {
                pre = {1.0, 1.0, 0.0}; post = {1.411, 1.628, -0.39}; g2 = {5.04, 1.0, 0.0};
                const double pi = std::atan(1.0) * 4.0;
                in2 = [pi](double x, double y) { return 2.9 - 0.2 * std::sin(2.0 * pi * (y + 0.25)) > x; };
                shocked = [](double cx) { return !(cx < 3.2); };
            }
            const double e_pre = sg_specific_ie(d.gamma1, d.pinf1, pre.p, pre.rho);
            const double e_post = sg_specific_ie(d.gamma1, d.pinf1, post.p, post.rho);
            const double e2 = sg_specific_ie(d.gamma2, d.pinf2, g2.p, g2.rho);
            for (int i = 0; i < f.H; ++i)
                for (int j = 0; j < f.W; ++j) {
                    const bool s = shocked(ccx(d, j));
                    const double rho1 = s ? post.rho : pre.rho, u = s ? post.u : pre.u, e1 = s ? e_post : e_pre;
                    const double z = 1.0 - f.sample_frac(i, j, 10, in2);
                    f.set_mixture(i, j, z, rho1, g2.rho, u, 0.0, e1, e2);
                }
        } else if (tc == "tin_air_implosion") {  // :877-990
            const double e_pre = sg_specific_ie(d.gamma1, d.pinf1, 1.0, 7.28);
            const double e_post = sg_specific_ie(d.gamma1, d.pinf1, 1000000.0, 11.84);
            const double e2 = sg_specific_ie(d.gamma2, d.pinf2, 1.0, 0.001);
            auto in_tin = [](double x, double y) { return std::sqrt(x * x + y * y) <= 24.0; };
            auto in_interface = [](double x, double y) {
                double theta = std::atan2(y, x);
                theta -= std::atan(1) * 2;
                const double r = 20.0 + 0.4 * std::cos(22 * theta) + 0.4 * std::cos(17 * theta) + 0.3 * std::cos(29 * theta);
                return std::sqrt(x * x + y * y) <= r;
            };
            for (int i = 0; i < f.H; ++i)
                for (int j = 0; j < f.W; ++j) {
                    const double in = f.sample_frac(i, j, 10, in_tin);
                    const double rho1 = in * 7.28 + (1.0 - in) * 11.84;
                    const double e1 = in * e_pre + (1.0 - in) * e_post;
                    const double z = 1.0 - f.sample_frac(i, j, 10, in_interface);
                    f.set_mixture(i, j, z, rho1, 0.001, 0.0, 0.0, e1, e2);
                }
        } else {  // shocked_bubbles
            double ze = sf.z_clip >= 0.0 ? sf.z_clip : 1e-6, zw = 1.0 - ze, ly = double(d.ny) / d.nx;
            d.z_lo = ze; d.z_hi = zw;
            cfdart_bubble_params p{};
            p.seed = sf.seed;
            p.n = sf.n_discs >= 0 ? sf.n_discs : 16;
            p.cx_lo = 0.15; p.cx_hi = 0.95; p.cy_lo = 0.05 * ly; p.cy_hi = 0.95 * ly;
            p.r_lo = 0.02; p.r_hi = 0.06;
            const double water[6] = {zw * 1000.0, (1 - zw) * 1.0, 0.0, 0.0, 1e5, zw};
            const double air[6] = {ze * 1000.0, (1 - ze) * 1.0, 0.0, 0.0, 1e5, ze};
            const double post[6] = {zw * 1323.65, (1 - zw) * 1.0, 681.58, 0.0, 1e9, zw};
            for (int k = 0; k < 6; ++k) { p.water[k] = water[k]; p.air[k] = air[k]; p.postshock[k] = post[k]; }
            p.shock_x = 0.1;
            if (cfdart_ic_shocked_bubbles(&d, &p, host.data())) return "shocked_bubbles failed";
        }
    }
    return "";
}

}  // namespace

GpuProblem::~GpuProblem()
{
    if (g_) cfdart_destroy(g_);
}

std::string GpuProblem::set_ICs(const Settings& sf)
{
    sf_ = sf;
    outname_ = sf.basename;
    cfl_ = sf.CFL;
    // the flux solver and its ghost width (euler_io.cpp:446-457, allaire_diffuse_io.cpp:412-423)
    if (sf.flux_solver != "MUSCL" && sf.flux_solver != "Godunov") return "flux_solver must be MUSCL or Godunov";
    if (sf.riemann_solver != "HLLC") return "riemann_solver must be HLLC";
    if (sf.splitting_scheme == "naive_splitting") split_ = CFDART_NAIVE;
    else if (sf.splitting_scheme == "strang_splitting") split_ = CFDART_STRANG;
    else if (sf.splitting_scheme == "alternating_splitting") split_ = CFDART_ALTERNATING;
    else return "unknown splitting_scheme " + sf.splitting_scheme;

    cfdart_desc& d = d_;
    const std::string err = build_case(sf, d, host_, T_);
    if (!err.empty()) return err;

    int rc = cfdart_create(&d, &g_);
    if (rc) return std::string("cfdart_create: ") + cfdart_last_error();
    rc = cfdart_upload(g_, host_.data());
    if (rc) return std::string("cfdart_upload: ") + cfdart_last_error();
    return "";
}

// ------------------------------------------------------------------ output

void GpuProblem::write_state(double t)
{
    std::ofstream out(outname_ + "-state-" + std::to_string(t) + ".dat");
    const cfdart_desc& d = d_;
    Field f(d, host_);
    for (int i = d.ng; i < d.ny + d.ng; ++i) {
        for (int j = d.ng; j < d.nx + d.ng; ++j) {
            const double* U = f.at(i, j);
            if (d.model == CFDART_EULER) {
                // euler_io.cpp:1196-1207 (its pressure call passes rho and e swapped)
                double rho = U[0], u = U[1] / rho, v = U[2] / rho, E = U[3];
                double e = E / rho - 0.5 * (u * u + v * v);
                double p = (d.gamma1 - 1.0) * e * rho - d.gamma1 * d.pinf1;
                out << ccx(d, j) << " " << ccy(d, i) << " " << rho << " " << u << " " << v << " " << e << " " << p << "\n";
            } else {
                // allaire_diffuse_io.cpp:1163-1176
                double rho = U[0] + U[1], u = U[2] / rho, v = U[3] / rho, E = U[4];
                double e = E / rho - 0.5 * (u * u + v * v), z = U[5];
                double eps = z / (d.gamma1 - 1.0) + (1.0 - z) / (d.gamma2 - 1.0);
                double k = (z * d.gamma1 * d.pinf1 / (d.gamma1 - 1.0) + (1.0 - z) * d.gamma2 * d.pinf2 / (d.gamma2 - 1.0)) / eps;
                double p = rho * e / eps - k;
                out << ccx(d, j) << " " << ccy(d, i) << " " << rho << " " << u << " " << v << " " << e << " " << z << " " << p << "\n";
            }
        }
        out << "\n";
    }
    frames_++;
}

void GpuProblem::write_schlieren(double t)
{
    // euler_io.cpp:1142-1185: |grad rho| / max |grad rho| by central differences, computed on
    // the device (cfdart_schlieren); the ghost frame is the mirror of the current state
    const cfdart_desc& d = d_;
    std::vector<double> g((size_t)d.nx * d.ny);
    if (cfdart_schlieren(g_, g.data(), nullptr) != 0) return;
    std::ofstream out(outname_ + "-schlieren-" + std::to_string(t) + ".dat");
    size_t c = 0;
    for (int i = d.ng; i < d.ny + d.ng; ++i) {
        for (int j = d.ng; j < d.nx + d.ng; ++j) out << ccx(d, j) << " " << ccy(d, i) << " " << g[c++] << "\n";
        out << "\n";
    }
}

void GpuProblem::write_masschange()
{
    // euler_io.cpp:1219-1233 (absolute); allaire_diffuse_io.cpp:1235-1249 (relative per phase)
    std::ofstream out(outname_ + "-masschange.dat");
    for (size_t k = 0; k < time_.size(); ++k) {
        if (d_.model == CFDART_EULER) out << time_[k] << " " << mass1_[k] - mass1_[0] << "\n";
        else out << time_[k] << " " << (mass1_[k] - mass1_[0]) / mass1_[0] << " " << (mass2_[k] - mass2_[0]) / mass2_[0] << "\n";
    }
}

void GpuProblem::output(int n, double t, bool force)
{
    // euler_io.cpp:1235-1261 cadence; frame_every_steps adds a fixed step cadence
    bool frame = false, final = (t == T_) || force;
    if (n == 0 || final) frame = true;
    else if (sf_.frame_every_steps > 0) frame = (n % sf_.frame_every_steps) == 0;
    else if (sf_.output_freq != 0 && std::isfinite(T_)) {
        const double interval = T_ / sf_.output_freq;
        if (t - last_output_time_ > interval) {
            frame = true;
            last_output_time_ = t;
        }
    }
    if (!frame) return;
    if (cfdart_download(g_, host_.data()) != 0) {
        std::cerr << "[cfdart] download failed: " << cfdart_last_error() << std::endl;
        return;
    }
    write_state(t);
    write_schlieren(t);
    if (final) write_masschange();
}

double GpuProblem::compute_dt(int n, double t)
{
    double dt = 0.0, mass[2] = {0.0, 0.0};
    if (cfdart_compute_dt(g_, n, t, T_, cfl_, &dt, mass) != 0) return -1.0;
    time_.push_back(t);
    mass1_.push_back(mass[0]);
    mass2_.push_back(mass[1]);
    return dt;
}

int GpuProblem::advance(int n, double dt) { return cfdart_step(g_, split_, n, dt); }

int GpuProblem::run_steps(int nsteps, double* t, int* n)
{
    // the device loop records compute_dt's mass history as it goes (cfdart_run_mass)
    std::vector<double> hist(3 * (size_t)std::max(nsteps, 1));
    const int n0 = *n;
    int rc = cfdart_run_mass(g_, split_, nsteps, INFINITY, sf_.CFL, t, n, hist.data());
    if (rc) return rc;
    for (int k = 0; k < *n - n0; ++k) {
        time_.push_back(hist[3 * (size_t)k]);
        mass1_.push_back(hist[3 * (size_t)k + 1]);
        mass2_.push_back(hist[3 * (size_t)k + 2]);
    }
    return 0;
}

// ------------------------------------------------------------------ time loop

int run_simulation(const Settings& sf, std::string* err)
{
    GpuProblem problem;
    std::string e = problem.set_ICs(sf);
    if (!e.empty()) {
        if (err) *err = e;
        return 1;
    }
    int n = 0;
    double t = 0.0;
    std::cout << "\n[simulation] Beginning time steps.." << std::endl;
    if (sf.num_steps >= 0) {
        // fixed step count: the loop stays on the device between frames
        long long left = sf.num_steps;
        problem.output(0, 0.0, false);
        int chunk = sf.frame_every_steps > 0 ? sf.frame_every_steps : (int)std::min<long long>(left, 1000);
        while (left > 0) {
            int k = (int)std::min<long long>(chunk, left);
            if (cfdart_set_time(problem.grid(), t, n) || problem.run_steps(k, &t, &n)) {
                if (err) *err = cfdart_last_error();
                return 1;
            }
            left -= k;
            std::cout << "[simulation] Time step " << n << " complete. Time t = " << t << "." << std::endl;
            problem.output(n, t, left == 0);
        }
    } else {
        // simulation.hpp:48-61
        while (t < problem.T()) {
            problem.output(n, t, false);
            double dt = problem.compute_dt(n, t);
            if (dt < 0.0 || problem.advance(n, dt) != 0) {
                if (err) *err = cfdart_last_error();
                return 1;
            }
            t += dt;
            n++;
            if (std::fabs(t - problem.T()) < 1e-12) t = problem.T();
            std::cout << "[simulation] Time step " << n << " complete. Time t = " << t << "." << std::endl;
        }
        std::cout << "[simulation] Time stepping complete." << std::endl;
        problem.output(n, t, true);
    }
    if (cfdart_sync(problem.grid()) != 0) {
        if (err) *err = cfdart_last_error();
        return 1;
    }
    return 0;
}

}  // namespace cfdhost

namespace cfd {
void set_last_error(const std::string& msg);  // cfdart_api.hip
}

extern "C" int cfdart_test_case(const char* problem, const char* test_case, int nx, int ny, const char* flux_solver,
                                cfdart_desc* d_out, double* host_state, size_t host_len, double* T_out)
{
    if (!problem || !test_case || !flux_solver || !d_out || nx < 2 || ny < 2) {
        cfd::set_last_error("bad argument");
        return CFDART_E_ARG;
    }
    cfdhost::Settings sf;
    sf.problem = problem;
    sf.test_case = test_case;
    sf.flux_solver = flux_solver;
    sf.Nx = nx;
    sf.Ny = ny;
    if (sf.problem == "allaire") sf.wave_speeds = "garrick";
    if (sf.flux_solver != "MUSCL" && sf.flux_solver != "Godunov") {
        cfd::set_last_error("flux_solver must be MUSCL or Godunov");
        return CFDART_E_ARG;
    }
    cfdart_desc d;
    std::vector<double> host;
    double T = 0.0;
    const std::string err = cfdhost::build_case(sf, d, host, T);
    if (!err.empty()) {
        cfd::set_last_error(err);
        return CFDART_E_ARG;
    }
    *d_out = d;
    if (T_out) *T_out = T;
    if (host_state) {
        if (host_len < host.size()) {
            cfd::set_last_error("host_state holds fewer than (nx+2ng)*(ny+2ng)*ncomp doubles");
            return CFDART_E_ARG;
        }
        std::copy(host.begin(), host.end(), host_state);
    }
    return 0;
}
