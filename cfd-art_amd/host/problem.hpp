// C++ host of the driver: the reference's settings_file, problem classes (initial conditions,
// frame output) and simulation loop, with every hot-path call going through the C-ABI of
// include/cfdart.h. Reference files mirrored:
//   settings_file     euler_equations/euler_simulation_code/headers/settings_file.hpp:19-72
//   problem_base      euler_equations/euler_simulation_code/headers/problem_base.hpp:18-40
//   simulation        euler_equations/euler_simulation_code/headers/simulation.hpp:23-68
//   Euler set_ICs     euler_equations/euler_simulation_code/sourcecode/euler_io.cpp:13-1139
//   Euler output      euler_equations/euler_simulation_code/sourcecode/euler_io.cpp:1142-1261
//   Allaire set_ICs   allaire_2d/sourcecode/allaire_diffuse_io.cpp:13-990
//   Allaire output    allaire_2d/sourcecode/allaire_diffuse_io.cpp:1109-1294
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "cfdart.h"
#include "cfdart_host.h"

namespace cfdhost {

// Key/value settings file. Reference keys: Nx Ny CFL splitting_scheme test_case output_freq
// outputpath flux_solver riemann_solver problem. Added keys (reference behaviour is the
// default): limiter, wave_speeds, num_steps, T, seed, n_discs, bc_left/right/bottom/top,
// z_clip, frame_every_steps, colsweep_uses_dx.
struct Settings {
    int Nx = 0, Ny = 0;
    double CFL = 0.0;
    std::string splitting_scheme, test_case, outputpath, basename, flux_solver, riemann_solver, problem;
    int output_freq = 0;
    std::string limiter = "minmod", wave_speeds = "garrick";
    long long num_steps = -1;  // < 0: run to T
    double T = -1.0;           // < 0: the test case's T
    unsigned long long seed = 1;
    int n_discs = -1;
    std::string bc[4];  // empty: the test case's boundary conditions
    double z_clip = -1.0;
    int frame_every_steps = 0;
    int colsweep_uses_dx = 1;

    // Returns "" or an error message. Unknown keys are ignored, as in the reference.
    std::string read(const std::string& filename);
};

// One simulation problem on the GPU (the role of problem_base + its sim_info).
class GpuProblem {
public:
    GpuProblem() = default;
    ~GpuProblem();

    // set_ICs: parameters of the test case, descriptor, initial state, device upload.
    std::string set_ICs(const Settings& sf);
    // output: the reference's frame cadence and gnuplot formats (n, t of the current state).
    void output(int n, double t, bool force);
    double compute_dt(int n, double t);
    int advance(int n, double dt);  // splitting_scheme::advance_timestep
    int run_steps(int nsteps, double* t, int* n);  // device-resident loop for fixed-step mode

    const cfdart_desc& desc() const { return d_; }
    double T() const { return T_; }
    int splitting() const { return split_; }
    std::vector<double>& host_state() { return host_; }
    cfdart_grid* grid() { return g_; }
    long long frames_written() const { return frames_; }

private:
    void write_state(double t);
    void write_schlieren(double t);
    void write_masschange();

    Settings sf_;
    cfdart_desc d_{};
    cfdart_grid* g_ = nullptr;
    std::vector<double> host_;
    double T_ = 0.0, cfl_ = 0.0;
    int split_ = CFDART_NAIVE;
    std::string outname_;
    double last_output_time_ = 0.0;  // per instance (the reference's is function-static)
    std::vector<double> time_, mass1_, mass2_;
    long long frames_ = 0;
};

// simulation::run_simulation: returns 0 on success.
int run_simulation(const Settings& sf, std::string* err);

}  // namespace cfdhost
