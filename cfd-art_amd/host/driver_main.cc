// Driver executable: `cfdart_run <settings file>` (euler_equations/euler_simulation_code/
// sourcecode/main.cpp:17-49, with problem euler | allaire selectable).
#include <chrono>
#include <iostream>

#include "problem.hpp"

int main(int argc, char* argv[])
{
    std::cout << "[main] Beginning simulation.." << std::endl;
    if (argc < 2) {
        std::cerr << "usage: " << argv[0] << " <settings file>" << std::endl;
        return 2;
    }
    cfdhost::Settings sf;
    std::string err = sf.read(argv[1]);
    if (!err.empty()) {
        std::cerr << "[main] " << err << std::endl;
        return 2;
    }
    auto t0 = std::chrono::steady_clock::now();
    int status = cfdhost::run_simulation(sf, &err);
    auto t1 = std::chrono::steady_clock::now();
    if (status) std::cerr << "[main] error: " << err << std::endl;
    std::cout << "[main] Simulation exited with status = " << status << std::endl;
    std::cout << "[main] The wall clock runtime was " << std::chrono::duration<double>(t1 - t0).count() << "s."
              << std::endl;
    return status;
}
