// gpesolver -- command-line entry point (the reference's main.cpp:4-16):
// reads ./gpe_params.json (or the directory given as argv[1]) and runs IT, RT
// or FS on the MI355X backend.
#include <cstdio>
#include <exception>
#include <iostream>

#include "gpe_host.hpp"

int main(int argc, char** argv) {
    try {
        gpe::host::SimulationManager mgr(argc > 1 ? argv[1] : ".");
        mgr.initialize();
        mgr.run_simulation();
        const auto& e = mgr.solver().last_energy();
        std::cout << "[INFO]    Final energy [meV]: " << gpe::host::Units::ene_au_to_meV(e.e_total) << std::endl;
    } catch (const std::exception& e) {
        std::cerr << "Error: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
