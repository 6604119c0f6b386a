// ref_driver.cpp -- TEST INFRASTRUCTURE, never part of the product.
//
// Drives the reference's own parameter loader (src/filemanager/file_manager.cpp),
// derived constants (src/parameters/parameters.cpp:37-59), coordinates
// (src/context/context.cpp:17-33), initial states and potentials
// (src/initializer/initializer.cpp, src/parameters/potentials.hpp) and file
// writers, compiled from the reference sources where they lie by oracle/ref.mk
// into oracle/_ref/gpe_ref_driver.  The tests compare this project's oracle and
// host restatements with what it writes.
//
//   gpe_ref_driver DIR init      DIR/gpe_params.json -> DIR/ref_derived.txt,
//                                ref_psi.raw (complex128), ref_pote.raw (float64),
//                                ref_pote_free.raw (MEXICAN_FREE), ref_edd.txt
//   gpe_ref_driver DIR validate  prints "ok" or the loader's exception text
//   gpe_ref_driver DIR save      DIR/in_psi.raw, in_pote.raw -> the reference's
//                                .gpe.bin / .gpe.dat / energy files in DIR
#include <unistd.h>

#include <cstdio>
#include <cstring>
#include <exception>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "context/context.hpp"
#include "filemanager/file_manager.hpp"
#include "initializer/initializer.hpp"
#include "manager/sim_mediator.hpp"
#include "parameters/parameters.hpp"

namespace {

class Capture : public AbstractSimulationMediator {
public:
    DataInitializer* init = nullptr;
    bool run_init = true;
    wavefunction_t psi;
    potential_t pote, pote_changed;
    void on_data_loaded(const wavefunction_t& w) override { psi = w; }
    void on_data_initialized(const wavefunction_t& w) override { psi = w; }
    void on_pote_initialized(const potential_t& p) override { pote = p; }
    void on_pote_changed(const potential_t& p) override { pote_changed = p; }
    void on_params_loaded() override {
        // the order of SimulationManager::on_params_loaded (src/manager/sim_manager.cpp:50-58)
        PhysicalParameters::getInstance()->init_parameters();
        SimulationContext::getInstance()->initialize();
        if (run_init) {
            init->initialize_wavefunction();
            init->initialize_potential();
        }
    }
    void request_load_from_binary(wavefunction_t&) override {}
    void request_load_from_text(wavefunction_t&) override {}
    void request_free_potential() override {}
    void save_data(const wavefunction_t&) override {}
    void save_checkpoint(const wavefunction_t&) override {}
    void save_initial_state(const wavefunction_t&) override {}
    void save_energies(const energies_container_t&) override {}
};

template <class T>
void write_raw(const char* name, const T* p, size_t n) {
    std::ofstream f(name, std::ios::binary);
    f.write(reinterpret_cast<const char*>(p), n * sizeof(T));
}

template <class T>
bool read_raw(const char* name, T* p, size_t n) {
    std::ifstream f(name, std::ios::binary);
    f.read(reinterpret_cast<char*>(p), n * sizeof(T));
    return (size_t)f.gcount() == n * sizeof(T);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: gpe_ref_driver DIR init|validate|save\n");
        return 2;
    }
    if (chdir(argv[1]) != 0) {
        std::perror("chdir");
        return 2;
    }
    const std::string mode = argv[2];
    Capture cap;
    cap.run_init = (mode == "init");
    DataInitializer init(&cap);
    cap.init = &init;
    FileManager fm(&cap);
    try {
        fm.load_params();
    } catch (const std::exception& e) {
        if (mode == "validate") {
            std::printf("error: %s\n", e.what());
            return 0;
        }
        std::fprintf(stderr, "load_params: %s\n", e.what());
        return 1;
    }
    if (mode == "validate") {
        std::printf("ok\n");
        return 0;
    }
    PhysicalParameters* p = PhysicalParameters::getInstance();
    const size_t n = (size_t)p->nx * p->ny * p->nz;
    if (mode == "init") {
        FILE* f = std::fopen("ref_derived.txt", "w");
        std::fprintf(f, "m %.17g\nn_atoms %.17g\ndx %.17g\ndy %.17g\ndz %.17g\n", p->m, p->n_atoms, p->dx, p->dy,
                     p->dz);
        std::fprintf(f, "omega_x %.17g\nomega_y %.17g\nomega_z %.17g\ndd %.17g\nedd %.17g\n", p->omega_x, p->omega_y,
                     p->omega_z, p->dd, p->edd);
        std::fprintf(f, "cdd %.17g\nggp11 %.17g\ngamma %.17g\nw_15 %.17g\naa %.17g\nb %.17g\ndxdydz %.17g\n", p->cdd,
                     p->ggp11, p->gamma, p->w_15, p->aa, p->b, p->get_dxdydz());
        std::fprintf(f, "imag_time_dt %.17g\nreal_time_dt %.17g\n", p->imag_time_dt, p->real_time_dt);
        std::fclose(f);
        write_raw("ref_psi.raw", reinterpret_cast<const double*>(cap.psi.get_data()), 2 * n);
        write_raw("ref_pote.raw", cap.pote.get_data(), n);
        // the FS switch (src/manager/sim_manager.cpp:78-80) reports the new map
        // through on_pote_initialized
        init.change_potential("MEXICAN_FREE");
        write_raw("ref_pote_free.raw", cap.pote.get_data(), n);
        if (!p->const_edd && p->iter_real > 0) {
            FILE* g = std::fopen("ref_edd.txt", "w");
            const unsigned its[] = {0u, 1u, 7u, (unsigned)(p->iter_real / 2), (unsigned)p->iter_real};
            for (unsigned it : its) {
                p->update_edd(it);
                std::fprintf(g, "%u %.17g %.17g %.17g\n", it, p->edd, p->ggp11, p->gamma);
            }
            std::fclose(g);
        }
        return 0;
    }
    if (mode == "save") {
        wavefunction_t psi(p->nx, p->ny, p->nz);
        potential_t pote(p->nx, p->ny, p->nz);
        if (!read_raw("in_psi.raw", reinterpret_cast<double*>(psi.get_data()), 2 * n) ||
            !read_raw("in_pote.raw", pote.get_data(), n)) {
            std::fprintf(stderr, "missing in_psi.raw / in_pote.raw\n");
            return 1;
        }
        fm.save_to_binary_file(psi, "ref_state");
        fm.save_to_text_file(psi, "ref_state");
        fm.save_pote_to_text_file(pote, "ref_pote");
        energies_container_t enes;
        for (int t = 0; t < 3; ++t) {
            BECEnergies e;
            e.e_kin = 1.25e-13 * (t + 1);
            e.e_pot = 2.5e-14 / (t + 1);
            e.e_int = -3.0e-15 * t;
            e.e_ext = 4.0e-16;
            e.e_bmf = 5.0e-17 * t;
            e.sum();
            enes.push_back(e);
        }
        fm.save_energies(enes);
        return 0;
    }
    std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
    return 2;
}
