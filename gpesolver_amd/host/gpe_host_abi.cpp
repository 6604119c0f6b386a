// C ABI of the host layer (include/gpesolver_host.h).
#include <cstring>
#include <exception>
#include <string>

#include "../../include/gpesolver_host.h"
#include "gpe_host.hpp"

namespace {
int report(const std::string& msg, char* err, int errlen) {
    if (err && errlen > 0) {
        std::strncpy(err, msg.c_str(), (size_t)errlen - 1);
        err[errlen - 1] = '\0';
    }
    return 1;
}
}  // namespace

extern "C" int gpes_validate_params_json(const char* json_text, char* err, int errlen) {
    if (!json_text) return report("null json text", err, errlen);
    try {
        gpe::host::PhysicalParameters p;
        gpe::host::FileManager fm(nullptr, &p);
        fm.load_all_v1(gpe::host::Json::parse(json_text));
        fm.check_params();
        return 0;
    } catch (const std::exception& e) {
        return report(e.what(), err, errlen);
    }
}

extern "C" int gpes_run(const char* workdir, double it_tol, double* energy_per_atom, long* it_steps, char* err,
                        int errlen) {
    try {
        gpe::host::SimulationManager mgr(workdir ? workdir : ".", true);
        mgr.solver().it_tolerance = it_tol;
        mgr.initialize();
        mgr.run_simulation();
        if (energy_per_atom) *energy_per_atom = mgr.solver().last_energy().e_total / mgr.parameters().n_atoms;
        if (it_steps) *it_steps = mgr.solver().it_steps_done();
        return 0;
    } catch (const std::exception& e) {
        return report(e.what(), err, errlen);
    }
}

extern "C" int gpes_write_files(const char* dir, const char* name, const double* psi, const char* pote_name,
                                const double* V, const double* energies, int n_energies, char* err, int errlen) {
    if (!dir || !name || !psi || !pote_name || !V || (n_energies > 0 && !energies))
        return report("null argument", err, errlen);
    try {
        gpe::host::PhysicalParameters p;
        gpe::host::FileManager fm(nullptr, &p, dir);
        fm.load_params();
        const size_t n = (size_t)p.nx * p.ny * p.nz;
        gpe::host::wavefunction_t w(n);
        for (size_t e = 0; e < n; ++e) w[e] = {psi[2 * e], psi[2 * e + 1]};
        gpe::host::potential_t pote(V, V + n);
        fm.save_to_binary_file(w, name);
        fm.save_to_text_file(w, name);
        fm.save_pote_to_text_file(pote, pote_name);
        gpe::host::energies_container_t enes;
        for (int t = 0; t < n_energies; ++t) {
            gpe::host::BECEnergies e{};
            e.e_kin = energies[5 * t];
            e.e_pot = energies[5 * t + 1];
            e.e_int = energies[5 * t + 2];
            e.e_ext = energies[5 * t + 3];
            e.e_bmf = energies[5 * t + 4];
            e.sum();
            enes.push_back(e);
        }
        fm.save_energies(enes);
        return 0;
    } catch (const std::exception& e) {
        return report(e.what(), err, errlen);
    }
}
