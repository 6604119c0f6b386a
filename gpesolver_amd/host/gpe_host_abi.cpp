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
