// gpe_host.hpp -- C++ host side of the MI355X backend.
//
// Mirrors the reference's solver / context / parameters interface so that the
// HIP backend drops in beside cpu_solver:
//   PhysicalParameters      src/parameters/parameters.hpp:10-123, parameters.cpp:37-59
//   units                   src/utils/units/units.cpp:3-41
//   potentials              src/parameters/potentials.hpp:76-157
//   initializer             src/initializer/initializer.cpp:46-248
//   FileManager (v1 JSON)   src/filemanager/file_manager.cpp:251-548
//   AbstractGrossPitaevskiSolver + 13 hooks   src/solver/solver.hpp:17-72, solver.cpp:15-153
//   SimulationManager       src/manager/sim_manager.cpp:7-116
// The solver hooks are implemented by HipGrossPitaevskiSolver on top of the C
// ABI of include/gpehip.h; nothing here touches HIP directly.
#pragma once

#include <complex>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gpehip.h"

namespace gpe {
namespace host {

// ------------------------------------------------------------------ units
struct Units {
    static double len_nm_to_au(double nm) { return nm * 18.897261260649092; }
    static double len_au_to_nm(double au) { return au * 0.05291772105; }
    static double mass_Da_to_au(double da) { return da / 0.000548579909; }
    static double mass_au_to_Da(double au) { return au * 0.000548579909; }
    static double freq_Hz_to_au(double hz) { return hz * (4.1356e-12 / 27211.6); }
    static double freq_au_to_Hz(double au) { return au / (4.1356e-12 / 27211.6); }
    static double ene_au_to_meV(double au) { return au * 27211.407953; }
};

enum class CalcStrategy { IMAGINARY_TIME, REAL_TIME, FULL };
enum class InitStrategy { COS, GAUSS, MULTIPLE_GAUSS, SETUP_GAUSS, FROM_BINARY_FILE, FROM_TEXT_FILE };

// ------------------------------------------------------------ parameters
struct PhysicalParameters {
    double omega_x = 0, omega_y = 0, omega_z = 0;
    double m = 0, n_atoms = 0;
    double aa = 0, dd = Units::len_nm_to_au(1500.0), b = 0, w_15 = 0;
    double ggp11 = 0, gamma = 0, cdd = 0, add = 131.0;
    double edd = 0;
    bool const_edd = true;
    double edd_start = 0, edd_stop = 0;
    int nx = 0, ny = 0, nz = 0;
    long iter_imag = 0, iter_real = 0;
    // The reference fixes both steps (parameters.hpp:70-72); they stay
    // configurable here with the reference values as defaults.
    double real_time_dt = 1.00e10;
    double imag_time_dt = 1.25e11;
    double dx = 0, dy = 0, dz = 0;
    int fftw_n_threads = 4;
    int n_gauss_max = 0;
    int bec_droplets_x = 0, bec_droplets_y = 0, bec_droplets_z = 0;
    std::string load_filename;
    CalcStrategy calc_strategy = CalcStrategy::FULL;
    InitStrategy init_strategy = InitStrategy::MULTIPLE_GAUSS;
    std::string pote_key;
    // extensions (no reference counterpart): GAUSS width in cells (0 = L/10),
    // seeded uniform phase noise added after initialisation, FP32/FP64 bulk.
    double gauss_sigma_cells = 0.0;
    double phase_noise = 0.0;
    unsigned long long noise_seed = 0;
    int precision = GPE_FP64;

    double get_dxdydz() const { return dx * dy * dz; }
    void init_parameters();             // parameters.cpp:37-51
    void update_edd(long iter);         // parameters.cpp:53-59
    gpe_params to_abi() const;
};

double coord(int n, double d, int i);
bool potential_known(const std::string& key);
double potential_value(const std::string& key, const PhysicalParameters& p, double x, double y, double z);

using wavefunction_t = std::vector<std::complex<double>>;
using potential_t = std::vector<double>;

struct BECEnergies {
    double e_kin = 0, e_pot = 0, e_int = 0, e_ext = 0, e_bmf = 0, e_total = 0;
    double mu = 0, residual = 0;
    void sum() { e_total = e_kin + e_pot + e_int + e_ext + e_bmf; }
};
using energies_container_t = std::vector<BECEnergies>;

void fill_potential(const std::string& key, const PhysicalParameters& p, potential_t& V);
void init_wavefunction(const PhysicalParameters& p, wavefunction_t& psi);
void add_phase_noise(wavefunction_t& psi, double amplitude, unsigned long long seed);

// ---------------------------------------------------------------- JSON
// Minimal JSON reader for the flat gpe_params.json object.
class Json {
public:
    enum class Kind { Null, Bool, Number, String, Object, Array };
    Kind kind = Kind::Null;
    double num = 0;
    bool b = false;
    std::string str;
    std::map<std::string, Json> obj;
    std::vector<Json> arr;

    static Json parse(const std::string& text);
    bool contains(const std::string& k) const { return kind == Kind::Object && obj.count(k) > 0; }
    const Json& operator[](const std::string& k) const;
    double as_number() const;
    int as_int() const;
    long as_long() const;
    const std::string& as_string() const;
};

// ---------------------------------------------------------------- mediator
class AbstractSimulationMediator {
public:
    virtual ~AbstractSimulationMediator() = default;
    virtual void on_data_loaded(const wavefunction_t&) = 0;
    virtual void on_data_initialized(const wavefunction_t&) = 0;
    virtual void on_pote_initialized(const potential_t&) = 0;
    virtual void on_params_loaded() = 0;
    virtual void request_load_from_binary(wavefunction_t&) = 0;
    virtual void request_load_from_text(wavefunction_t&) = 0;
    virtual void request_free_potential() = 0;
    virtual void save_data(const wavefunction_t&) = 0;
    virtual void save_checkpoint(const wavefunction_t&) = 0;
    virtual void save_initial_state(const wavefunction_t&) = 0;
    virtual void save_energies(const energies_container_t&) = 0;
};

// ---------------------------------------------------------------- files
class FileManager {
public:
    FileManager(AbstractSimulationMediator* mediator, PhysicalParameters* params, std::string dir = ".")
        : mediator(mediator), params(params), dir(std::move(dir)) {}
    void load_params();                       // file_manager.cpp:56-79 (v1 only)
    void load_all_v1(const Json& j);          // file_manager.cpp:543-548
    void load_initialization(const Json& j);  // :436-466
    void load_simulation(const Json& j);      // :507-541
    void load_box(const Json& j);             // :491-505
    void load_potential(const Json& j);       // :468-489
    void check_params();                      // :216-235

    void save_to_binary_file(const wavefunction_t& psi, std::string name);   // :251-288
    void load_from_binary_file(std::string name, wavefunction_t& out);       // :290-335
    void save_to_text_file(const wavefunction_t& psi, std::string name);     // :81-112
    void load_from_text_file(std::string name, wavefunction_t& out);         // :114-162
    void save_energies(const energies_container_t& e);                       // :387-401
    void save_pote_to_text_file(const potential_t& V, std::string name);    // :403-434
    std::string path(const std::string& name) const { return dir + "/" + name; }

private:
    AbstractSimulationMediator* mediator;
    PhysicalParameters* params;
    std::string dir;
};

// ---------------------------------------------------------------- solver
class AbstractGrossPitaevskiSolver {
public:
    AbstractGrossPitaevskiSolver(AbstractSimulationMediator* mediator, PhysicalParameters* params)
        : params(params), p_mediator(mediator) {}
    virtual ~AbstractGrossPitaevskiSolver() = default;
    void solve();       // solver.cpp:25-46
    void initialize();  // solver.cpp:15-23
    void load_buffer(const wavefunction_t&);
    void load_pote(const potential_t&);
    const energies_container_t& energies() const { return enes; }
    const BECEnergies& last_energy() const { return ene; }
    const wavefunction_t& buffer() const { return cpsi; }
    long it_steps_done() const { return it_done; }

    // Convergence test (no reference counterpart): stop IT once the relative
    // change of E/N between two energy checks is below tol (0 = run all steps).
    double it_tolerance = 0.0;

protected:
    PhysicalParameters* params;
    AbstractSimulationMediator* p_mediator;
    energies_container_t enes;
    BECEnergies ene;
    double xnorma = 0.0;
    wavefunction_t cpsi, cpsii;
    potential_t pote;
    long it_done = 0;

    void calc_initial_state();  // solver.cpp:48-70
    void calc_evolution();      // solver.cpp:72-97
    void free_potential_well();
    void imag_time_iter();      // solver.cpp:99-105
    void real_time_iter();      // solver.cpp:107-112
    // n consecutive iterations; subclasses may fuse them.
    virtual void imag_time_iters(long n);
    virtual void real_time_iters(long first_iter, long n);
    // Output cadence without stopping the loop: snapshot_begin enqueues a
    // copy of the state and its energies behind the launched steps (false:
    // not supported, the loop then exports and computes them in place);
    // snapshot_end makes them cpsi and ene.
    virtual bool snapshot_begin() { return false; }
    virtual void snapshot_end() {}
    void finish_snapshot();

    virtual void calc_energy() = 0;
    virtual void init_containers() = 0;
    virtual void calc_fi3d() = 0;
    virtual void calc_norm() = 0;
    virtual void normalize() = 0;
    virtual void imag_iter_linear_step() = 0;
    virtual void imag_iter_nonlinear_step() = 0;
    virtual void real_fft_potential_half_step() = 0;
    virtual void real_fft_kinetic_step() = 0;
    virtual void prepare_fft() = 0;
    virtual void export_data() = 0;
    virtual void import_data() = 0;
    virtual void import_pote() = 0;
};

class HipGrossPitaevskiSolver : public AbstractGrossPitaevskiSolver {
public:
    using AbstractGrossPitaevskiSolver::AbstractGrossPitaevskiSolver;
    ~HipGrossPitaevskiSolver() override;
    // The generated initial states (COS, GAUSS, MULTIPLE_GAUSS, SETUP_GAUSS,
    // with the phase noise) and the potentials built on the device
    // (gpe_init_wavefunction / gpe_init_potential) instead of on the host:
    // load_buffer / load_pote semantics, the host copies (cpsi, pote) are
    // read back for the files the reference writes.
    void init_state_on_device(const PhysicalParameters& p);
    void init_pote_on_device(const std::string& key, potential_t& V);

protected:
    gpe_handle* h = nullptr;
    void check(int rc) const;
    void imag_time_iters(long n) override;
    void real_time_iters(long first_iter, long n) override;
    void calc_energy() override;
    void init_containers() override;
    void calc_fi3d() override;
    void calc_norm() override;
    void normalize() override;
    void imag_iter_linear_step() override;
    void imag_iter_nonlinear_step() override;
    void real_fft_potential_half_step() override;
    void real_fft_kinetic_step() override;
    void prepare_fft() override;
    void export_data() override;
    void import_data() override;
    void import_pote() override;
    bool snapshot_begin() override;
    void snapshot_end() override;
};

// ---------------------------------------------------------------- manager
class SimulationManager : public AbstractSimulationMediator {
public:
    explicit SimulationManager(std::string dir = ".", bool quiet = false);
    void initialize();
    void run_simulation();
    PhysicalParameters& parameters() { return params; }
    AbstractGrossPitaevskiSolver& solver() { return *m_solver; }

    void on_data_loaded(const wavefunction_t&) override;
    void on_data_initialized(const wavefunction_t&) override;
    void on_pote_initialized(const potential_t&) override;
    void on_params_loaded() override;
    void request_load_from_binary(wavefunction_t&) override;
    void request_load_from_text(wavefunction_t&) override;
    void request_free_potential() override;
    void save_data(const wavefunction_t&) override;
    void save_checkpoint(const wavefunction_t&) override;
    void save_initial_state(const wavefunction_t&) override;
    void save_energies(const energies_container_t&) override;

private:
    PhysicalParameters params;
    std::string dir;
    bool quiet;
    int checkpoint_counter = 0;
    std::unique_ptr<FileManager> m_file_manager;
    std::unique_ptr<AbstractGrossPitaevskiSolver> m_solver;
    void change_potential(const std::string& key);
    void check_load_filename();
};

}  // namespace host
}  // namespace gpe
