// gpe_host.cpp -- C++ host side: parameters, set-up, files, solver template
// method and the HIP-backed hook implementations.  File:line citations refer
// to the reference sources restated (declarations in gpe_host.hpp).
#include "gpe_host.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <regex>
#include <sstream>

namespace gpe {
namespace host {

// ------------------------------------------------------------ parameters
void PhysicalParameters::init_parameters() {
    aa = m * std::pow(omega_x, 2) / 4. / std::pow(dd, 2);
    b = 0.5 * m * std::pow(omega_x, 2);
    add = 131.;
    cdd = 12. * M_PI * add / m;
    const double a = add / edd;
    ggp11 = 4. * M_PI * a / m;
    gamma = 128. * std::sqrt(M_PI) * std::pow(a, 2.5) / 3. / m * (1. + 1.5 * std::pow(edd, 2));
    w_15 = std::pow(n_atoms, 1.5);
}

void PhysicalParameters::update_edd(long iter) {
    edd = edd_start + (edd_stop - edd_start) * (double)iter / (double)iter_real;
    const double a = add / edd;
    ggp11 = 4 * M_PI * a / m;
    gamma = 128. * std::sqrt(M_PI) * std::pow(a, 2.5) / 3. / m * (1. + 1.5 * std::pow(edd, 2));
}

gpe_params PhysicalParameters::to_abi() const {
    gpe_params g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.dx = dx;
    g.dy = dy;
    g.dz = dz;
    g.m = m;
    g.n_atoms = n_atoms;
    g.cdd = cdd;
    g.ggp11 = ggp11;
    g.gamma = gamma;
    g.w_15 = w_15;
    g.imag_time_dt = imag_time_dt;
    g.real_time_dt = real_time_dt;
    return g;
}

double coord(int n, double d, int i) { return (double)(i - n / 2) * d; }

static const char* kPotentials[] = {"HARMONIC", "MEXICAN", "MEXICAN_FREE", "CYLINDRICAL", "FREE",
                                    "MEXICAN_ASYMETRIC", "MEXICAN_ASYMETRIC_FREE"};

bool potential_known(const std::string& key) {
    for (const char* k : kPotentials)
        if (key == k) return true;
    return false;
}

double potential_value(const std::string& key, const PhysicalParameters& p, double x, double y, double z) {
    const double vy = 0.5 * p.m * y * y * p.omega_y * p.omega_y;
    const double vz = 0.5 * p.m * z * z * p.omega_z * p.omega_z;
    const double slope = (0.25 / (27211.4 * 11.6 * 1e9)) / p.dd;
    if (key == "HARMONIC") return 0.5 * p.m * x * x * p.omega_x * p.omega_x + vy + vz;
    if (key == "MEXICAN") return -p.b * x * x + p.aa * std::pow(x, 4) + vy + vz;
    if (key == "MEXICAN_FREE") return p.aa * std::pow(x, 4) + vy + vz;
    if (key == "CYLINDRICAL") {
        const double w = (p.omega_y + p.omega_x) / 2.;
        return 0.5 * p.m * (x * x + y * y) * w * w + vz;
    }
    if (key == "FREE") return vz;
    if (key == "MEXICAN_ASYMETRIC") return -p.b * x * x + p.aa * std::pow(x, 4) - slope * x + vy + vz;
    if (key == "MEXICAN_ASYMETRIC_FREE") return p.aa * std::pow(x, 4) - slope * x + vy + vz;
    throw std::runtime_error("Not know potential option");
}

// initializer.cpp:229-248.  The reference samples the potential through a
// std::function<double(int, int, int)> (initializer.hpp:37): the coordinates
// are truncated toward zero to whole atomic units on the way in.
void fill_potential(const std::string& key, const PhysicalParameters& p, potential_t& V) {
    V.assign((size_t)p.nx * p.ny * p.nz, 0.0);
    auto t = [](double c) { return (double)(int)c; };
    for (int i = 0; i < p.nx; i++)
        for (int j = 0; j < p.ny; j++)
            for (int k = 0; k < p.nz; k++)
                V[((size_t)i * p.ny + j) * p.nz + k] = potential_value(
                    key, p, t(coord(p.nx, p.dx, i)), t(coord(p.ny, p.dy, j)), t(coord(p.nz, p.dz, k)));
}

void init_wavefunction(const PhysicalParameters& p, wavefunction_t& psi) {
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    psi.assign((size_t)nx * ny * nz, {0.0, 0.0});
    std::vector<double> cx, cy, cz;
    const double xt = nx * p.dx, yt = ny * p.dy, zt = nz * p.dz;
    if (p.init_strategy == InitStrategy::MULTIPLE_GAUSS) {
        const int n = p.n_gauss_max;
        for (int idx = 0; idx < n; idx++) {
            cx.push_back(idx % 2 ? p.dd : -p.dd);
            if (n % 2 == 1)
                cy.push_back((idx - n / 2) * (ny * p.dy) / 2. / (n + 1.));
            else
                cy.push_back(coord(ny, p.dy, 0) + (idx / 2 + 1) * (ny * p.dy) / (n / 2 + 1));
            cz.push_back(0.0);
        }
    } else if (p.init_strategy == InitStrategy::SETUP_GAUSS) {
        for (int a = 0; a < p.bec_droplets_x; a++)
            for (int b = 0; b < p.bec_droplets_y; b++)
                for (int c = 0; c < p.bec_droplets_z; c++) {
                    cx.push_back(-xt / 2 + xt / (double)(p.bec_droplets_x + 1) * (a + 1));
                    cy.push_back(-yt / 2 + yt / (double)(p.bec_droplets_y + 1) * (b + 1));
                    cz.push_back(-zt / 2 + zt / (double)(p.bec_droplets_z + 1) * (c + 1));
                }
    }
    for (int i = 1; i < nx - 1; i++)
        for (int j = 1; j < ny - 1; j++)
            for (int k = 1; k < nz - 1; k++) {
                const double x = coord(nx, p.dx, i), y = coord(ny, p.dy, j), z = coord(nz, p.dz, k);
                double val = 0.0;
                switch (p.init_strategy) {
                case InitStrategy::COS: {
                    const double rrr = (nx / 2) * p.dx;
                    val = std::cos(2 * M_PI * (x - p.dd) / rrr) * std::cos(2 * M_PI * y / rrr) *
                          std::cos(1 * M_PI * z / rrr);
                    break;
                }
                case InitStrategy::GAUSS: {
                    double sx = xt / 10., sy = yt / 10., sz = zt / 10.;
                    if (p.gauss_sigma_cells > 0) {
                        sx = p.gauss_sigma_cells * p.dx;
                        sy = p.gauss_sigma_cells * p.dy;
                        sz = p.gauss_sigma_cells * p.dz;
                    }
                    val = std::exp(-0.5 * x * x / (sx * sx)) * std::exp(-0.5 * y * y / (sy * sy)) *
                          std::exp(-0.5 * z * z / (sz * sz));
                    break;
                }
                case InitStrategy::MULTIPLE_GAUSS:
                case InitStrategy::SETUP_GAUSS: {
                    const double sx = xt / 20., sy = yt / 20., sz = zt / 20.;
                    const bool zc = p.init_strategy == InitStrategy::SETUP_GAUSS;
                    for (size_t m = 0; m < cx.size(); m++) {
                        const double xc = x - cx[m], yc = y - cy[m], zz = zc ? z - cz[m] : z;
                        val += std::exp(-0.5 * xc * xc / (sx * sx)) * std::exp(-0.5 * yc * yc / (sy * sy)) *
                               std::exp(-0.5 * zz * zz / (sz * sz));
                    }
                    break;
                }
                default:
                    throw std::runtime_error("init_wavefunction: file strategies load through FileManager");
                }
                psi[((size_t)i * ny + j) * nz + k] = {val, 0.0};
            }
}

// SplitMix64, one draw per grid point in layout order (same stream as
// gpesolver_amd.params.splitmix64).
void add_phase_noise(wavefunction_t& psi, double amplitude, unsigned long long seed) {
    if (amplitude == 0.0) return;
    unsigned long long s = seed;
    for (auto& v : psi) {
        s += 0x9E3779B97F4A7C15ULL;
        unsigned long long z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z = z ^ (z >> 31);
        const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
        const double th = amplitude * (2.0 * u - 1.0);
        v *= std::complex<double>(std::cos(th), std::sin(th));
    }
}

// ---------------------------------------------------------------- JSON
namespace {
struct Parser {
    const std::string& s;
    size_t i = 0;
    explicit Parser(const std::string& s) : s(s) {}
    void ws() {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\n' || s[i] == '\t' || s[i] == '\r')) ++i;
    }
    [[noreturn]] void err(const std::string& m) { throw std::runtime_error("JSON parse error at " + std::to_string(i) + ": " + m); }
    Json value() {
        ws();
        if (i >= s.size()) err("unexpected end");
        Json j;
        const char c = s[i];
        if (c == '{') {
            j.kind = Json::Kind::Object;
            ++i;
            ws();
            if (i < s.size() && s[i] == '}') {
                ++i;
                return j;
            }
            while (true) {
                ws();
                if (i >= s.size() || s[i] != '"') err("expected key");
                std::string k = string();
                ws();
                if (i >= s.size() || s[i] != ':') err("expected ':'");
                ++i;
                j.obj[k] = value();
                ws();
                if (i < s.size() && s[i] == ',') {
                    ++i;
                    continue;
                }
                if (i < s.size() && s[i] == '}') {
                    ++i;
                    break;
                }
                err("expected ',' or '}'");
            }
        } else if (c == '[') {
            j.kind = Json::Kind::Array;
            ++i;
            ws();
            if (i < s.size() && s[i] == ']') {
                ++i;
                return j;
            }
            while (true) {
                j.arr.push_back(value());
                ws();
                if (i < s.size() && s[i] == ',') {
                    ++i;
                    continue;
                }
                if (i < s.size() && s[i] == ']') {
                    ++i;
                    break;
                }
                err("expected ',' or ']'");
            }
        } else if (c == '"') {
            j.kind = Json::Kind::String;
            j.str = string();
        } else if (s.compare(i, 4, "true") == 0) {
            j.kind = Json::Kind::Bool;
            j.b = true;
            i += 4;
        } else if (s.compare(i, 5, "false") == 0) {
            j.kind = Json::Kind::Bool;
            i += 5;
        } else if (s.compare(i, 4, "null") == 0) {
            i += 4;
        } else {
            const char* begin = s.c_str() + i;
            char* end = nullptr;
            j.num = std::strtod(begin, &end);
            if (end == begin) err("bad value");
            i += (size_t)(end - begin);
            j.kind = Json::Kind::Number;
        }
        return j;
    }
    std::string string() {
        ++i;  // opening quote
        std::string out;
        while (i < s.size() && s[i] != '"') {
            if (s[i] == '\\' && i + 1 < s.size()) {
                ++i;
                const char e = s[i];
                out.push_back(e == 'n' ? '\n' : e == 't' ? '\t' : e);
            } else {
                out.push_back(s[i]);
            }
            ++i;
        }
        if (i >= s.size()) err("unterminated string");
        ++i;
        return out;
    }
};
}  // namespace

Json Json::parse(const std::string& text) {
    Parser p(text);
    Json j = p.value();
    p.ws();
    if (p.i != text.size()) p.err("trailing characters");
    return j;
}

const Json& Json::operator[](const std::string& k) const {
    auto it = obj.find(k);
    if (it == obj.end()) throw std::runtime_error("Missing required parameter: " + k);
    return it->second;
}
double Json::as_number() const {
    if (kind == Kind::Number) return num;
    if (kind == Kind::Bool) return b ? 1.0 : 0.0;
    throw std::runtime_error("JSON value is not a number");
}
int Json::as_int() const { return (int)as_number(); }
long Json::as_long() const { return (long)as_number(); }
const std::string& Json::as_string() const {
    if (kind != Kind::String) throw std::runtime_error("JSON value is not a string");
    return str;
}

// ---------------------------------------------------------------- files
#define CHECK_REQUIRED(J, KEY) \
    if (!(J).contains(KEY)) throw std::runtime_error(std::string("Missing required parameter: ") + (KEY));

void FileManager::load_initialization(const Json& j) {
    CHECK_REQUIRED(j, "init_strategy");
    const std::string s = j["init_strategy"].as_string();
    if (s == "COS")
        params->init_strategy = InitStrategy::COS;
    else if (s == "GAUSS")
        params->init_strategy = InitStrategy::GAUSS;
    else if (s == "MULTIPLE_GAUSS")
        params->init_strategy = InitStrategy::MULTIPLE_GAUSS;
    else if (s == "SETUP_GAUSS")
        params->init_strategy = InitStrategy::SETUP_GAUSS;
    else if (s == "BINARY_FILE")
        params->init_strategy = InitStrategy::FROM_BINARY_FILE;
    else if (s == "TEXT_FILE")
        params->init_strategy = InitStrategy::FROM_TEXT_FILE;
    else
        throw std::runtime_error("Unknown initialization strategy: " + s);
    if (params->init_strategy == InitStrategy::MULTIPLE_GAUSS) {
        CHECK_REQUIRED(j, "initial_maximas");
        params->n_gauss_max = j["initial_maximas"].as_int();
        if (params->n_gauss_max <= 0) throw std::runtime_error("Number of Gaussian maxima must be positive.");
    }
    if (params->init_strategy == InitStrategy::SETUP_GAUSS) {
        CHECK_REQUIRED(j, "bec_droplets_x");
        CHECK_REQUIRED(j, "bec_droplets_y");
        CHECK_REQUIRED(j, "bec_droplets_z");
        params->bec_droplets_x = j["bec_droplets_x"].as_int();
        params->bec_droplets_y = j["bec_droplets_y"].as_int();
        params->bec_droplets_z = j["bec_droplets_z"].as_int();
    }
    if (params->init_strategy == InitStrategy::FROM_BINARY_FILE ||
        params->init_strategy == InitStrategy::FROM_TEXT_FILE) {
        CHECK_REQUIRED(j, "load_filename");
        params->load_filename = j["load_filename"].as_string();
    }
    // extensions
    if (j.contains("gauss_sigma_cells")) params->gauss_sigma_cells = j["gauss_sigma_cells"].as_number();
    if (j.contains("phase_noise")) params->phase_noise = j["phase_noise"].as_number();
    if (j.contains("noise_seed")) params->noise_seed = (unsigned long long)j["noise_seed"].as_number();
}

void FileManager::load_potential(const Json& j) {
    CHECK_REQUIRED(j, "pote_strategy");
    params->pote_key = j["pote_strategy"].as_string();
    if (!potential_known(params->pote_key)) throw std::runtime_error("Not know potential option");
    const std::string& k = params->pote_key;
    if (k == "MEXICAN" || k == "MEXICAN_FREE" || k == "MEXICAN_ASYMETRIC" || k == "MEXICAN_ASYMETRIC_FREE") {
        CHECK_REQUIRED(j, "dd");
        params->dd = Units::len_nm_to_au(j["dd"].as_number());
    }
    CHECK_REQUIRED(j, "omega_x");
    CHECK_REQUIRED(j, "omega_y");
    CHECK_REQUIRED(j, "omega_z");
    params->omega_x = Units::freq_Hz_to_au(j["omega_x"].as_number());
    params->omega_y = Units::freq_Hz_to_au(j["omega_y"].as_number());
    params->omega_z = Units::freq_Hz_to_au(j["omega_z"].as_number());
}

void FileManager::load_box(const Json& j) {
    for (const char* k : {"dx", "dy", "dz", "nx", "ny", "nz"}) CHECK_REQUIRED(j, k);
    params->dx = Units::len_nm_to_au(j["dx"].as_number());
    params->dy = Units::len_nm_to_au(j["dy"].as_number());
    params->dz = Units::len_nm_to_au(j["dz"].as_number());
    params->nx = j["nx"].as_int();
    params->ny = j["ny"].as_int();
    params->nz = j["nz"].as_int();
}

void FileManager::load_simulation(const Json& j) {
    for (const char* k : {"calc_strategy", "iter_imag", "iter_real", "n_atoms", "m"}) CHECK_REQUIRED(j, k);
    // Unknown strings leave the default FS (complex_param.hpp:38-47).
    const std::string cs = j["calc_strategy"].as_string();
    if (cs == "IT") params->calc_strategy = CalcStrategy::IMAGINARY_TIME;
    if (cs == "RT") params->calc_strategy = CalcStrategy::REAL_TIME;
    if (cs == "FS") params->calc_strategy = CalcStrategy::FULL;
    params->iter_imag = j["iter_imag"].as_long();
    params->iter_real = j["iter_real"].as_long();
    params->n_atoms = j["n_atoms"].as_number();
    params->m = Units::mass_Da_to_au(j["m"].as_number());
    if (params->calc_strategy == CalcStrategy::IMAGINARY_TIME) CHECK_REQUIRED(j, "edd");
    if (j.contains("edd")) {
        params->const_edd = true;
        params->edd = j["edd"].as_number();
    } else {
        CHECK_REQUIRED(j, "edd_start");
        CHECK_REQUIRED(j, "edd_stop");
        params->const_edd = false;
        params->edd_start = j["edd_start"].as_number();
        params->edd_stop = j["edd_stop"].as_number();
        // The reference leaves edd unset here (file_manager.cpp:520-535); the
        // imaginary-time part of FS then uses edd_start.
        params->edd = params->edd_start;
    }
    params->fftw_n_threads = j.contains("fftw_n_threads") ? j["fftw_n_threads"].as_int() : 4;
    if (params->fftw_n_threads <= 0) throw std::runtime_error("FFTW number of threads must be positive.");
    // extensions
    if (j.contains("real_time_dt")) params->real_time_dt = j["real_time_dt"].as_number();
    if (j.contains("imag_time_dt")) params->imag_time_dt = j["imag_time_dt"].as_number();
    if (j.contains("precision")) {
        const std::string pr = j["precision"].as_string();
        if (pr == "fp32")
            params->precision = GPE_FP32;
        else if (pr == "fp64")
            params->precision = GPE_FP64;
        else
            throw std::runtime_error("precision must be \"fp32\" or \"fp64\"");
    }
}

void FileManager::load_all_v1(const Json& j) {
    load_initialization(j);
    load_simulation(j);
    load_box(j);
    load_potential(j);
}

void FileManager::check_params() {
    if (params->nx <= 0 || params->ny <= 0 || params->nz <= 0)
        throw std::runtime_error("Grid dimensions must be positive.");
    if (params->dx <= 0 || params->dy <= 0 || params->dz <= 0)
        throw std::runtime_error("Grid spacings must be positive.");
    if (params->n_atoms <= 0) throw std::runtime_error("Number of atoms must be positive.");
    if (params->m <= 0) throw std::runtime_error("Mass must be positive.");
}

void FileManager::load_params() {
    std::ifstream f(path("gpe_params.json"));
    if (!f.is_open()) throw std::runtime_error("Could not open " + path("gpe_params.json"));
    std::stringstream ss;
    ss << f.rdbuf();
    const Json j = Json::parse(ss.str());
    load_all_v1(j);
    check_params();
    if (mediator) mediator->on_params_loaded();
}

void FileManager::save_to_binary_file(const wavefunction_t& psi, std::string name) {
    std::ofstream f(path(name + ".gpe.bin"), std::ios::binary);
    if (!f.is_open()) throw std::runtime_error("Could not open last state file for writing.");
    int n[3] = {params->nx, params->ny, params->nz};
    double d[3] = {Units::len_au_to_nm(params->dx), Units::len_au_to_nm(params->dy), Units::len_au_to_nm(params->dz)};
    f.write(reinterpret_cast<const char*>(n), sizeof(n));
    f.write(reinterpret_cast<const char*>(d), sizeof(d));
    f.write(reinterpret_cast<const char*>(psi.data()), psi.size() * sizeof(std::complex<double>));
}

void FileManager::load_from_binary_file(std::string name, wavefunction_t& out) {
    std::ifstream f(path(name + ".gpe.bin"), std::ios::binary);
    if (!f.is_open()) throw std::runtime_error("Could not open last state file for reading.");
    int n[3];
    double d[3];
    f.read(reinterpret_cast<char*>(n), sizeof(n));
    if (n[0] != params->nx || n[1] != params->ny || n[2] != params->nz)
        throw std::runtime_error("Grid dimensions in the file do not match current parameters.");
    f.read(reinterpret_cast<char*>(d), sizeof(d));
    if (std::abs(Units::len_nm_to_au(d[0]) - params->dx) > 1e-3 ||
        std::abs(Units::len_nm_to_au(d[1]) - params->dy) > 1e-3 ||
        std::abs(Units::len_nm_to_au(d[2]) - params->dz) > 1e-3)
        throw std::runtime_error("Grid size in the file do not match current parameters.");
    out.assign((size_t)n[0] * n[1] * n[2], {0.0, 0.0});
    f.read(reinterpret_cast<char*>(out.data()), out.size() * sizeof(std::complex<double>));
    if (!f) throw std::runtime_error("Truncated binary wavefunction file.");
}

void FileManager::save_to_text_file(const wavefunction_t& psi, std::string name) {
    std::ofstream f(path(name + ".gpe.dat"));
    if (!f.is_open()) throw std::runtime_error("Can't open initial state file for writing.");
    f << params->nx << "\n" << params->ny << "\n" << params->nz << "\n";
    f << Units::len_au_to_nm(params->dx) << "\n"
      << Units::len_au_to_nm(params->dy) << "\n"
      << Units::len_au_to_nm(params->dz) << "\n";
    for (const auto& v : psi) f << v.real() << "\t" << v.imag() << "\n";
}

void FileManager::load_from_text_file(std::string name, wavefunction_t& out) {
    std::ifstream f(path(name + ".gpe.dat"));
    if (!f.is_open()) throw std::runtime_error("Could not open initial state file for reading.");
    int nx, ny, nz;
    f >> nx >> ny >> nz;
    if (nx != params->nx || ny != params->ny || nz != params->nz)
        throw std::runtime_error("Grid dimensions in the file do not match current parameters.");
    double dx, dy, dz;
    f >> dx >> dy >> dz;
    if (std::abs(Units::len_nm_to_au(dx) - params->dx) > 1e-3 || std::abs(Units::len_nm_to_au(dy) - params->dy) > 1e-3 ||
        std::abs(Units::len_nm_to_au(dz) - params->dz) > 1e-3)
        throw std::runtime_error("Grid size in the file do not match current parameters.");
    out.assign((size_t)nx * ny * nz, {0.0, 0.0});
    for (auto& v : out) {
        double re, im;
        f >> re >> im;
        v = {re, im};
    }
}

void FileManager::save_energies(const energies_container_t& e) {
    std::ofstream f(path("energy.gpe.dat"));
    int iter = 0;
    for (const auto& x : e)
        f << iter++ << "\t" << x.e_kin << "\t" << x.e_pot << "\t" << x.e_int << "\t" << x.e_ext << "\t" << x.e_bmf
          << "\t" << x.e_total << std::endl;
}

void FileManager::save_pote_to_text_file(const potential_t& V, std::string name) {
    std::ofstream f(path(name + ".gpe.dat"));
    if (!f.is_open()) throw std::runtime_error("Can't open potential file for writing.");
    f << "# nx = " << params->nx << "\n# ny = " << params->ny << "\n# nz = " << params->nz << "\n";
    f << "# dx = " << Units::len_au_to_nm(params->dx) << "\n# dy = " << Units::len_au_to_nm(params->dy)
      << "\n# dz = " << Units::len_au_to_nm(params->dz) << "\n";
    for (double v : V) f << v << "\n";
}

// ---------------------------------------------------------------- solver
void AbstractGrossPitaevskiSolver::initialize() {
    const size_t n = (size_t)params->nx * params->ny * params->nz;
    cpsi.assign(n, {0.0, 0.0});
    cpsii.assign(n, {0.0, 0.0});
    pote.assign(n, 0.0);
    init_containers();
    prepare_fft();
}

void AbstractGrossPitaevskiSolver::solve() {
    switch (params->calc_strategy) {
    case CalcStrategy::IMAGINARY_TIME:
        calc_initial_state();
        p_mediator->save_initial_state(cpsi);
        break;
    case CalcStrategy::REAL_TIME:
        calc_evolution();
        break;
    case CalcStrategy::FULL:
        calc_initial_state();
        p_mediator->save_initial_state(cpsi);
        free_potential_well();
        calc_evolution();
        break;
    }
}

void AbstractGrossPitaevskiSolver::imag_time_iter() {
    calc_fi3d();
    imag_iter_linear_step();
    imag_iter_nonlinear_step();
    calc_norm();
    normalize();
}

void AbstractGrossPitaevskiSolver::real_time_iter() {
    real_fft_potential_half_step();
    calc_fi3d();
    real_fft_kinetic_step();
    real_fft_potential_half_step();
}

void AbstractGrossPitaevskiSolver::imag_time_iters(long n) {
    for (long t = 0; t < n; ++t) imag_time_iter();
}

void AbstractGrossPitaevskiSolver::real_time_iters(long first_iter, long n) {
    for (long t = 0; t < n; ++t) {
        if (!params->const_edd) params->update_edd(first_iter + t);
        real_time_iter();
    }
}

// A snapshot begun at a checkpoint, finished once the next chunk of steps is
// launched: its copy runs beside that chunk.
void AbstractGrossPitaevskiSolver::finish_snapshot() {
    snapshot_end();
    p_mediator->save_checkpoint(cpsi);
    enes.push_back(ene);
}

// Loop of solver.cpp:48-70: iterations 0..iter_imag inclusive, checkpoint +
// energy every 100.  Iterations between two checkpoints run as one fused call;
// the checkpoint's copy and energies run beside the next call (synchronously
// when the convergence stop needs the energy at once).
void AbstractGrossPitaevskiSolver::calc_initial_state() {
    long iter = 0;
    double last = 0.0;
    bool have_last = false, pending = false;
    it_done = 0;
    while (iter <= params->iter_imag) {
        const long next_cp = (iter % 100 == 0) ? iter : (iter / 100 + 1) * 100;
        const long upto = std::min(next_cp, params->iter_imag);
        imag_time_iters(upto - iter + 1);
        if (pending) {
            finish_snapshot();
            pending = false;
        }
        it_done += upto - iter + 1;
        iter = upto;
        if (iter % 100 == 0) {
            if (it_tolerance <= 0 && snapshot_begin()) {
                pending = true;
            } else {
                export_data();
                p_mediator->save_checkpoint(cpsi);
                calc_energy();
                const double e = ene.e_total / params->n_atoms;
                if (it_tolerance > 0 && have_last && std::abs(e - last) < it_tolerance * std::abs(e)) break;
                last = e;
                have_last = true;
            }
        }
        ++iter;
    }
    if (pending) finish_snapshot();
    // As in the reference, cpsi keeps the state of the last checkpoint.
    p_mediator->save_energies(enes);
}

// Loop of solver.cpp:72-97: iterations 0..iter_real inclusive, update_edd
// before each step, checkpoint + energy every 1000 (their copy runs beside the
// next 1000 steps).
void AbstractGrossPitaevskiSolver::calc_evolution() {
    long iter = 0;
    bool pending = false;
    while (iter <= params->iter_real) {
        const long next_cp = (iter % 1000 == 0) ? iter : (iter / 1000 + 1) * 1000;
        const long upto = std::min(next_cp, params->iter_real);
        real_time_iters(iter, upto - iter + 1);
        if (pending) {
            finish_snapshot();
            pending = false;
        }
        iter = upto;
        if (iter % 1000 == 0) {
            if (snapshot_begin()) {
                pending = true;
            } else {
                export_data();
                p_mediator->save_checkpoint(cpsi);
                calc_energy();
            }
        }
        ++iter;
    }
    if (pending) finish_snapshot();
    p_mediator->save_data(cpsi);
    p_mediator->save_energies(enes);
}

void AbstractGrossPitaevskiSolver::free_potential_well() { p_mediator->request_free_potential(); }

void AbstractGrossPitaevskiSolver::load_buffer(const wavefunction_t& w) {
    cpsi = w;
    cpsii = w;
    import_data();
    calc_norm();
    normalize();
}

void AbstractGrossPitaevskiSolver::load_pote(const potential_t& V) {
    pote = V;
    import_pote();
}

// ---------------------------------------------------------------- HIP hooks
HipGrossPitaevskiSolver::~HipGrossPitaevskiSolver() {
    if (h) gpe_destroy(h);
}

void HipGrossPitaevskiSolver::check(int rc) const {
    if (rc != 0) throw std::runtime_error(std::string("gpehip: ") + gpe_last_error(h));
}

void HipGrossPitaevskiSolver::init_containers() {
    if (h) gpe_destroy(h);
    h = nullptr;
    const gpe_params g = params->to_abi();
    if (gpe_create(&g, params->precision, &h) != 0)
        throw std::runtime_error(std::string("gpehip: ") + gpe_last_error(nullptr));
}
void HipGrossPitaevskiSolver::prepare_fft() {}  // tables are built by gpe_create
void HipGrossPitaevskiSolver::calc_fi3d() { check(gpe_calc_fi3d(h)); }
void HipGrossPitaevskiSolver::calc_norm() { check(gpe_calc_norm(h, &xnorma)); }
void HipGrossPitaevskiSolver::normalize() { check(gpe_normalize(h)); }
void HipGrossPitaevskiSolver::imag_iter_linear_step() { check(gpe_imag_iter_step(h)); }
void HipGrossPitaevskiSolver::imag_iter_nonlinear_step() {}  // fused into the linear step
void HipGrossPitaevskiSolver::real_fft_potential_half_step() {
    check(gpe_set_coupling(h, params->ggp11, params->gamma));
    check(gpe_real_potential_half_step(h));
}
void HipGrossPitaevskiSolver::real_fft_kinetic_step() { check(gpe_real_kinetic_step(h)); }
void HipGrossPitaevskiSolver::export_data() {
    check(gpe_export_psi(h, reinterpret_cast<double*>(cpsi.data())));
}
void HipGrossPitaevskiSolver::import_data() {
    check(gpe_load_buffer(h, reinterpret_cast<const double*>(cpsi.data())));
}
void HipGrossPitaevskiSolver::import_pote() { check(gpe_load_pote(h, pote.data())); }

static gpe_trap trap_of(const PhysicalParameters& p) {
    gpe_trap t;
    t.omega_x = p.omega_x;
    t.omega_y = p.omega_y;
    t.omega_z = p.omega_z;
    t.dd = p.dd;
    return t;
}

void HipGrossPitaevskiSolver::init_state_on_device(const PhysicalParameters& p) {
    int strategy;
    switch (p.init_strategy) {
    case InitStrategy::COS: strategy = GPE_INIT_COS; break;
    case InitStrategy::GAUSS: strategy = GPE_INIT_GAUSS; break;
    case InitStrategy::MULTIPLE_GAUSS: strategy = GPE_INIT_MULTIPLE_GAUSS; break;
    case InitStrategy::SETUP_GAUSS: strategy = GPE_INIT_SETUP_GAUSS; break;
    default: throw std::runtime_error("init_state_on_device: file strategies load through FileManager");
    }
    const gpe_trap t = trap_of(p);
    // the host initializer applies gauss_sigma_cells to GAUSS only
    const double sigma = p.init_strategy == InitStrategy::GAUSS ? p.gauss_sigma_cells : 0.0;
    check(gpe_init_wavefunction(h, strategy, &t, p.n_gauss_max, p.bec_droplets_x, p.bec_droplets_y, p.bec_droplets_z,
                                sigma, p.phase_noise, p.noise_seed));
    cpsi.assign((size_t)p.nx * p.ny * p.nz, {0.0, 0.0});
    check(gpe_export_psi(h, reinterpret_cast<double*>(cpsi.data())));
    cpsii = cpsi;
}

void HipGrossPitaevskiSolver::init_pote_on_device(const std::string& key, potential_t& V) {
    static const char* keys[] = {"HARMONIC", "MEXICAN", "MEXICAN_FREE", "CYLINDRICAL", "FREE", "MEXICAN_ASYMETRIC",
                                 "MEXICAN_ASYMETRIC_FREE"};
    int k = GPE_POT_FREE;  // an unknown key falls back to FREE (potentials.hpp:48,53)
    for (int q = 0; q < 7; ++q)
        if (key == keys[q]) k = q;
    const gpe_trap t = trap_of(*params);
    check(gpe_init_potential(h, k, &t));
    V.assign((size_t)params->nx * params->ny * params->nz, 0.0);
    check(gpe_export_pote(h, V.data()));
    pote = V;
}

void HipGrossPitaevskiSolver::imag_time_iters(long n) {
    check(gpe_set_coupling(h, params->ggp11, params->gamma));
    check(gpe_imag_time_iters(h, n));
}

void HipGrossPitaevskiSolver::real_time_iters(long first_iter, long n) {
    if (params->const_edd) {
        check(gpe_set_coupling(h, params->ggp11, params->gamma));
        check(gpe_real_time_iters(h, n, nullptr, nullptr));
        return;
    }
    std::vector<double> g(n), gm(n);
    for (long t = 0; t < n; ++t) {
        params->update_edd(first_iter + t);
        g[t] = params->ggp11;
        gm[t] = params->gamma;
    }
    check(gpe_real_time_iters(h, n, g.data(), gm.data()));
}

bool HipGrossPitaevskiSolver::snapshot_begin() {
    check(gpe_set_coupling(h, params->ggp11, params->gamma));
    check(gpe_snapshot_begin(h));
    return true;
}

void HipGrossPitaevskiSolver::snapshot_end() {
    gpe_energies e;
    check(gpe_snapshot_end(h, reinterpret_cast<double*>(cpsi.data()), &e));
    ene.e_kin = e.e_kin;
    ene.e_pot = e.e_pot;
    ene.e_int = e.e_int;
    ene.e_ext = e.e_ext;
    ene.e_bmf = e.e_bmf;
    ene.mu = e.mu;
    ene.residual = e.residual;
    ene.sum();
}

void HipGrossPitaevskiSolver::calc_energy() {
    gpe_energies e;
    check(gpe_set_coupling(h, params->ggp11, params->gamma));
    check(gpe_calc_energy(h, &e));
    ene.e_kin = e.e_kin;
    ene.e_pot = e.e_pot;
    ene.e_int = e.e_int;
    ene.e_ext = e.e_ext;
    ene.e_bmf = e.e_bmf;
    ene.mu = e.mu;
    ene.residual = e.residual;
    ene.sum();
    enes.push_back(ene);
}

// ---------------------------------------------------------------- manager
SimulationManager::SimulationManager(std::string dir, bool quiet)
    : dir(std::move(dir)), quiet(quiet), m_file_manager(new FileManager(this, &params, this->dir)),
      m_solver(new HipGrossPitaevskiSolver(this, &params)) {}

void SimulationManager::initialize() { m_file_manager->load_params(); }
void SimulationManager::run_simulation() { m_solver->solve(); }

void SimulationManager::on_params_loaded() {
    params.init_parameters();
    m_solver->initialize();
    auto* dev = dynamic_cast<HipGrossPitaevskiSolver*>(m_solver.get());
    wavefunction_t psi;
    if (params.init_strategy == InitStrategy::FROM_BINARY_FILE) {
        request_load_from_binary(psi);
    } else if (params.init_strategy == InitStrategy::FROM_TEXT_FILE) {
        request_load_from_text(psi);
    } else if (dev) {
        dev->init_state_on_device(params);  // generated on the device
    } else {
        init_wavefunction(params, psi);
    }
    if (!psi.empty() || !dev) {
        add_phase_noise(psi, params.phase_noise, params.noise_seed);
        on_data_initialized(psi);
    }
    potential_t V;
    fill_potential(params.pote_key, params, V);  // FP64 values for the file the reference writes
    if (dev) {
        potential_t Vd;
        dev->init_pote_on_device(params.pote_key, Vd);
        m_file_manager->save_pote_to_text_file(V, "initial_potential");
    } else {
        on_pote_initialized(V);
    }
}

void SimulationManager::on_data_loaded(const wavefunction_t& w) { m_solver->load_buffer(w); }
void SimulationManager::on_data_initialized(const wavefunction_t& w) { m_solver->load_buffer(w); }
void SimulationManager::on_pote_initialized(const potential_t& V) {
    m_file_manager->save_pote_to_text_file(V, "initial_potential");
    m_solver->load_pote(V);
}

void SimulationManager::change_potential(const std::string& key) {
    params.pote_key = key;
    potential_t V;
    fill_potential(key, params, V);  // FP64 values for the file the reference writes
    if (auto* dev = dynamic_cast<HipGrossPitaevskiSolver*>(m_solver.get())) {
        potential_t Vd;
        dev->init_pote_on_device(key, Vd);
        m_file_manager->save_pote_to_text_file(V, "changed_potential");
        return;
    }
    m_file_manager->save_pote_to_text_file(V, "changed_potential");
    m_solver->load_pote(V);
}

void SimulationManager::request_free_potential() { change_potential("MEXICAN_FREE"); }

void SimulationManager::check_load_filename() {
    std::smatch m;
    static const std::regex pat("wavefunction_(\\d+)");
    if (std::regex_search(params.load_filename, m, pat)) checkpoint_counter = std::stoi(m[1].str()) + 1;
}

void SimulationManager::request_load_from_binary(wavefunction_t& w) {
    check_load_filename();
    m_file_manager->load_from_binary_file(params.load_filename, w);
}
void SimulationManager::request_load_from_text(wavefunction_t& w) {
    check_load_filename();
    m_file_manager->load_from_text_file(params.load_filename, w);
}
void SimulationManager::save_data(const wavefunction_t& w) { m_file_manager->save_to_binary_file(w, "bec_wavefunction"); }
void SimulationManager::save_checkpoint(const wavefunction_t& w) {
    m_file_manager->save_to_binary_file(w, "wavefunction_" + std::to_string(checkpoint_counter));
    checkpoint_counter++;
    if (!quiet) std::cout << "[INFO]    checkpoint " << checkpoint_counter - 1 << std::endl;
}
void SimulationManager::save_initial_state(const wavefunction_t& w) {
    m_file_manager->save_to_text_file(w, "initial_state");
}
void SimulationManager::save_energies(const energies_container_t& e) { m_file_manager->save_energies(e); }

}  // namespace host
}  // namespace gpe
