/*
 * gpe_oracle.h -- FP64 CPU restatement of GPESolver's cpu_solver hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP backend:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it.  The product (gpesolver_amd) never links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the GPESolver source tree).  Layout everywhere is the reference's StdMat3D
 * layout: row-major, index (i*ny + j)*nz + k, k fastest
 * (src/utils/types/mat3d/stdmat3d.hpp:26-28); complex values are interleaved
 * (re, im) float64 pairs.
 */
#ifndef GPE_ORACLE_H
#define GPE_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Physical parameters in atomic units (src/parameters/parameters.hpp:10-123). */
typedef struct {
    int nx, ny, nz;
    double dx, dy, dz;
    double m, n_atoms;
    double omega_x, omega_y, omega_z;
    double dd;
    double edd, edd_start, edd_stop;
    int const_edd;
    long iter_imag, iter_real;
    double imag_time_dt, real_time_dt;
    /* derived by orc_init_parameters */
    double aa, b, dxdydz, add, cdd, ggp11, gamma, w_15;
} orc_params;

/* Potential keys in registration order (src/parameters/potentials.hpp:76-157). */
enum {
    ORC_POT_HARMONIC = 0,
    ORC_POT_MEXICAN = 1,
    ORC_POT_MEXICAN_FREE = 2,
    ORC_POT_CYLINDRICAL = 3,
    ORC_POT_FREE = 4,
    ORC_POT_MEXICAN_ASYMETRIC = 5,
    ORC_POT_MEXICAN_ASYMETRIC_FREE = 6
};

/* Initial-state strategies (src/parameters/complex_param.hpp, initializer.cpp). */
enum { ORC_INIT_COS = 0, ORC_INIT_GAUSS = 1, ORC_INIT_MULTIPLE_GAUSS = 2, ORC_INIT_SETUP_GAUSS = 3 };

/* Unit conversions (src/utils/units/units.cpp:3-41). */
double orc_len_nm_to_au(double nm);
double orc_len_au_to_nm(double au);
double orc_mass_Da_to_au(double da);
double orc_freq_Hz_to_au(double hz);
double orc_ene_au_to_meV(double au);

void orc_init_parameters(orc_params *p);
void orc_update_edd(orc_params *p, long iter);
double orc_coord(int n, double d, int i);
double orc_potential_value(int key, const orc_params *p, double x, double y, double z);
void orc_fill_potential(int key, const orc_params *p, double *V);
int orc_init_wavefunction(int strategy, const orc_params *p, int n_gauss_max, int drop_x,
                          int drop_y, int drop_z, double sigma_cells, double *psi);

/* FFT: unnormalised DFT, sign -1 forward / +1 backward, any length (radix 2,3,4,5
 * plus a direct DFT for other prime factors). */
void orc_fft_1d(double *x, int n, int sign);
void orc_fft_3d(double *x, int nx, int ny, int nz, int sign);
void orc_set_threads(int n);
int orc_get_threads(void);

/* Solver state: the four CPUSolverData arrays (src/solver/solver_data/cpu_solver_data.hpp:11-31)
 * plus the two FFT helper objects. */
typedef struct orc_solver orc_solver;

orc_solver *orc_solver_create(const orc_params *p);
void orc_solver_destroy(orc_solver *s);
orc_params *orc_solver_params(orc_solver *s);
double *orc_psi(orc_solver *s);
double *orc_psii(orc_solver *s);
double *orc_fi3d(orc_solver *s);
double *orc_pote(orc_solver *s);
double orc_xnorma(orc_solver *s);
const double *orc_vdip_table(orc_solver *s);
const double *orc_kinetic_table(orc_solver *s);
void orc_refresh_derived(orc_solver *s);

void orc_load_buffer(orc_solver *s, const double *psi);
void orc_load_pote(orc_solver *s, const double *V);
void orc_calc_fi3d(orc_solver *s);
void orc_calc_norm(orc_solver *s);
void orc_normalize(orc_solver *s);
void orc_imag_iter_linear_step(orc_solver *s);
void orc_imag_iter_nonlinear_step(orc_solver *s);
void orc_real_fft_potential_half_step(orc_solver *s);
void orc_real_fft_kinetic_step(orc_solver *s);
void orc_imag_time_iter(orc_solver *s);
void orc_real_time_iter(orc_solver *s);
void orc_calc_energy(orc_solver *s, double out[8]);
void orc_it_steps(orc_solver *s, long n);
void orc_rt_steps(orc_solver *s, long first_iter, long n);

#ifdef __cplusplus
}
#endif
#endif
