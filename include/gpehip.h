/*
 * gpehip.h -- C ABI of the MI355X (gfx950) Gross-Pitaevskii solver backend.
 *
 * One handle = one solver instance on one HIP stream: the state of
 * GpuGrossPitaevskiSolver / CpuGrossPitaevskiSolver in the reference
 * (src/solver/cuda_solver/gpu_solver.hpp, src/solver/cpu_solver/cpu_solver.hpp),
 * with the device buffers owned by the library and every host buffer owned by
 * the caller.  No C++ exception crosses this boundary: every call returns 0 on
 * success and a non-zero status otherwise; gpe_last_error() gives the message.
 *
 * Host arrays use the reference's StdMat3D layout
 * (src/utils/types/mat3d/stdmat3d.hpp:26-28): row-major, index (i*ny+j)*nz+k,
 * complex values interleaved (re, im) float64.
 *
 * Each entry point names the reference hook it replaces (file:line).
 */
#ifndef GPEHIP_H
#define GPEHIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gpe_handle gpe_handle;

/* Arithmetic of the bulk arrays.  Norm, energy and chemical-potential
 * reductions are FP64 in both. */
enum { GPE_FP32 = 32, GPE_FP64 = 64 };

/* Solver constants in atomic units, computed on the host exactly as
 * PhysicalParameters::init_parameters does (src/parameters/parameters.cpp:37-51),
 * plus the two time steps (src/parameters/parameters.hpp:70-72). */
typedef struct {
    int nx, ny, nz;
    double dx, dy, dz;
    double m;
    double n_atoms;
    double cdd;
    double ggp11;
    double gamma;
    double w_15;
    double imag_time_dt;
    double real_time_dt;
} gpe_params;

/* Energy record: BECEnergies (src/context/context.hpp:16-27) plus the two
 * quantities the convergence test uses.  Totals in hartree. */
typedef struct {
    double e_kin, e_pot, e_int, e_ext, e_bmf, e_total;
    double mu;       /* <psi|H|psi>, H = the imaginary-time operator */
    double residual; /* sqrt(dV * sum |H psi - mu psi|^2) over the interior */
} gpe_energies;

/* Create / destroy.  Replaces GpuGrossPitaevskiSolver::GpuGrossPitaevskiSolver +
 * init_containers + prepare_fft (src/solver/cuda_solver/gpu_solver.cpp:4-38,
 * src/solver/solver.cpp:15-23).  Grid lengths of the form 2^a 3^b 5^c 7^d;
 * planes whose Poisson tile exceeds one block's LDS (256x256x64, say) run
 * the pencil plane pass, with two extra complex fields of workspace. */
int gpe_create(const gpe_params *p, int precision, gpe_handle **out);
void gpe_destroy(gpe_handle *h);
/* Message of the last failed call on h (h == NULL: the calling thread's last
 * failed gpe_create or gpe_batch_create). */
const char *gpe_last_error(const gpe_handle *h);
/* Library version string. */
const char *gpe_version(void);

/* AbstractGrossPitaevskiSolver::load_buffer (src/solver/solver.cpp:118-125):
 * cpsi = cpsii = psi, then calc_norm + normalize. */
int gpe_load_buffer(gpe_handle *h, const double *psi);
/* AbstractGrossPitaevskiSolver::load_pote (src/solver/solver.cpp:127-131). */
int gpe_load_pote(gpe_handle *h, const double *V);
/* export_data (src/solver/cuda_solver/gpu_solver.cpp:136-139): normalised cpsi. */
int gpe_export_psi(gpe_handle *h, double *psi);
/* fi3d (the dipolar potential buffer) to / from the host. */
int gpe_export_fi3d(gpe_handle *h, double *fi3d);
int gpe_import_fi3d(gpe_handle *h, const double *fi3d);
/* PhysicalParameters::update_edd result (src/parameters/parameters.cpp:53-59):
 * new contact and LHY couplings for the following steps. */
int gpe_set_coupling(gpe_handle *h, double ggp11, double gamma);

/* On-device initial states and potentials, per handle (and so per batch
 * instance).  Keys in the reference's registration order
 * (src/parameters/potentials.hpp:76-157, src/parameters/complex_param.hpp). */
enum {
    GPE_POT_HARMONIC = 0,
    GPE_POT_MEXICAN = 1,
    GPE_POT_MEXICAN_FREE = 2,
    GPE_POT_CYLINDRICAL = 3,
    GPE_POT_FREE = 4,
    GPE_POT_MEXICAN_ASYMETRIC = 5,
    GPE_POT_MEXICAN_ASYMETRIC_FREE = 6
};
enum { GPE_INIT_COS = 0, GPE_INIT_GAUSS = 1, GPE_INIT_MULTIPLE_GAUSS = 2, GPE_INIT_SETUP_GAUSS = 3 };
/* Trap of PhysicalParameters (src/parameters/parameters.hpp:10-123), a.u. */
typedef struct {
    double omega_x, omega_y, omega_z;
    double dd;
} gpe_trap;
/* DataInitializer::init_pote (src/initializer/initializer.cpp:229-248):
 * V on every grid point, computed on the device.  Replaces building V on the
 * host and gpe_load_pote. */
int gpe_init_potential(gpe_handle *h, int pote_key, const gpe_trap *trap);
/* DataInitializer::init_wavefunction for COS / GAUSS / MULTIPLE_GAUSS /
 * SETUP_GAUSS (src/initializer/initializer.cpp:46-219) on the device, then
 * load_buffer (src/solver/solver.cpp:118-125): cpsi = cpsii = psi_0, norm,
 * normalize.  sigma_cells > 0 sets the GAUSS / MULTIPLE_GAUSS width to that
 * many cells per axis (the reference's L/10 and L/20 otherwise);
 * phase_noise > 0 multiplies psi_0 by exp(i phase_noise (2u - 1)), u the
 * e-th SplitMix64 output of noise_seed at layout index e. */
int gpe_init_wavefunction(gpe_handle *h, int init_strategy, const gpe_trap *trap, int n_gauss_max, int drops_x,
                          int drops_y, int drops_z, double sigma_cells, double phase_noise,
                          unsigned long long noise_seed);
/* The potential as the solver holds it (the reference keeps it in
 * CPUSolverData::pote, src/solver/solver_data/cpu_solver_data.hpp:11-31). */
int gpe_export_pote(gpe_handle *h, double *V);

/* Hooks (src/solver/solver.hpp:55-63). */
int gpe_calc_fi3d(gpe_handle *h);                   /* calc_fi3d: fftw_poisson_solver.cpp:51-96 */
int gpe_calc_norm(gpe_handle *h, double *xnorma);   /* calc_norm: cpu_solver.cpp:19-34 */
int gpe_normalize(gpe_handle *h);                   /* normalize: cpu_solver.cpp:36-48 */
int gpe_imag_iter_step(gpe_handle *h);              /* imag_iter_linear_step + imag_iter_nonlinear_step:
                                                       cpu_solver.cpp:50-97 */
int gpe_real_potential_half_step(gpe_handle *h);    /* cpu_solver.cpp:99-131 */
int gpe_real_kinetic_step(gpe_handle *h);           /* fftw_split_solver.cpp:44-86 */
int gpe_calc_energy(gpe_handle *h, gpe_energies *out); /* cpu_solver.cpp:137-194 (+ mu, residual) */
/* The output cadence of the solver loops without stopping them: the reference
 * copies the state and computes the energies synchronously every 100 IT /
 * 1000 RT steps (export_data + save_checkpoint + calc_energy,
 * src/solver/solver.cpp:54-59,81-86).  gpe_snapshot_begin enqueues, behind
 * the work already launched, a device copy of the state and its energies
 * (gpe_calc_energy's values), and their copy to pinned host memory on a second
 * stream; later launches are not held back.  gpe_snapshot_end waits for that
 * copy only and returns the state (as gpe_export_psi; psi may be NULL) and the
 * energies (e may be NULL).  One snapshot per handle in flight. */
int gpe_snapshot_begin(gpe_handle *h);
int gpe_snapshot_end(gpe_handle *h, double *psi, gpe_energies *e);
/* No counterpart in the reference.  For callers that cannot allocate (code
 * under hipGraph capture, tight serving loops): gpe_calc_energy with no
 * allocation, host copy or synchronisation.  The eight values are written,
 * in gpe_energies order, to out_dev (8 doubles of device memory) on the
 * handle's stream; the per-plane partials live in a static device array of
 * the library (4096 planes at most: larger nx is rejected).  Every call on
 * one GPU shares that array: calls that overlap on one GPU race. */
int gpe_calc_energy_noalloc(gpe_handle *h, double *out_dev);

/* Fused loops: n consecutive imag_time_iter (src/solver/solver.cpp:99-105) /
 * real_time_iter (src/solver/solver.cpp:107-112) calls.  For the real-time
 * loop, ggp11/gamma may give one coupling pair per step (the update_edd ramp
 * of src/solver/solver.cpp:76-77); NULL keeps the current coupling. */
int gpe_imag_time_iters(gpe_handle *h, long n);
int gpe_real_time_iters(gpe_handle *h, long n, const double *ggp11, const double *gamma);
/* The calc_initial_state loop (src/solver/solver.cpp:48-70) with the energy
 * check it prints every 100 steps (solver.cpp:58-60) used as a stop test:
 * imaginary time until |E/N(k) - E/N(k - check_every)| < tol |E/N(k)| or
 * max_steps (tol <= 0: max_steps).  steps_out / energy_per_atom_out (may be
 * NULL) receive the steps run and the last E/N.  Same test as
 * gpe_batch_it_until. */
int gpe_imag_time_until(gpe_handle *h, long max_steps, long check_every, double tol, long *steps_out,
                        double *energy_per_atom_out);

/* Batch (no reference counterpart: the reference is single-instance with
 * singleton parameters, src/parameters/parameters.hpp:14-20).  n instances,
 * each with its own constants and grid (grids may differ), share one stream;
 * in FP32 every pass of a step is one launch over all active instances, each
 * block running the single-instance kernel body of its instance's grid (so
 * results equal a single handle's bit for bit; one launch per grid shape in
 * FP64 and for run-time-size grids), and the steady-state steps are replayed
 * as hipGraphs.  gpe_batch_instance gives the per-instance handle for
 * load/export/energy calls. */
typedef struct gpe_batch gpe_batch;
int gpe_batch_create(const gpe_params *params, int n, int precision, gpe_batch **out);
void gpe_batch_destroy(gpe_batch *b);
gpe_handle *gpe_batch_instance(gpe_batch *b, int i);
const char *gpe_batch_last_error(const gpe_batch *b);
/* Imaginary time on every instance until |E/N(k) - E/N(k - check_every)| <
 * tol |E/N(k)| (energies checked every check_every steps) or max_steps; a
 * converged instance leaves the launch set, and the last instance left runs
 * its handle's own fused loop (gpe_imag_time_iters; GPE_BATCH_SINGLE=0 keeps
 * the batch's launches).  steps_out / energy_per_atom_out (length n, may be
 * NULL) receive each instance's step count and final E/N. */
int gpe_batch_it_until(gpe_batch *b, long max_steps, long check_every, double tol, long *steps_out,
                       double *energy_per_atom_out);
/* Fallback of a warm-started instance (no reference counterpart; the
 * fallback state enters through load_buffer, src/solver/solver.cpp:118-125):
 * psi (interleaved complex128, the StdMat3D layout; NULL clears it) is the
 * state gpe_batch_it_until restarts instance i from, once, when at a check
 * its energy or residual is not finite (a NaN or zero state given to
 * gpe_load_buffer), its E/N has risen since the previous check by more
 * than tol relative (1e-9 when tol is 0: imaginary time on a stable map only
 * lowers it), or dtau mu >= 1 (the map's factor 1 - dtau mu is not positive
 * on the state: a mode an unstable map has grown).  Its step count keeps
 * running.  gpe_batch_fallback_taken tells whether the last
 * gpe_batch_it_until took it. */
int gpe_batch_set_fallback(gpe_batch *b, int i, const double *psi);
int gpe_batch_fallback_taken(const gpe_batch *b, int i);

/* Learned warm start (no reference counterpart): a Fourier neural operator
 * mapping per-point parameter fields [B][nx*ny*nz][cin] to a field
 * [B][nx*ny*nz] (the ground-state amplitude, used as the IT initial state via
 * gpe_load_buffer).  Every layer runs as BF16-MFMA GEMMs with FP32
 * accumulation; weights are FP32 masters rounded to BF16 at use. */
typedef struct {
    int nx, ny, nz;
    int cin;       /* input channels */
    int width;     /* channel width of the spectral layers */
    int layers;    /* number of spectral layers */
    int hidden;    /* projection hidden width */
    int modes;     /* retained modes per axis (capped at nx, ny and nz/2+1) */
    int max_batch; /* samples per call */
} gpe_fno_config;
typedef struct gpe_fno gpe_fno;
int gpe_fno_create(const gpe_fno_config *c, unsigned long long seed, gpe_fno **out);
void gpe_fno_destroy(gpe_fno *f);
const char *gpe_fno_last_error(const gpe_fno *f);
long gpe_fno_num_params(const gpe_fno *f);
int gpe_fno_get_params(gpe_fno *f, float *host);
int gpe_fno_set_params(gpe_fno *f, const float *host);
int gpe_fno_get_grads(gpe_fno *f, float *host);
int gpe_fno_forward(gpe_fno *f, const float *x, int batch, float *y_out);
int gpe_fno_loss_grad(gpe_fno *f, const float *x, const float *target, int batch, double *loss);
int gpe_fno_adam_step(gpe_fno *f, float lr);
/* Training set resident on the device: n samples, inputs [n][P][cin] and
 * targets [n][P] (host, FP32), copied once; gpe_fno_loss_grad_idx then takes
 * the batch idx[0..batch) from it with device copies, as gpe_fno_loss_grad
 * does from host arrays. */
int gpe_fno_set_data(gpe_fno *f, const float *x, const float *target, long n);
int gpe_fno_loss_grad_idx(gpe_fno *f, const int *idx, int batch, double *loss);
/* The network's input and output buffers on the device ([max_batch][P][cin]
 * and [max_batch][P], FP32; valid until gpe_fno_destroy) and a forward pass
 * of the first `batch` inputs that leaves its output there: with
 * gpe_warm_features and gpe_load_amplitude the warm start never leaves the
 * device. */
float *gpe_fno_input_dev(gpe_fno *f);
const float *gpe_fno_output_dev(gpe_fno *f);
int gpe_fno_forward_dev(gpe_fno *f, int batch);
/* Warm start on the device (no reference counterpart).  gpe_warm_features
 * writes the network input of instance h, [P][cin] at x_dev (a device
 * pointer, e.g. gpe_fno_input_dev + slot P cin): channel 0 = V / v_scale from
 * the instance's potential, channels 1 .. cin-1 = consts[0 .. cin-2].
 * gpe_load_amplitude loads psi0 = max(a, 0) on the interior, zero on the
 * boundary, from the predicted amplitude a at amp_dev ([P], device; with
 * mirror, a is first averaged with its mirror images along x, y, z over the
 * interior pairs), then normalises as gpe_load_buffer does
 * (src/solver/solver.cpp:118-125); *norm_out receives dV sum |psi0|^2 before
 * normalising, NaN or 0 for a prediction the caller must replace. */
int gpe_warm_features(gpe_handle *h, float *x_dev, int cin, double v_scale, const float *consts);
int gpe_load_amplitude(gpe_handle *h, const float *amp_dev, int mirror, double *norm_out);
/* C = op(A) op(B) on the MFMA GEMM (row-major host matrices; imaginary parts
 * may be NULL; Ci NULL keeps only the real part).  Test entry point. */
int gpe_gemm_test(const float *Ar, const float *Ai, const float *Br, const float *Bi, float *Cr, float *Ci, int M,
                  int N, int K, int conjA, int conjB);
/* C = A op(B) (+ bias) + beta C for many rows and N, K <= 64 (the network's
 * pointwise layers) on the row kernel (rows = 1) or the tile kernel
 * (rows = 0); host row-major matrices, B [K][N] or (transB) [N][K], C in /
 * out.  Test entry point: the two kernels agree bit for bit. */
int gpe_gemm_rows_test(const float *A, const float *B, const float *bias, float *C, long M, int N, int K, int transB,
                       float beta, int rows);
/* Weight-gradient partials P[z] = A_z^T D_z of `chunks` chunks of 512 rows
 * (A [chunks 512][K], D [chunks 512][N], K, N <= 64, host row-major) on the
 * chunk kernel (wg = 1) or the tile kernel (wg = 0).  Test entry point: the
 * two agree bit for bit. */
int gpe_gemm_wgrad_test(const float *A, const float *D, float *P, int chunks, int K, int N, int wg);

/* Autotuner (no reference counterpart).  Times the launch-shape variants of
 * the fused loops (one or two blocks per plane in the steady-state plane pass;
 * register-resident or LDS x pass; graph-replayed or directly launched
 * steady state) over `steps` steps on a scratch instance of h's grid and
 * precision, and keeps the fastest for h (cached per grid and precision for
 * the process).  config receives bit 0 = split plane pass, bit 1 = register
 * x pass, bit 2 = graph replay; rt_step_us / it_step_us the tuned step times.
 * Every variant computes the same transforms; the two x passes round
 * differently, so a tuned handle is deterministic but may differ in the last
 * bits from an untuned one (defaults: split, register x pass). */
int gpe_autotune(gpe_handle *h, long steps, int *config, double *rt_step_us, double *it_step_us);

/* Launch tiles (no reference counterpart): kz columns per Poisson tile (cp)
 * and k columns per kinetic tile (ck) of the LDS x pass, and the pencil plane
 * pass's column tile (pen_c, LDS y passes), row tile (pen_r), register y
 * passes (pen_yreg, 0/1) and register z passes of its row tiles (pen_zreg,
 * 0/1; nz = 64 with 32-row tiles).  gpe_autotune searches them after the
 * launch variant (cp / ck when it picks the LDS x pass, the pencil ones on
 * pencil grids); gpe_set_tiles forces them (a value <= 0, or pen_yreg /
 * pen_zreg < 0, keeps the current one) within the bounds the kernels were
 * built for, which gpe_get_tiles reports: cur[6] and bound[6] in the order
 * cp, ck, pen_c, pen_r, pen_yreg, pen_zreg (pencil entries 0 on single-block
 * grids).  Tiles change only how lines are grouped into blocks: the LDS x
 * pass gives the same values at every cp / ck, the pencil pass the same up
 * to the summation order of its FP64 norm partials (and the rounding of the
 * register z passes against the LDS ones). */
int gpe_set_tiles(gpe_handle *h, int cp, int ck, int pen_c, int pen_r, int pen_yreg, int pen_zreg);
int gpe_get_tiles(const gpe_handle *h, int *cur, int *bound);

/* Batch autotuner (no reference counterpart): per x length of the batch's
 * grids (one x-pass launch per length), the register x pass where it exists
 * and the LDS x pass at each cp tile are timed as whole batched IT steps
 * (`steps` steps each) on a scratch batch of the same grids and precision,
 * one length at a time, and the fastest is kept for every instance of that
 * length.  config (length n, may be NULL) receives per instance bit 1 =
 * register x pass and bits 8.. = cp. */
int gpe_batch_autotune(gpe_batch *b, long steps, int *config);

/* Wait for all work of h. */
int gpe_synchronize(gpe_handle *h);

/* Timing of the fused loops on the handle's stream: average duration per
 * launch (microseconds) of the line pass and the plane pass over the last
 * gpe_*_iters call when profiling is on.  Used by the benchmark. */
int gpe_set_timing(gpe_handle *h, int on);
int gpe_get_timing(gpe_handle *h, double *plane_us, double *line_us, double *total_ms);

#ifdef __cplusplus
}
#endif
#endif
