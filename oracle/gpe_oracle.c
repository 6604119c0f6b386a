/*
 * gpe_oracle.c -- FP64 CPU restatement of GPESolver's cpu_solver hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see gpe_oracle.h).  It is the checker for the HIP
 * backend and is never linked into, or called by, the product.
 *
 * The reference's own CPU path cannot be built in this image (it includes
 * <fftw3.h>, which is absent), so this file restates its algorithms from the
 * reference sources, cited per function.  The FFT is written here from the
 * textbook Stockham autosort formulation; it is pinned against numpy.fft and a
 * direct DFT in tests/test_oracle_fft.py.
 */
#include "gpe_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

/* ---------------------------------------------------------------- units */
/* src/utils/units/units.cpp:3-41 */
double orc_len_nm_to_au(double nm) { return nm * 18.897261260649092; }
double orc_len_au_to_nm(double au) { return au * 0.05291772105; }
double orc_mass_Da_to_au(double da) { return da / 0.000548579909; }
double orc_freq_Hz_to_au(double hz) { return hz * (4.1356e-12 / 27211.6); }
double orc_ene_au_to_meV(double au) { return au * 27211.407953; }

/* ----------------------------------------------------------- parameters */
/* src/parameters/parameters.cpp:37-51 (init_parameters). */
void orc_init_parameters(orc_params *p) {
    p->aa = p->m * pow(p->omega_x, 2) / 4. / pow(p->dd, 2);
    p->b = 0.5 * p->m * pow(p->omega_x, 2);
    p->dxdydz = p->dx * p->dy * p->dz;
    p->add = 131.;
    p->cdd = 12. * M_PI * p->add / p->m;
    double a = p->add / p->edd;
    p->ggp11 = 4. * M_PI * a / p->m;
    p->gamma = 128. * sqrt(M_PI) * pow(a, 2.5) / 3. / p->m * (1. + 1.5 * pow(p->edd, 2));
    p->w_15 = pow(p->n_atoms, 1.5);
}

/* src/parameters/parameters.cpp:53-59 (update_edd): only g and gamma change. */
void orc_update_edd(orc_params *p, long iter) {
    p->edd = p->edd_start + (p->edd_stop - p->edd_start) * (double)iter / (double)p->iter_real;
    double a = p->add / p->edd;
    p->ggp11 = 4 * M_PI * a / p->m;
    p->gamma = 128. * sqrt(M_PI) * pow(a, 2.5) / 3. / p->m * (1. + 1.5 * pow(p->edd, 2));
}

/* src/context/context.cpp:17-33: x_i = (i - floor(n/2)) * d. */
double orc_coord(int n, double d, int i) { return (double)(i - (n / 2)) * d; }

/* ------------------------------------------------------------ potentials */
/* src/parameters/potentials.hpp:76-157. */
double orc_potential_value(int key, const orc_params *p, double x, double y, double z) {
    const double vy = 0.5 * p->m * pow(y, 2) * pow(p->omega_y, 2);
    const double vz = 0.5 * p->m * pow(z, 2) * pow(p->omega_z, 2);
    const double conv = 27211.4 * 11.6 * 1e9;
    const double slope = (0.25 / conv) / p->dd;
    switch (key) {
    case ORC_POT_HARMONIC:
        return 0.5 * p->m * pow(x, 2) * pow(p->omega_x, 2) + vy + vz;
    case ORC_POT_MEXICAN:
        return -p->b * pow(x, 2) + p->aa * pow(x, 4) + vy + vz;
    case ORC_POT_MEXICAN_FREE:
        return p->aa * pow(x, 4) + vy + vz;
    case ORC_POT_CYLINDRICAL: {
        double r = sqrt(x * x + y * y);
        return 0.5 * p->m * pow(r, 2) * pow((p->omega_y + p->omega_x) / 2., 2) + vz;
    }
    case ORC_POT_FREE:
        return vz;
    case ORC_POT_MEXICAN_ASYMETRIC:
        return -p->b * pow(x, 2) + p->aa * pow(x, 4) - slope * x + vy + vz;
    case ORC_POT_MEXICAN_ASYMETRIC_FREE:
        return p->aa * pow(x, 4) - slope * x + vy + vz;
    default:
        return vz; /* unknown id falls back to FREE (potentials.hpp:48,53) */
    }
}

/* src/initializer/initializer.cpp:229-248 (init_pote): V sampled on every point.
 * The sampler goes through DataInitializer::_pote_func, declared
 * std::function<double(int, int, int)> (src/initializer/initializer.hpp:37):
 * the coordinates reach the potential truncated toward zero to whole atomic
 * units of length. */
static double orc_trunc_coord(double x) { return (double)(int)x; }
void orc_fill_potential(int key, const orc_params *p, double *V) {
    for (int i = 0; i < p->nx; i++)
        for (int j = 0; j < p->ny; j++)
            for (int k = 0; k < p->nz; k++)
                V[((size_t)i * p->ny + j) * p->nz + k] =
                    orc_potential_value(key, p, orc_trunc_coord(orc_coord(p->nx, p->dx, i)),
                                        orc_trunc_coord(orc_coord(p->ny, p->dy, j)),
                                        orc_trunc_coord(orc_coord(p->nz, p->dz, k)));
}

/* ---------------------------------------------------------- initializer */
/* src/initializer/initializer.cpp:46-219.  Interior points 1..n-2 only; the
 * boundary stays zero.  sigma_cells > 0 replaces the reference's GAUSS width
 * L/10 by sigma_cells grid cells per axis (config option; 0 = reference). */
int orc_init_wavefunction(int strategy, const orc_params *p, int n_gauss_max, int drop_x,
                          int drop_y, int drop_z, double sigma_cells, double *psi) {
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    memset(psi, 0, sizeof(double) * 2 * (size_t)nx * ny * nz);
    int nmax = 0;
    double *cx = NULL, *cy = NULL, *cz = NULL;
    if (strategy == ORC_INIT_MULTIPLE_GAUSS) {
        /* initializer.cpp:167-190 */
        nmax = n_gauss_max;
        if (nmax <= 0) return -1;
        cx = calloc(nmax, sizeof(double));
        cy = calloc(nmax, sizeof(double));
        cz = calloc(nmax, sizeof(double));
        for (int idx = 0; idx < nmax; idx++) {
            cx[idx] = (idx % 2) ? p->dd : -p->dd;
            if (nmax % 2 == 1) {
                int center_y_idx = nmax / 2;
                cy[idx] = (idx - center_y_idx) * (ny * p->dy) / 2. / (nmax + 1.);
            } else {
                int y_idx = idx / 2 + 1;
                int y_maximas = nmax / 2 + 1;
                cy[idx] = orc_coord(ny, p->dy, 0) + y_idx * (ny * p->dy) / y_maximas;
            }
        }
    } else if (strategy == ORC_INIT_SETUP_GAUSS) {
        /* initializer.cpp:101-130 */
        nmax = drop_x * drop_y * drop_z;
        if (nmax <= 0) return -1;
        cx = calloc(nmax, sizeof(double));
        cy = calloc(nmax, sizeof(double));
        cz = calloc(nmax, sizeof(double));
        double xt = nx * p->dx, yt = ny * p->dy, zt = nz * p->dz;
        for (int a = 0; a < drop_x; a++)
            for (int b = 0; b < drop_y; b++)
                for (int c = 0; c < drop_z; c++) {
                    int idx = a * drop_y * drop_z + b * drop_z + c;
                    cx[idx] = -xt / 2 + xt / (double)(drop_x + 1) * (a + 1);
                    cy[idx] = -yt / 2 + yt / (double)(drop_y + 1) * (b + 1);
                    cz[idx] = -zt / 2 + zt / (double)(drop_z + 1) * (c + 1);
                }
    } else if (strategy != ORC_INIT_COS && strategy != ORC_INIT_GAUSS) {
        return -1;
    }
    for (int i = 1; i < nx - 1; i++)
        for (int j = 1; j < ny - 1; j++)
            for (int k = 1; k < nz - 1; k++) {
                double x = orc_coord(nx, p->dx, i), y = orc_coord(ny, p->dy, j),
                       z = orc_coord(nz, p->dz, k);
                double val = 0.0;
                if (strategy == ORC_INIT_COS) {
                    /* initializer.cpp:46-71 */
                    double xs = x - p->dd;
                    double rrr = (nx / 2) * p->dx;
                    val = cos(2 * M_PI * xs / rrr) * cos(2 * M_PI * y / rrr) * cos(1 * M_PI * z / rrr);
                } else if (strategy == ORC_INIT_GAUSS) {
                    /* initializer.cpp:73-99 */
                    double sx = (nx * p->dx) / 10., sy = (ny * p->dy) / 10., sz = (nz * p->dz) / 10.;
                    if (sigma_cells > 0) {
                        sx = sigma_cells * p->dx;
                        sy = sigma_cells * p->dy;
                        sz = sigma_cells * p->dz;
                    }
                    val = exp(-0.5 * (x * x) / (sx * sx)) * exp(-0.5 * (y * y) / (sy * sy)) *
                          exp(-0.5 * (z * z) / (sz * sz));
                } else {
                    /* initializer.cpp:133-159 and 192-218: sigma = L/20 per axis
                     * (sigma_cells > 0 overrides it for MULTIPLE_GAUSS) */
                    double sx = (nx * p->dx) / 20., sy = (ny * p->dy) / 20., sz = (nz * p->dz) / 20.;
                    if (sigma_cells > 0 && strategy == ORC_INIT_MULTIPLE_GAUSS) {
                        sx = sigma_cells * p->dx;
                        sy = sigma_cells * p->dy;
                        sz = sigma_cells * p->dz;
                    }
                    for (int m = 0; m < nmax; m++) {
                        double xc = x - cx[m], yc = y - cy[m];
                        double zc = (strategy == ORC_INIT_SETUP_GAUSS) ? z - cz[m] : z;
                        val += exp(-0.5 * (xc * xc) / (sx * sx)) * exp(-0.5 * (yc * yc) / (sy * sy)) *
                               exp(-0.5 * (zc * zc) / (sz * sz));
                    }
                }
                size_t id = ((size_t)i * ny + j) * nz + k;
                psi[2 * id] = val;
                psi[2 * id + 1] = 0.0;
            }
    free(cx);
    free(cy);
    free(cz);
    return 0;
}

/* ------------------------------------------------------------------- FFT */
static int g_threads = 1;
void orc_set_threads(int n) {
    g_threads = n > 0 ? n : 1;
#ifdef _OPENMP
    omp_set_num_threads(g_threads);
#endif
}
int orc_get_threads(void) { return g_threads; }

typedef struct {
    int n;
    int nstage;
    int radix[64];
    cplx *tw_fwd; /* exp(-2 pi i m / n) */
    cplx *tw_bwd; /* exp(+2 pi i m / n) */
} fft_plan;

#define PLAN_CACHE 32
static fft_plan g_plans[PLAN_CACHE];
static int g_nplans = 0;

static fft_plan *get_plan(int n) {
    for (int i = 0; i < g_nplans; i++)
        if (g_plans[i].n == n) return &g_plans[i];
    fft_plan *pl = NULL;
#pragma omp critical(orc_plan)
    {
        for (int i = 0; i < g_nplans; i++)
            if (g_plans[i].n == n) pl = &g_plans[i];
        if (!pl) {
            if (g_nplans == PLAN_CACHE) abort();
            pl = &g_plans[g_nplans];
            pl->n = n;
            int rem = n, ns = 0;
            while (rem % 4 == 0) { pl->radix[ns++] = 4; rem /= 4; }
            while (rem % 2 == 0) { pl->radix[ns++] = 2; rem /= 2; }
            while (rem % 3 == 0) { pl->radix[ns++] = 3; rem /= 3; }
            while (rem % 5 == 0) { pl->radix[ns++] = 5; rem /= 5; }
            for (int f = 7; rem > 1; f += 2)
                while (rem % f == 0) { pl->radix[ns++] = f; rem /= f; }
            pl->nstage = ns;
            pl->tw_fwd = malloc(sizeof(cplx) * n);
            pl->tw_bwd = malloc(sizeof(cplx) * n);
            for (int m = 0; m < n; m++) {
                double a = 2.0 * M_PI * (double)m / (double)n;
                pl->tw_fwd[m] = cos(a) - I * sin(a);
                pl->tw_bwd[m] = cos(a) + I * sin(a);
            }
            __atomic_store_n(&g_nplans, g_nplans + 1, __ATOMIC_RELEASE);
        }
    }
    return pl;
}

/* In-register DFT of length r with sign s (-1 forward). */
static void small_dft(cplx *v, int r, int s, const cplx *tw, int n) {
    if (r == 2) {
        cplx a = v[0], b = v[1];
        v[0] = a + b;
        v[1] = a - b;
    } else if (r == 4) {
        cplx a = v[0] + v[2], b = v[0] - v[2], c = v[1] + v[3], d = (v[1] - v[3]) * (s < 0 ? -I : I);
        v[0] = a + c;
        v[2] = a - c;
        v[1] = b + d;
        v[3] = b - d;
    } else {
        cplx out[64];
        for (int q = 0; q < r; q++) {
            cplx acc = 0;
            for (int t = 0; t < r; t++) acc += v[t] * tw[(size_t)((long)q * t % r) * (n / r)];
            out[q] = acc;
        }
        for (int q = 0; q < r; q++) v[q] = out[q];
    }
}

/* Stockham autosort FFT of one contiguous line, work buffer w of length n. */
static void fft_line(cplx *x, cplx *w, int n, int sign) {
    if (n <= 1) return;
    fft_plan *pl = get_plan(n);
    const cplx *tw = sign < 0 ? pl->tw_fwd : pl->tw_bwd;
    cplx *src = x, *dst = w;
    int ns = 1;
    for (int st = 0; st < pl->nstage; st++) {
        int r = pl->radix[st];
        int stride = n / r;
        for (int j = 0; j < stride; j++) {
            int k = j % ns;
            cplx v[64];
            for (int q = 0; q < r; q++) {
                cplx a = src[j + q * stride];
                if (k) a *= tw[(size_t)((long)q * k % (ns * r)) * (n / (ns * r))];
                v[q] = a;
            }
            small_dft(v, r, sign, tw, n);
            int base = (j / ns) * ns * r + k;
            for (int q = 0; q < r; q++) dst[base + q * ns] = v[q];
        }
        ns *= r;
        cplx *t = src;
        src = dst;
        dst = t;
    }
    if (src != x) memcpy(x, src, sizeof(cplx) * n);
}

void orc_fft_1d(double *x, int n, int sign) {
    cplx *w = malloc(sizeof(cplx) * n);
    fft_line((cplx *)x, w, n, sign);
    free(w);
}

/* 3D transform on [nx][ny][nz] complex array: z lines, then y, then x. */
void orc_fft_3d(double *xd, int nx, int ny, int nz, int sign) {
    cplx *x = (cplx *)xd;
    get_plan(nx);
    get_plan(ny);
    get_plan(nz);
    /* z: contiguous lines */
#pragma omp parallel num_threads(g_threads)
    {
        cplx *w = malloc(sizeof(cplx) * nz);
#pragma omp for schedule(static)
        for (long l = 0; l < (long)nx * ny; l++) fft_line(x + l * nz, w, nz, sign);
        free(w);
    }
    /* y: for each (i, k) line, stride nz */
#pragma omp parallel num_threads(g_threads)
    {
        cplx *buf = malloc(sizeof(cplx) * ny);
        cplx *w = malloc(sizeof(cplx) * ny);
#pragma omp for schedule(static)
        for (long l = 0; l < (long)nx * nz; l++) {
            long i = l / nz, k = l % nz;
            cplx *base = x + i * (long)ny * nz + k;
            for (int j = 0; j < ny; j++) buf[j] = base[(long)j * nz];
            fft_line(buf, w, ny, sign);
            for (int j = 0; j < ny; j++) base[(long)j * nz] = buf[j];
        }
        free(buf);
        free(w);
    }
    /* x: stride ny*nz */
#pragma omp parallel num_threads(g_threads)
    {
        cplx *buf = malloc(sizeof(cplx) * nx);
        cplx *w = malloc(sizeof(cplx) * nx);
#pragma omp for schedule(static)
        for (long l = 0; l < (long)ny * nz; l++) {
            cplx *base = x + l;
            long st = (long)ny * nz;
            for (int i = 0; i < nx; i++) buf[i] = base[i * st];
            fft_line(buf, w, nx, sign);
            for (int i = 0; i < nx; i++) base[i * st] = buf[i];
        }
        free(buf);
        free(w);
    }
}

/* ---------------------------------------------------------------- solver */
struct orc_solver {
    orc_params p;
    size_t n;        /* nx*ny*nz */
    cplx *cpsi, *cpsii;
    double *fi3d, *pote;
    double xnorma;
    /* Poisson (fftw_poisson_solver.cpp): padded grid (2nx,2ny,2nz) */
    double *vdip;    /* real part of h_Vdip_k, [2nx][2ny][nz+1] */
    cplx *pad;       /* full complex padded grid */
    /* split solver (fftw_split_solver.cpp) */
    double *kin;     /* h_kinetic_factor [nx][ny][nz] */
    cplx *kbuf;
};

#define IDX(s, i, j, k) (((size_t)(i) * (s)->p.ny + (j)) * (s)->p.nz + (k))

/* src/solver/fft_solvers/cpu/fftw_poisson_solver.cpp:12-49 (prepare_containers). */
static void prepare_vdip(orc_solver *s) {
    int nx = 2 * s->p.nx, ny = 2 * s->p.ny, nz = 2 * s->p.nz;
    double dkx = 2. * M_PI / (nx * s->p.dx);
    double dky = 2. * M_PI / (ny * s->p.dy);
    double dkz = 2. * M_PI / (nz * s->p.dz);
    int nzh = nz / 2 + 1;
    for (int i = 0; i < nx; ++i) {
        double kx = (i <= nx / 2) ? i * dkx : (i - nx) * dkx;
        for (int j = 0; j < ny; ++j) {
            double ky = (j <= ny / 2) ? j * dky : (j - ny) * dky;
            for (int k = 0; k < nzh; ++k) {
                size_t idx = ((size_t)i * ny + j) * nzh + k;
                double kz = (k <= nz / 2) ? k * dkz : (k - nz) * dkz;
                double k2 = kx * kx + ky * ky + kz * kz;
                s->vdip[idx] = (k2 > 1e-30) ? (kz * kz / k2) : 0.0;
            }
        }
    }
}

/* src/solver/fft_solvers/cpu/fftw_split_solver.cpp:11-42 (prepare_containers). */
static void prepare_kinetic(orc_solver *s) {
    int nx = s->p.nx, ny = s->p.ny, nz = s->p.nz;
    double dt = s->p.real_time_dt;
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                double kx = (i <= nx / 2) ? (2.0 * M_PI * i / (nx * s->p.dx)) : (2.0 * M_PI * (i - nx) / (nx * s->p.dx));
                double ky = (j <= ny / 2) ? (2.0 * M_PI * j / (ny * s->p.dy)) : (2.0 * M_PI * (j - ny) / (ny * s->p.dy));
                double kz = (k <= nz / 2) ? (2.0 * M_PI * k / (nz * s->p.dz)) : (2.0 * M_PI * (k - nz) / (nz * s->p.dz));
                double k_sq = kx * kx + ky * ky + kz * kz;
                s->kin[((size_t)i * ny + j) * nz + k] = -dt / (2.0 * s->p.m) * k_sq;
            }
}

orc_solver *orc_solver_create(const orc_params *p) {
    orc_solver *s = calloc(1, sizeof(orc_solver));
    s->p = *p;
    s->n = (size_t)p->nx * p->ny * p->nz;
    s->cpsi = calloc(s->n, sizeof(cplx));
    s->cpsii = calloc(s->n, sizeof(cplx));
    s->fi3d = calloc(s->n, sizeof(double));
    s->pote = calloc(s->n, sizeof(double));
    size_t np = 8 * s->n;
    s->vdip = calloc((size_t)(2 * p->nx) * (2 * p->ny) * (p->nz + 1), sizeof(double));
    s->pad = calloc(np, sizeof(cplx));
    s->kin = calloc(s->n, sizeof(double));
    s->kbuf = calloc(s->n, sizeof(cplx));
    prepare_vdip(s);
    prepare_kinetic(s);
    /* solver.cpp:15-23: initialize() ends with calc_norm + normalize on the
     * zero buffers; the reference then produces NaN, which load_buffer later
     * overwrites.  Nothing observable depends on it, so it is skipped here. */
    return s;
}

void orc_solver_destroy(orc_solver *s) {
    if (!s) return;
    free(s->cpsi);
    free(s->cpsii);
    free(s->fi3d);
    free(s->pote);
    free(s->vdip);
    free(s->pad);
    free(s->kin);
    free(s->kbuf);
    free(s);
}

orc_params *orc_solver_params(orc_solver *s) { return &s->p; }
double *orc_psi(orc_solver *s) { return (double *)s->cpsi; }
double *orc_psii(orc_solver *s) { return (double *)s->cpsii; }
double *orc_fi3d(orc_solver *s) { return s->fi3d; }
double *orc_pote(orc_solver *s) { return s->pote; }
double orc_xnorma(orc_solver *s) { return s->xnorma; }
const double *orc_vdip_table(orc_solver *s) { return s->vdip; }
const double *orc_kinetic_table(orc_solver *s) { return s->kin; }

/* Re-derive tables after the caller edited parameters (e.g. dt). */
void orc_refresh_derived(orc_solver *s) {
    prepare_vdip(s);
    prepare_kinetic(s);
}

/* src/solver/solver.cpp:118-125 (load_buffer): cpsi = cpsii = wvf; norm; normalize. */
void orc_load_buffer(orc_solver *s, const double *psi) {
    memcpy(s->cpsi, psi, sizeof(cplx) * s->n);
    memcpy(s->cpsii, psi, sizeof(cplx) * s->n);
    orc_calc_norm(s);
    orc_normalize(s);
}

/* src/solver/solver.cpp:127-131 (load_pote). */
void orc_load_pote(orc_solver *s, const double *V) { memcpy(s->pote, V, sizeof(double) * s->n); }

/* src/solver/fft_solvers/cpu/fftw_poisson_solver.cpp:51-96 (execute):
 * rho = N|psi|^2 on the [0,n)^3 corner of the (2nx,2ny,2nz) grid, zero
 * elsewhere; R2C; multiply by Vdip; C2R; fi3d = corner / Np.
 * The R2C/C2R pair is done here as a complex transform of real data: with the
 * real, even multiplier the spectrum stays Hermitian, so the result equals the
 * real transform pair. */
void orc_calc_fi3d(orc_solver *s) {
    int nx = 2 * s->p.nx, ny = 2 * s->p.ny, nz = 2 * s->p.nz;
    size_t np = (size_t)nx * ny * nz;
    int nzh = nz / 2 + 1;
    memset(s->pad, 0, sizeof(cplx) * np);
    for (int i = 0; i < nx / 2; i++)
        for (int j = 0; j < ny / 2; j++)
            for (int k = 0; k < nz / 2; k++) {
                cplx v = s->cpsi[IDX(s, i, j, k)];
                double re = creal(v), im = cimag(v);
                s->pad[((size_t)i * ny + j) * nz + k] = (re * re + im * im) * s->p.n_atoms;
            }
    orc_fft_3d((double *)s->pad, nx, ny, nz, -1);
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                /* Vdip(k) for the upper half follows from evenness: the
                 * stored half-spectrum entry at (-i,-j,nz-k). */
                size_t idx = ((size_t)i * ny + j) * nz + k;
                double v;
                if (k < nzh)
                    v = s->vdip[((size_t)i * ny + j) * nzh + k];
                else
                    v = s->vdip[((size_t)((nx - i) % nx) * ny + (ny - j) % ny) * nzh + (nz - k)];
                s->pad[idx] *= v;
            }
    orc_fft_3d((double *)s->pad, nx, ny, nz, +1);
    double norm_factor = 1.0 / (double)np;
    for (int i = 0; i < nx / 2; i++)
        for (int j = 0; j < ny / 2; j++)
            for (int k = 0; k < nz / 2; k++)
                s->fi3d[IDX(s, i, j, k)] = creal(s->pad[((size_t)i * ny + j) * nz + k]) * norm_factor;
}

/* src/solver/cpu_solver/cpu_solver.cpp:19-34: xnorma = dV * sum over ALL points. */
void orc_calc_norm(orc_solver *s) {
    double acc = 0.0;
    for (size_t t = 0; t < s->n; t++) {
        double re = creal(s->cpsi[t]), im = cimag(s->cpsi[t]);
        acc += re * re + im * im;
    }
    s->xnorma = acc * s->p.dxdydz;
}

/* src/solver/cpu_solver/cpu_solver.cpp:36-48: psi /= sqrt(xnorma) over ALL points. */
void orc_normalize(orc_solver *s) {
    double r = sqrt(s->xnorma);
    for (size_t t = 0; t < s->n; t++) s->cpsi[t] /= r;
}

/* src/solver/cpu_solver/cpu_solver.cpp:50-73: interior explicit-Euler linear part. */
void orc_imag_iter_linear_step(orc_solver *s) {
    const int nx = s->p.nx, ny = s->p.ny, nz = s->p.nz;
    const double cx = -0.5 / (s->p.m * pow(s->p.dx, 2));
    const double cy = -0.5 / (s->p.m * pow(s->p.dy, 2));
    const double cz = -0.5 / (s->p.m * pow(s->p.dz, 2));
    for (int i = 1; i < nx - 1; i++)
        for (int j = 1; j < ny - 1; j++)
            for (int k = 1; k < nz - 1; k++) {
                cplx c = s->cpsi[IDX(s, i, j, k)];
                double v = s->pote[IDX(s, i, j, k)];
                cplx c1 = cx * (s->cpsi[IDX(s, i - 1, j, k)] + s->cpsi[IDX(s, i + 1, j, k)] - 2. * c) +
                          cy * (s->cpsi[IDX(s, i, j - 1, k)] + s->cpsi[IDX(s, i, j + 1, k)] - 2. * c) +
                          cz * (s->cpsi[IDX(s, i, j, k - 1)] + s->cpsi[IDX(s, i, j, k + 1)] - 2. * c) +
                          c * (v + s->p.cdd * s->fi3d[IDX(s, i, j, k)]);
                s->cpsii[IDX(s, i, j, k)] = c - s->p.imag_time_dt * c1;
            }
}

/* src/solver/cpu_solver/cpu_solver.cpp:75-97: interior nonlinear part from the
 * OLD psi, then cpsi = cpsii over the whole array (boundary taken from cpsii). */
void orc_imag_iter_nonlinear_step(orc_solver *s) {
    const int nx = s->p.nx, ny = s->p.ny, nz = s->p.nz;
    const double w = s->p.n_atoms;
    for (int i = 1; i < nx - 1; i++)
        for (int j = 1; j < ny - 1; j++)
            for (int k = 1; k < nz - 1; k++) {
                cplx c = s->cpsi[IDX(s, i, j, k)];
                double nrm = creal(c) * creal(c) + cimag(c) * cimag(c);
                double ab = cabs(c);
                s->cpsii[IDX(s, i, j, k)] =
                    s->cpsii[IDX(s, i, j, k)] -
                    s->p.imag_time_dt * ((s->p.ggp11 - s->p.cdd / 3) * nrm * c * w +
                                         s->p.gamma * (ab * ab * ab) * c * s->p.w_15);
            }
    memcpy(s->cpsi, s->cpsii, sizeof(cplx) * s->n);
}

/* src/solver/cpu_solver/cpu_solver.cpp:99-131: interior phase exp(-i dt/2 Veff). */
void orc_real_fft_potential_half_step(orc_solver *s) {
    const int nx = s->p.nx, ny = s->p.ny, nz = s->p.nz;
    const double w = s->p.n_atoms;
    const double dt_factor = s->p.real_time_dt / 2.;
    for (int i = 1; i < nx - 1; i++)
        for (int j = 1; j < ny - 1; j++)
            for (int k = 1; k < nz - 1; k++) {
                size_t id = IDX(s, i, j, k);
                cplx c = s->cpsi[id];
                double re = creal(c), im = cimag(c);
                double density = re * re + im * im;
                double ab = cabs(c);
                double v_int = (s->p.ggp11 - s->p.cdd / 3) * density * w + s->p.gamma * (ab * ab * ab) * s->p.w_15;
                double total = s->pote[id] + s->p.cdd * s->fi3d[id] + v_int;
                double sn = sin(-dt_factor * total), cs = cos(-dt_factor * total);
                s->cpsi[id] = (re * cs - im * sn) + I * (re * sn + im * cs);
            }
}

/* src/solver/fft_solvers/cpu/fftw_split_solver.cpp:44-86: FFT, multiply by
 * exp(i f_k), inverse FFT, divide by N (whole periodic grid). */
void orc_real_fft_kinetic_step(orc_solver *s) {
    const int nx = s->p.nx, ny = s->p.ny, nz = s->p.nz;
    memcpy(s->kbuf, s->cpsi, sizeof(cplx) * s->n);
    orc_fft_3d((double *)s->kbuf, nx, ny, nz, -1);
    for (size_t t = 0; t < s->n; t++) {
        double f = s->kin[t];
        s->kbuf[t] *= cos(f) + I * sin(f);
    }
    orc_fft_3d((double *)s->kbuf, nx, ny, nz, +1);
    double nf = 1.0 / (double)s->n;
    for (size_t t = 0; t < s->n; t++) s->cpsi[t] = s->kbuf[t] * nf;
}

/* src/solver/solver.cpp:99-105. */
void orc_imag_time_iter(orc_solver *s) {
    orc_calc_fi3d(s);
    orc_imag_iter_linear_step(s);
    orc_imag_iter_nonlinear_step(s);
    orc_calc_norm(s);
    orc_normalize(s);
}

/* src/solver/solver.cpp:107-112. */
void orc_real_time_iter(orc_solver *s) {
    orc_real_fft_potential_half_step(s);
    orc_calc_fi3d(s);
    orc_real_fft_kinetic_step(s);
    orc_real_fft_potential_half_step(s);
}

/* src/solver/solver.cpp:51-53 loop body without the I/O cadence. */
void orc_it_steps(orc_solver *s, long n) {
    for (long t = 0; t < n; t++) orc_imag_time_iter(s);
}

/* src/solver/solver.cpp:75-79: update_edd(iter) before each step when the
 * ramp is active. */
void orc_rt_steps(orc_solver *s, long first_iter, long n) {
    for (long t = 0; t < n; t++) {
        if (!s->p.const_edd) orc_update_edd(&s->p, first_iter + t);
        orc_real_time_iter(s);
    }
}

/* src/solver/cpu_solver/cpu_solver.cpp:137-194 (calc_energy), interior sums,
 * totals in hartree.  out = {e_kin, e_pot, e_int, e_ext, e_bmf, e_total, mu,
 * residual}.  mu and residual have no counterpart in the reference:
 *   H psi = -(1/2m) Lap_FD psi + (V + cdd Phi) psi + (g - cdd/3) N |psi|^2 psi
 *           + gamma N^1.5 |psi|^3 psi     (the operator of cpu_solver.cpp:50-97)
 *   mu = dV * sum_interior Re(conj(psi) H psi)
 *   residual = sqrt(dV * sum_interior |H psi - mu psi|^2). */
void orc_calc_energy(orc_solver *s, double out[8]) {
    const int nx = s->p.nx, ny = s->p.ny, nz = s->p.nz;
    const orc_params *p = &s->p;
    double e_kin = 0, e_pot = 0, e_int = 0, e_ext = 0, e_bmf = 0, mu = 0;
    const double cx = -0.5 / (p->m * pow(p->dx, 2));
    const double cy = -0.5 / (p->m * pow(p->dy, 2));
    const double cz = -0.5 / (p->m * pow(p->dz, 2));
    for (int i = 1; i < nx - 1; i++)
        for (int j = 1; j < ny - 1; j++)
            for (int k = 1; k < nz - 1; k++) {
                size_t id = IDX(s, i, j, k);
                cplx c = s->cpsi[id];
                cplx gx = -(s->cpsi[IDX(s, i + 1, j, k)] + s->cpsi[IDX(s, i - 1, j, k)] - 2. * c) / pow(p->dx, 2);
                cplx gy = -(s->cpsi[IDX(s, i, j + 1, k)] + s->cpsi[IDX(s, i, j - 1, k)] - 2. * c) / pow(p->dy, 2);
                cplx gz = -(s->cpsi[IDX(s, i, j, k + 1)] + s->cpsi[IDX(s, i, j, k - 1)] - 2. * c) / pow(p->dz, 2);
                double nrm = creal(c) * creal(c) + cimag(c) * cimag(c);
                double ab = sqrt(nrm);
                e_kin += creal((gx + gy + gz) * conj(c));
                e_pot += s->pote[id] * nrm;
                e_int += 0.5 * p->ggp11 * nrm * nrm;
                e_ext += 0.5 * p->cdd * s->fi3d[id] * nrm * p->n_atoms;
                e_ext -= p->cdd / 3. * pow(nrm, 2) / 2. * pow(p->n_atoms, 2);
                e_bmf += 2. / 5. * p->gamma * pow(ab, 5);
            }
    e_kin *= p->dxdydz / (2 * p->m) * p->n_atoms;
    e_pot *= p->dxdydz * p->n_atoms;
    e_int *= p->dxdydz * pow(p->n_atoms, 2);
    e_ext *= p->dxdydz;
    e_bmf *= p->dxdydz * pow(p->n_atoms, 2.5);
    double e_total = e_kin + e_pot + e_int + e_ext + e_bmf;
    /* mu = <psi|H|psi> and the residual, straight from the operator */
    double res = 0.0;
    for (int pass = 0; pass < 2; pass++) {
        double acc = 0.0;
        for (int i = 1; i < nx - 1; i++)
            for (int j = 1; j < ny - 1; j++)
                for (int k = 1; k < nz - 1; k++) {
                    size_t id = IDX(s, i, j, k);
                    cplx c = s->cpsi[id];
                    double nrm = creal(c) * creal(c) + cimag(c) * cimag(c);
                    double ab = sqrt(nrm);
                    cplx h = cx * (s->cpsi[IDX(s, i - 1, j, k)] + s->cpsi[IDX(s, i + 1, j, k)] - 2. * c) +
                             cy * (s->cpsi[IDX(s, i, j - 1, k)] + s->cpsi[IDX(s, i, j + 1, k)] - 2. * c) +
                             cz * (s->cpsi[IDX(s, i, j, k - 1)] + s->cpsi[IDX(s, i, j, k + 1)] - 2. * c) +
                             c * (s->pote[id] + p->cdd * s->fi3d[id]) +
                             c * ((p->ggp11 - p->cdd / 3) * nrm * p->n_atoms + p->gamma * ab * ab * ab * p->w_15);
                    if (pass == 0) {
                        acc += creal(conj(c) * h);
                    } else {
                        cplx r = h - mu * c;
                        acc += creal(r) * creal(r) + cimag(r) * cimag(r);
                    }
                }
        if (pass == 0)
            mu = acc * p->dxdydz;
        else
            res = sqrt(acc * p->dxdydz);
    }
    out[0] = e_kin;
    out[1] = e_pot;
    out[2] = e_int;
    out[3] = e_ext;
    out[4] = e_bmf;
    out[5] = e_total;
    out[6] = mu;
    out[7] = res;
}
