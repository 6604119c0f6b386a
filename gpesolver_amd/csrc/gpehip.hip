// gpehip.hip -- MI355X (gfx950) Gross-Pitaevskii solver backend: kernels + C ABI.
//
// Data layout in HBM (one solver instance, grid nx x ny x nz, N = nx*ny*nz):
//   store[2]  complex T [nx][ny][nz]   phi = sqrt(dV) * psi (ping-pong for IT)
//   psii      complex T [nx][ny][nz]   the reference's cpsii (its boundary is
//                                      what every IT step writes there)
//   V, fi3d   T [nx][ny][nz]           potential and dipolar potential (a.u.)
//   S1        complex T [2ny][nx][nz]  Poisson spectrum after the z (R2C) and
//                                      y passes, x still in real space, for
//                                      kz = 1..nz: the kz = 0 column of
//                                      rho^ * kz^2/k^2 is identically zero, so
//                                      it is neither stored nor transformed
//                                      (and rows are 2^a-aligned)
//   Tk        complex T [ny][nx][nz]   kinetic split-step work array
//   (the real-time plane pass ping-pongs S1 / Tk with a second pair)
// The field is stored scaled, phi = sqrt(dV) psi, so that sum |phi|^2 = the
// reference's xnorma and no FP32 value becomes subnormal (|psi|^5 ~ 1e-42 in
// a.u. would); the nonlinear couplings absorb the powers of dV.
//
// Kernels (T = float or double):
//   k_plane  one block (RT steady state: two) per x-plane i: finishes the dipolar Poisson solve of the
//            plane (inverse y FFT, z C2R), applies the IT stencil or the RT
//            potential phase, starts the next Poisson solve (z R2C, y FFT of
//            N|psi|^2) and runs the kinetic propagation along z and y.
//   k_xpass  one block per tile of x-lines: Poisson x FFT, x Vdd(k), inverse
//            x FFT on S1, and the kinetic propagation along x on Tk (the
//            register x pass of 128 x 128 x 32 and 256 x 64 x 16 grids also
//            runs the z propagation there, see kz_in_x).
// One IT or RT step is one k_xpass plus one k_plane launch.  The kinetic
// propagator exp(-i dt k^2/2m) = prod_ax exp(-i dt k_ax^2/2m) is separable, so
// the 3D split step is three 1D (FFT, phase, inverse FFT) passes; the Poisson
// multiplier kz^2/k^2 is not separable and keeps the 3-stage pipeline.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "../../include/gpehip.h"
#include "gpe_fft.hpp"
#include "gpe_fft_ct.hpp"
#include "gpe_sincos.hpp"

using gpe::cplx;
using gpe::LinePlan;

// Reciprocal: the hardware v_rcp_f32 (1 ulp) in FP32, a division in FP64.
__device__ __forceinline__ float rcp_t(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double rcp_t(double x) { return 1.0 / x; }

// Write-through store of one complex value, for the arrays one pass of a
// step writes and the next pass reads (S1, Tk, the IT state, the pencil
// intermediates): global_store with the sc0 sc1 cache policy writes the line
// through the XCD's L2 instead of leaving it dirty there.  The next pass reads
// it from blocks on other XCDs, whose L2s hold no copy anyway, and a launch
// ends by writing back its XCD's dirty lines; without them that write-back is
// short.  Measured at 128x128x32 real time: FP32 x pass 12.0 -> 9.4 us, plane
// pass 13.1 -> 12.8 us against nontemporal stores (plain stores: 14.5 us
// plane pass); FP64 x pass 16.8 -> 15.2 us.  A vector store of more than 8
// bytes needs one wait state before its data registers are rewritten, which
// the compiler cannot see through inline assembly: the 16-byte FP64 store
// carries its own s_nop.
template <typename T>
__device__ __forceinline__ void store_wt(cplx<T>* p, cplx<T> v) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    const v2 t = {v.x, v.y};
    if constexpr (sizeof(T) == 4)
        asm volatile("global_store_dwordx2 %0, %1, off sc0 sc1" ::"v"(p), "v"(t) : "memory");
    else
        asm volatile("global_store_dwordx4 %0, %1, off sc0 sc1\n\ts_nop 1" ::"v"(p), "v"(t) : "memory");
}

namespace {

// Threads of a plane block.  The compiled FP32 shapes run 16 waves per
// plane (more loads in flight per CU, and they fit in 128 VGPRs); FP64 and
// the run-time-size kernel run 8 waves, since they need more than 128 VGPRs.
// The batch launches the same kernels per grid shape.
template <typename T, bool COMPILED>
struct PlaneThreads {
    static constexpr int value = (sizeof(T) == 4 && COMPILED) ? 1024 : 512;
};
constexpr int kLineThreads = 256;
constexpr int kMaxPlanePts = 8192;  // ny*nz bound of the single-block plane path
constexpr int kLineMaxE = 4096;     // complex elements of one x-line tile

// Plane shapes (ny, nz) and x lengths with compile-time kernels.
#define GPE_PLANE_SHAPES(X) X(128, 32) X(64, 16) X(96, 24) X(16, 8) X(24, 12) X(12, 6)
#define GPE_LINE_SHAPES(X) X(128) X(64) X(96) X(16) X(24) X(32) X(256)

template <typename T>
struct Lim;
template <>
struct Lim<float> {
    static constexpr int maxe = 20224;  // 158 KiB of complex<float> (static LDS uses the rest)
};
template <>
struct Lim<double> {
    static constexpr int maxe = 10112;  // 158 KiB of complex<double>
};

// Mode bits of k_plane.
enum : int {
    M_PINV = 1,        // finish Poisson from S1 -> Phi (registers)
    M_PHI_LOAD = 2,    // Phi from the fi3d array
    M_PHI_STORE = 4,   // write Phi to the fi3d array
    M_IN_TK = 8,       // psi plane comes from Tk (else from `in`, canonical layout)
    M_PHASE_HALF = 16, // exp(-i dt/2 Veff) on the interior
    M_PHASE_FULL = 32, // two merged half steps (end of one step + start of next)
    M_RHO_FWD = 64,    // start Poisson of |psi|^2 -> S1
    M_KIN = 128,       // z,y kinetic propagation, result to Tk
    M_OUT = 256,       // write psi plane to `out` (canonical)
    M_IT = 512,        // imaginary-time stencil from `in` to `out`
    M_NORM_OUT = 1024, // per-plane sum |psi_out|^2 -> partials_out[i]
    M_NORM_IN = 2048,  // scale input by 1/sqrt(sum partials_in)
    M_NO_PSI_OUT = 4096, // IT: the stencil result is not written (Poisson-start block of a split pass)
};

template <typename T>
struct PlaneArgs {
    int nx, ny, nz, mode;
    int generic;  // 1: run-time-size kernels even for a compiled shape
    const cplx<T>* in;
    cplx<T>* out;
    const cplx<T>* psii;
    cplx<T>* tk;        // kinetic work array, read (M_IN_TK)
    cplx<T>* s1;        // Poisson spectrum, read (M_PINV)
    cplx<T>* tko;       // written by M_KIN (== tk unless the buffers ping-pong)
    cplx<T>* s1o;       // written by M_RHO_FWD (== s1 unless the buffers ping-pong)
    int split;          // 1: two blocks per plane (grid 2 nx), see k_plane
    int xcd;            // 1: contiguous plane ranges per XCD (IT, nx % 8 == 0), see k_plane
    int kz_skip;        // M_KIN leaves the z propagation to the x pass (LineArgs::kinz)
    const T* V;
    T* phi;
    const double* part_in;
    double* part_out;
    const cplx<T>* tw;
    int tw_n;           // entries of the twiddle table (staged in LDS by the compiled shapes)
    const cplx<T>* ez;  // (cos fz(k) - 1, sin fz(k)) of the z kinetic phase fz
    const cplx<T>* ey;  // (cos fy(j) - 1, sin fy(j))
    const cplx<T>* twly;  // lo parts of W_ny^m (FP32 transforms of psi)
    const cplx<T>* twlz;  // lo parts of W_nz^m
    LinePlan pz2, py2, pz, py;
    double pinv_scale;  // (N_atoms / dV) / (8 N)
    double cdd;
    // IT stencil (phi units)
    double dtau, cx, cy, cz, gq, gmq;
    // RT phase: theta = -hdt * (cv*(V + cdd*Phi) + pg*rho + pgm*rho^1.5)
    double hdt, cv, pg, pgm;
    // Optional per-step coupling table (scaled g', gamma' pairs) indexed by
    // *cbase + cidx; when set it overrides pg / pgm (the update_edd ramp).
    const double* coef;
    const int* cbase;
    int cidx;
    // pencil pass (planes beyond one block's LDS): per-plane intermediates in
    // HBM, [x][y][c] like psi, and the tile shapes
    cplx<T>* pen_gw;  // G (inverse y2 of S1), then W (z-propagated psi)
    cplx<T>* pen_h;   // H (z R2C of rho, kz = 1..nz)
    int pen_c;        // columns per column tile
    int pen_r;        // rows per row tile
    double* pen_part; // FP64 norm partial per (plane, row tile)
};

template <typename T>
struct LineArgs {
    int nx, ny, nz;
    int generic;
    int ntile_p;  // Poisson tiles (first blocks)
    int cp;       // kz per Poisson tile
    int ck;       // k per kinetic tile
    int tiles_kz;  // ceil(nz/cp): S1 columns kz = 1..nz
    int tiles_k;   // ceil(nz/ck)
    cplx<T>* s1;
    cplx<T>* tk;
    const cplx<T>* tw;
    const cplx<T>* ex;  // (cos fx(i) - 1, sin fx(i))
    const cplx<T>* twlx;  // lo parts of W_nx^m (FP32 transforms of psi)
    const double* kx2;  // padded lattice kx^2, [2nx]
    const double* ky2;  // padded lattice ky^2, [2ny]
    const double* kz2;  // padded lattice kz^2, [nz+1]
    LinePlan px2, px;
    const cplx<T>* ez;  // (cos fz(k) - 1, sin fz(k)) of the z kinetic phase fz
    int kinz;           // kinetic tiles also run the z propagation (one tile = one y row, all z)
};

// |c|^2 in FP64 from a value of precision T: exact products of FP32
// components and one FP64 rounding, so a large nonlinear phase (tens of
// radians per step) does not inherit the FP32 rounding of the density.
template <typename T>
__device__ __forceinline__ double dens64(cplx<T> c) {
    return (double)c.x * (double)c.x + (double)c.y * (double)c.y;
}

// Deterministic block sum of one double per thread (fixed tree).
template <int NT>
__device__ double block_sum(double v, double* red) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
        for (int w = 0; w < NT / 64; ++w) r += red[w];
        red[NT / 64] = r;
    }
    __syncthreads();
    r = red[NT / 64];
    __syncthreads();
    return r;
}

// Sum of the per-plane partials, same order in every block.
// Lane l of the first wave adds p[l], p[l+64], ... and a fixed shuffle tree
// combines the lanes.
// The first wave's loads of the first 64 kPartQ partials, issued together
// (reduce_partials adds them, then any further ones, in index order).
constexpr int kPartQ = 4;
__device__ __forceinline__ void load_partials(const double* p, int n, double (&v)[kPartQ]) {
#pragma unroll
    for (int u = 0; u < kPartQ; ++u) {
        v[u] = 0.0;
        if (threadIdx.x < 64 && (int)threadIdx.x + 64 * u < n) v[u] = p[threadIdx.x + 64 * u];
    }
}
__device__ double reduce_partials(const double* p, int n, const double (&v)[kPartQ], double* red) {
    if (threadIdx.x < 64) {
        double s = 0.0;
#pragma unroll
        for (int u = 0; u < kPartQ; ++u)
            if ((int)threadIdx.x + 64 * u < n) s += v[u];
        for (int t = threadIdx.x + 64 * kPartQ; t < n; t += 64) s += p[t];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
        if (threadIdx.x == 0) red[0] = s;
    }
    __syncthreads();
    double r = red[0];
    __syncthreads();
    return r;
}
__device__ double sum_partials(const double* p, int n, double* red) {
    double v[kPartQ];
    load_partials(p, n, v);
    return reduce_partials(p, n, v, red);
}

// ---------------------------------------------------------------------------
// Size policies: the kernels are written once; a policy supplies the grid
// sizes and the line FFTs, either at run time (any 2^a 3^b grid, generic
// radix-4/2/3 engine) or as compile-time constants (large-radix engine).
struct PlanePolGeneric {
    static constexpr bool kStage = false;  // tables stay in global memory
    static constexpr int kNY = 0, kNZ = 0;
    static constexpr bool kRegY = false, kRegZ = false, kRegY2 = false;
    template <class A>
    __device__ static int ny(const A& a) { return a.ny; }
    template <class A>
    __device__ static int nz(const A& a) { return a.nz; }
    template <int NT>
    static constexpr int ppt() { return kMaxPlanePts / NT; }
    template <typename T, int NT>
    static constexpr int rq() { return (Lim<T>::maxe / 2 + NT - 1) / NT; }
    template <int NT>
    static constexpr int rq2() { return 1; }  // unused: no staged S1 load
    template <typename T, int NT, int S>
    __device__ static void fft_y2(cplx<T>* b, int nl, int es, int ls, const PlaneArgs<T>& a, const cplx<T>* tw) {
        gpe::lds_fft<T, NT, Lim<T>::maxe>(b, a.py2, nl, es, ls, S, tw);
    }
    template <typename T, int NT, int S>
    __device__ static void fft_z2(cplx<T>* b, int nl, int es, int ls, const PlaneArgs<T>& a, const cplx<T>* tw) {
        gpe::lds_fft<T, NT, Lim<T>::maxe>(b, a.pz2, nl, es, ls, S, tw);
    }
    template <typename T, int NT, int S>
    __device__ static void fft_y(cplx<T>* b, int nl, int es, int ls, const PlaneArgs<T>& a, const cplx<T>* tw,
                                 const cplx<T>* twl) {
        gpe::lds_fft<T, NT, Lim<T>::maxe>(b, a.py, nl, es, ls, S, tw, twl);
    }
    template <typename T, int NT, int S>
    __device__ static void fft_z(cplx<T>* b, int nl, int es, int ls, const PlaneArgs<T>& a, const cplx<T>* tw,
                                 const cplx<T>* twl) {
        gpe::lds_fft<T, NT, Lim<T>::maxe>(b, a.pz, nl, es, ls, S, tw, twl);
    }
};

template <int NY, int NZ>
struct PlanePolCT {
    // twiddle / kinetic tables and the psi, V loads are staged at block start
    static constexpr bool kStage = true;
    static constexpr int kNY = NY, kNZ = NZ;
    // register-resident kinetic y pass (four-step, NY = 16 B with B | 16)
    static constexpr bool kRegY = (NY == 128 || NY == 64) && NY / 16 * NZ <= 512;
    // register-resident kinetic z pass (one or two lanes per row)
    static constexpr bool kRegZ = NZ <= 16 || NZ == 32 || NZ == 24;
    // register-resident y transforms of the Poisson stages
    static constexpr bool kRegY2 = (NY == 128 || NY == 64) && 2 * NY / 16 * NZ <= 512;
    template <class A>
    __device__ static constexpr int ny(const A&) { return NY; }
    template <class A>
    __device__ static constexpr int nz(const A&) { return NZ; }
    template <int NT>
    static constexpr int ppt() { return (NY * NZ + NT - 1) / NT; }
    template <typename T, int NT>
    static constexpr int rq() { return (NY * (NZ + 1) + NT - 1) / NT; }
    template <int NT>
    static constexpr int rq2() { return (2 * NY * NZ + NT - 1) / NT; }
    template <typename T, int NT, int S>
    __device__ static void fft_y2(cplx<T>* b, int nl, int es, int ls, const PlaneArgs<T>& a, const cplx<T>* tw) {
        gpe::ct::fft<T, 2 * NY, NT, NZ + 1, S>(b, nl, es, ls, tw + a.py2.tw);
    }
    template <typename T, int NT, int S>
    __device__ static void fft_z2(cplx<T>* b, int nl, int es, int ls, const PlaneArgs<T>& a, const cplx<T>* tw) {
        gpe::ct::fft<T, 2 * NZ, NT, NY, S>(b, nl, es, ls, tw + a.pz2.tw);
    }
    // transforms of psi: twiddles with their lo parts (twl), EX constants
    template <typename T, int NT, int S>
    __device__ static void fft_y(cplx<T>* b, int nl, int es, int ls, const PlaneArgs<T>& a, const cplx<T>* tw,
                                 const cplx<T>* twl) {
        gpe::ct::fft<T, NY, NT, NZ, S, true>(b, nl, es, ls, tw + a.py.tw, twl);
    }
    template <typename T, int NT, int S>
    __device__ static void fft_z(cplx<T>* b, int nl, int es, int ls, const PlaneArgs<T>& a, const cplx<T>* tw,
                                 const cplx<T>* twl) {
        gpe::ct::fft<T, NZ, NT, NY, S, true>(b, nl, es, ls, tw + a.pz.tw, twl);
    }
};

struct LinePolGeneric {
    template <class A>
    __device__ static int nx(const A& a) { return a.nx; }
    template <typename T, int NT, int S>
    __device__ static void fft_x2(cplx<T>* b, int nl, int es, int ls, const LineArgs<T>& a) {
        gpe::lds_fft<T, NT, kLineMaxE>(b, a.px2, nl, es, ls, S, a.tw);
    }
    template <typename T, int NT, int S>
    __device__ static void fft_x(cplx<T>* b, int nl, int es, int ls, const LineArgs<T>& a) {
        gpe::lds_fft<T, NT, kLineMaxE>(b, a.px, nl, es, ls, S, a.tw, a.twlx);
    }
};

template <int NX>
struct LinePolCT {
    template <class A>
    __device__ static constexpr int nx(const A&) { return NX; }
    template <typename T, int NT, int S>
    __device__ static void fft_x2(cplx<T>* b, int nl, int es, int ls, const LineArgs<T>& a) {
        gpe::ct::fft<T, 2 * NX, NT, kLineMaxE / (2 * NX), S>(b, nl, es, ls, a.tw + a.px2.tw);
    }
    template <typename T, int NT, int S>
    __device__ static void fft_x(cplx<T>* b, int nl, int es, int ls, const LineArgs<T>& a) {
        gpe::ct::fft<T, NX, NT, kLineMaxE / NX, S, true>(b, nl, es, ls, a.tw + a.px.tw, a.twlx);
    }
};

// ---------------------------------------------------------------------------
// Register-resident kinetic propagation inside the plane kernel (compiled
// shapes).  Each 1-D line runs forward DFT -> x phase table -> inverse DFT
// without LDS round trips between the stages.

// Value of the neighbouring lane (lane ^ 1) through a DPP quad permutation
// [1, 0, 3, 2]: a VALU operand modifier, no LDS round trip.
__device__ __forceinline__ float dpp_xor1(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));
}
__device__ __forceinline__ double dpp_xor1(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), 0xB1, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0xB1, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <typename T>
__device__ __forceinline__ cplx<T> shfl_xor1(cplx<T> v) {
    return gpe::mk(dpp_xor1(v.x), dpp_xor1(v.y));
}

// z: rows Q[j][0..NZ) (row stride qs), j < ny.  NZ <= 16: one thread per
// row.  NZ = 2A (A = 16 or 12): two lanes per row, lane h holding
// x[2a + h]; DFT_A over a, twiddle W_NZ^(h k1) (compile-time constants),
// and the radix-2 stage across the lane pair through shuffles (thread h then
// holds X[k1 + A h]); the inverse runs the same steps backwards.
// ez = (cos f - 1, sin f) of the z kinetic phase f.
template <typename T, int NZ, int NT>
__device__ __forceinline__ void kin_z_reg(cplx<T>* Q, int ny, int qs, const cplx<T>* ez) {
    if constexpr (NZ <= 16) {
        for (int j = threadIdx.x; j < ny; j += NT) {
            cplx<T> v[NZ];
#pragma unroll
            for (int k = 0; k < NZ; ++k) v[k] = Q[j * qs + k];
            gpe::ct::Dft<T, NZ, -1, true>::run(v);
#pragma unroll
            for (int k = 0; k < NZ; ++k) v[k] = gpe::kin_mul(v[k], ez[k], 1.0 / NZ);
            gpe::ct::Dft<T, NZ, +1, true>::run(v);
#pragma unroll
            for (int k = 0; k < NZ; ++k) Q[j * qs + k] = v[k];
        }
    } else {
        constexpr int A = NZ / 2;
        static_assert(NZ == 2 * A && (A == 16 || A == 12), "z length of the paired register pass");
        const int nthr = 2 * ny;
        for (int t0 = 0; t0 < nthr; t0 += NT) {
            if (t0 + (int)(threadIdx.x & ~63u) >= nthr) continue;  // whole waves only: shuffles below
            const int t = t0 + threadIdx.x;
            const bool act = t < nthr;
            const int j = act ? (t >> 1) : 0, h = t & 1;
            cplx<T> v[A];
#pragma unroll
            for (int q = 0; q < A; ++q) v[q] = act ? Q[j * qs + 2 * q + h] : gpe::mk(T(0), T(0));
            gpe::ct::Dft<T, A, -1, true>::run(v);
            if (h) {
#pragma unroll
                for (int k1 = 1; k1 < A; ++k1) v[k1] = gpe::ct::mul_w<T, NZ, -1, true>(v[k1], k1);
            }
#pragma unroll
            for (int k1 = 0; k1 < A; ++k1) {
                const cplx<T> o = shfl_xor1(v[k1]);
                v[k1] = h ? gpe::csub(o, v[k1]) : gpe::cadd(v[k1], o);
                v[k1] = gpe::kin_mul(v[k1], ez[k1 + A * h], 1.0 / NZ);
            }
#pragma unroll
            for (int k1 = 0; k1 < A; ++k1) {
                const cplx<T> o = shfl_xor1(v[k1]);
                v[k1] = h ? gpe::csub(o, v[k1]) : gpe::cadd(v[k1], o);
            }
            if (h) {
#pragma unroll
                for (int k1 = 1; k1 < A; ++k1) v[k1] = gpe::ct::mul_w<T, NZ, +1, true>(v[k1], k1);
            }
            gpe::ct::Dft<T, A, +1, true>::run(v);
            if (act) {
#pragma unroll
                for (int q = 0; q < A; ++q) Q[j * qs + 2 * q + h] = v[q];
            }
        }
    }
}

// a * tw[m] (SIGN -1) or a * conj(tw[m]) (SIGN +1) in a transform of psi:
// in FP32 with the lo part twl[m], one rounding of the exact product.
template <typename T, int SIGN>
__device__ __forceinline__ cplx<T> tw_psi(cplx<T> a, const cplx<T>* tw, const cplx<T>* twl, int m) {
    if constexpr (sizeof(T) == 4)
        return SIGN < 0 ? gpe::cmul2(a, tw[m], twl[m]) : gpe::cmulc2(a, tw[m], twl[m]);
    else
        return SIGN < 0 ? gpe::cmul(a, tw[m]) : gpe::cmulc(a, tw[m]);
}

// y: columns c < NZ of Q[j][c] (row stride qs), length NY = 16 B (B = 4 or
// 8), four-step with one LDS exchange each way: thread (b, c) holds rows
// j = B a + b; DFT_16 over a, twiddle W_NY^(b k1), exchange, DFT_B over b for
// its rows k1 = b + B r, y phase ey, inverse DFT_B, conjugate twiddle,
// exchange back, inverse DFT_16; the result goes straight from registers
// to Tk[j][i][c] (global).  The exchange area aliases Q.  Every thread of the
// block reaches the barriers.
template <typename T, int NY, int NZ, int NT>
__device__ __forceinline__ void kin_y_reg_store(cplx<T>* Q, int qs, const cplx<T>* ey, const cplx<T>* twy,
                                                const cplx<T>* twyl, cplx<T>* tk_plane, int nx) {
    constexpr int A = 16, B = NY / A, R = A / B;
    static_assert(NY == A * B && A % B == 0 && B * NZ <= NT, "y length of the register pass");
    constexpr int ES = NZ + 1;  // exchange row stride (odd: distinct banks)
    const bool act = threadIdx.x < B * NZ;
    const int c = threadIdx.x % NZ, b = threadIdx.x / NZ;
    cplx<T> v[A];
    if (act) {
#pragma unroll
        for (int q = 0; q < A; ++q) v[q] = Q[(B * q + b) * qs + c];
        gpe::ct::Dft<T, A, -1, true>::run(v);
#pragma unroll
        for (int k1 = 1; k1 < A; ++k1) v[k1] = tw_psi<T, -1>(v[k1], twy, twyl, (b * k1) % NY);
    }
    __syncthreads();  // Q is read: the exchange area may overwrite it
    cplx<T>* E = Q;
    if (act) {
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) E[(k1 * B + b) * ES + c] = v[k1];
    }
    __syncthreads();
    cplx<T> w[R][B];
    if (act) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int k1 = b + B * r;
#pragma unroll
            for (int bb = 0; bb < B; ++bb) w[r][bb] = E[(k1 * B + bb) * ES + c];
            gpe::ct::Dft<T, B, -1, true>::run(w[r]);
#pragma unroll
            for (int k2 = 0; k2 < B; ++k2) w[r][k2] = gpe::kin_mul(w[r][k2], ey[k1 + A * k2], 1.0 / NY);
            gpe::ct::Dft<T, B, +1, true>::run(w[r]);
#pragma unroll
            for (int bb = 1; bb < B; ++bb) w[r][bb] = tw_psi<T, +1>(w[r][bb], twy, twyl, (bb * k1) % NY);
        }
    }
    __syncthreads();
    if (act) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int k1 = b + B * r;
#pragma unroll
            for (int bb = 0; bb < B; ++bb) E[(k1 * B + bb) * ES + c] = w[r][bb];
        }
    }
    __syncthreads();
    if (act) {
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) v[k1] = E[(k1 * B + b) * ES + c];
        gpe::ct::Dft<T, A, +1, true>::run(v);
#pragma unroll
        for (int q = 0; q < A; ++q) store_wt(&tk_plane[(size_t)(B * q + b) * nx * NZ + c], v[q]);
    }
}

// y propagation on twice the threads of kin_y_reg_store: A = 8 values per
// thread, B = NY / 8 threads per column.  Thread (b, c) holds rows
// j = B a + b: DFT_8 over a, twiddle W_NY^(b k1), LDS exchange.  The DFT_B
// over b of row k1 then runs on a lane pair (lane bit h, holding
// b = 2 m + h): DFT_B/2 over m, twiddle W_B^(h k), and the radix-2 stage
// across the pair through DPP, so lane h holds X[k1 + 8 (k + h B/2)]; the y
// phase ey, then the same steps backwards and the result straight from
// registers to Tk[j][i][c].  The exchange area aliases Q.  Every thread of
// the block reaches the barriers.
template <typename T, int NY, int NZ, int NT>
__device__ __forceinline__ void kin_y_pair_store(cplx<T>* Q, int qs, const cplx<T>* ey, const cplx<T>* twy,
                                                 const cplx<T>* twyl, cplx<T>* tk_plane, int nx) {
    constexpr int A = 8, B = NY / A, H = B / 2, R = A / H, NTH = B * NZ;
    static_assert(NY == A * B && B % 2 == 0 && A % H == 0 && NTH <= NT && NTH % 64 == 0,
                  "y length of the paired register pass");
    constexpr int ES = NZ + 1;  // exchange row stride (odd: distinct banks)
    const bool act = threadIdx.x < NTH;  // whole waves: the DPP stage needs both lanes of a pair
    const int c = threadIdx.x % NZ, b = threadIdx.x / NZ;
    cplx<T> v[A];
    if (act) {
#pragma unroll
        for (int q = 0; q < A; ++q) v[q] = Q[(B * q + b) * qs + c];
        gpe::ct::Dft<T, A, -1, true>::run(v);
#pragma unroll
        for (int k1 = 1; k1 < A; ++k1) v[k1] = tw_psi<T, -1>(v[k1], twy, twyl, (b * k1) % NY);
    }
    __syncthreads();  // Q is read: the exchange area may overwrite it
    cplx<T>* E = Q;
    if (act) {
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) E[(k1 * B + b) * ES + c] = v[k1];
    }
    __syncthreads();
    // lane pair (h) of column c2 takes rows k1 = g + H r
    const int h = threadIdx.x & 1, u = threadIdx.x >> 1;
    const int c2 = u % NZ, g = u / NZ;
    cplx<T> w[R][H];
    if (act) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int k1 = g + H * r;
#pragma unroll
            for (int m = 0; m < H; ++m) w[r][m] = E[(k1 * B + 2 * m + h) * ES + c2];
            gpe::ct::Dft<T, H, -1, true>::run(w[r]);
            if (h) {
#pragma unroll
                for (int k = 1; k < H; ++k) w[r][k] = gpe::ct::mul_w<T, B, -1, true>(w[r][k], k);
            }
#pragma unroll
            for (int k = 0; k < H; ++k) {
                const cplx<T> o = shfl_xor1(w[r][k]);
                w[r][k] = h ? gpe::csub(o, w[r][k]) : gpe::cadd(w[r][k], o);
                w[r][k] = gpe::kin_mul(w[r][k], ey[k1 + A * (k + H * h)], 1.0 / NY);
            }
#pragma unroll
            for (int k = 0; k < H; ++k) {
                const cplx<T> o = shfl_xor1(w[r][k]);
                w[r][k] = h ? gpe::csub(o, w[r][k]) : gpe::cadd(w[r][k], o);
            }
            if (h) {
#pragma unroll
                for (int k = 1; k < H; ++k) w[r][k] = gpe::ct::mul_w<T, B, +1, true>(w[r][k], k);
            }
            gpe::ct::Dft<T, H, +1, true>::run(w[r]);
        }
    }
    __syncthreads();  // every read of E is done before it is rewritten
    if (act) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int k1 = g + H * r;
#pragma unroll
            for (int m = 0; m < H; ++m) E[(k1 * B + 2 * m + h) * ES + c2] = w[r][m];
        }
    }
    __syncthreads();
    if (act) {
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) v[k1] = E[(k1 * B + b) * ES + c];
#pragma unroll
        for (int k1 = 1; k1 < A; ++k1) v[k1] = tw_psi<T, +1>(v[k1], twy, twyl, (b * k1) % NY);
        gpe::ct::Dft<T, A, +1, true>::run(v);
#pragma unroll
        for (int q = 0; q < A; ++q) store_wt(&tk_plane[(size_t)(B * q + b) * nx * NZ + c], v[q]);
    }
}

// ---------------------------------------------------------------------------
// The 2NY-point y transforms of the Poisson stages in registers (compiled
// shapes with NY = 64 or 128): four-step with A = 16 values per thread and
// B = 2NY/16 threads per kz column (kz = c + 1, c < NZ), one LDS exchange
// (aliasing the plane tile).  Thread (b, c) holds y = B a + b on input and
// rows k1 = b + B r (r < 16/B), outputs k1 + 16 k2, after the exchange.
template <int NY, int NZ>
struct Y2Reg {
    static constexpr int L = 2 * NY, A = 16, B = L / A, R = A / B, ES = NZ + 1, NTH = B * NZ;
    static_assert(L == A * B && A % B == 0 && NY == 8 * B, "y length of the register Poisson pass");
};

// Inverse y of S1 column c (v[a] = S1[B a + b][i][c], loaded by the caller),
// then the z C2R input rows of the packed pairs (j, j + NY/2) straight from
// registers: P[j][kz] = X_j + i X_j+NY/2 and its Hermitian mirror at 2NZ - kz
// (real DC and Nyquist terms; the kz = 0 column is zero).  tw2 = W_2NY^m.
// The caller places a barrier before the z transform.
template <typename T, int NY, int NZ>
__device__ __forceinline__ void y2_inv_reg(cplx<T>* v, cplx<T>* P, int zs, const cplx<T>* tw2) {
    using Y = Y2Reg<NY, NZ>;
    const bool act = threadIdx.x < Y::NTH;
    const int c = threadIdx.x % NZ, b = threadIdx.x / NZ;
    __syncthreads();  // the staged twiddles are published
    if (act) {
        gpe::ct::Dft<T, Y::A, +1>::run(v);
#pragma unroll
        for (int k1 = 1; k1 < Y::A; ++k1) v[k1] = gpe::cmulc(v[k1], tw2[(b * k1) % Y::L]);
#pragma unroll
        for (int k1 = 0; k1 < Y::A; ++k1) P[(k1 * Y::B + b) * Y::ES + c] = v[k1];
    }
    __syncthreads();
    cplx<T> w[Y::R][Y::B];
    if (act) {
#pragma unroll
        for (int r = 0; r < Y::R; ++r) {
            const int k1 = b + Y::B * r;
#pragma unroll
            for (int bb = 0; bb < Y::B; ++bb) w[r][bb] = P[(k1 * Y::B + bb) * Y::ES + c];
            gpe::ct::Dft<T, Y::B, +1>::run(w[r]);
        }
    }
    __syncthreads();  // every exchange read is done before the packed rows overwrite it
    if (act) {
        const int kz = c + 1;
#pragma unroll
        for (int r = 0; r < Y::R; ++r) {
            const int k1 = b + Y::B * r;
#pragma unroll
            for (int k2 = 0; k2 < Y::B / 4; ++k2) {
                const int j = k1 + 16 * k2;  // j < NY/2; its partner j + NY/2 is output k2 + B/4
                const cplx<T> xa = w[r][k2], xb = w[r][k2 + Y::B / 4];
                if (kz == NZ) {
                    P[j * zs + NZ] = gpe::mk(xa.x, xb.x);
                } else {
                    P[j * zs + kz] = gpe::mk(xa.x - xb.y, xa.y + xb.x);
                    P[j * zs + (2 * NZ - kz)] = gpe::mk(xa.x + xb.y, xb.x - xa.y);
                }
                if (c == 0) P[j * zs] = gpe::mk(T(0), T(0));
            }
        }
    }
}

// Forward y after the packed z R2C (P rows m < NY/2 hold rho_m + i rho_m+NY/2
// transformed along z): separates the two spectra per element, transforms
// the column (rows y >= NY are the zero padding) and stores S1 of plane i
// straight from registers (s1_plane = S1 + i NZ, row stride nx NZ).
template <typename T, int NY, int NZ>
__device__ __forceinline__ void y2_fwd_reg_store(cplx<T>* P, int zs, const cplx<T>* tw2, cplx<T>* s1_plane, int nx) {
    using Y = Y2Reg<NY, NZ>;
    const bool act = threadIdx.x < Y::NTH;
    const int c = threadIdx.x % NZ, b = threadIdx.x / NZ;
    const int kz = c + 1;
    cplx<T> v[Y::A];
    if (act) {
#pragma unroll
        for (int a = 0; a < Y::A; ++a) {
            v[a] = gpe::mk(T(0), T(0));
            if (a < 8) {  // y = B a + b < NY
                const int y = Y::B * a + b;
                const int m = a < 4 ? y : y - NY / 2;
                const cplx<T> zk = P[m * zs + kz];
                const cplx<T> zm = P[m * zs + (2 * NZ - kz)];
                v[a] = a < 4 ? gpe::mk(T(0.5) * (zk.x + zm.x), T(0.5) * (zk.y - zm.y))
                             : gpe::mk(T(0.5) * (zk.y + zm.y), T(0.5) * (zm.x - zk.x));
            }
        }
        gpe::ct::Dft<T, Y::A, -1>::run(v);
#pragma unroll
        for (int k1 = 1; k1 < Y::A; ++k1) v[k1] = gpe::cmul(v[k1], tw2[(b * k1) % Y::L]);
    }
    __syncthreads();  // the z spectra are read: the exchange may overwrite them
    if (act) {
#pragma unroll
        for (int k1 = 0; k1 < Y::A; ++k1) P[(k1 * Y::B + b) * Y::ES + c] = v[k1];
    }
    __syncthreads();
    if (act) {
#pragma unroll
        for (int r = 0; r < Y::R; ++r) {
            const int k1 = b + Y::B * r;
            cplx<T> w[Y::B];
#pragma unroll
            for (int bb = 0; bb < Y::B; ++bb) w[bb] = P[(k1 * Y::B + bb) * Y::ES + c];
            gpe::ct::Dft<T, Y::B, -1>::run(w);
#pragma unroll
            for (int k2 = 0; k2 < Y::B; ++k2) store_wt(&s1_plane[(size_t)(k1 + 16 * k2) * nx * NZ + c], w[k2]);
        }
    }
}

// ---------------------------------------------------------------------------
// Plane kernel.  Dynamic LDS: the plane tile, max(2ny*(nz+1), ny*(2nz+1))
// complex T, followed for the compiled shapes by the staged tables: the
// twiddle table (tw_n entries), ez (nz), ey (ny) and the lo parts of the
// W_ny (ny) and W_nz (nz) twiddles.
template <typename T, int NT, class Pol>
__device__ __forceinline__ void plane_body(const PlaneArgs<T>& a, const int i, const int mode, unsigned char* lds_raw,
                                           double* red) {
    constexpr int PPT = Pol::template ppt<NT>();  // points per thread
    constexpr int RQ = Pol::template rq<T, NT>();
    constexpr bool kStage = Pol::kStage;
    constexpr int PRE = kStage ? PPT : 1;
    cplx<T>* P = reinterpret_cast<cplx<T>*>(lds_raw);
    const int nx = a.nx, ny = Pol::ny(a), nz = Pol::nz(a);
    const int nzh = nz + 1, nz2 = 2 * nz, ny2 = 2 * ny;
    // LDS row strides padded to an odd number of complex elements so that
    // consecutive rows start on different banks
    const int zs = nz2 + 1;  // z-line rows of the R2C/C2R stage
    const int qs = nz + 1;   // psi plane rows
    const int npl = ny * nz;
    const size_t plane_off = (size_t)i * npl;
    const bool is_it = (mode & M_IT) != 0;
    const bool have_psi = (mode & (M_IT | M_PHASE_HALF | M_PHASE_FULL | M_KIN | M_OUT | M_RHO_FWD)) != 0;
    const bool need_v = (mode & (M_IT | M_PHASE_HALF | M_PHASE_FULL)) != 0;

    // ---- 0. staging (compiled shapes): the loads of this plane's S1
    // column block, psi and V are issued first, then the tables go to LDS
    // while they are in flight.  Every FFT below follows a barrier, which
    // publishes the tables.
    const cplx<T>* tw = a.tw;
    const cplx<T>* ez = a.ez;
    const cplx<T>* ey = a.ey;
    const cplx<T>* twly = a.twly;
    const cplx<T>* twlz = a.twlz;
    constexpr bool kRegY2 = Pol::kRegY2;
    constexpr int R2 = (kStage && !kRegY2) ? Pol::template rq2<NT>() : 1;
    cplx<T> st[R2];        // S1 plane (Poisson finish input), element order
    cplx<T> sy[kRegY2 ? 16 : 1];  // S1 column of the register y pass: S1[B a + b][i][c]
    cplx<T> pre[PRE];
    T vpre[PRE];
    // tables staged to LDS: every load of this thread is issued before its
    // first LDS store, so they cost one memory round trip (a load / store loop
    // per table would wait for each table in turn)
    constexpr int TQ = kStage ? 2 : 1;  // twiddles per thread in that round trip
    cplx<T> tq[TQ], kq[4];
    if constexpr (kStage) {
        // psi and V first: the phase terms without Phi are formed from them
        // while the S1 loads are still in flight
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
            const int p = threadIdx.x + q * NT;
            pre[q] = gpe::mk(T(0), T(0));
            vpre[q] = T(0);
            if (p < npl) {
                if (have_psi) {
                    if (!is_it && (mode & M_IN_TK)) {
                        const int j = p / nz, k = p - j * nz;
                        pre[q] = a.tk[((size_t)j * nx + i) * nz + k];
                    } else {
                        pre[q] = a.in[plane_off + p];
                    }
                }
                if (need_v) vpre[q] = a.V[plane_off + p];
            }
        }
        if (mode & M_PINV) {
            if constexpr (kRegY2) {
                constexpr int B = Y2Reg<Pol::kNY, Pol::kNZ>::B;
                if (threadIdx.x < Y2Reg<Pol::kNY, Pol::kNZ>::NTH) {
                    const int c = threadIdx.x % nz, b = threadIdx.x / nz;
#pragma unroll
                    for (int q = 0; q < 16; ++q) sy[q] = a.s1[((size_t)(B * q + b) * nx + i) * nz + c];
                }
            } else {
#pragma unroll
                for (int q = 0; q < R2; ++q) {
                    const int e = threadIdx.x + q * NT;
                    if (e < ny2 * nz) st[q] = a.s1[((size_t)(e / nz) * nx + i) * nz + e % nz];
                }
            }
        }
        static_assert(Pol::kNY <= NT && Pol::kNZ <= NT, "one kinetic table element per thread");
        const int t = threadIdx.x;
        const bool kin = (mode & M_KIN) != 0;
#pragma unroll
        for (int q = 0; q < TQ; ++q) {
            tq[q] = gpe::mk(T(0), T(0));
            if (t + q * NT < a.tw_n) tq[q] = a.tw[t + q * NT];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) kq[q] = gpe::mk(T(0), T(0));
        if (kin) {
            if (t < nz) {
                kq[0] = a.ez[t];
                kq[3] = a.twlz[t];
            }
            if (t < ny) {
                kq[1] = a.ey[t];
                kq[2] = a.twly[t];
            }
        }
    }

    // partial norms of the previous pass, loaded with the staging loads
    double pn[kPartQ];
    if (mode & M_NORM_IN) load_partials(a.part_in, nx, pn);

    // phase couplings of this step, issued while the staging loads are in
    // flight (the table holds one step more than is read, so the four values
    // load unconditionally; with a device step base it is a dependent pair)
    double pg = a.pg, pgm = a.pgm;
    if (a.coef && (mode & (M_PHASE_HALF | M_PHASE_FULL))) {
        const int ts = (a.cbase ? *a.cbase : 0) + a.cidx;
        const double* cp = a.coef + 2 * ts;
        const double c0 = cp[0], c1 = cp[1], c2 = cp[2], c3 = cp[3];
        const bool full = (mode & M_PHASE_FULL) != 0;
        pg = full ? c0 + c2 : c0;
        pgm = full ? c1 + c3 : c1;
    }

    // The phase argument without its Phi term, from the staged psi and V
    // (the plane as the phase reads it: Tk, or an input without the norm
    // scale), formed while the S1 loads are in flight.  FP32: tv V + tg |psi|^2
    // + tgm |psi|^3 of the small-phase path, which adds tphi Phi.  FP64:
    // cv V + pg |psi|^2 + pgm |psi|^3, to which the phase adds cv cdd Phi.
    const bool th0_on = kStage && (mode & (M_PHASE_HALF | M_PHASE_FULL)) &&
                        ((mode & M_IN_TK) || !(mode & M_NORM_IN));
    T th0[PRE];
    if constexpr (kStage) {
        if (th0_on) {
            if constexpr (sizeof(T) == 4) {
                const float tv = (float)(a.hdt * a.cv), tg = (float)(a.hdt * pg), tgm = (float)(a.hdt * pgm);
#pragma unroll
                for (int q = 0; q < PRE; ++q) {
                    const float nrm = pre[q].x * pre[q].x + pre[q].y * pre[q].y;
                    th0[q] = tv * vpre[q] + tg * nrm + tgm * (nrm * sqrtf(nrm));
                }
            } else {
#pragma unroll
                for (int q = 0; q < PRE; ++q) {
                    const double dn = dens64(pre[q]);
                    th0[q] = a.cv * vpre[q] + pg * dn + pgm * (dn * sqrt(dn));
                }
            }
        }
    }

    if constexpr (kStage) {
        const int t = threadIdx.x;
        const bool kin = (mode & M_KIN) != 0;
        cplx<T>* TW = P + max(ny2 * nzh, ny * zs);
#pragma unroll
        for (int q = 0; q < TQ; ++q)
            if (t + q * NT < a.tw_n) TW[t + q * NT] = tq[q];
        for (int m = t + TQ * NT; m < a.tw_n; m += NT) TW[m] = a.tw[m];  // long x lines only
        tw = TW;
        if (kin) {
            cplx<T>* EZ = TW + a.tw_n;
            cplx<T>* EY = EZ + nz;
            cplx<T>* TWLY = EY + ny;
            cplx<T>* TWLZ = TWLY + ny;
            if (t < nz) {
                EZ[t] = kq[0];
                TWLZ[t] = kq[3];
            }
            if (t < ny) {
                EY[t] = kq[1];
                TWLY[t] = kq[2];
            }
            ez = EZ;
            ey = EY;
            twly = TWLY;
            twlz = TWLZ;
        }
    }

    T phi_r[PPT];
    T rho_r[PPT];

    double s_in = 1.0;
    if (mode & M_NORM_IN) s_in = 1.0 / sqrt(reduce_partials(a.part_in, nx, pn, red));

    // ---- 1. Phi for this plane -------------------------------------------
    if ((mode & M_PINV) && kRegY2) {
        const T sc = (T)(a.pinv_scale * s_in * s_in);
        if constexpr (kRegY2) {
            y2_inv_reg<T, Pol::kNY, Pol::kNZ>(sy, P, zs, tw + a.py2.tw);
            __syncthreads();
            Pol::template fft_z2<T, NT, +1>(P, ny / 2, 1, zs, a, tw);
#pragma unroll
            for (int q = 0; q < PPT; ++q) {
                const int p = threadIdx.x + q * NT;
                if (p < npl) {
                    const int j = p / nz, k = p - j * nz;
                    const int hj = j >= ny / 2;
                    const cplx<T> v = P[(j - hj * (ny / 2)) * zs + k];
                    phi_r[q] = (hj ? v.y : v.x) * sc;
                    if (mode & M_PHI_STORE) a.phi[plane_off + p] = phi_r[q];
                }
            }
        }
        __syncthreads();
    } else if (mode & M_PINV) {
        // load S1[ky][i][kz-1] -> P[ky][kz], kz = 1..nz; the kz = 0 column
        // of rho^ Vdd is identically zero (Vdd = kz^2/k^2) and is not stored
        if constexpr (kStage) {
#pragma unroll
            for (int q = 0; q < R2; ++q) {
                const int e = threadIdx.x + q * NT;
                if (e < ny2 * nz) P[(e / nz) * nzh + e % nz + 1] = st[q];
            }
        } else {
            for (int e = threadIdx.x; e < ny2 * nz; e += NT) {
                const int ky = e / nz, c = e - ky * nz;
                P[ky * nzh + c + 1] = a.s1[((size_t)ky * nx + i) * nz + c];
            }
        }
        for (int ky = threadIdx.x; ky < ny2; ky += NT) P[ky * nzh] = gpe::mk(T(0), T(0));
        __syncthreads();
        Pol::template fft_y2<T, NT, +1>(P + 1, nz, nzh, 1, a, tw);
        const T sc = (T)(a.pinv_scale * s_in * s_in);
        if constexpr (kStage) {
            // z C2R, two rows per complex transform (ny is even for the
            // compiled shapes): row m carries X_m + i X_m+ny/2, each
            // Hermitian-extended with real DC and Nyquist terms, so the real
            // and imaginary parts of its inverse are the two real rows
            constexpr int RQH = (RQ + 1) / 2;
            const int nw = (ny / 2) * nzh;
            cplx<T> xa[RQH], xb[RQH];
#pragma unroll
            for (int q = 0; q < RQH; ++q) {
                const int w = threadIdx.x + q * NT;
                if (w < nw) {
                    const int m = w / nzh, kz = w - m * nzh;
                    xa[q] = P[m * nzh + kz];
                    xb[q] = P[(m + ny / 2) * nzh + kz];
                }
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < RQH; ++q) {
                const int w = threadIdx.x + q * NT;
                if (w < nw) {
                    const int m = w / nzh, kz = w - m * nzh;
                    if (kz == 0 || kz == nz) {
                        P[m * zs + kz] = gpe::mk(xa[q].x, xb[q].x);
                    } else {
                        P[m * zs + kz] = gpe::mk(xa[q].x - xb[q].y, xa[q].y + xb[q].x);
                        P[m * zs + (nz2 - kz)] = gpe::mk(xa[q].x + xb[q].y, xb[q].x - xa[q].y);
                    }
                }
            }
            __syncthreads();
            Pol::template fft_z2<T, NT, +1>(P, ny / 2, 1, zs, a, tw);
#pragma unroll
            for (int q = 0; q < PPT; ++q) {
                const int p = threadIdx.x + q * NT;
                if (p < npl) {
                    const int j = p / nz, k = p - j * nz;
                    const int hj = j >= ny / 2;
                    const cplx<T> v = P[(j - hj * (ny / 2)) * zs + k];
                    phi_r[q] = (hj ? v.y : v.x) * sc;
                    if (mode & M_PHI_STORE) a.phi[plane_off + p] = phi_r[q];
                }
            }
        } else {
            // rows j < ny -> Z[j][0..2nz) with Hermitian extension, via registers
            cplx<T> tmp[RQ];
#pragma unroll
            for (int q = 0; q < RQ; ++q) {
                const int e = threadIdx.x + q * NT;
                if (e < ny * nzh) tmp[q] = P[e];
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < RQ; ++q) {
                const int e = threadIdx.x + q * NT;
                if (e < ny * nzh) {
                    const int j = e / nzh, kz = e - j * nzh;
                    P[j * zs + kz] = tmp[q];
                    if (kz >= 1 && kz < nz) P[j * zs + (nz2 - kz)] = gpe::mk(tmp[q].x, -tmp[q].y);
                }
            }
            __syncthreads();
            Pol::template fft_z2<T, NT, +1>(P, ny, 1, zs, a, tw);
#pragma unroll
            for (int q = 0; q < PPT; ++q) {
                const int p = threadIdx.x + q * NT;
                if (p < npl) {
                    const int j = p / nz, k = p - j * nz;
                    phi_r[q] = P[j * zs + k].x * sc;
                    if (mode & M_PHI_STORE) a.phi[plane_off + p] = phi_r[q];
                }
            }
        }
        __syncthreads();
    } else if (mode & M_PHI_LOAD) {
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
            const int p = threadIdx.x + q * NT;
            if (p < npl) phi_r[q] = a.phi[plane_off + p];
        }
    }

    // ---- 2. psi update: IT stencil or RT phase ------------------------------
    double nacc = 0.0;
    cplx<T>* Q = P;  // aliases P: P is free here
    if (is_it) {
        const cplx<T>* in = a.in;
        if constexpr (kStage) {
#pragma unroll
            for (int q = 0; q < PPT; ++q) {
                const int p = threadIdx.x + q * NT;
                if (p < npl) Q[(p / nz) * qs + p % nz] = gpe::mk((T)(pre[q].x * s_in), (T)(pre[q].y * s_in));
            }
        } else {
            for (int p = threadIdx.x; p < npl; p += NT) {
                cplx<T> v = in[plane_off + p];
                Q[(p / nz) * qs + p % nz] = gpe::mk((T)(v.x * s_in), (T)(v.y * s_in));
            }
        }
        __syncthreads();
        const T sv = (T)s_in;
        const T dtau = (T)a.dtau, cx = (T)a.cx, cy = (T)a.cy, cz = (T)a.cz, gq = (T)a.gq, gmq = (T)a.gmq,
                cdd = (T)a.cdd;
        const bool iint = (i >= 1 && i <= nx - 2);
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
            const int p = threadIdx.x + q * NT;
            if (p < npl) {
                const int j = p / nz, k = p - j * nz;
                cplx<T> r;
                if (iint && j >= 1 && j <= ny - 2 && k >= 1 && k <= nz - 2) {
                    const int qi = j * qs + k;
                    const cplx<T> c = Q[qi];
                    const cplx<T> xm = in[plane_off - npl + p], xp = in[plane_off + npl + p];
                    const cplx<T> ym = Q[qi - qs], yp = Q[qi + qs], zm = Q[qi - 1], zp = Q[qi + 1];
                    const T v = (kStage ? vpre[q < PRE ? q : 0] : a.V[plane_off + p]) + cdd * phi_r[q];
                    cplx<T> c1;
                    c1.x = cx * (xm.x * sv + xp.x * sv - T(2) * c.x) + cy * (ym.x + yp.x - T(2) * c.x) +
                           cz * (zm.x + zp.x - T(2) * c.x) + c.x * v;
                    c1.y = cx * (xm.y * sv + xp.y * sv - T(2) * c.y) + cy * (ym.y + yp.y - T(2) * c.y) +
                           cz * (zm.y + zp.y - T(2) * c.y) + c.y * v;
                    const T nrm = c.x * c.x + c.y * c.y;
                    const T ab = sqrt(nrm);
                    const T f = gq * nrm + gmq * (ab * ab * ab);
                    r.x = c.x - dtau * c1.x;
                    r.y = c.y - dtau * c1.y;
                    r.x = r.x - dtau * (f * c.x);
                    r.y = r.y - dtau * (f * c.y);
                } else {
                    r = a.psii[plane_off + p];
                }
                if (!(mode & M_NO_PSI_OUT)) store_wt(&a.out[plane_off + p], r);
                const T n2 = r.x * r.x + r.y * r.y;
                rho_r[q] = n2;
                if (mode & M_NORM_OUT) nacc += (double)r.x * (double)r.x + (double)r.y * (double)r.y;
            }
        }
        __syncthreads();
    } else if (have_psi) {
        // psi plane into Q
        if constexpr (kStage) {
#pragma unroll
            for (int q = 0; q < PPT; ++q) {
                const int p = threadIdx.x + q * NT;
                if (p < npl) {
                    cplx<T> v = pre[q];
                    if (!(mode & M_IN_TK) && (mode & M_NORM_IN)) v = gpe::mk((T)(v.x * s_in), (T)(v.y * s_in));
                    Q[(p / nz) * qs + p % nz] = v;
                }
            }
        } else if (mode & M_IN_TK) {
            for (int p = threadIdx.x; p < npl; p += NT) {
                const int j = p / nz, k = p - j * nz;
                Q[j * qs + k] = a.tk[((size_t)j * nx + i) * nz + k];
            }
        } else {
            for (int p = threadIdx.x; p < npl; p += NT) {
                cplx<T> v = a.in[plane_off + p];
                if (mode & M_NORM_IN) v = gpe::mk((T)(v.x * s_in), (T)(v.y * s_in));
                Q[(p / nz) * qs + p % nz] = v;
            }
        }
        __syncthreads();
        const bool iint = (i >= 1 && i <= nx - 2);
        // FP32: the phase argument in FP32 where it is small (|theta| <= 1/16,
        // absolute error ~1e-9 rad, no reduction needed); larger phases take
        // the FP64 argument and quarter-turn reduction
        const float tv = (float)(a.hdt * a.cv), tphi = (float)(a.hdt * a.cv * a.cdd), tg = (float)(a.hdt * pg),
                    tgm = (float)(a.hdt * pgm);
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
            const int p = threadIdx.x + q * NT;
            if (p < npl) {
                const int qi = (p / nz) * qs + p % nz;
                cplx<T> c = Q[qi];
                T nrm = c.x * c.x + c.y * c.y;
                if (mode & (M_PHASE_HALF | M_PHASE_FULL)) {
                    const int j = p / nz, k = p - j * nz;
                    if (iint && j >= 1 && j <= ny - 2 && k >= 1 && k <= nz - 2) {
                        bool done = false;
                        if constexpr (sizeof(T) == 4) {
                            const float vv = kStage ? vpre[q < PRE ? q : 0] : a.V[plane_off + p];
                            const float t32 = th0_on ? th0[q < PRE ? q : 0] + tphi * phi_r[q]
                                                     : tv * vv + tphi * phi_r[q] + tg * nrm + tgm * (nrm * sqrtf(nrm));
                            if (fabsf(t32) <= 0.0625f) {
                                c = gpe::phase_rot_small(c, -t32);
                                done = true;
                            }
                        }
                        if (!done) {
                            // phase argument in FP64, reduced in FP64 before the
                            // sin / cos of the bulk precision, so a large phase (dense
                            // states, long steps) keeps its absolute accuracy
                            double veff;
                            if (sizeof(T) == 8 && th0_on) {
                                veff = (double)th0[q < PRE ? q : 0] + (a.cv * a.cdd) * (double)phi_r[q];
                            } else {
                                const double dn = dens64(c);
                                const double vv =
                                    kStage ? (double)vpre[q < PRE ? q : 0] : (double)a.V[plane_off + p];
                                veff = a.cv * (vv + a.cdd * (double)phi_r[q]) + pg * dn + pgm * (dn * sqrt(dn));
                            }
                            if constexpr (kStage || sizeof(T) == 4)
                                c = gpe::phase_rot(c, -a.hdt * veff);
                            else
                                c = gpe::phase_rot_lib(c, -a.hdt * veff);
                        }
                        Q[qi] = c;
                    }
                }
                // the density of the Poisson start: |psi|^2 before the phase,
                // which leaves it unchanged (the split pass's Poisson-start
                // block never applies the phase, and both agree bit for bit)
                rho_r[q] = nrm;
                if (mode & M_NORM_OUT) nacc += (double)c.x * (double)c.x + (double)c.y * (double)c.y;
            }
        }
        __syncthreads();
        if (mode & M_OUT) {
            for (int p = threadIdx.x; p < npl; p += NT) store_wt(&a.out[plane_off + p], Q[(p / nz) * qs + p % nz]);
        }
        if ((mode & M_KIN) && !a.kz_skip) {
            if (mode & M_OUT) __syncthreads();  // the psi output above has read Q
            // z propagation: rows j, length nz (unless the x pass runs it)
            if constexpr (Pol::kRegZ) {
                // tables named through the LDS base, so their reads compile to LDS loads
                const cplx<T>* TWs = P + max(ny2 * nzh, ny * zs);
                kin_z_reg<T, Pol::kNZ, NT>(Q, ny, qs, TWs + a.tw_n);
                __syncthreads();
            } else {
                Pol::template fft_z<T, NT, -1>(Q, ny, 1, qs, a, tw, twlz);
                for (int p = threadIdx.x; p < npl; p += NT) {
                    const int qi = (p / nz) * qs + p % nz;
                    Q[qi] = gpe::kin_mul(Q[qi], ez[p % nz], 1.0 / nz);
                }
                __syncthreads();
                Pol::template fft_z<T, NT, +1>(Q, ny, 1, qs, a, tw, twlz);
            }
        }
        if ((mode & M_KIN) && Pol::kRegY) {
            // y propagation in registers, result straight to Tk
            if constexpr (Pol::kRegY) {
                const cplx<T>* TWs = P + max(ny2 * nzh, ny * zs);
                // the paired pass where its thread count fits the block
                constexpr int BP = Pol::kNY / 8;
                constexpr bool kPair = BP * Pol::kNZ <= NT && (BP * Pol::kNZ) % 64 == 0 && 8 % (BP / 2) == 0;
                if constexpr (kPair)
                    kin_y_pair_store<T, Pol::kNY, Pol::kNZ, NT>(Q, qs, TWs + a.tw_n + nz, TWs + a.py.tw,
                                                                TWs + a.tw_n + nz + ny, a.tko + (size_t)i * nz, nx);
                else
                    kin_y_reg_store<T, Pol::kNY, Pol::kNZ, NT>(Q, qs, TWs + a.tw_n + nz, TWs + a.py.tw,
                                                               TWs + a.tw_n + nz + ny, a.tko + (size_t)i * nz, nx);
            }
        } else if (mode & M_KIN) {
            // y propagation: columns k, length ny
            Pol::template fft_y<T, NT, -1>(Q, nz, qs, 1, a, tw, twly);
            for (int p = threadIdx.x; p < npl; p += NT) {
                const int qi = (p / nz) * qs + p % nz;
                Q[qi] = gpe::kin_mul(Q[qi], ey[p / nz], 1.0 / ny);
            }
            __syncthreads();
            Pol::template fft_y<T, NT, +1>(Q, nz, qs, 1, a, tw, twly);
            for (int p = threadIdx.x; p < npl; p += NT) {
                const int j = p / nz, k = p - j * nz;
                store_wt(&a.tko[((size_t)j * nx + i) * nz + k], Q[j * qs + k]);
            }
        }
        __syncthreads();
    }

    if (mode & M_NORM_OUT) {
        const double s = block_sum<NT>(nacc, red);
        if (threadIdx.x == 0) a.part_out[i] = s;
    }

    // ---- 3. start the Poisson solve of rho = |psi|^2 ----------------------
    if (mode & M_RHO_FWD) {
        cplx<T> tmp[RQ];
        if constexpr (kStage) {
            // z R2C, two rows per complex transform: row m = rho_m + i rho_m+ny/2
            // (zero for k in [nz, 2nz)); the spectra separate as
            // A = (Z_k + conj Z_-k) / 2, B = (Z_k - conj Z_-k) / 2i
#pragma unroll
            for (int q = 0; q < PPT; ++q) {
                const int p = threadIdx.x + q * NT;
                if (p < npl) {
                    const int j = p / nz, k = p - j * nz;
                    const int hj = j >= ny / 2;
                    reinterpret_cast<T*>(&P[(j - hj * (ny / 2)) * zs + k])[hj] = rho_r[q];
                    if (!hj) P[j * zs + nz + k] = gpe::mk(T(0), T(0));
                }
            }
            __syncthreads();
            Pol::template fft_z2<T, NT, -1>(P, ny / 2, 1, zs, a, tw);
            if constexpr (kRegY2) {
                // y in registers, S1 stored from registers: the plane is done
                y2_fwd_reg_store<T, Pol::kNY, Pol::kNZ>(P, zs, tw + a.py2.tw, a.s1o + (size_t)i * nz, nx);
                return;
            }
#pragma unroll
            for (int q = 0; q < RQ; ++q) {
                const int e = threadIdx.x + q * NT;
                if (e < ny * nzh) {
                    const int j = e / nzh, kz = e - j * nzh;
                    const int m = j >= ny / 2 ? j - ny / 2 : j;
                    const cplx<T> zk = P[m * zs + kz];
                    const cplx<T> zm = P[m * zs + (kz == 0 ? 0 : nz2 - kz)];
                    if (j >= ny / 2)
                        tmp[q] = gpe::mk(T(0.5) * (zk.y + zm.y), T(0.5) * (zm.x - zk.x));
                    else
                        tmp[q] = gpe::mk(T(0.5) * (zk.x + zm.x), T(0.5) * (zk.y - zm.y));
                }
            }
        } else {
            // Z[j][k] = rho for k < nz, 0 for k in [nz, 2nz)
#pragma unroll
            for (int q = 0; q < PPT; ++q) {
                const int p = threadIdx.x + q * NT;
                if (p < npl) {
                    const int j = p / nz, k = p - j * nz;
                    P[j * zs + k] = gpe::mk(rho_r[q], T(0));
                    P[j * zs + nz + k] = gpe::mk(T(0), T(0));
                }
            }
            __syncthreads();
            Pol::template fft_z2<T, NT, -1>(P, ny, 1, zs, a, tw);
#pragma unroll
            for (int q = 0; q < RQ; ++q) {
                const int e = threadIdx.x + q * NT;
                if (e < ny * nzh) {
                    const int j = e / nzh, kz = e - j * nzh;
                    tmp[q] = P[j * zs + kz];
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < RQ; ++q) {
            const int e = threadIdx.x + q * NT;
            if (e < ny * nzh) P[e] = tmp[q];
        }
        for (int e = ny * nzh + threadIdx.x; e < ny2 * nzh; e += NT) P[e] = gpe::mk(T(0), T(0));
        __syncthreads();
        Pol::template fft_y2<T, NT, -1>(P + 1, nz, nzh, 1, a, tw);
        for (int e = threadIdx.x; e < ny2 * nz; e += NT) {
            const int ky = e / nz, c = e - ky * nz;
            store_wt(&a.s1o[((size_t)ky * nx + i) * nz + c], P[ky * nzh + c + 1]);
        }
    }
}

// ---------------------------------------------------------------------------
// Pencil plane pass: planes whose Poisson tile (2ny x (nz+1) complex) does
// not fit one block's LDS.  The modes of plane_body run as three launches
// over LDS tiles, one block per (x-plane, tile), with the plane-sized
// intermediates in HBM:
//   k_pencil_ci (column tiles): inverse y2 of S1 -> G[y][kz-1], y < ny
//   k_pencil_r  (row tiles): z C2R of G -> Phi; psi update (IT stencil or RT
//       phase); z R2C of rho -> H[y][kz-1]; z kinetic of psi -> W (over G's
//       rows); FP64 norm partial per tile
//   k_pencil_cf (column tiles): y2 of H -> S1o; y kinetic of W -> Tk; the
//       per-plane norm partial as a fixed-order sum of the tile partials
// Same arithmetic as plane_body's run-time-size path.
constexpr int kPencilThreads = 512;
constexpr int kPencilPPT = 4;  // row-tile points per thread

template <typename T>
struct PencilLim;  // complex elements of one tile
template <>
struct PencilLim<float> {
    static constexpr int maxe = 8192;
};
template <>
struct PencilLim<double> {
    static constexpr int maxe = 4096;
};

// Tile shapes and line FFTs of the pencil kernels: run-time sizes (radix-4/2/
// 3/5/7 engine, tiles from the arguments) or compile-time shapes (large-radix
// engine, fixed tiles of C columns and R rows).
struct PencilPolGeneric {
    template <class A>
    __device__ static int C(const A& a) { return a.pen_c; }
    template <class A>
    __device__ static int R(const A& a) { return a.pen_r; }
    template <typename T, int NT, int S>
    __device__ static void y2(cplx<T>* b, int nl, int es, int ls, const PlaneArgs<T>& a) {
        gpe::lds_fft<T, NT, PencilLim<T>::maxe>(b, a.py2, nl, es, ls, S, a.tw);
    }
    template <typename T, int NT, int S>
    __device__ static void z2(cplx<T>* b, int nl, int es, int ls, const PlaneArgs<T>& a) {
        gpe::lds_fft<T, NT, PencilLim<T>::maxe>(b, a.pz2, nl, es, ls, S, a.tw);
    }
    template <typename T, int NT, int S>
    __device__ static void y(cplx<T>* b, int nl, int es, int ls, const PlaneArgs<T>& a) {
        gpe::lds_fft<T, NT, PencilLim<T>::maxe>(b, a.py, nl, es, ls, S, a.tw, a.twly);
    }
    template <typename T, int NT, int S>
    __device__ static void z(cplx<T>* b, int nl, int es, int ls, const PlaneArgs<T>& a) {
        gpe::lds_fft<T, NT, PencilLim<T>::maxe>(b, a.pz, nl, es, ls, S, a.tw, a.twlz);
    }
};

template <int NY, int NZ, int CC, int RR>
struct PencilPolCT {
    // CC / RR bound the tiles (the line bounds of the FFTs below); the tiles
    // themselves come at run time (gpe_set_tiles / the autotuner may lower them)
    static constexpr int kC = CC, kR = RR;
    template <class A>
    __device__ static int C(const A& a) { return a.pen_c; }
    template <class A>
    __device__ static int R(const A& a) { return a.pen_r; }
    template <typename T, int NT, int S>
    __device__ static void y2(cplx<T>* b, int nl, int es, int ls, const PlaneArgs<T>& a) {
        gpe::ct::fft<T, 2 * NY, NT, CC, S>(b, nl, es, ls, a.tw + a.py2.tw);
    }
    template <typename T, int NT, int S>
    __device__ static void z2(cplx<T>* b, int nl, int es, int ls, const PlaneArgs<T>& a) {
        gpe::ct::fft<T, 2 * NZ, NT, RR, S>(b, nl, es, ls, a.tw + a.pz2.tw);
    }
    template <typename T, int NT, int S>
    __device__ static void y(cplx<T>* b, int nl, int es, int ls, const PlaneArgs<T>& a) {
        gpe::ct::fft<T, NY, NT, CC, S, true>(b, nl, es, ls, a.tw + a.py.tw, a.twly);
    }
    template <typename T, int NT, int S>
    __device__ static void z(cplx<T>* b, int nl, int es, int ls, const PlaneArgs<T>& a) {
        gpe::ct::fft<T, NZ, NT, RR, S, true>(b, nl, es, ls, a.tw + a.pz.tw, a.twlz);
    }
};
// Pencil shapes with compile-time kernels: (ny, nz, columns, rows per tile).
#define GPE_PENCIL_SHAPES(X) X(256, 64, 8, 32)

// phase couplings of this step (as in plane_body)
template <typename T>
__device__ __forceinline__ void pencil_couplings(const PlaneArgs<T>& a, int mode, double* pg, double* pgm) {
    *pg = a.pg;
    *pgm = a.pgm;
    if (a.coef && (mode & (M_PHASE_HALF | M_PHASE_FULL))) {
        const int t = (a.cbase ? *a.cbase : 0) + a.cidx;
        *pg = a.coef[2 * t];
        *pgm = a.coef[2 * t + 1];
        if (mode & M_PHASE_FULL) {
            *pg += a.coef[2 * t + 2];
            *pgm += a.coef[2 * t + 3];
        }
    }
}

// inverse y2 of the S1 columns c0 .. c0+cw (kz = c + 1) of plane i -> G
template <typename T, int NT, class Pol>
__global__ void __launch_bounds__(NT) k_pencil_ci(PlaneArgs<T> a) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    cplx<T>* X = reinterpret_cast<cplx<T>*>(lds_raw);
    const int nx = a.nx, ny = a.ny, nz = a.nz, ny2 = 2 * ny, C = Pol::C(a);
    const int nct = (nz + C - 1) / C;
    const int i = blockIdx.x / nct, c0 = (blockIdx.x - i * nct) * C;
    const int cw = min(C, nz - c0), cs = C + 1;
    cplx<T>* gw = a.pen_gw + (size_t)i * ny * nz;
    for (int e = threadIdx.x; e < ny2 * cw; e += NT) {
        const int ky = e / cw, c = e - ky * cw;
        X[ky * cs + c] = a.s1[((size_t)ky * nx + i) * nz + c0 + c];
    }
    __syncthreads();
    Pol::template y2<T, NT, +1>(X, cw, cs, 1, a);
    for (int e = threadIdx.x; e < ny * cw; e += NT) {
        const int y = e / cw, c = e - y * cw;
        gw[(size_t)y * nz + c0 + c] = X[y * cs + c];
    }
}

// rows y0 .. y0+rw of plane i
template <typename T, int NT, class Pol>
__global__ void __launch_bounds__(NT) k_pencil_r(PlaneArgs<T> a) {
    constexpr int PPT = kPencilPPT;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ double red[NT / 64 + 1];
    cplx<T>* X = reinterpret_cast<cplx<T>*>(lds_raw);
    const int mode = a.mode;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int nz2 = 2 * nz, npl = ny * nz, R = Pol::R(a);
    const int nrt = (ny + R - 1) / R;
    const int i = blockIdx.x / nrt, tr = blockIdx.x - i * nrt, y0 = tr * R;
    const int rw = min(R, ny - y0), np = rw * nz;
    const size_t plane_off = (size_t)i * npl;
    const bool is_it = (mode & M_IT) != 0;
    const bool have_psi = (mode & (M_IT | M_PHASE_HALF | M_PHASE_FULL | M_KIN | M_OUT | M_RHO_FWD)) != 0;
    cplx<T>* gw = a.pen_gw + plane_off;  // [y][c]
    cplx<T>* hs = a.pen_h + plane_off;   // [y][kz - 1]
    double pg, pgm;
    pencil_couplings(a, mode, &pg, &pgm);
    double s_in = 1.0;
    if (mode & M_NORM_IN) s_in = 1.0 / sqrt(sum_partials(a.part_in, nx, red));
    const int zs = nz2 + 1, qs = nz + 1;
    const bool iint = (i >= 1 && i <= nx - 2);
    double nacc = 0.0;
    T phi_r[PPT], rho_r[PPT];
    cplx<T> cv[PPT];
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        phi_r[q] = T(0);
        rho_r[q] = T(0);
        cv[q] = gpe::mk(T(0), T(0));
    }
    // Phi of the tile rows: z C2R of G (Hermitian extension, kz = 0 zero).
    // The first GQ loads of each thread are issued together (one round trip).
    if (mode & M_PINV) {
        constexpr int GQ = 4;
        const int ne = rw * (nz + 1);
        cplx<T> gv[GQ];
#pragma unroll
        for (int q = 0; q < GQ; ++q) {
            const int e = threadIdx.x + q * NT;
            const int r = e / (nz + 1), kz = e - r * (nz + 1);
            gv[q] = (e < ne && kz) ? gw[(size_t)(y0 + r) * nz + kz - 1] : gpe::mk(T(0), T(0));
        }
        auto put = [&](int e, cplx<T> v) {
            const int r = e / (nz + 1), kz = e - r * (nz + 1);
            X[r * zs + kz] = v;
            if (kz >= 1 && kz < nz) X[r * zs + nz2 - kz] = gpe::mk(v.x, -v.y);
        };
#pragma unroll
        for (int q = 0; q < GQ; ++q)
            if (threadIdx.x + q * NT < ne) put(threadIdx.x + q * NT, gv[q]);
        for (int e = threadIdx.x + GQ * NT; e < ne; e += NT) {
            const int r = e / (nz + 1), kz = e - r * (nz + 1);
            put(e, kz ? gw[(size_t)(y0 + r) * nz + kz - 1] : gpe::mk(T(0), T(0)));
        }
        __syncthreads();
        Pol::template z2<T, NT, +1>(X, rw, 1, zs, a);
        const T sc = (T)(a.pinv_scale * s_in * s_in);
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
            const int p = threadIdx.x + q * NT;
            if (p < np) {
                const int r = p / nz, z = p - r * nz;
                phi_r[q] = X[r * zs + z].x * sc;
                if (mode & M_PHI_STORE) a.phi[plane_off + (size_t)(y0 + r) * nz + z] = phi_r[q];
            }
        }
        __syncthreads();
    } else if (mode & M_PHI_LOAD) {
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
            const int p = threadIdx.x + q * NT;
            if (p < np) phi_r[q] = a.phi[plane_off + (size_t)y0 * nz + p];
        }
    }
    // psi update
    if (is_it) {
        const cplx<T>* in = a.in;
        const T sv = (T)s_in;
        const T dtau = (T)a.dtau, cx = (T)a.cx, cy = (T)a.cy, cz = (T)a.cz, gq = (T)a.gq, gmq = (T)a.gmq,
                cdd = (T)a.cdd;
        auto ld = [&](size_t e) {
            const cplx<T> v = in[e];
            return gpe::mk((T)(v.x * s_in), (T)(v.y * s_in));
        };
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
            const int p = threadIdx.x + q * NT;
            if (p < np) {
                const int r = p / nz, k = p - r * nz, j = y0 + r;
                const size_t e = plane_off + (size_t)j * nz + k;
                cplx<T> o;
                if (iint && j >= 1 && j <= ny - 2 && k >= 1 && k <= nz - 2) {
                    const cplx<T> c = ld(e);
                    const cplx<T> xm = in[e - npl], xp = in[e + npl];
                    const cplx<T> ym = ld(e - nz), yp = ld(e + nz), zm = ld(e - 1), zp = ld(e + 1);
                    const T v = a.V[e] + cdd * phi_r[q];
                    cplx<T> c1;
                    c1.x = cx * (xm.x * sv + xp.x * sv - T(2) * c.x) + cy * (ym.x + yp.x - T(2) * c.x) +
                           cz * (zm.x + zp.x - T(2) * c.x) + c.x * v;
                    c1.y = cx * (xm.y * sv + xp.y * sv - T(2) * c.y) + cy * (ym.y + yp.y - T(2) * c.y) +
                           cz * (zm.y + zp.y - T(2) * c.y) + c.y * v;
                    const T nrm = c.x * c.x + c.y * c.y;
                    const T ab = sqrt(nrm);
                    const T f = gq * nrm + gmq * (ab * ab * ab);
                    o.x = c.x - dtau * c1.x;
                    o.y = c.y - dtau * c1.y;
                    o.x = o.x - dtau * (f * c.x);
                    o.y = o.y - dtau * (f * c.y);
                } else {
                    o = a.psii[e];
                }
                a.out[e] = o;
                cv[q] = o;
                rho_r[q] = o.x * o.x + o.y * o.y;
                if (mode & M_NORM_OUT) nacc += (double)o.x * (double)o.x + (double)o.y * (double)o.y;
            }
        }
    } else if (have_psi) {
        // FP32: the phase argument in FP32 where it is small (as in the
        // single-block plane pass), FP64 argument and reduction otherwise
        const float tv = (float)(a.hdt * a.cv), tphi = (float)(a.hdt * a.cv * a.cdd), tg = (float)(a.hdt * pg),
                    tgm = (float)(a.hdt * pgm);
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
            const int p = threadIdx.x + q * NT;
            if (p < np) {
                const int r = p / nz, k = p - r * nz, j = y0 + r;
                const size_t e = plane_off + (size_t)j * nz + k;
                cplx<T> c;
                if (mode & M_IN_TK) {
                    c = a.tk[((size_t)j * nx + i) * nz + k];
                } else {
                    c = a.in[e];
                    if (mode & M_NORM_IN) c = gpe::mk((T)(c.x * s_in), (T)(c.y * s_in));
                }
                T nrm = c.x * c.x + c.y * c.y;
                if ((mode & (M_PHASE_HALF | M_PHASE_FULL)) && iint && j >= 1 && j <= ny - 2 && k >= 1 &&
                    k <= nz - 2) {
                    bool done = false;
                    if constexpr (sizeof(T) == 4) {
                        const float t32 = tv * a.V[e] + tphi * phi_r[q] + tg * nrm + tgm * (nrm * sqrtf(nrm));
                        if (fabsf(t32) <= 0.0625f) {
                            c = gpe::phase_rot_small(c, -t32);
                            done = true;
                        }
                    }
                    if (!done) {
                        const double dn = dens64(c);
                        const double veff =
                            a.cv * ((double)a.V[e] + a.cdd * (double)phi_r[q]) + pg * dn + pgm * (dn * sqrt(dn));
                        c = gpe::phase_rot(c, -a.hdt * veff);
                    }
                    nrm = c.x * c.x + c.y * c.y;
                }
                cv[q] = c;
                rho_r[q] = nrm;
                if (mode & M_NORM_OUT) nacc += (double)c.x * (double)c.x + (double)c.y * (double)c.y;
                if (mode & M_OUT) a.out[e] = c;
            }
        }
    }
    // z R2C of rho -> H
    if (mode & M_RHO_FWD) {
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
            const int p = threadIdx.x + q * NT;
            if (p < np) {
                const int r = p / nz, z = p - r * nz;
                X[r * zs + z] = gpe::mk(rho_r[q], T(0));
                X[r * zs + nz + z] = gpe::mk(T(0), T(0));
            }
        }
        __syncthreads();
        Pol::template z2<T, NT, -1>(X, rw, 1, zs, a);
        for (int e = threadIdx.x; e < np; e += NT) {
            const int r = e / nz, c = e - r * nz;
            hs[(size_t)(y0 + r) * nz + c] = X[r * zs + c + 1];
        }
        __syncthreads();
    }
    // z kinetic of psi -> W (the G rows of this tile are consumed above)
    if (mode & M_KIN) {
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
            const int p = threadIdx.x + q * NT;
            if (p < np) X[(p / nz) * qs + p % nz] = cv[q];
        }
        __syncthreads();
        Pol::template z<T, NT, -1>(X, rw, 1, qs, a);
        for (int e = threadIdx.x; e < np; e += NT) {
            const int qi = (e / nz) * qs + e % nz;
            X[qi] = gpe::kin_mul(X[qi], a.ez[e % nz], 1.0 / nz);
        }
        __syncthreads();
        Pol::template z<T, NT, +1>(X, rw, 1, qs, a);
        for (int e = threadIdx.x; e < np; e += NT) gw[(size_t)y0 * nz + e] = X[(e / nz) * qs + e % nz];
    }
    if (mode & M_NORM_OUT) {
        const double sum = block_sum<NT>(nacc, red);
        if (threadIdx.x == 0) a.pen_part[(size_t)i * nrt + tr] = sum;
    }
}

// The row-tile step of the pencil pass with its z transforms in registers
// (nz = 64, 32 rows y0 .. y0 + 31 of plane i per 256-thread block).  Thread
// (r, b), b < 8, works on row y0 + r:
//   Poisson finish: the 128-point inverse of the Hermitian-extended G row as
//     a four-step 16 x 8 (x[8 a + b] loaded straight from G, DFT_16,
//     twiddle, one LDS exchange, DFT_8 of rows k1 = b, b + 8), which leaves
//     the thread holding z = k1 + 16 k2 = 8 a' + b, a' = k1 / 8 + 2 k2 < 8,
//     of the real output;
//   the stencil or the phase at those eight points (psi and V loaded at the
//     kernel start);
//   Poisson start: the forward 128-point transform of the zero-padded
//     density, whose four-step input is exactly those eight points (x[8 a' +
//     b], a' < 8), one exchange, H from registers;
//   kinetic z: the 64-point four-step 8 x 8 of the same eight points (two
//     exchanges), W from registers.
// Four LDS exchanges in all, against eight Stockham passes of the LDS tile.
constexpr int kPzRows = 32;
template <typename T>
struct PzReg {
    static constexpr int NZ = 64, L = 128, B = 8, LS = kPzRows + 1;
    static constexpr int EXA = L * LS;                      // exchange area (complex)
    static constexpr int TAB = L + 3 * NZ;                  // W_128, W_64, lo parts of W_64, ez
    static constexpr size_t lds_bytes() { return (size_t)(EXA + TAB) * sizeof(cplx<T>); }
};

template <typename T>
__global__ void __launch_bounds__(256) k_pencil_rz(PlaneArgs<T> a) {
    using Z = PzReg<T>;
    constexpr int NZ = Z::NZ, L = Z::L, B = Z::B, LS = Z::LS;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ double red[256 / 64 + 1];
    cplx<T>* E = reinterpret_cast<cplx<T>*>(lds_raw);
    cplx<T>* TW2 = E + Z::EXA;  // W_128^m
    cplx<T>* TW1 = TW2 + L;     // W_64^m
    cplx<T>* TWL = TW1 + NZ;    // lo parts of W_64^m
    cplx<T>* EZ = TWL + NZ;     // z kinetic phase
    const int mode = a.mode;
    const int nx = a.nx, ny = a.ny, nz = NZ;
    const int npl = ny * nz;
    const int nrt = ny / kPzRows;
    const int i = blockIdx.x / nrt, tr = blockIdx.x - i * nrt, y0 = tr * kPzRows;
    const int b = threadIdx.x & 7, r = threadIdx.x >> 3;
    const int j = y0 + r;
    const size_t plane_off = (size_t)i * npl;
    const size_t row_off = plane_off + (size_t)j * nz;
    const bool is_it = (mode & M_IT) != 0;
    const bool have_psi = (mode & (M_IT | M_PHASE_HALF | M_PHASE_FULL | M_KIN | M_OUT | M_RHO_FWD)) != 0;
    const bool need_v = (mode & (M_IT | M_PHASE_HALF | M_PHASE_FULL)) != 0;
    cplx<T>* gw = a.pen_gw + plane_off;  // [y][c]
    cplx<T>* hs = a.pen_h + plane_off;   // [y][kz - 1]
    // the thread's points z = 8 q + b, q < 8
    // ---- loads first: G row (with its Hermitian mirror), psi, V ----------
    cplx<T> v[16];
    if (mode & M_PINV) {
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int kz = 8 * q + b;  // 0 .. 127
            cplx<T> g = gpe::mk(T(0), T(0));
            if (kz >= 1 && kz <= NZ) {
                g = gw[(size_t)j * nz + kz - 1];
            } else if (kz > NZ) {
                const cplx<T> m = gw[(size_t)j * nz + (L - kz) - 1];
                g = gpe::mk(m.x, -m.y);
            }
            v[q] = g;
        }
    }
    cplx<T> c[8];
    T vv[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int z = 8 * q + b;
        c[q] = gpe::mk(T(0), T(0));
        vv[q] = T(0);
        if (have_psi) {
            if (!is_it && (mode & M_IN_TK))
                c[q] = a.tk[((size_t)j * nx + i) * nz + z];
            else
                c[q] = a.in[row_off + z];
        }
        if (need_v) vv[q] = a.V[row_off + z];
    }
    for (int m = threadIdx.x; m < L; m += 256) TW2[m] = a.tw[a.pz2.tw + m];
    if (threadIdx.x < NZ) {
        TW1[threadIdx.x] = a.tw[a.pz.tw + threadIdx.x];
        TWL[threadIdx.x] = a.twlz[threadIdx.x];
        EZ[threadIdx.x] = a.ez[threadIdx.x];
    }
    double pg, pgm;
    pencil_couplings(a, mode, &pg, &pgm);
    double s_in = 1.0;
    if (mode & M_NORM_IN) s_in = 1.0 / sqrt(sum_partials(a.part_in, nx, red));
    __syncthreads();  // tables published
    // ---- Poisson finish: Phi at z = 8 q + b ---------------------------------
    T phi[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) phi[q] = T(0);
    if (mode & M_PINV) {
        gpe::ct::Dft<T, 16, +1>::run(v);
#pragma unroll
        for (int k1 = 1; k1 < 16; ++k1) v[k1] = gpe::cmulc(v[k1], TW2[(b * k1) % L]);
#pragma unroll
        for (int k1 = 0; k1 < 16; ++k1) E[(k1 * B + b) * LS + r] = v[k1];
        __syncthreads();
        const T sc = (T)(a.pinv_scale * s_in * s_in);
#pragma unroll
        for (int rr = 0; rr < 2; ++rr) {
            const int k1 = b + B * rr;
            cplx<T> w[B];
#pragma unroll
            for (int bb = 0; bb < B; ++bb) w[bb] = E[(k1 * B + bb) * LS + r];
            gpe::ct::Dft<T, B, +1>::run(w);
            // output z = k1 + 16 k2 = 8 (rr + 2 k2) + b, k2 < 4 (z < 64)
#pragma unroll
            for (int k2 = 0; k2 < 4; ++k2) phi[rr + 2 * k2] = w[k2].x * sc;
        }
        if (mode & M_PHI_STORE) {
#pragma unroll
            for (int q = 0; q < 8; ++q) a.phi[row_off + 8 * q + b] = phi[q];
        }
        __syncthreads();  // the exchange area is read before it is reused
    } else if (mode & M_PHI_LOAD) {
#pragma unroll
        for (int q = 0; q < 8; ++q) phi[q] = a.phi[row_off + 8 * q + b];
    }
    // ---- psi update ---------------------------------------------------------
    const bool iint = (i >= 1 && i <= nx - 2);
    const bool jint = (j >= 1 && j <= ny - 2);
    double nacc = 0.0;
    T rho[8];
    if (is_it) {
        const cplx<T>* in = a.in;
        const T sv = (T)s_in;
        const T dtau = (T)a.dtau, cx = (T)a.cx, cy = (T)a.cy, cz = (T)a.cz, gq = (T)a.gq, gmq = (T)a.gmq,
                cdd = (T)a.cdd;
        auto ld = [&](size_t e) {
            const cplx<T> t = in[e];
            return gpe::mk((T)(t.x * s_in), (T)(t.y * s_in));
        };
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int k = 8 * q + b;
            const size_t e = row_off + k;
            cplx<T> o;
            if (iint && jint && k >= 1 && k <= nz - 2) {
                const cplx<T> cc = gpe::mk((T)(c[q].x * s_in), (T)(c[q].y * s_in));
                const cplx<T> xm = in[e - npl], xp = in[e + npl];
                const cplx<T> ym = ld(e - nz), yp = ld(e + nz), zm = ld(e - 1), zp = ld(e + 1);
                const T vt = vv[q] + cdd * phi[q];
                cplx<T> c1;
                c1.x = cx * (xm.x * sv + xp.x * sv - T(2) * cc.x) + cy * (ym.x + yp.x - T(2) * cc.x) +
                       cz * (zm.x + zp.x - T(2) * cc.x) + cc.x * vt;
                c1.y = cx * (xm.y * sv + xp.y * sv - T(2) * cc.y) + cy * (ym.y + yp.y - T(2) * cc.y) +
                       cz * (zm.y + zp.y - T(2) * cc.y) + cc.y * vt;
                const T nrm = cc.x * cc.x + cc.y * cc.y;
                const T ab = sqrt(nrm);
                const T f = gq * nrm + gmq * (ab * ab * ab);
                o.x = cc.x - dtau * c1.x;
                o.y = cc.y - dtau * c1.y;
                o.x = o.x - dtau * (f * cc.x);
                o.y = o.y - dtau * (f * cc.y);
            } else {
                o = a.psii[e];
            }
            store_wt(&a.out[e], o);
            c[q] = o;
            rho[q] = o.x * o.x + o.y * o.y;
            if (mode & M_NORM_OUT) nacc += (double)o.x * (double)o.x + (double)o.y * (double)o.y;
        }
    } else {
        const float tv = (float)(a.hdt * a.cv), tphi = (float)(a.hdt * a.cv * a.cdd), tg = (float)(a.hdt * pg),
                    tgm = (float)(a.hdt * pgm);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int k = 8 * q + b;
            cplx<T> cc = c[q];
            if (!(mode & M_IN_TK) && (mode & M_NORM_IN)) cc = gpe::mk((T)(cc.x * s_in), (T)(cc.y * s_in));
            const T nrm = cc.x * cc.x + cc.y * cc.y;
            if ((mode & (M_PHASE_HALF | M_PHASE_FULL)) && iint && jint && k >= 1 && k <= nz - 2) {
                bool done = false;
                if constexpr (sizeof(T) == 4) {
                    const float t32 = tv * vv[q] + tphi * phi[q] + tg * nrm + tgm * (nrm * sqrtf(nrm));
                    if (fabsf(t32) <= 0.0625f) {
                        cc = gpe::phase_rot_small(cc, -t32);
                        done = true;
                    }
                }
                if (!done) {
                    const double dn = dens64(cc);
                    const double veff =
                        a.cv * ((double)vv[q] + a.cdd * (double)phi[q]) + pg * dn + pgm * (dn * sqrt(dn));
                    cc = gpe::phase_rot(cc, -a.hdt * veff);
                }
            }
            c[q] = cc;
            rho[q] = nrm;  // the density before the phase, which leaves it unchanged
            if (mode & M_NORM_OUT) nacc += (double)cc.x * (double)cc.x + (double)cc.y * (double)cc.y;
            if (mode & M_OUT) store_wt(&a.out[row_off + k], cc);
        }
    }
    // ---- Poisson start: H = forward transform of the zero-padded density ---
    if (mode & M_RHO_FWD) {
#pragma unroll
        for (int q = 0; q < 16; ++q) v[q] = gpe::mk(q < 8 ? rho[q] : T(0), T(0));
        gpe::ct::Dft<T, 16, -1>::run(v);
#pragma unroll
        for (int k1 = 1; k1 < 16; ++k1) v[k1] = gpe::cmul(v[k1], TW2[(b * k1) % L]);
#pragma unroll
        for (int k1 = 0; k1 < 16; ++k1) E[(k1 * B + b) * LS + r] = v[k1];
        __syncthreads();
#pragma unroll
        for (int rr = 0; rr < 2; ++rr) {
            const int k1 = b + B * rr;
            cplx<T> w[B];
#pragma unroll
            for (int bb = 0; bb < B; ++bb) w[bb] = E[(k1 * B + bb) * LS + r];
            gpe::ct::Dft<T, B, -1>::run(w);
#pragma unroll
            for (int k2 = 0; k2 < 5; ++k2) {
                const int kz = k1 + 16 * k2;  // H holds kz = 1 .. nz
                if (kz >= 1 && kz <= NZ) store_wt(&hs[(size_t)j * nz + kz - 1], w[k2]);
            }
        }
        __syncthreads();
    }
    // ---- kinetic z: 64-point four-step 8 x 8 on z = 8 q + b -----------------
    if (mode & M_KIN) {
        cplx<T> u[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) u[q] = c[q];
        gpe::ct::Dft<T, 8, -1, true>::run(u);
#pragma unroll
        for (int k1 = 1; k1 < 8; ++k1) u[k1] = tw_psi<T, -1>(u[k1], TW1, TWL, (b * k1) % NZ);
#pragma unroll
        for (int k1 = 0; k1 < 8; ++k1) E[(k1 * B + b) * LS + r] = u[k1];
        __syncthreads();
        // row k1 = b: X[b + 8 k2]
        cplx<T> w[8];
#pragma unroll
        for (int bb = 0; bb < 8; ++bb) w[bb] = E[(b * B + bb) * LS + r];
        gpe::ct::Dft<T, 8, -1, true>::run(w);
#pragma unroll
        for (int k2 = 0; k2 < 8; ++k2) w[k2] = gpe::kin_mul(w[k2], EZ[b + 8 * k2], 1.0 / NZ);
        gpe::ct::Dft<T, 8, +1, true>::run(w);
#pragma unroll
        for (int bb = 1; bb < 8; ++bb) w[bb] = tw_psi<T, +1>(w[bb], TW1, TWL, (bb * b) % NZ);
        // row b's slots are read and rewritten by this thread only
#pragma unroll
        for (int bb = 0; bb < 8; ++bb) E[(b * B + bb) * LS + r] = w[bb];
        __syncthreads();
#pragma unroll
        for (int k1 = 0; k1 < 8; ++k1) u[k1] = E[(k1 * B + b) * LS + r];
        gpe::ct::Dft<T, 8, +1, true>::run(u);
#pragma unroll
        for (int q = 0; q < 8; ++q) store_wt(&gw[(size_t)j * nz + 8 * q + b], u[q]);
    }
    if (mode & M_NORM_OUT) {
        const double sum = block_sum<256>(nacc, red);
        if (threadIdx.x == 0) a.pen_part[(size_t)i * nrt + tr] = sum;
    }
}

// blocks [0, nx*nct): y2 of H -> S1o (M_RHO_FWD); [nx*nct, 2 nx*nct): y
// kinetic of W -> Tk (M_KIN); the first block of each plane also sums the
// plane's tile norm partials in order (M_NORM_OUT)
template <typename T, int NT, class Pol>
__global__ void __launch_bounds__(NT) k_pencil_cf(PlaneArgs<T> a) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    cplx<T>* X = reinterpret_cast<cplx<T>*>(lds_raw);
    const int mode = a.mode;
    const int nx = a.nx, ny = a.ny, nz = a.nz, ny2 = 2 * ny, C = Pol::C(a);
    const int nct = (nz + C - 1) / C;
    int b = blockIdx.x;
    const bool kin = b >= nx * nct;
    if (kin) b -= nx * nct;
    const int i = b / nct, c0 = (b - i * nct) * C;
    const int cw = min(C, nz - c0), cs = C + 1;
    const size_t plane_off = (size_t)i * ny * nz;
    if (!kin && c0 == 0 && (mode & M_NORM_OUT) && threadIdx.x == 0) {
        const int nrt = (ny + Pol::R(a) - 1) / Pol::R(a);
        double sum = 0.0;
        for (int t = 0; t < nrt; ++t) sum += a.pen_part[(size_t)i * nrt + t];
        a.part_out[i] = sum;
    }
    if (!kin && (mode & M_RHO_FWD)) {
        const cplx<T>* hs = a.pen_h + plane_off;
        for (int e = threadIdx.x; e < ny2 * cw; e += NT) {
            const int ky = e / cw, c = e - ky * cw;
            X[ky * cs + c] = ky < ny ? hs[(size_t)ky * nz + c0 + c] : gpe::mk(T(0), T(0));
        }
        __syncthreads();
        Pol::template y2<T, NT, -1>(X, cw, cs, 1, a);
        for (int e = threadIdx.x; e < ny2 * cw; e += NT) {
            const int ky = e / cw, c = e - ky * cw;
            a.s1o[((size_t)ky * nx + i) * nz + c0 + c] = X[ky * cs + c];
        }
    }
    if (kin && (mode & M_KIN)) {
        const cplx<T>* gw = a.pen_gw + plane_off;
        for (int e = threadIdx.x; e < ny * cw; e += NT) {
            const int y = e / cw, c = e - y * cw;
            X[y * cs + c] = gw[(size_t)y * nz + c0 + c];
        }
        __syncthreads();
        Pol::template y<T, NT, -1>(X, cw, cs, 1, a);
        for (int e = threadIdx.x; e < ny * cw; e += NT) {
            const int y = e / cw, c = e - y * cw;
            X[y * cs + c] = gpe::kin_mul(X[y * cs + c], a.ey[y], 1.0 / ny);
        }
        __syncthreads();
        Pol::template y<T, NT, +1>(X, cw, cs, 1, a);
        for (int e = threadIdx.x; e < ny * cw; e += NT) {
            const int y = e / cw, c = e - y * cw;
            a.tko[((size_t)y * nx + i) * nz + c0 + c] = X[y * cs + c];
        }
    }
}

// ---------------------------------------------------------------------------
// Line kernel: x-direction passes.  Dynamic LDS: max(2nx*cp, nx*ck) complex T.
template <typename T, int NT, class Pol>
__device__ __forceinline__ void line_body(const LineArgs<T>& a, int b, unsigned char* lds_raw) {
    cplx<T>* X = reinterpret_cast<cplx<T>*>(lds_raw);
    const int nx = Pol::nx(a), nzs = a.nz;  // S1 rows hold kz = 1..nz
    if (b < a.ntile_p) {
        const int ky = b / a.tiles_kz;
        const int kz0 = (b - ky * a.tiles_kz) * a.cp;  // column index, kz = kz0 + c + 1
        const int C = min(a.cp, nzs - kz0);
        const int nx2 = 2 * nx;
        const int CS = C + 1;  // padded LDS stride between x points
        cplx<T>* base = a.s1 + (size_t)ky * nx * nzs + kz0;
        for (int e = threadIdx.x; e < nx2 * C; e += NT) {
            const int ii = e / C, c = e - ii * C;
            X[ii * CS + c] = (ii < nx) ? base[(size_t)ii * nzs + c] : gpe::mk(T(0), T(0));
        }
        __syncthreads();
        Pol::template fft_x2<T, NT, -1>(X, C, CS, 1, a);
        // Vdd(k) = kz^2 / k^2 from the FP64 per-axis tables, formed in T
        const T ky2 = (T)a.ky2[ky];
        for (int e = threadIdx.x; e < nx2 * C; e += NT) {
            const int ii = e / C, c = e - ii * C;
            const T kz2 = (T)a.kz2[kz0 + c + 1];
            const T k2 = (T)a.kx2[ii] + ky2 + kz2;
            const T v = (k2 > T(1e-30)) ? kz2 / k2 : T(0);
            X[ii * CS + c] = gpe::cscale(X[ii * CS + c], v);
        }
        __syncthreads();
        Pol::template fft_x2<T, NT, +1>(X, C, CS, 1, a);
        for (int e = threadIdx.x; e < nx * C; e += NT) {
            const int ii = e / C, c = e - ii * C;
            store_wt(&base[(size_t)ii * nzs + c], X[ii * CS + c]);
        }
    } else {
        b -= a.ntile_p;
        const int nz = a.nz;
        const int j = b / a.tiles_k;
        const int k0 = (b - j * a.tiles_k) * a.ck;
        const int C = min(a.ck, nz - k0);
        const int CS = C + 1;
        cplx<T>* base = a.tk + (size_t)j * nx * nz + k0;
        for (int e = threadIdx.x; e < nx * C; e += NT) {
            const int ii = e / C, c = e - ii * C;
            X[ii * CS + c] = base[(size_t)ii * nz + c];
        }
        __syncthreads();
        Pol::template fft_x<T, NT, -1>(X, C, CS, 1, a);
        for (int e = threadIdx.x; e < nx * C; e += NT) {
            const int ii = e / C, c = e - ii * C;
            X[ii * CS + c] = gpe::kin_mul(X[ii * CS + c], a.ex[ii], 1.0 / nx);
        }
        __syncthreads();
        Pol::template fft_x<T, NT, +1>(X, C, CS, 1, a);
        for (int e = threadIdx.x; e < nx * C; e += NT) {
            const int ii = e / C, c = e - ii * C;
            store_wt(&base[(size_t)ii * nz + c], X[ii * CS + c]);
        }
    }
}

// One block per x-plane, or (split) two.  Block i finishes the Poisson solve
// and applies the phase (real time) or the stencil (imaginary time) of plane
// i, then runs the kinetic z / y propagation (and any psi, fi3d or norm
// output); block i + nx starts the next Poisson solve.  In imaginary time it
// also finishes the Poisson solve and runs the stencil, since the density is
// that of the updated psi; in real time the phase leaves |psi|^2 unchanged,
// so it starts from |psi|^2 of the plane as read.  The two halves of the
// plane work run on different CUs, at the price of reading the plane twice;
// the outputs go to the other buffer of a ping-pong pair (s1o, tko) so no
// block overwrites an input another block may not have read yet.
template <typename T, int NT, class Pol>
__global__ void __launch_bounds__(NT) k_plane(PlaneArgs<T> a) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ double red[NT / 64 + 1];
    int i = blockIdx.x, mode = a.mode;
    if (a.xcd) {
        // Workgroups go to the eight XCDs in turn (block b to XCD b % 8), and
        // each XCD has its own L2.  Block b = 8 s + x takes plane
        // x (nx / 8) + s % (nx / 8), in role s / (nx / 8) of a split pass: XCD x
        // holds planes [x nx / 8, (x + 1) nx / 8) in both roles, so the x
        // neighbours an imaginary-time stencil reads, and the plane both roles
        // read, are mostly in its L2.  Role A blocks still come first.
        const int p8 = a.nx >> 3, x = blockIdx.x & 7, s = blockIdx.x >> 3;
        i = (s / p8) * a.nx + x * p8 + s % p8;
    }
    if (a.split) {
        if (i < a.nx) {
            mode &= ~M_RHO_FWD;
        } else {
            mode = (mode & ~(M_KIN | M_OUT | M_PHI_STORE | M_NORM_OUT)) | M_NO_PSI_OUT;
            if (mode & (M_PHASE_FULL | M_PHASE_HALF)) mode &= ~(M_PINV | M_PHASE_FULL | M_PHASE_HALF);
        }
        i = i < a.nx ? i : i - a.nx;
    }
    plane_body<T, NT, Pol>(a, i, mode, lds_raw, red);
}

template <typename T, int NT, class Pol>
__global__ void __launch_bounds__(NT) k_xpass(LineArgs<T> a) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    line_body<T, NT, Pol>(a, blockIdx.x, lds_raw);
}

// ---------------------------------------------------------------------------
// Register-resident x pass (power-of-two nx).  Each x line of length
// L = A*B is held by B threads, A values each (A = 16):
//   x[B a + b] --DFT_A over a--> Y[b][k1] * W_L^{b k1}  --LDS exchange-->
//   thread t: rows k1 = t + B r --DFT_B over b--> X[k1 + A k2]
// the multiplier (Vdd or the kinetic phase) is applied there, and the
// inverse transform runs the same steps backwards from the same registers,
// so a forward / multiply / inverse line costs two LDS exchanges.  Global
// loads and stores go straight to / from registers; lines are the fastest
// lane index, so neighbouring lanes touch neighbouring kz / k (contiguous).
// SPLIT: the two exchanges move the real parts, then the imaginary parts,
// through one area of L (LPB + 1) reals (half the LDS of a complex area).
template <typename T, int L, int LPB, bool POISSON, int A = 16, bool SPLIT = false>
__device__ __forceinline__ void xline_reg(const LineArgs<T>& a, int blk, cplx<T>* E, cplx<T>* TB) {
    constexpr int B = L / A, R = A / B;
    constexpr int LS = LPB + 1;  // padded row stride of the exchange area
    static_assert(L % A == 0 && A % B == 0, "line split");
    const int nx = a.nx, nz = a.nz;
    const int nlines = POISSON ? 2 * a.ny * nz : a.ny * nz;
    const int ll = threadIdx.x % LPB, b = threadIdx.x / LPB;
    const int line = blk * LPB + ll;
    const bool valid = line < nlines;
    cplx<T>* base;
    int estride, ky = 0, kz = 0;
    if (POISSON) {
        ky = valid ? line / nz : 0;
        const int c = valid ? line - ky * nz : 0;  // S1 rows hold kz = 1..nz
        kz = c + 1;
        base = a.s1 + (size_t)ky * nx * nz + c;
        estride = nz;
    } else {
        const int j = valid ? line / nz : 0;
        const int k = valid ? line - j * nz : 0;
        base = a.tk + (size_t)j * nx * nz + k;
        estride = nz;
    }
    // issue the line loads first, then stage the tables while they are in
    // flight: TB[0, L) = W_L^m, then (Poisson) kx^2 as T or (kinetic) the
    // x phase table and the lo parts of W_L^m
    constexpr int AL = POISSON ? A / 2 : A;  // Poisson: x >= nx is the zero padding
    cplx<T> v[A];
#pragma unroll
    for (int q = 0; q < A; ++q) {
        v[q] = gpe::mk(T(0), T(0));
        if (q < AL && valid) v[q] = base[(size_t)(B * q + b) * estride];
    }
    T ky2 = T(0), kz2 = T(0);
    if (POISSON) {
        ky2 = (T)a.ky2[ky];
        kz2 = (T)a.kz2[kz];
    }
    const cplx<T>* tw = TB;
    T* kx2s = reinterpret_cast<T*>(TB + L);
    const cplx<T>* exs = TB + L;
    const cplx<T>* twl = TB + 2 * L;
    for (int m = threadIdx.x; m < L; m += 256) {
        TB[m] = a.tw[(POISSON ? a.px2.tw : a.px.tw) + m];
        if (POISSON) {
            kx2s[m] = (T)a.kx2[m];
        } else {
            TB[L + m] = a.ex[m];
            TB[2 * L + m] = a.twlx[m];
            if (m < nz) TB[3 * L + m] = a.ez[m];
        }
    }
    __syncthreads();
    // forward, stage 1 (kinetic lines are transforms of psi: EX constants and
    // twiddles with their lo parts)
    gpe::ct::Dft<T, A, -1, !POISSON>::run(v);
#pragma unroll
    for (int k1 = 1; k1 < A; ++k1) {
        if constexpr (POISSON)
            v[k1] = gpe::cmul(v[k1], tw[(b * k1) % L]);
        else
            v[k1] = tw_psi<T, -1>(v[k1], tw, twl, (b * k1) % L);
    }
    // forward stage 2, multiply, inverse stage 1 of row k1 = b + B r (w[r] loaded)
    auto stage2 = [&](cplx<T>(&wr)[B], int r) {
        const int k1 = b + B * r;
        gpe::ct::Dft<T, B, -1, !POISSON>::run(wr);
        if (POISSON) {
#pragma unroll
            for (int k2 = 0; k2 < B; ++k2) {
                const T k2s = kx2s[k1 + A * k2] + ky2 + kz2;
                const T vd = (k2s > T(1e-30)) ? kz2 * rcp_t(k2s) : T(0);
                wr[k2] = gpe::cscale(wr[k2], vd);
            }
        } else {
#pragma unroll
            for (int k2 = 0; k2 < B; ++k2) wr[k2] = gpe::kin_mul(wr[k2], exs[k1 + A * k2], 1.0 / L);
        }
        gpe::ct::Dft<T, B, +1, !POISSON>::run(wr);
#pragma unroll
        for (int bb = 1; bb < B; ++bb) {
            if constexpr (POISSON)
                wr[bb] = gpe::cmulc(wr[bb], tw[(bb * k1) % L]);
            else
                wr[bb] = tw_psi<T, +1>(wr[bb], tw, twl, (bb * k1) % L);
        }
    };
    cplx<T> w[R][B];
    if constexpr (SPLIT) {
        static_assert(POISSON, "split exchange: Poisson tiles");
        T* ER = reinterpret_cast<T*>(E);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (h) __syncthreads();  // the real parts are read before the imaginary parts land
#pragma unroll
            for (int k1 = 0; k1 < A; ++k1) ER[(k1 * B + b) * LS + ll] = h ? v[k1].y : v[k1].x;
            __syncthreads();
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int bb = 0; bb < B; ++bb) {
                    const T e = ER[((b + B * r) * B + bb) * LS + ll];
                    if (h) w[r][bb].y = e; else w[r][bb].x = e;
                }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) stage2(w[r], r);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            __syncthreads();  // every read of the area is done before it is rewritten
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int bb = 0; bb < B; ++bb) ER[((b + B * r) * B + bb) * LS + ll] = h ? w[r][bb].y : w[r][bb].x;
            __syncthreads();
#pragma unroll
            for (int k1 = 0; k1 < A; ++k1) {
                const T e = ER[(k1 * B + b) * LS + ll];
                if (h) v[k1].y = e; else v[k1].x = e;
            }
        }
    } else {
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) E[(k1 * B + b) * LS + ll] = v[k1];
        __syncthreads();
        if constexpr (!POISSON && (LPB == 32 || LPB == 16)) {
            // the z propagation of the kinetic step: with LPB = nz the tile holds
            // whole z rows, row k1 B + b of E being element k1 B + b of every z
            // line of one y row after the first x stage; exp(-i dt kz^2/2m)
            // commutes with the x transform, so it applies to these rows as to
            // psi itself
            if (a.kinz) {
                kin_z_reg<T, LPB, 256>(E, L, LS, TB + 3 * L);
                __syncthreads();
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int k1 = b + B * r;
#pragma unroll
            for (int bb = 0; bb < B; ++bb) w[r][bb] = E[(k1 * B + bb) * LS + ll];
            stage2(w[r], r);
        }
        __syncthreads();  // every stage-2 read of E is done before it is rewritten
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int k1 = b + B * r;
#pragma unroll
            for (int bb = 0; bb < B; ++bb) E[(k1 * B + bb) * LS + ll] = w[r][bb];
        }
        __syncthreads();
        // inverse stage 2: thread b gathers Z[k1][b] for every k1
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) v[k1] = E[(k1 * B + b) * LS + ll];
    }
    gpe::ct::Dft<T, A, +1, !POISSON>::run(v);
    if (valid) {
        // write-through stores: the next pass reads them on other XCDs, and
        // lines that do not linger dirty in L2 shorten the end-of-kernel flush
#pragma unroll
        for (int q = 0; q < AL; ++q) store_wt(&base[(size_t)(B * q + b) * estride], v[q]);
    }
}

// Register x pass shapes: A values per thread, B = L / A threads per line
// (A % B == 0).  The 512-point Poisson lines of nx = 256 take A = 32.
template <int NX>
struct XReg {
    static constexpr int LP = 2 * NX, LK = NX;            // Poisson / kinetic line lengths
    static constexpr int AP = LP / 16 > 16 ? 32 : 16, AK = 16;
    static constexpr int LPB_P = 256 / (LP / AP);         // lines per 256-thread block
    static constexpr int LPB_K = 256 / (LK / AK);
    // the 512-point Poisson lines exchange real and imaginary parts in turn,
    // so their area is no larger than the kinetic tiles' (3 blocks per CU in
    // FP32 instead of 2; in FP64 the registers allow one, and the complex
    // exchange with fewer barriers is faster)
    static constexpr bool SPLIT_P = NX == 256;
    template <typename T>
    static constexpr bool split_p() { return SPLIT_P && sizeof(T) == 4; }
    // exchange area (complex units): L (LPB + 1) complex, or reals when split
    template <typename T>
    static constexpr int exa() {
        const int p = split_p<T>() ? (LP * (LPB_P + 1) + 1) / 2 : LP * (LPB_P + 1);
        return p > LK * (LPB_K + 1) ? p : LK * (LPB_K + 1);
    }
};

// LDS: exchange area (XReg::exa complex units) + tables (Poisson 2 * 2NX,
// kinetic 3 * NX + nz complex).
// The kinetic tiles (twice the loads of a zero-padded Poisson tile) take the
// first block indices so that they are dispatched first.
template <typename T, int NX>
__global__ void __launch_bounds__(256) k_xpass_reg(LineArgs<T> a, int nblk_k) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    cplx<T>* E = reinterpret_cast<cplx<T>*>(lds_raw);
    using X = XReg<NX>;
    cplx<T>* TB = E + X::template exa<T>();
    if ((int)blockIdx.x < nblk_k)
        xline_reg<T, X::LK, X::LPB_K, false, X::AK>(a, blockIdx.x, E, TB);
    else
        xline_reg<T, X::LP, X::LPB_P, true, X::AP, X::template split_p<T>()>(a, blockIdx.x - nblk_k, E, TB);
}

// Register y passes of the pencil plane pass (compiled ny): the
// four-step of the register x pass on y lines (kz columns of one plane, LPB
// columns per 256-thread block), one direction at a time: YR_FWD forward
// (inputs from index nin on are zero), YR_INV inverse (outputs from index
// nout on are not stored), YR_KIN forward, kinetic phase, inverse (psi
// transforms: compensated twiddles and constants).  in / out point at the
// thread's line; istr / ostr are the element strides.  tw, twl, ek are LDS
// tables of the line length (twl and ek: YR_KIN only).
enum { YR_FWD = 0, YR_INV = 1, YR_KIN = 2 };
template <typename T, int L, int LPB, int DIR, int A>
__device__ __forceinline__ void yline_reg(const cplx<T>* in, size_t istr, int nin, cplx<T>* out, size_t ostr, int nout,
                                          bool valid, T* E, const cplx<T>* tw, const cplx<T>* twl,
                                          const cplx<T>* ek) {
    constexpr int B = L / A, R = A / B, LS = LPB + 1;
    constexpr bool EX = DIR == YR_KIN;
    static_assert(L % A == 0 && A % B == 0 && LPB * B == 256, "y line split");
    const int ll = threadIdx.x % LPB, b = threadIdx.x / LPB;
    cplx<T> w[R][B];
    // The exchange moves the real parts, then the imaginary parts, through one
    // area of L (LPB + 1) reals: half the LDS of a complex area, so more blocks
    // fit on a CU.  Slot (k1 B + bb) LS + ll holds element k1 B + bb of line ll.
    // v -> w: thread b gives elements k1 B + b, takes rows k1 = b + B r.
    auto v_to_w = [&](const cplx<T>(&v)[A]) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (h) __syncthreads();  // the real parts are read before the imaginary parts land
#pragma unroll
            for (int k1 = 0; k1 < A; ++k1) E[(k1 * B + b) * LS + ll] = h ? v[k1].y : v[k1].x;
            __syncthreads();
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int bb = 0; bb < B; ++bb) {
                    const T e = E[((b + B * r) * B + bb) * LS + ll];
                    if (h) w[r][bb].y = e; else w[r][bb].x = e;
                }
        }
    };
    // w -> v, the reverse
    auto w_to_v = [&](cplx<T>(&v)[A]) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (h) __syncthreads();
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int bb = 0; bb < B; ++bb) E[((b + B * r) * B + bb) * LS + ll] = h ? w[r][bb].y : w[r][bb].x;
            __syncthreads();
#pragma unroll
            for (int k1 = 0; k1 < A; ++k1) {
                const T e = E[(k1 * B + b) * LS + ll];
                if (h) v[k1].y = e; else v[k1].x = e;
            }
        }
    };
    if constexpr (DIR != YR_INV) {
        cplx<T> v[A];
#pragma unroll
        for (int q = 0; q < A; ++q) {
            v[q] = gpe::mk(T(0), T(0));
            if (valid && B * q + b < nin) v[q] = in[(size_t)(B * q + b) * istr];
        }
        gpe::ct::Dft<T, A, -1, EX>::run(v);
#pragma unroll
        for (int k1 = 1; k1 < A; ++k1) {
            if constexpr (EX)
                v[k1] = tw_psi<T, -1>(v[k1], tw, twl, (b * k1) % L);
            else
                v[k1] = gpe::cmul(v[k1], tw[(b * k1) % L]);
        }
        v_to_w(v);
#pragma unroll
        for (int r = 0; r < R; ++r) gpe::ct::Dft<T, B, -1, EX>::run(w[r]);
        if constexpr (DIR == YR_FWD) {
            if (valid) {
#pragma unroll
                for (int r = 0; r < R; ++r)
#pragma unroll
                    for (int k2 = 0; k2 < B; ++k2) store_wt(&out[(size_t)(b + B * r + A * k2) * ostr], w[r][k2]);
            }
            return;
        }
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int k2 = 0; k2 < B; ++k2) w[r][k2] = gpe::kin_mul(w[r][k2], ek[b + B * r + A * k2], 1.0 / L);
        __syncthreads();  // every read of E is done before it is rewritten
    } else {
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int k2 = 0; k2 < B; ++k2)
                w[r][k2] = valid ? in[(size_t)(b + B * r + A * k2) * istr] : gpe::mk(T(0), T(0));
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int k1 = b + B * r;
        gpe::ct::Dft<T, B, +1, EX>::run(w[r]);
#pragma unroll
        for (int bb = 1; bb < B; ++bb) {
            if constexpr (EX)
                w[r][bb] = tw_psi<T, +1>(w[r][bb], tw, twl, (bb * k1) % L);
            else
                w[r][bb] = gpe::cmulc(w[r][bb], tw[(bb * k1) % L]);
        }
    }
    cplx<T> v[A];
    w_to_v(v);
    gpe::ct::Dft<T, A, +1, EX>::run(v);
    if (valid) {
#pragma unroll
        for (int q = 0; q < A; ++q)
            if (B * q + b < nout) store_wt(&out[(size_t)(B * q + b) * ostr], v[q]);
    }
}

// Shapes of the register pencil y passes: 2NY-point lines as 32 x 16, NY-point
// lines as 16 x 16, 16 kz columns per block either way.
template <int NY>
struct PenY {
    static constexpr int L2 = 2 * NY, L1 = NY, A2 = 32, A1 = 16, LPB = 16;
    static constexpr int EXR = L2 * (LPB + 1);  // exchange area (reals)
    static constexpr int TAB = L2 > 3 * L1 ? L2 : 3 * L1;  // tables (complex)
    template <typename T>
    static constexpr size_t lds_bytes() { return EXR * sizeof(T) + TAB * sizeof(cplx<T>); }
};

// inverse y2 of S1 columns -> G (the k_pencil_ci step)
template <typename T, int NY>
__global__ void __launch_bounds__(256) k_pencil_yci(PlaneArgs<T> a) {
    using Y = PenY<NY>;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    T* E = reinterpret_cast<T*>(lds_raw);
    cplx<T>* TB = reinterpret_cast<cplx<T>*>(E + Y::EXR);
    const int nx = a.nx, nz = a.nz, ncb = (nz + Y::LPB - 1) / Y::LPB;
    const int i = blockIdx.x / ncb, c = (blockIdx.x - i * ncb) * Y::LPB + (int)threadIdx.x % Y::LPB;
    const bool valid = c < nz;
    for (int m = threadIdx.x; m < Y::L2; m += 256) TB[m] = a.tw[a.py2.tw + m];
    __syncthreads();
    const int cc = valid ? c : 0;
    yline_reg<T, Y::L2, Y::LPB, YR_INV, Y::A2>(a.s1 + (size_t)i * nz + cc, (size_t)nx * nz, Y::L2,
                                              a.pen_gw + (size_t)i * NY * nz + cc, nz, NY, valid, E, TB, nullptr,
                                              nullptr);
}

// blocks [0, nx ncb): y2 of H -> S1o (M_RHO_FWD); [nx ncb, 2 nx ncb): y
// kinetic of W -> Tk (M_KIN); the first block of each plane also sums the
// plane's row-tile norm partials in order (M_NORM_OUT).  The k_pencil_cf step.
template <typename T, int NY>
__global__ void __launch_bounds__(256) k_pencil_ycf(PlaneArgs<T> a) {
    using Y = PenY<NY>;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    T* E = reinterpret_cast<T*>(lds_raw);
    cplx<T>* TB = reinterpret_cast<cplx<T>*>(E + Y::EXR);
    const int mode = a.mode;
    const int nx = a.nx, nz = a.nz, ncb = (nz + Y::LPB - 1) / Y::LPB;
    int blk = blockIdx.x;
    const bool kin = blk >= nx * ncb;
    if (kin) blk -= nx * ncb;
    const int i = blk / ncb, c0 = (blk - i * ncb) * Y::LPB, c = c0 + (int)threadIdx.x % Y::LPB;
    const bool valid = c < nz;
    const int cc = valid ? c : 0;
    const size_t plane_off = (size_t)i * NY * nz;
    if (!kin && c0 == 0 && (mode & M_NORM_OUT) && threadIdx.x == 0) {
        const int nrt = (NY + a.pen_r - 1) / a.pen_r;
        double sum = 0.0;
        for (int t = 0; t < nrt; ++t) sum += a.pen_part[(size_t)i * nrt + t];
        a.part_out[i] = sum;
    }
    if (!kin && (mode & M_RHO_FWD)) {
        for (int m = threadIdx.x; m < Y::L2; m += 256) TB[m] = a.tw[a.py2.tw + m];
        __syncthreads();
        yline_reg<T, Y::L2, Y::LPB, YR_FWD, Y::A2>(a.pen_h + plane_off + cc, nz, NY, a.s1o + (size_t)i * nz + cc,
                                                  (size_t)nx * nz, Y::L2, valid, E, TB, nullptr, nullptr);
    }
    if (kin && (mode & M_KIN)) {
        for (int m = threadIdx.x; m < Y::L1; m += 256) {
            TB[m] = a.tw[a.py.tw + m];
            TB[Y::L1 + m] = a.twly[m];
            TB[2 * Y::L1 + m] = a.ey[m];
        }
        __syncthreads();
        yline_reg<T, Y::L1, Y::LPB, YR_KIN, Y::A1>(a.pen_gw + plane_off + cc, nz, Y::L1, a.tko + (size_t)i * nz + cc,
                                                  (size_t)nx * nz, Y::L1, valid, E, TB, TB + Y::L1, TB + 2 * Y::L1);
    }
}

// Batched launches, one per grid shape: one block per work item (instance,
// plane) / (instance, tile); each instance brings its own argument block.
// A launch per shape keeps the single-instance kernels' register budget (the
// compiled FP32 plane pass runs 1024 threads in 128 VGPRs; one kernel
// holding every shape would not).
template <typename T, int NT, class Pol>
__global__ void __launch_bounds__(NT) k_plane_group(const PlaneArgs<T>* __restrict__ args,
                                                    const int2* __restrict__ work) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ double red[NT / 64 + 1];
    const int2 w = work[blockIdx.x];
    const PlaneArgs<T> a = args[w.x];
    plane_body<T, NT, Pol>(a, w.y, a.mode, lds_raw, red);
}

template <typename T, int NT, class Pol>
__global__ void __launch_bounds__(NT) k_xpass_group(const LineArgs<T>* __restrict__ args,
                                                    const int2* __restrict__ work) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int2 w = work[blockIdx.x];
    const LineArgs<T> a = args[w.x];
    line_body<T, NT, Pol>(a, w.y, lds_raw);
}

// Register x pass, Poisson tiles only (imaginary time).
template <typename T, int NX>
__global__ void __launch_bounds__(256) k_xpass_reg_group(const LineArgs<T>* __restrict__ args,
                                                         const int2* __restrict__ work) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    cplx<T>* E = reinterpret_cast<cplx<T>*>(lds_raw);
    using X = XReg<NX>;
    cplx<T>* TB = E + X::template exa<T>();
    const int2 w = work[blockIdx.x];
    const LineArgs<T> a = args[w.x];
    xline_reg<T, X::LP, X::LPB_P, true, X::AP, X::template split_p<T>()>(a, w.y, E, TB);
}

// One launch per pass over the instances of several grid shapes (a batch):
// work item (instance, index | slot << kSlotShift), where slot names the
// compiled kernel body that item runs.  Every compiled FP32 plane shape runs
// blocks of the same size within 128 VGPRs, so one launch carries them all
// and their blocks interleave on the GPU; the x pass likewise over the
// register Poisson tiles of nx = 128 and 64 and the LDS line shapes below
// 256 (nx = 256 needs more VGPRs, and a kernel's register count is that of
// its largest body).  FP32 only: in FP64 the joint plane body needs more
// registers than its 512-thread blocks have and spills.
constexpr int kSlotShift = 20;
constexpr int kSlotMask = (1 << kSlotShift) - 1;

template <typename T, int NT>
__global__ void __launch_bounds__(NT) k_plane_multi(const PlaneArgs<T>* __restrict__ args,
                                                    const int2* __restrict__ work) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ double red[NT / 64 + 1];
    const int2 w = work[blockIdx.x];
    const PlaneArgs<T> a = args[w.x];
    const int i = w.y & kSlotMask, slot = w.y >> kSlotShift;
    int s = 0;
#define GPE_PM(NY, NZ)                                                        \
    if (slot == s++) {                                                      \
        plane_body<T, NT, PlanePolCT<NY, NZ>>(a, i, a.mode, lds_raw, red);  \
        return;                                                             \
    }
    GPE_PLANE_SHAPES(GPE_PM)
#undef GPE_PM
}

// x-pass slots: 0 and 1 the register Poisson tiles of nx = 128 and 64, then
// the LDS pass of each x length of kMultiLineShapes.
#define GPE_MULTI_LINE_SHAPES(X) X(128) X(64) X(96) X(16) X(24) X(32)
template <typename T>
__global__ void __launch_bounds__(kLineThreads) k_line_multi(const LineArgs<T>* __restrict__ args,
                                                            const int2* __restrict__ work) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int2 w = work[blockIdx.x];
    const LineArgs<T> a = args[w.x];
    const int t = w.y & kSlotMask, slot = w.y >> kSlotShift;
    if (slot < 2) {
        cplx<T>* E = reinterpret_cast<cplx<T>*>(lds_raw);
        if (slot == 0) {
            using X = XReg<128>;
            xline_reg<T, X::LP, X::LPB_P, true, X::AP, X::template split_p<T>()>(a, t, E, E + X::template exa<T>());
        } else {
            using X = XReg<64>;
            xline_reg<T, X::LP, X::LPB_P, true, X::AP, X::template split_p<T>()>(a, t, E, E + X::template exa<T>());
        }
        return;
    }
    int s = 2;
#define GPE_LM(NX)                                                    \
    if (slot == s++) {                                              \
        line_body<T, kLineThreads, LinePolCT<NX>>(a, t, lds_raw);   \
        return;                                                     \
    }
    GPE_MULTI_LINE_SHAPES(GPE_LM)
#undef GPE_LM
}

// ---------------------------------------------------------------------------
// Elementwise / reduction kernels.

// psi *= s over every point, s = 1/sqrt(sum partials) (normalize).
template <typename T>
__global__ void k_scale(cplx<T>* psi, size_t n, const double* part, int np) {
    __shared__ double sred[2];
    const T sc = (T)(1.0 / sqrt(sum_partials(part, np, sred)));
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        cplx<T> v = psi[e];
        psi[e] = gpe::mk(v.x * sc, v.y * sc);
    }
}

// Step counter of the graph-replayed loops (one thread).
__global__ void k_advance(int* base, int inc) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *base += inc;
}

// Per-plane sum |psi|^2 (calc_norm), one block per plane.
template <typename T, int NT>
__global__ void __launch_bounds__(NT) k_norm_planes(const cplx<T>* psi, int npl, double* part) {
    __shared__ double red[NT / 64 + 1];
    const cplx<T>* p = psi + (size_t)blockIdx.x * npl;
    double acc = 0.0;
    for (int e = threadIdx.x; e < npl; e += NT) {
        cplx<T> v = p[e];
        acc += (double)v.x * v.x + (double)v.y * v.y;
    }
    const double s = block_sum<NT>(acc, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// RT potential half step, interior only (cpu_solver.cpp:99-131).
template <typename T>
__global__ void k_potential_half(cplx<T>* psi, const T* V, const T* phi, int nx, int ny, int nz, double hdt,
                                 double cdd, double pg, double pgm) {
    const size_t n = (size_t)nx * ny * nz;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const int k = e % nz;
        const int j = (e / nz) % ny;
        const int i = e / ((size_t)ny * nz);
        if (i < 1 || i > nx - 2 || j < 1 || j > ny - 2 || k < 1 || k > nz - 2) continue;
        cplx<T> c = psi[e];
        const double dn = dens64(c);
        const double veff = (double)V[e] + cdd * (double)phi[e] + pg * dn + pgm * (dn * sqrt(dn));
        psi[e] = gpe::phase_rot(c, -hdt * veff);
    }
}

// Energies (cpu_solver.cpp:137-194) and <psi|H|psi>, per plane, FP64 sums.
// pass 0: out[i*8 + {0..5}] = {sum Re(-Lap psi conj psi) (no 1/2m), sum V|phi|^2,
//          sum |phi|^4, sum Phi|phi|^2, sum |phi|^5, sum Re(conj phi H phi)}
// pass 1: out[i*8 + 6] = sum |H phi - mu phi|^2
template <typename T, int NT>
__device__ __forceinline__ void energy_plane(const cplx<T>* psi, const T* V, const T* phi, int i, int nx, int ny,
                                             int nz, double ix2, double iy2, double iz2, double m, double cdd,
                                             double gq, double gmq, double mu, int pass, double* out, double* red) {
    const int npl = ny * nz;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    if (i >= 1 && i <= nx - 2) {
        const size_t off = (size_t)i * npl;
        for (int p = threadIdx.x; p < npl; p += NT) {
            const int j = p / nz, k = p - j * nz;
            if (j < 1 || j > ny - 2 || k < 1 || k > nz - 2) continue;
            const size_t e = off + p;
            const cplx<T> cT = psi[e];
            const double cr = cT.x, ci = cT.y;
            auto ld = [&](size_t t, double& r, double& im) {
                cplx<T> v = psi[t];
                r = v.x;
                im = v.y;
            };
            double xmr, xmi, xpr, xpi, ymr, ymi, ypr, ypi, zmr, zmi, zpr, zpi;
            ld(e - npl, xmr, xmi);
            ld(e + npl, xpr, xpi);
            ld(e - nz, ymr, ymi);
            ld(e + nz, ypr, ypi);
            ld(e - 1, zmr, zmi);
            ld(e + 1, zpr, zpi);
            // -Lap psi
            const double lr = -(xmr + xpr - 2 * cr) * ix2 - (ymr + ypr - 2 * cr) * iy2 - (zmr + zpr - 2 * cr) * iz2;
            const double li = -(xmi + xpi - 2 * ci) * ix2 - (ymi + ypi - 2 * ci) * iy2 - (zmi + zpi - 2 * ci) * iz2;
            const double nrm = cr * cr + ci * ci;
            const double ab = sqrt(nrm);
            const double v = (double)V[e];
            const double ph = (double)phi[e];
            const double f = v + cdd * ph + gq * nrm + gmq * nrm * ab;
            const double hr = lr / (2 * m) + f * cr;
            const double hi = li / (2 * m) + f * ci;
            if (pass == 0) {
                acc[0] += lr * cr + li * ci;
                acc[1] += v * nrm;
                acc[2] += nrm * nrm;
                acc[3] += ph * nrm;
                acc[4] += nrm * nrm * ab;
                acc[5] += cr * hr + ci * hi;
            } else {
                const double rr = hr - mu * cr, ri = hi - mu * ci;
                acc[0] += rr * rr + ri * ri;
            }
        }
    }
    const int nv = pass == 0 ? 6 : 1;
    for (int t = 0; t < nv; ++t) {
        const double s = block_sum<NT>(acc[t], red);
        if (threadIdx.x == 0) out[(size_t)i * 8 + (pass == 0 ? t : 6)] = s;
    }
}

template <typename T, int NT>
__global__ void __launch_bounds__(NT) k_energy(const cplx<T>* psi, const T* V, const T* phi, int nx, int ny, int nz,
                                               double ix2, double iy2, double iz2, double m, double cdd, double gq,
                                               double gmq, double mu, int pass, double* out,
                                               const double* mu_dev = nullptr) {
    __shared__ double red[NT / 64 + 1];
    if (mu_dev) mu = *mu_dev;
    energy_plane<T, NT>(psi, V, phi, blockIdx.x, nx, ny, nz, ix2, iy2, iz2, m, cdd, gq, gmq, mu, pass, out, red);
}

// Per-instance arguments of the batched energy check: the state (scaled in
// place by 1/sqrt(sum of its norm partials), as k_scale does), the k_energy
// inputs, the partials (the instance's epart), and sums / out (8 doubles
// each) of k_energy_sums.
template <typename T>
struct EArgs {
    cplx<T>* psi[2];  // the two state buffers; the batch's parity picks one
    const T* V;
    const T* phi;
    const double* npart[2];  // per-plane sum |phi|^2 of psi, by parity
    double* part;
    double* sums;
    double* out;
    size_t n;
    int nx, ny, nz;
    double ix2, iy2, iz2, m, cdd, gq, gmq;
    double N, dV, g, gamma, dv15, n25;  // dv15 = dV^1.5, n25 = N^2.5 (host pow)
};

constexpr int kScaleChunk = 16384;  // elements per block of k_scale_group

// k_scale over the active instances of a batch: work (instance, chunk).
template <typename T>
__global__ void __launch_bounds__(256) k_scale_group(const EArgs<T>* __restrict__ args, const int2* __restrict__ work,
                                                     int cur, int pcur) {
    __shared__ double sred[2];
    const int2 w = work[blockIdx.x];
    const EArgs<T>& a = args[w.x];
    const T sc = (T)(1.0 / sqrt(sum_partials(a.npart[pcur], a.nx, sred)));
    cplx<T>* psi = a.psi[cur];
    const size_t e1 = min(a.n, (size_t)(w.y + 1) * kScaleChunk);
    for (size_t e = (size_t)w.y * kScaleChunk + threadIdx.x; e < e1; e += 256) {
        cplx<T> v = psi[e];
        psi[e] = gpe::mk(v.x * sc, v.y * sc);
    }
}

// k_energy over the active instances: work (instance, plane); pass 1 takes
// mu from the instance's sums[5].
template <typename T, int NT>
__global__ void __launch_bounds__(NT) k_energy_group(const EArgs<T>* __restrict__ args, const int2* __restrict__ work,
                                                     int pass, int cur) {
    __shared__ double red[NT / 64 + 1];
    const int2 w = work[blockIdx.x];
    const EArgs<T>& a = args[w.x];
    const double mu = pass ? a.sums[5] : 0.0;
    energy_plane<T, NT>(a.psi[cur], a.V, a.phi, w.y, a.nx, a.ny, a.nz, a.ix2, a.iy2, a.iz2, a.m, a.cdd, a.gq, a.gmq, mu,
                        pass, a.part, red);
}

// Energies from the per-plane partials of k_energy, on the device: pass 0
// adds the six pass-0 sums in plane order into sums[0..5] (mu = sums[5]);
// pass 1 adds the residual sums and writes the eight outputs (gpe_energies
// order) with the formulas and operation order of do_energy.
__device__ __forceinline__ void energy_sums(const double* part, int nx, int pass, double* sums, double* out, double m,
                                            double N, double dV, double g, double gamma, double cdd, double dv15,
                                            double n25) {
#pragma clang fp contract(off)
    if (pass == 0) {
        double s[6] = {0, 0, 0, 0, 0, 0};
        for (int i = 0; i < nx; ++i)
            for (int t = 0; t < 6; ++t) s[t] += part[(size_t)i * 8 + t];
        for (int t = 0; t < 6; ++t) sums[t] = s[t];
        return;
    }
    double r2 = 0;
    for (int i = 0; i < nx; ++i) r2 += part[(size_t)i * 8 + 6];
    const double* s = sums;
    const double e_kin = s[0] / (2 * m) * N;
    const double e_pot = s[1] * N;
    const double e_int = 0.5 * g * s[2] / dV * N * N;
    const double e_ext = 0.5 * cdd * s[3] * N - cdd / 3. * s[2] / dV / 2. * N * N;
    const double e_bmf = 2. / 5. * gamma * s[4] / dv15 * n25;
    out[0] = e_kin;
    out[1] = e_pot;
    out[2] = e_int;
    out[3] = e_ext;
    out[4] = e_bmf;
    out[5] = e_kin + e_pot + e_int + e_ext + e_bmf;
    out[6] = s[5];
    out[7] = sqrt(r2);
}

__global__ void k_energy_sums(const double* part, int nx, int pass, double* sums, double* out, double m, double N,
                              double dV, double g, double gamma, double cdd, double dv15, double n25) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    energy_sums(part, nx, pass, sums, out, m, N, dV, g, gamma, cdd, dv15, n25);
}

// k_energy_sums of every active instance: block q, one thread.
template <typename T>
__global__ void k_energy_sums_group(const EArgs<T>* __restrict__ args, int pass) {
    if (threadIdx.x != 0) return;
    const EArgs<T>& a = args[blockIdx.x];
    energy_sums(a.part, a.nx, pass, a.sums, a.out, a.m, a.N, a.dV, a.g, a.gamma, a.cdd, a.dv15, a.n25);
}

// ---------------------------------------------------------------------------
// On-device initial states and potentials (src/initializer/initializer.cpp:46-248,
// src/parameters/potentials.hpp:76-157).  Coordinates x_i = (i - n/2) d
// (src/context/context.cpp:17-33); the potential sees them truncated toward
// zero to whole atomic units (DataInitializer::_pote_func is a
// std::function<double(int, int, int)>, src/initializer/initializer.hpp:37).
constexpr int kMaxCenters = 64;
struct InitArgs {
    int nx, ny, nz;
    double dx, dy, dz;
    int key;  // potential key (GPE_POT_*) or initial state (GPE_INIT_*)
    // potential: m, b = m wx^2/2, aa = m wx^2/(4 dd^2), trap frequencies, slope
    double m, b, aa, omx, omy, omz, slope;
    // Gaussian sums: centres, widths, whether z is taken about the centre
    int ncent, zcent;
    double cx[kMaxCenters], cy[kMaxCenters], cz[kMaxCenters];
    double sx, sy, sz;
    double cos_dd, cos_r;  // COS: x offset dd, rrr = (nx/2) dx
    double scale;          // sqrt(dV): the state is stored as phi = sqrt(dV) psi
    double noise;          // phase noise amplitude (rad)
    unsigned long long seed;
};

__device__ __forceinline__ double coord_of(int i, int n, double d) { return (double)(i - n / 2) * d; }

template <typename T>
__global__ void k_init_pote(InitArgs a, T* V) {
    const size_t n = (size_t)a.nx * a.ny * a.nz;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(e % a.nz), j = (int)((e / a.nz) % a.ny), i = (int)(e / ((size_t)a.ny * a.nz));
        const double x = trunc(coord_of(i, a.nx, a.dx)), y = trunc(coord_of(j, a.ny, a.dy)),
                     z = trunc(coord_of(k, a.nz, a.dz));
        const double vy = 0.5 * a.m * (y * y) * (a.omy * a.omy), vz = 0.5 * a.m * (z * z) * (a.omz * a.omz);
        const double x2 = x * x, x4 = x2 * x2;
        double v;
        switch (a.key) {
        case GPE_POT_HARMONIC: v = 0.5 * a.m * x2 * (a.omx * a.omx) + vy + vz; break;
        case GPE_POT_MEXICAN: v = -a.b * x2 + a.aa * x4 + vy + vz; break;
        case GPE_POT_MEXICAN_FREE: v = a.aa * x4 + vy + vz; break;
        case GPE_POT_CYLINDRICAL: {
            const double w = (a.omy + a.omx) / 2.;
            v = 0.5 * a.m * (x2 + y * y) * (w * w) + vz;
            break;
        }
        case GPE_POT_MEXICAN_ASYMETRIC: v = -a.b * x2 + a.aa * x4 - a.slope * x + vy + vz; break;
        case GPE_POT_MEXICAN_ASYMETRIC_FREE: v = a.aa * x4 - a.slope * x + vy + vz; break;
        default: v = vz; break;  // FREE
        }
        V[e] = (T)v;
    }
}

// SplitMix64 (public-domain generator of Steele, Lea and Flood): the e-th
// output of the stream that starts at `seed`, as a double in [0, 1).
__device__ __forceinline__ double splitmix_u01(unsigned long long seed, unsigned long long e) {
    unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (e + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

// psi_0 on the interior points (boundary zero), times the phase noise
// exp(i noise (2u - 1)), stored as phi = sqrt(dV) psi into both buffers
// (cpsi = cpsii = wvf, src/solver/solver.cpp:118-125).
template <typename T>
__global__ void k_init_psi(InitArgs a, cplx<T>* psi, cplx<T>* psii) {
    const size_t n = (size_t)a.nx * a.ny * a.nz;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(e % a.nz), j = (int)((e / a.nz) % a.ny), i = (int)(e / ((size_t)a.ny * a.nz));
        double val = 0.0;
        if (i >= 1 && i <= a.nx - 2 && j >= 1 && j <= a.ny - 2 && k >= 1 && k <= a.nz - 2) {
            const double x = coord_of(i, a.nx, a.dx), y = coord_of(j, a.ny, a.dy), z = coord_of(k, a.nz, a.dz);
            if (a.key == GPE_INIT_COS) {
                val = cos(2 * M_PI * (x - a.cos_dd) / a.cos_r) * cos(2 * M_PI * y / a.cos_r) * cos(1 * M_PI * z / a.cos_r);
            } else {
                for (int c = 0; c < a.ncent; ++c) {
                    const double xc = x - a.cx[c], yc = y - a.cy[c], zc = a.zcent ? z - a.cz[c] : z;
                    val += exp(-0.5 * (xc * xc) / (a.sx * a.sx)) * exp(-0.5 * (yc * yc) / (a.sy * a.sy)) *
                           exp(-0.5 * (zc * zc) / (a.sz * a.sz));
                }
            }
        }
        double re = val, im = 0.0;
        if (a.noise != 0.0) {
            const double th = a.noise * (2.0 * splitmix_u01(a.seed, e) - 1.0);
            re = val * cos(th);
            im = val * sin(th);
        }
        const cplx<T> c = gpe::mk((T)(re * a.scale), (T)(im * a.scale));
        psi[e] = c;
        psii[e] = c;
    }
}

// Scratch of gpe_calc_energy_noalloc: per-plane partials (8 doubles per
// plane) and the pass-0 sums.  Host code never uses this name as an address:
// it takes the per-device address from hipGetSymbolAddress (energy_scratch).
constexpr int kNoAllocPlanes = 4096;
static __device__ __align__(256) double scratch[kNoAllocPlanes * 8 + 8];

// Unfused IT step (the two hooks imag_iter_linear_step + imag_iter_nonlinear_step):
// out interior from in (already normalised), out boundary from psii.
template <typename T>
__global__ void k_it_stencil(const cplx<T>* in, cplx<T>* out, const cplx<T>* psii, const T* V, const T* phi, int nx,
                             int ny, int nz, double dtau_, double cx_, double cy_, double cz_, double cdd_, double gq_,
                             double gmq_) {
    const size_t n = (size_t)nx * ny * nz;
    const int npl = ny * nz;
    const T dtau = (T)dtau_, cx = (T)cx_, cy = (T)cy_, cz = (T)cz_, cdd = (T)cdd_, gq = (T)gq_, gmq = (T)gmq_;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const int k = e % nz;
        const int j = (e / nz) % ny;
        const int i = e / ((size_t)npl);
        if (i < 1 || i > nx - 2 || j < 1 || j > ny - 2 || k < 1 || k > nz - 2) {
            out[e] = psii[e];
            continue;
        }
        const cplx<T> c = in[e];
        const cplx<T> xm = in[e - npl], xp = in[e + npl], ym = in[e - nz], yp = in[e + nz], zm = in[e - 1],
                      zp = in[e + 1];
        const T v = V[e] + cdd * phi[e];
        cplx<T> c1;
        c1.x = cx * (xm.x + xp.x - T(2) * c.x) + cy * (ym.x + yp.x - T(2) * c.x) + cz * (zm.x + zp.x - T(2) * c.x) +
               c.x * v;
        c1.y = cx * (xm.y + xp.y - T(2) * c.y) + cy * (ym.y + yp.y - T(2) * c.y) + cz * (zm.y + zp.y - T(2) * c.y) +
               c.y * v;
        const T nrm = c.x * c.x + c.y * c.y;
        const T ab = sqrt(nrm);
        const T f = gq * nrm + gmq * (ab * ab * ab);
        cplx<T> r;
        r.x = c.x - dtau * c1.x;
        r.y = c.y - dtau * c1.y;
        r.x = r.x - dtau * (f * c.x);
        r.y = r.y - dtau * (f * c.y);
        out[e] = r;
    }
}

}  // namespace

// ===========================================================================
// Host side
// ===========================================================================

#define GPE_CHECK(call)                                                                             \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess) {                                                                     \
            return fail(h, std::string(#call) + ": " + hipGetErrorString(e_));                      \
        }                                                                                           \
    } while (0)

struct gpe_handle {
    gpe_params p;
    int prec;
    int dev = 0;  // the device current at gpe_create
    size_t N;
    hipStream_t stream = nullptr;
    bool own_stream = true;
    std::string err;
    // device buffers (bytes allocated per precision)
    void* store[2] = {nullptr, nullptr};
    int cur = 0;
    void* psii = nullptr;
    void* V = nullptr;
    void* phi = nullptr;
    void* s1 = nullptr;
    void* tk = nullptr;
    void* s1b = nullptr;  // second S1 / Tk of the split plane pass (allocated on first use)
    void* tkb = nullptr;
    void* tw = nullptr;
    void* twl = nullptr;  // lo parts of the W_nx, W_ny, W_nz twiddles, [nx][ny][nz]
    void* ex = nullptr;
    void* ey = nullptr;
    void* ez = nullptr;
    double* kx2 = nullptr;
    double* ky2 = nullptr;
    double* kz2 = nullptr;
    double* part[2] = {nullptr, nullptr};
    double* epart = nullptr;
    int pcur = 0;
    bool norm_pending = false;  // true: normalised psi = store[cur]/sqrt(sum part[pcur])
    bool force_generic = false; // GPE_FORCE_GENERIC=1: run-time-size kernels only
    bool line_reg = true;       // register-resident x pass (GPE_LINE_LDS=1 disables)
    bool plane_split = true;    // two blocks per plane in steady RT / IT steps (GPE_NO_SPLIT=1 disables)
    bool pencil = false;        // plane beyond one block's LDS: the pencil plane pass
    void* pen_gw = nullptr;     // its per-plane intermediates (N complex each)
    void* pen_h = nullptr;
    double* pen_part = nullptr;  // norm partial per (plane, row tile)
    int pen_c = 0, pen_r = 0;   // column / row tile shapes
    int pen_c_max = 0, pen_r_max = 0;  // the bounds the pencil kernels were built for
    bool pen_yreg = true;       // register y passes of the pencil pass (GPE_PENCIL_YREG=0: LDS column tiles)
    bool pen_zreg = true;       // register z passes of its row tiles (GPE_PENCIL_ZREG=0: LDS row tiles)
    size_t pen_lds = 0;
    LinePlan px, py, pz, px2, py2, pz2;
    // launch shapes: kz / k columns per LDS x-pass tile, and their bounds
    int cp = 0, ck = 0;
    int cp_max = 0, ck_max = 0;
    size_t plane_lds = 0, line_lds = 0;
    size_t plane_stage = 0;  // extra LDS of the compiled plane shapes (staged tables)
    int tw_n = 0;            // entries of the twiddle table
    // scaled constants
    double sqdv = 1.0;
    double ggp11, gamma;
    // graph-replayed steady state (chunks of kGraphSteps steps)
    double* coef = nullptr;      // per-step coupling table (2 doubles per step)
    size_t coef_cap = 0;         // steps the table holds
    double* coef_host = nullptr; // pinned staging of the table (coef_cap + 2 pairs)
    hipEvent_t coef_ev = nullptr; // recorded after the last upload from coef_host
    bool coef_const = false;     // the table holds coef_n copies of coef_cval
    double coef_cval[2] = {0, 0};
    size_t coef_n = 0;
    // Real time: the closing half phase of the last step of a call is left
    // pending and merged into the first step of the next call (the two half
    // steps between consecutive steps share Phi and the density).  The state
    // is then the x-transformed Poisson spectrum and kinetic array of parity
    // rt_par; materialize() applies the half phase and writes psi and fi3d.
    bool rt_pending = false;
    int rt_par = 0;
    double rt_pend[2] = {0, 0};  // scaled couplings (g', gamma') of that half phase
    // Snapshots (gpe_snapshot_begin / _end): the state and its energies are
    // copied on the device behind the launched work, then to pinned host
    // memory on a second stream, so the loop keeps running meanwhile.
    hipStream_t copy_stream = nullptr;
    hipEvent_t snap_ev = nullptr, copy_ev = nullptr;
    void* snap = nullptr;           // device copy of the state (bulk precision, phi units)
    double* snap_dev = nullptr;     // device: sums[8], energies[8]
    void* snap_host = nullptr;      // pinned host copy of the state
    double* snap_host_e = nullptr;  // pinned host copy of the energies
    bool snap_in_flight = false;
    int* cbase = nullptr;        // device step counter read by graph nodes
    hipGraphExec_t rt_graph = nullptr;
    hipGraphExec_t it_graph[4] = {nullptr, nullptr, nullptr, nullptr};
    double it_graph_key[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    bool use_graphs = true;      // GPE_NO_GRAPH=1 disables
    // timing
    bool timing = false;
    std::vector<hipEvent_t> ev;
    // Per-launch start / stop events of the steady-state plane (0) and line
    // (1) launches, attached to the dispatches themselves
    // (hipExtLaunchKernelGGL), so they time exactly the kernel.
    bool tim_on = false;
    std::vector<hipEvent_t> kev[2][2];
    size_t kn[2] = {0, 0};
    double t_plane_us = 0, t_line_us = 0, t_total_ms = 0;
};

// Message of the last failed call that has no handle (create, batch
// create), per host thread, so concurrent creates do not mix messages.
static thread_local std::string g_create_err;

// Device address of `scratch` per device, set once per device at gpe_create
// (with that device current): each device holds its own copy of the array.
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch[kMaxDevices] = {};
static hipError_t init_scratch(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[dev]) return hipSuccess;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e == hipSuccess) g_scratch[dev] = static_cast<double*>(ptr);
    return e;
}
static double* energy_scratch(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch[dev];
}


static int fail(gpe_handle* h, const std::string& msg) {
    if (h)
        h->err = msg;
    else
        g_create_err = msg;
    return 1;
}

static bool make_plan(int L, LinePlan& pl) {
    pl.L = L;
    pl.nst = 0;
    int r = L;
    while (r % 4 == 0 && pl.nst < gpe::kMaxStages) {
        pl.radix[pl.nst++] = 4;
        r /= 4;
    }
    while (r % 2 == 0 && pl.nst < gpe::kMaxStages) {
        pl.radix[pl.nst++] = 2;
        r /= 2;
    }
    while (r % 3 == 0 && pl.nst < gpe::kMaxStages) {
        pl.radix[pl.nst++] = 3;
        r /= 3;
    }
    while (r % 5 == 0 && pl.nst < gpe::kMaxStages) {
        pl.radix[pl.nst++] = 5;
        r /= 5;
    }
    while (r % 7 == 0 && pl.nst < gpe::kMaxStages) {
        pl.radix[pl.nst++] = 7;
        r /= 7;
    }
    return r == 1;
}

static size_t csize(const gpe_handle* h) { return h->prec == GPE_FP64 ? 16 : 8; }
static size_t rsize(const gpe_handle* h) { return h->prec == GPE_FP64 ? 8 : 4; }

// Host FP64 array -> device array of the handle's precision (optionally scaled).
static int upload_real(gpe_handle* h, void* dst, const double* src, size_t n, double scale) {
    if (h->prec == GPE_FP64) {
        std::vector<double> b(n);
        for (size_t t = 0; t < n; ++t) b[t] = src[t] * scale;
        GPE_CHECK(hipMemcpyAsync(dst, b.data(), n * 8, hipMemcpyHostToDevice, h->stream));
        GPE_CHECK(hipStreamSynchronize(h->stream));
    } else {
        std::vector<float> b(n);
        for (size_t t = 0; t < n; ++t) b[t] = (float)(src[t] * scale);
        GPE_CHECK(hipMemcpyAsync(dst, b.data(), n * 4, hipMemcpyHostToDevice, h->stream));
        GPE_CHECK(hipStreamSynchronize(h->stream));
    }
    return 0;
}

static int download_real(gpe_handle* h, double* dst, const void* src, size_t n, double scale) {
    if (h->prec == GPE_FP64) {
        std::vector<double> b(n);
        GPE_CHECK(hipMemcpyAsync(b.data(), src, n * 8, hipMemcpyDeviceToHost, h->stream));
        GPE_CHECK(hipStreamSynchronize(h->stream));
        for (size_t t = 0; t < n; ++t) dst[t] = b[t] * scale;
    } else {
        std::vector<float> b(n);
        GPE_CHECK(hipMemcpyAsync(b.data(), src, n * 4, hipMemcpyDeviceToHost, h->stream));
        GPE_CHECK(hipStreamSynchronize(h->stream));
        for (size_t t = 0; t < n; ++t) dst[t] = (double)b[t] * scale;
    }
    return 0;
}

// Table entry for the complex constant c + i s.  FP64: rounded to nearest.
// FP32: among the FP32 pairs within one unit in the last place of (c, s),
// the one whose magnitude is closest to sqrt(c^2 + s^2), ties and near ties
// settled by `wang` times the angle error.  Every table entry enters a
// unitary product (a transform pass, a phase factor); nearest rounding
// leaves |w| short of its exact value for most angles used (cos = sin at
// pi/4, for instance), and the products then lose norm every step.
template <typename T>
static cplx<T> fit_unit(double c, double s, double wang) {
    cplx<T> r;
    r.x = (T)c;
    r.y = (T)s;
    if (sizeof(T) == 8) return r;
    const double m2 = c * c + s * s, m = sqrt(m2);
    double best = 1e300;
    const float c0 = (float)c, s0 = (float)s;
    const float cs[3] = {nextafterf(c0, -INFINITY), c0, nextafterf(c0, INFINITY)};
    const float ss[3] = {nextafterf(s0, -INFINITY), s0, nextafterf(s0, INFINITY)};
    for (float cc : cs)
        for (float sv : ss) {
            const double dm = fabs(((double)cc * cc + (double)sv * sv) - m2) / (2.0 * m2);
            // angle error: component of the error orthogonal to (c, s), relative to |w|
            const double da = fabs(((double)sv - s) * c - ((double)cc - c) * s) / m2;
            const double cost = dm + wang * da;
            if (cost < best) {
                best = cost;
                r.x = (T)cc;
                r.y = (T)sv;
            }
        }
    return r;
}

template <typename T>
static void fill_phase(std::vector<cplx<T>>& out, int n, double d, double dt, double m) {
    out.resize(n);
    for (int t = 0; t < n; ++t) {
        const double k = (t <= n / 2) ? 2.0 * M_PI * t / (n * d) : 2.0 * M_PI * (t - n) / (n * d);
        const double f = -dt / (2.0 * m) * k * k;
        // (cos f - 1, sin f); the kernels apply (1 + e) / n
        const double sh = sin(0.5 * f);
        out[t].x = (T)(-2.0 * sh * sh);
        out[t].y = (T)sin(f);
    }
}

template <typename T>
static int setup_tables(gpe_handle* h) {
    const gpe_params& p = h->p;
    // twiddles for the six line lengths
    const int lens[6] = {2 * p.nx, 2 * p.ny, 2 * p.nz, p.nx, p.ny, p.nz};
    LinePlan* pls[6] = {&h->px2, &h->py2, &h->pz2, &h->px, &h->py, &h->pz};
    std::vector<cplx<T>> tw;
    for (int t = 0; t < 6; ++t) {
        if (!make_plan(lens[t], *pls[t])) return fail(h, "grid length " + std::to_string(lens[t]) +
                                                             " is not of the form 2^a 3^b 5^c 7^d");
        pls[t]->tw = (int)tw.size();
        for (int m = 0; m < lens[t]; ++m) {
            const double ang = 2.0 * M_PI * m / lens[t];
            // a twiddle's angle error largely cancels between a forward pass
            // and the inverse pass with the conjugate twiddle; its magnitude
            // error does not
            tw.push_back(fit_unit<T>(cos(ang), -sin(ang), 0.25));
        }
    }
    h->tw_n = (int)tw.size();
    GPE_CHECK(hipMalloc(&h->tw, tw.size() * sizeof(cplx<T>)));
    GPE_CHECK(hipMemcpy(h->tw, tw.data(), tw.size() * sizeof(cplx<T>), hipMemcpyHostToDevice));
    // lo parts W_L^m - tw[m] of the psi transform lengths (FP32: the
    // transforms of psi multiply by tw + twl, see tw_psi)
    std::vector<cplx<T>> twl;
    for (int t = 3; t < 6; ++t) {
        for (int m = 0; m < lens[t]; ++m) {
            const double ang = 2.0 * M_PI * m / lens[t];
            const cplx<T> hi = tw[pls[t]->tw + m];
            cplx<T> lo;
            lo.x = (T)(cos(ang) - (double)hi.x);
            lo.y = (T)(-sin(ang) - (double)hi.y);
            twl.push_back(lo);
        }
    }
    GPE_CHECK(hipMalloc(&h->twl, twl.size() * sizeof(cplx<T>)));
    GPE_CHECK(hipMemcpy(h->twl, twl.data(), twl.size() * sizeof(cplx<T>), hipMemcpyHostToDevice));
    // kinetic phase tables (fftw_split_solver.cpp:11-42, factorised per axis)
    std::vector<cplx<T>> e;
    fill_phase<T>(e, p.nx, p.dx, p.real_time_dt, p.m);
    GPE_CHECK(hipMalloc(&h->ex, e.size() * sizeof(cplx<T>)));
    GPE_CHECK(hipMemcpy(h->ex, e.data(), e.size() * sizeof(cplx<T>), hipMemcpyHostToDevice));
    fill_phase<T>(e, p.ny, p.dy, p.real_time_dt, p.m);
    GPE_CHECK(hipMalloc(&h->ey, e.size() * sizeof(cplx<T>)));
    GPE_CHECK(hipMemcpy(h->ey, e.data(), e.size() * sizeof(cplx<T>), hipMemcpyHostToDevice));
    fill_phase<T>(e, p.nz, p.dz, p.real_time_dt, p.m);
    GPE_CHECK(hipMalloc(&h->ez, e.size() * sizeof(cplx<T>)));
    GPE_CHECK(hipMemcpy(h->ez, e.data(), e.size() * sizeof(cplx<T>), hipMemcpyHostToDevice));
    // padded-lattice k^2 per axis (fftw_poisson_solver.cpp:12-49), FP64
    const int NX = 2 * p.nx, NY = 2 * p.ny, NZ = 2 * p.nz;
    std::vector<double> kx(NX), ky(NY), kz(p.nz + 1);
    const double dkx = 2. * M_PI / (NX * p.dx), dky = 2. * M_PI / (NY * p.dy), dkz = 2. * M_PI / (NZ * p.dz);
    for (int t = 0; t < NX; ++t) {
        const double k = (t <= NX / 2) ? t * dkx : (t - NX) * dkx;
        kx[t] = k * k;
    }
    for (int t = 0; t < NY; ++t) {
        const double k = (t <= NY / 2) ? t * dky : (t - NY) * dky;
        ky[t] = k * k;
    }
    for (int t = 0; t <= p.nz; ++t) {
        const double k = t * dkz;
        kz[t] = k * k;
    }
    GPE_CHECK(hipMalloc(&h->kx2, NX * 8));
    GPE_CHECK(hipMalloc(&h->ky2, NY * 8));
    GPE_CHECK(hipMalloc(&h->kz2, (p.nz + 1) * 8));
    GPE_CHECK(hipMemcpy(h->kx2, kx.data(), NX * 8, hipMemcpyHostToDevice));
    GPE_CHECK(hipMemcpy(h->ky2, ky.data(), NY * 8, hipMemcpyHostToDevice));
    GPE_CHECK(hipMemcpy(h->kz2, kz.data(), (p.nz + 1) * 8, hipMemcpyHostToDevice));
    return 0;
}

// The z propagation of the kinetic step runs in the register x pass when its
// kinetic tiles hold whole z rows (LPB_K = nz: nx = 128 with nz = 32, nx = 256
// with nz = 16) and the plane pass is a compiled single-block shape; the
// plane pass then leaves it out.  Plane and x pass arguments both read this,
// so the pair of launches of one step always agrees.
static bool kz_in_x(const gpe_handle* h) {
    if (h->force_generic || !h->line_reg || h->pencil) return false;
    const int nx = h->p.nx, ny = h->p.ny, nz = h->p.nz;
    if (!((nx == 128 && nz == 32) || (nx == 256 && nz == 16))) return false;
#define GPE_PL(NY, NZ) \
    if (ny == NY && nz == NZ) return true;
    GPE_PLANE_SHAPES(GPE_PL)
#undef GPE_PL
    return false;
}

// XCD-aware plane order of k_plane (GPE_XCD_MAP=0 turns it off).
static bool xcd_map_on() {
    static const bool on = [] {
        const char* e = getenv("GPE_XCD_MAP");
        return !(e && e[0] == '0');
    }();
    return on;
}

template <typename T>
static PlaneArgs<T> plane_args(gpe_handle* h, int mode) {
    PlaneArgs<T> a;
    memset(&a, 0, sizeof(a));
    const gpe_params& p = h->p;
    const double dV = p.dx * p.dy * p.dz;
    a.nx = p.nx;
    a.ny = p.ny;
    a.nz = p.nz;
    a.mode = mode;
    a.generic = h->force_generic ? 1 : 0;
    a.psii = (const cplx<T>*)h->psii;
    a.tk = a.tko = (cplx<T>*)h->tk;
    a.s1 = a.s1o = (cplx<T>*)h->s1;
    a.split = 0;
    // (imaginary time only: its stencil reads the neighbour planes; real time
    // measured no gain)
    a.xcd = (p.nx % 8 == 0 && (mode & M_IT) && xcd_map_on()) ? 1 : 0;
    a.kz_skip = kz_in_x(h) ? 1 : 0;
    a.V = (const T*)h->V;
    a.phi = (T*)h->phi;
    a.tw = (const cplx<T>*)h->tw;
    a.tw_n = h->tw_n;
    a.ez = (const cplx<T>*)h->ez;
    a.ey = (const cplx<T>*)h->ey;
    a.twly = (const cplx<T>*)h->twl + p.nx;
    a.twlz = (const cplx<T>*)h->twl + p.nx + p.ny;
    a.pen_gw = (cplx<T>*)h->pen_gw;
    a.pen_h = (cplx<T>*)h->pen_h;
    a.pen_c = h->pen_c;
    a.pen_r = h->pen_r;
    a.pen_part = h->pen_part;
    a.pz2 = h->pz2;
    a.py2 = h->py2;
    a.pz = h->pz;
    a.py = h->py;
    a.pinv_scale = (p.n_atoms / dV) / (8.0 * (double)h->N);
    a.cdd = p.cdd;
    a.dtau = p.imag_time_dt;
    a.cx = -0.5 / (p.m * p.dx * p.dx);
    a.cy = -0.5 / (p.m * p.dy * p.dy);
    a.cz = -0.5 / (p.m * p.dz * p.dz);
    a.gq = (h->ggp11 - p.cdd / 3) * p.n_atoms / dV;
    a.gmq = h->gamma * p.w_15 / pow(dV, 1.5);
    a.hdt = p.real_time_dt / 2.0;
    a.cv = 1.0;
    a.pg = a.gq;
    a.pgm = a.gmq;
    return a;
}

template <typename T>
static LineArgs<T> line_args(gpe_handle* h, bool poisson, bool kinetic) {
    LineArgs<T> a;
    memset(&a, 0, sizeof(a));
    const gpe_params& p = h->p;
    a.nx = p.nx;
    a.ny = p.ny;
    a.nz = p.nz;
    a.generic = h->force_generic ? 1 : 0;
    a.cp = h->cp;
    a.ck = h->ck;
    a.tiles_kz = (p.nz + h->cp - 1) / h->cp;  // S1 columns kz = 1..nz
    a.tiles_k = (p.nz + h->ck - 1) / h->ck;
    a.ntile_p = poisson ? 2 * p.ny * a.tiles_kz : 0;
    a.s1 = (cplx<T>*)h->s1;
    a.tk = (cplx<T>*)h->tk;
    a.tw = (const cplx<T>*)h->tw;
    a.ex = (const cplx<T>*)h->ex;
    a.twlx = (const cplx<T>*)h->twl;
    a.kx2 = h->kx2;
    a.ky2 = h->ky2;
    a.kz2 = h->kz2;
    a.px2 = h->px2;
    a.px = h->px;
    a.ez = (const cplx<T>*)h->ez;
    a.kinz = kz_in_x(h) ? 1 : 0;
    (void)kinetic;
    return a;
}

// Start / stop events for the next timed launch of kind k (0 plane, 1 line),
// or nulls when this launch is not timed.
static void kernel_events(gpe_handle* h, int k, hipEvent_t* e0, hipEvent_t* e1) {
    *e0 = *e1 = nullptr;
    if (!h->timing || !h->tim_on) return;
    const size_t i = h->kn[k]++;
    for (int w = 0; w < 2; ++w)
        while (h->kev[k][w].size() <= i) {
            hipEvent_t e;
            (void)hipEventCreate(&e);
            h->kev[k][w].push_back(e);
        }
    *e0 = h->kev[k][0][i];
    *e1 = h->kev[k][1][i];
}

// The plane shape has a compiled kernel (PlanePolCT) that the launches use.
static bool plane_compiled(const gpe_handle* h) {
    if (h->force_generic) return false;
    const int ny = h->p.ny, nz = h->p.nz;
#define GPE_PL(NY, NZ) \
    if (ny == NY && nz == NZ) return true;
    GPE_PLANE_SHAPES(GPE_PL)
#undef GPE_PL
    return false;
}

template <typename T, class Pol>
static int launch_plane_pol(gpe_handle* h, const PlaneArgs<T>& a) {
    constexpr int NT = PlaneThreads<T, Pol::kStage>::value;
    const size_t lds = h->plane_lds + (Pol::kStage ? h->plane_stage : 0);
    hipEvent_t e0, e1;
    kernel_events(h, 0, &e0, &e1);
    hipExtLaunchKernelGGL((k_plane<T, NT, Pol>), dim3(h->p.nx * (a.split ? 2 : 1)), dim3(NT), lds, h->stream, e0, e1,
                          0, a);
    GPE_CHECK(hipGetLastError());
    return 0;
}


// The pencil plane pass: three launches; the timing events bracket all three.
template <typename T, class Pol>
static int launch_pencil(gpe_handle* h, const PlaneArgs<T>& a) {
    hipEvent_t e0, e1;
    kernel_events(h, 0, &e0, &e1);
    const int nx = h->p.nx, ny = h->p.ny, nz = h->p.nz;
    const int nct = (nz + h->pen_c - 1) / h->pen_c, nrt = (ny + h->pen_r - 1) / h->pen_r;
    const bool ci = (a.mode & M_PINV) != 0;
    const bool cf = (a.mode & (M_RHO_FWD | M_KIN | M_NORM_OUT)) != 0;
    // register y passes: compiled pencil shapes with ny = 256
    constexpr bool kYReg = !std::is_same<Pol, PencilPolGeneric>::value;
    const bool yreg = kYReg && h->pen_yreg && ny == 256;
    using Y = PenY<256>;
    const int ncb = (nz + Y::LPB - 1) / Y::LPB;
    const size_t ylds = Y::template lds_bytes<T>();
    if (ci) {
        if (yreg)
            hipExtLaunchKernelGGL((k_pencil_yci<T, 256>), dim3(nx * ncb), dim3(256), ylds, h->stream, e0, nullptr, 0,
                                  a);
        else
            hipExtLaunchKernelGGL((k_pencil_ci<T, kPencilThreads, Pol>), dim3(nx * nct), dim3(kPencilThreads),
                                  h->pen_lds, h->stream, e0, nullptr, 0, a);
        GPE_CHECK(hipGetLastError());
    }
    // register z passes: compiled pencil shapes with nz = 64 and 32-row tiles
    const bool zreg = kYReg && h->pen_zreg && nz == PzReg<T>::NZ && h->pen_r == kPzRows && ny % kPzRows == 0;
    if (zreg)
        hipExtLaunchKernelGGL((k_pencil_rz<T>), dim3(nx * nrt), dim3(256), PzReg<T>::lds_bytes(), h->stream,
                              ci ? nullptr : e0, cf ? nullptr : e1, 0, a);
    else
        hipExtLaunchKernelGGL((k_pencil_r<T, kPencilThreads, Pol>), dim3(nx * nrt), dim3(kPencilThreads),
                              h->pen_lds, h->stream, ci ? nullptr : e0, cf ? nullptr : e1, 0, a);
    GPE_CHECK(hipGetLastError());
    if (cf) {
        if (yreg)
            hipExtLaunchKernelGGL((k_pencil_ycf<T, 256>), dim3(2 * nx * ncb), dim3(256), ylds, h->stream, nullptr, e1, 0,
                                  a);
        else
            hipExtLaunchKernelGGL((k_pencil_cf<T, kPencilThreads, Pol>), dim3(2 * nx * nct), dim3(kPencilThreads),
                                  h->pen_lds, h->stream, nullptr, e1, 0, a);
        GPE_CHECK(hipGetLastError());
    }
    return 0;
}

template <typename T>
static int launch_plane(gpe_handle* h, const PlaneArgs<T>& a) {
    const int ny = h->p.ny, nz = h->p.nz;
    if (h->pencil) {
        if (!h->force_generic) {
#define GPE_PP(NY, NZ, CC, RR) \
    if (ny == NY && nz == NZ) return launch_pencil<T, PencilPolCT<NY, NZ, CC, RR>>(h, a);
            GPE_PENCIL_SHAPES(GPE_PP)
#undef GPE_PP
        }
        return launch_pencil<T, PencilPolGeneric>(h, a);
    }
    if (!h->force_generic) {
#define GPE_PL(NY, NZ) \
    if (ny == NY && nz == NZ) return launch_plane_pol<T, PlanePolCT<NY, NZ>>(h, a);
        GPE_PLANE_SHAPES(GPE_PL)
#undef GPE_PL
    }
    return launch_plane_pol<T, PlanePolGeneric>(h, a);
}

template <typename T, class Pol>
static int launch_line_pol(gpe_handle* h, const LineArgs<T>& a, int nblk) {
    hipEvent_t e0, e1;
    kernel_events(h, 1, &e0, &e1);
    hipExtLaunchKernelGGL((k_xpass<T, kLineThreads, Pol>), dim3(nblk), dim3(kLineThreads), h->line_lds, h->stream, e0,
                          e1, 0, a);
    GPE_CHECK(hipGetLastError());
    return 0;
}

template <typename T, int NX>
static int launch_line_reg(gpe_handle* h, const LineArgs<T>& a, bool poisson, bool kinetic) {
    using X = XReg<NX>;
    const int np = poisson ? (2 * h->p.ny * h->p.nz + X::LPB_P - 1) / X::LPB_P : 0;
    const int nk = kinetic ? (h->p.ny * h->p.nz + X::LPB_K - 1) / X::LPB_K : 0;
    if (np + nk == 0) return 0;
    const size_t lds = (size_t)(X::template exa<T>() + 4 * NX) * sizeof(cplx<T>);
    hipEvent_t e0, e1;
    kernel_events(h, 1, &e0, &e1);
    hipExtLaunchKernelGGL((k_xpass_reg<T, NX>), dim3(np + nk), dim3(256), lds, h->stream, e0, e1, 0, a, nk);
    GPE_CHECK(hipGetLastError());
    return 0;
}

template <typename T>
static int launch_line(gpe_handle* h, bool poisson, bool kinetic, void* s1 = nullptr, void* tk = nullptr) {
    LineArgs<T> a = line_args<T>(h, poisson, kinetic);
    if (s1) a.s1 = (cplx<T>*)s1;
    if (tk) a.tk = (cplx<T>*)tk;
    const int nblk = a.ntile_p + (kinetic ? h->p.ny * a.tiles_k : 0);
    if (nblk == 0) return 0;
    const int nx = h->p.nx;
    if (!h->force_generic && h->line_reg) {
        if (nx == 128) return launch_line_reg<T, 128>(h, a, poisson, kinetic);
        if (nx == 64) return launch_line_reg<T, 64>(h, a, poisson, kinetic);
        if (nx == 256) return launch_line_reg<T, 256>(h, a, poisson, kinetic);
    }
    if (!h->force_generic) {
#define GPE_LN(NX) \
    if (nx == NX) return launch_line_pol<T, LinePolCT<NX>>(h, a, nblk);
        GPE_LINE_SHAPES(GPE_LN)
#undef GPE_LN
    }
    return launch_line_pol<T, LinePolGeneric>(h, a, nblk);
}

// S1 / Tk buffers of step parity par: the split plane pass reads pair par
// and writes pair 1 - par (the unsplit pass works in place on s1 / tk).
// Every loop starts on parity 0 (its prologue writes s1).
static void* rt_s1(gpe_handle* h, int par) { return (h->plane_split && par) ? h->s1b : h->s1; }
static void* rt_tk(gpe_handle* h, int par) { return (h->plane_split && par) ? h->tkb : h->tk; }

static int ensure_split_buffers(gpe_handle* h, bool tk) {
    if (!h->plane_split) return 0;
    const size_t cs = csize(h);
    if (!h->s1b) GPE_CHECK(hipMalloc(&h->s1b, (size_t)2 * h->p.ny * h->p.nx * h->p.nz * cs));
    if (tk && !h->tkb) GPE_CHECK(hipMalloc(&h->tkb, h->N * cs));
    return 0;
}

// Apply a pending real-time half phase: finish the Poisson solve of the last
// step's Phi, rotate psi by exp(-i dt/2 Veff) and write psi and fi3d.
template <typename T>
static int flush_rt(gpe_handle* h) {
    if (!h->rt_pending) return 0;
    h->rt_pending = false;
    PlaneArgs<T> a = plane_args<T>(h, M_PINV | M_PHI_STORE | M_IN_TK | M_PHASE_HALF | M_OUT);
    a.s1 = a.s1o = (cplx<T>*)rt_s1(h, h->rt_par);
    a.tk = a.tko = (cplx<T>*)rt_tk(h, h->rt_par);
    a.out = (cplx<T>*)h->store[h->cur];
    a.pg = h->rt_pend[0];
    a.pgm = h->rt_pend[1];
    return launch_plane<T>(h, a);
}

// Bring the state to the reference's form: psi in store[cur] (normalised),
// fi3d in phi.
template <typename T>
static int materialize(gpe_handle* h) {
    if (flush_rt<T>(h)) return 1;
    if (!h->norm_pending) return 0;
    hipLaunchKernelGGL((k_scale<T>), dim3(1024), dim3(256), 0, h->stream, (cplx<T>*)h->store[h->cur], h->N,
                       h->part[h->pcur], h->p.nx);
    GPE_CHECK(hipGetLastError());
    h->norm_pending = false;
    return 0;
}

template <typename T>
static int do_calc_norm(gpe_handle* h, double* xnorma) {
    if (materialize<T>(h)) return 1;
    hipLaunchKernelGGL((k_norm_planes<T, 256>), dim3(h->p.nx), dim3(256), 0, h->stream,
                       (const cplx<T>*)h->store[h->cur], h->p.ny * h->p.nz, h->part[h->pcur]);
    GPE_CHECK(hipGetLastError());
    std::vector<double> part(h->p.nx);
    GPE_CHECK(hipMemcpyAsync(part.data(), h->part[h->pcur], h->p.nx * 8, hipMemcpyDeviceToHost, h->stream));
    GPE_CHECK(hipStreamSynchronize(h->stream));
    double s = 0;
    for (double v : part) s += v;
    if (xnorma) *xnorma = s;  // = dV * sum |psi|^2
    return 0;
}

template <typename T>
static int do_normalize(gpe_handle* h) {
    // normalize() divides by sqrt of the norm computed by the last calc_norm
    // (part[pcur]); the IT loop leaves the partials of the new state there.
    if (flush_rt<T>(h)) return 1;
    hipLaunchKernelGGL((k_scale<T>), dim3(1024), dim3(256), 0, h->stream, (cplx<T>*)h->store[h->cur], h->N,
                       h->part[h->pcur], h->p.nx);
    GPE_CHECK(hipGetLastError());
    h->norm_pending = false;
    return 0;
}

template <typename T>
static int do_calc_fi3d(gpe_handle* h) {
    if (materialize<T>(h)) return 1;
    PlaneArgs<T> a = plane_args<T>(h, M_RHO_FWD);
    a.in = (const cplx<T>*)h->store[h->cur];
    if (launch_plane<T>(h, a)) return 1;
    if (launch_line<T>(h, true, false)) return 1;
    a = plane_args<T>(h, M_PINV | M_PHI_STORE);
    return launch_plane<T>(h, a);
}

template <typename T>
static int do_it_hook(gpe_handle* h) {
    if (materialize<T>(h)) return 1;
    const gpe_params& p = h->p;
    const double dV = p.dx * p.dy * p.dz;
    hipLaunchKernelGGL((k_it_stencil<T>), dim3(2048), dim3(256), 0, h->stream, (const cplx<T>*)h->store[h->cur],
                       (cplx<T>*)h->store[1 - h->cur], (const cplx<T>*)h->psii, (const T*)h->V, (const T*)h->phi,
                       p.nx, p.ny, p.nz, p.imag_time_dt, -0.5 / (p.m * p.dx * p.dx), -0.5 / (p.m * p.dy * p.dy),
                       -0.5 / (p.m * p.dz * p.dz), p.cdd, (h->ggp11 - p.cdd / 3) * p.n_atoms / dV,
                       h->gamma * p.w_15 / pow(dV, 1.5));
    GPE_CHECK(hipGetLastError());
    h->cur = 1 - h->cur;
    return 0;
}

template <typename T>
static int do_potential_half(gpe_handle* h) {
    if (materialize<T>(h)) return 1;
    const gpe_params& p = h->p;
    const double dV = p.dx * p.dy * p.dz;
    hipLaunchKernelGGL((k_potential_half<T>), dim3(2048), dim3(256), 0, h->stream, (cplx<T>*)h->store[h->cur],
                       (const T*)h->V, (const T*)h->phi, p.nx, p.ny, p.nz, p.real_time_dt / 2.0, p.cdd,
                       (h->ggp11 - p.cdd / 3) * p.n_atoms / dV, h->gamma * p.w_15 / pow(dV, 1.5));
    GPE_CHECK(hipGetLastError());
    return 0;
}

template <typename T>
static int do_kinetic(gpe_handle* h) {
    if (materialize<T>(h)) return 1;
    PlaneArgs<T> a = plane_args<T>(h, M_KIN);
    a.in = (const cplx<T>*)h->store[h->cur];
    if (launch_plane<T>(h, a)) return 1;
    if (launch_line<T>(h, false, true)) return 1;
    a = plane_args<T>(h, M_IN_TK | M_OUT);
    a.out = (cplx<T>*)h->store[h->cur];
    return launch_plane<T>(h, a);
}

static void ev_record(gpe_handle* h, size_t idx) {
    if (!h->timing) return;
    while (h->ev.size() <= idx) {
        hipEvent_t e;
        (void)hipEventCreate(&e);
        h->ev.push_back(e);
    }
    (void)hipEventRecord(h->ev[idx], h->stream);
}

static int finish_timing(gpe_handle* h, size_t nev) {
    h->tim_on = false;
    if (!h->timing || nev < 2) return 0;
    GPE_CHECK(hipEventSynchronize(h->ev[nev - 1]));
    double t[2] = {0, 0};
    for (int k = 0; k < 2; ++k) {
        for (size_t i = 0; i < h->kn[k]; ++i) {
            float ms = 0;
            GPE_CHECK(hipEventElapsedTime(&ms, h->kev[k][0][i], h->kev[k][1][i]));
            t[k] += ms;
        }
        t[k] = h->kn[k] ? t[k] * 1e3 / h->kn[k] : 0;
    }
    float tot = 0;
    GPE_CHECK(hipEventElapsedTime(&tot, h->ev[0], h->ev[nev - 1]));
    h->t_plane_us = t[0];
    h->t_line_us = t[1];
    h->t_total_ms = tot;
    return 0;
}

constexpr int kGraphSteps = 64;  // steady-state steps per captured graph (even)
static_assert(kGraphSteps % 2 == 0, "a replay must end on the S1 / Tk parity it started on");

// Coupling table of `pairs` (g', gamma') pairs on the device, with its pinned
// host staging buffer.
static int ensure_coef(gpe_handle* h, size_t pairs) {
    if (pairs <= h->coef_cap) return 0;
    if (h->rt_graph) {
        (void)hipGraphExecDestroy(h->rt_graph);
        h->rt_graph = nullptr;
    }
    if (h->coef_ev) GPE_CHECK(hipEventSynchronize(h->coef_ev));
    if (h->coef) GPE_CHECK(hipFree(h->coef));
    if (h->coef_host) GPE_CHECK(hipHostFree(h->coef_host));
    h->coef = nullptr;
    h->coef_host = nullptr;
    h->coef_cap = 0;
    h->coef_const = false;
    const size_t cap = std::max<size_t>(pairs, 4096);
    GPE_CHECK(hipMalloc(&h->coef, cap * 2 * sizeof(double)));
    GPE_CHECK(hipHostMalloc((void**)&h->coef_host, cap * 2 * sizeof(double), hipHostMallocDefault));
    if (!h->coef_ev) GPE_CHECK(hipEventCreateWithFlags(&h->coef_ev, hipEventDisableTiming));
    h->coef_cap = cap;
    return 0;
}

// Scaled couplings (phi units) of one RT step with couplings g, gamma.
static void rt_coupling(const gpe_handle* h, double g, double gm, double* out) {
    const gpe_params& p = h->p;
    const double dV = p.dx * p.dy * p.dz;
    out[0] = (g - p.cdd / 3) * p.n_atoms / dV;
    out[1] = gm * p.w_15 / pow(dV, 1.5);
}

// Capture `body` (launches on h->stream) into an executable graph.
template <class F>
static int capture(gpe_handle* h, hipGraphExec_t* out, F&& body) {
    hipGraph_t g = nullptr;
    GPE_CHECK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    const int rc = body();
    hipError_t e = hipStreamEndCapture(h->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return 1;
    }
    if (e != hipSuccess) return fail(h, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(h, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    return 0;
}

// One IT step: Poisson x pass, then the plane pass (finish Poisson, stencil,
// norm partials, start next Poisson; split into two blocks per plane except
// on the last step, which starts no Poisson solve).  Flips cur / pcur and,
// split, the S1 parity.
template <typename T>
static PlaneArgs<T> it_plane_args(gpe_handle* h, bool last, int par) {
    int mode = M_PINV | M_NORM_IN | M_IT | M_NORM_OUT | M_RHO_FWD;
    if (last) mode = (mode & ~M_RHO_FWD) | M_PHI_STORE;
    PlaneArgs<T> a = plane_args<T>(h, mode);
    a.in = (const cplx<T>*)h->store[h->cur];
    a.out = (cplx<T>*)h->store[1 - h->cur];
    a.part_in = h->part[h->pcur];
    a.part_out = h->part[1 - h->pcur];
    a.s1 = a.s1o = (cplx<T>*)rt_s1(h, par);
    if (h->plane_split && !last) {
        a.split = 1;
        a.s1o = (cplx<T>*)rt_s1(h, 1 - par);
    }
    return a;
}

template <typename T>
static int it_step_launch(gpe_handle* h, bool last, int& par) {
    if (launch_line<T>(h, true, false, rt_s1(h, par))) return 1;
    if (launch_plane<T>(h, it_plane_args<T>(h, last, par))) return 1;
    h->cur = 1 - h->cur;
    h->pcur = 1 - h->pcur;
    par = 1 - par;
    return 0;
}

template <typename T>
static int do_it_iters(gpe_handle* h, long n) {
    if (n <= 0) return 0;
    if (flush_rt<T>(h)) return 1;
    if (ensure_split_buffers(h, false)) return 1;
    int par = 0;
    // prologue: partials of the current state + Poisson start of |store|^2
    PlaneArgs<T> a = plane_args<T>(h, M_RHO_FWD | M_NORM_OUT);
    a.in = (const cplx<T>*)h->store[h->cur];
    a.part_out = h->part[h->pcur];
    if (launch_plane<T>(h, a)) return 1;
    size_t ne = 0;
    ev_record(h, ne++);
    long steady = n - 1;
    if (h->use_graphs && !h->timing && steady >= kGraphSteps) {
        const int gi = h->cur * 2 + h->pcur;
        if (h->it_graph[gi] && (h->it_graph_key[gi][0] != h->ggp11 || h->it_graph_key[gi][1] != h->gamma)) {
            (void)hipGraphExecDestroy(h->it_graph[gi]);
            h->it_graph[gi] = nullptr;
        }
        if (!h->it_graph[gi]) {
            const int c0 = h->cur, p0 = h->pcur;
            const int rc = capture(h, &h->it_graph[gi], [&]() -> int {
                int cpar = par;  // kGraphSteps is even: a replay ends on the parity it started on
                for (int t = 0; t < kGraphSteps; ++t)
                    if (it_step_launch<T>(h, false, cpar)) return 1;
                return 0;
            });
            h->cur = c0;
            h->pcur = p0;
            if (rc) return 1;
            h->it_graph_key[gi][0] = h->ggp11;
            h->it_graph_key[gi][1] = h->gamma;
        }
        while (steady >= kGraphSteps) {
            GPE_CHECK(hipGraphLaunch(h->it_graph[gi], h->stream));
            steady -= kGraphSteps;  // an even number of steps: cur / pcur unchanged
        }
    }
    h->kn[0] = h->kn[1] = 0;
    h->tim_on = true;  // time the steady-state launches
    for (long t = 0; t < steady; ++t) {
        if (launch_line<T>(h, true, false, rt_s1(h, par))) return 1;
        ev_record(h, ne++);
        if (launch_plane<T>(h, it_plane_args<T>(h, false, par))) return 1;
        ev_record(h, ne++);
        h->cur = 1 - h->cur;
        h->pcur = 1 - h->pcur;
        par = 1 - par;
    }
    h->tim_on = false;
    if (it_step_launch<T>(h, true, par)) return 1;
    h->norm_pending = true;
    return finish_timing(h, ne);
}

// Steady-state RT plane pass merging the closing half phase of one step with
// the opening half phase of the next (coupling pairs cidx, cidx + 1 of the
// table, offset by *cbase when set); reads S1 / Tk of parity par and writes
// parity 1 - par when split.
template <typename T>
static PlaneArgs<T> rt_steady_args(gpe_handle* h, int cidx, const int* cbase, int par) {
    PlaneArgs<T> a = plane_args<T>(h, M_PINV | M_IN_TK | M_PHASE_FULL | M_RHO_FWD | M_KIN);
    a.cv = 2.0;
    a.coef = h->coef;
    a.cbase = cbase;
    a.cidx = cidx;
    a.s1 = a.s1o = (cplx<T>*)rt_s1(h, par);
    a.tk = a.tko = (cplx<T>*)rt_tk(h, par);
    if (h->plane_split) {
        a.split = 1;
        a.s1o = (cplx<T>*)rt_s1(h, 1 - par);
        a.tko = (cplx<T>*)rt_tk(h, 1 - par);
    }
    return a;
}

// One pipelined RT step: the steady plane pass, then the x pass of the new
// Poisson spectrum and kinetic array (parity 1 - par when split).
template <typename T>
static int rt_step_launch(gpe_handle* h, int cidx, const int* cbase, int& par) {
    if (launch_plane<T>(h, rt_steady_args<T>(h, cidx, cbase, par))) return 1;
    const int np = h->plane_split ? 1 - par : par;
    if (launch_line<T>(h, true, true, rt_s1(h, np), rt_tk(h, np))) return 1;
    par = np;
    return 0;
}

__global__ void k_set_int(int* p, int v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// n real_time_iter steps (solver.cpp:107-112), with per-step couplings g, gm
// (the update_edd ramp; null: the current couplings).  Pipelined: step t's
// plane pass applies the closing half phase of step t-1 and the opening half
// phase of step t as one phase (both use Phi of step t-1: real_time_iter
// computes Phi once, after its first half step), starts Phi of step t and runs
// the z / y kinetic propagation; the x pass then does the x parts of both.
// The closing half phase of the last step is left pending (flush_rt).
template <typename T>
static int do_rt_iters(gpe_handle* h, long n, const double* g, const double* gm) {
    if (n <= 0) return 0;
    if (h->norm_pending && materialize<T>(h)) return 1;  // (rt_pending and norm_pending never hold together)
    const bool pend = h->rt_pending;
    const size_t off = pend ? 1 : 0;
    // coupling table: [pending pair], pairs of steps 0..n-1, one spare (a
    // steady pass loads two pairs)
    const size_t pairs = (size_t)n + off + 1;
    if (ensure_coef(h, pairs)) return 1;
    const bool constant = !g && !gm;
    double cc[2];
    rt_coupling(h, h->ggp11, h->gamma, cc);
    const bool pend_same = !pend || (h->rt_pend[0] == cc[0] && h->rt_pend[1] == cc[1]);
    const bool cached = constant && pend_same && h->coef_const && h->coef_cval[0] == cc[0] &&
                        h->coef_cval[1] == cc[1] && h->coef_n >= pairs;
    if (!cached) {
        // the previous upload from the staging buffer must be complete
        GPE_CHECK(hipEventSynchronize(h->coef_ev));
        double* cf = h->coef_host;
        const size_t fill = constant ? h->coef_cap : pairs;
        for (size_t t = 0; t < fill; ++t) {
            const long s = std::min<long>((long)t - (long)off, n - 1);
            if (constant || s < 0) {
                cf[2 * t] = cc[0];
                cf[2 * t + 1] = cc[1];
            } else {
                rt_coupling(h, g ? g[s] : h->ggp11, gm ? gm[s] : h->gamma, cf + 2 * t);
            }
        }
        if (pend) {
            cf[0] = h->rt_pend[0];
            cf[1] = h->rt_pend[1];
        }
        GPE_CHECK(hipMemcpyAsync(h->coef, cf, fill * 2 * sizeof(double), hipMemcpyHostToDevice, h->stream));
        GPE_CHECK(hipEventRecord(h->coef_ev, h->stream));
        h->coef_const = constant && pend_same;
        h->coef_cval[0] = cc[0];
        h->coef_cval[1] = cc[1];
        h->coef_n = fill;
    }
    if (ensure_split_buffers(h, true)) return 1;
    size_t ne = 0;
    ev_record(h, ne++);
    int par = pend ? h->rt_par : 0;
    long t = 0;  // steps launched
    if (!pend) {
        // opening half phase of step 0 with the current fi3d, Poisson start,
        // z / y kinetic, then its x pass
        PlaneArgs<T> a = plane_args<T>(h, M_PHI_LOAD | M_PHASE_HALF | M_RHO_FWD | M_KIN);
        a.in = (const cplx<T>*)h->store[h->cur];
        a.s1 = a.s1o = (cplx<T>*)rt_s1(h, 0);
        a.tk = a.tko = (cplx<T>*)rt_tk(h, 0);
        a.coef = h->coef;
        a.cidx = 0;
        if (launch_plane<T>(h, a)) return 1;
        if (launch_line<T>(h, true, true, rt_s1(h, 0), rt_tk(h, 0))) return 1;
        t = 1;
    }
    // step t merges pairs (t - 1 + off, t + off)
    if (h->use_graphs && !h->timing && n - t >= kGraphSteps + (par ? 1 : 0)) {
        if (par) {  // the captured chunk starts on parity 0
            if (rt_step_launch<T>(h, (int)(t - 1 + off), nullptr, par)) return 1;
            ++t;
        }
        if (!h->rt_graph) {
            // kGraphSteps is even: a replay ends on the buffer parity it started on
            const int rc = capture(h, &h->rt_graph, [&]() -> int {
                int cpar = 0;
                for (int s = 0; s < kGraphSteps; ++s)
                    if (rt_step_launch<T>(h, s, h->cbase, cpar)) return 1;
                hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, h->stream, h->cbase, kGraphSteps);
                GPE_CHECK(hipGetLastError());
                return 0;
            });
            if (rc) return 1;
        }
        hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, h->stream, h->cbase, (int)(t - 1 + off));
        GPE_CHECK(hipGetLastError());
        while (n - t >= kGraphSteps) {
            GPE_CHECK(hipGraphLaunch(h->rt_graph, h->stream));
            t += kGraphSteps;
        }
    }
    h->kn[0] = h->kn[1] = 0;
    h->tim_on = true;  // time the steady-state launches
    for (; t < n; ++t)
        if (rt_step_launch<T>(h, (int)(t - 1 + off), nullptr, par)) return 1;
    h->tim_on = false;
    ev_record(h, ne++);
    h->rt_pending = true;
    h->rt_par = par;
    rt_coupling(h, g ? g[n - 1] : h->ggp11, gm ? gm[n - 1] : h->gamma, h->rt_pend);
    if (g) h->ggp11 = g[n - 1];
    if (gm) h->gamma = gm[n - 1];
    return finish_timing(h, ne);
}

template <typename T>
static int do_energy(gpe_handle* h, gpe_energies* out) {
    if (materialize<T>(h)) return 1;
    const gpe_params& p = h->p;
    const double dV = p.dx * p.dy * p.dz;
    const double gq = (h->ggp11 - p.cdd / 3) * p.n_atoms / dV, gmq = h->gamma * p.w_15 / pow(dV, 1.5);
    auto run = [&](int pass, double mu) -> int {
        hipLaunchKernelGGL((k_energy<T, 256>), dim3(p.nx), dim3(256), 0, h->stream, (const cplx<T>*)h->store[h->cur],
                           (const T*)h->V, (const T*)h->phi, p.nx, p.ny, p.nz, 1.0 / (p.dx * p.dx),
                           1.0 / (p.dy * p.dy), 1.0 / (p.dz * p.dz), p.m, p.cdd, gq, gmq, mu, pass, h->epart);
        GPE_CHECK(hipGetLastError());
        return 0;
    };
    GPE_CHECK(hipMemsetAsync(h->epart, 0, (size_t)p.nx * 8 * 8, h->stream));
    if (run(0, 0.0)) return 1;
    std::vector<double> ep((size_t)p.nx * 8);
    GPE_CHECK(hipMemcpyAsync(ep.data(), h->epart, ep.size() * 8, hipMemcpyDeviceToHost, h->stream));
    GPE_CHECK(hipStreamSynchronize(h->stream));
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < p.nx; ++i)
        for (int t = 0; t < 6; ++t) s[t] += ep[(size_t)i * 8 + t];
    const double mu = s[5];
    if (run(1, mu)) return 1;
    GPE_CHECK(hipMemcpyAsync(ep.data(), h->epart, ep.size() * 8, hipMemcpyDeviceToHost, h->stream));
    GPE_CHECK(hipStreamSynchronize(h->stream));
    double r2 = 0;
    for (int i = 0; i < p.nx; ++i) r2 += ep[(size_t)i * 8 + 6];
    const double N = p.n_atoms;
    // phi = sqrt(dV) psi: dV*sum f(psi) = sum f(phi) for quadratic f.
    out->e_kin = s[0] / (2 * p.m) * N;
    out->e_pot = s[1] * N;
    out->e_int = 0.5 * h->ggp11 * s[2] / dV * N * N;
    out->e_ext = 0.5 * p.cdd * s[3] * N - p.cdd / 3. * s[2] / dV / 2. * N * N;
    out->e_bmf = 2. / 5. * h->gamma * s[4] / pow(dV, 1.5) * pow(N, 2.5);
    out->e_total = out->e_kin + out->e_pot + out->e_int + out->e_ext + out->e_bmf;
    out->mu = mu;
    out->residual = sqrt(r2);
    return 0;
}


// The energies of the materialized state enqueued on the handle's stream:
// the two k_energy passes (partials in part, nx * 8 doubles), their sums in
// sums (8 doubles; pass 1 reads mu from sums[5]) and the eight values, in
// gpe_energies order, to out_dev.  No host copy, no synchronisation.
template <typename T>
static int enqueue_energy(gpe_handle* h, double* part, double* sums, double* out_dev) {
    const gpe_params& p = h->p;
    const double dV = p.dx * p.dy * p.dz;
    const double gq = (h->ggp11 - p.cdd / 3) * p.n_atoms / dV, gmq = h->gamma * p.w_15 / pow(dV, 1.5);
    for (int pass = 0; pass < 2; ++pass) {
        hipLaunchKernelGGL((k_energy<T, 256>), dim3(p.nx), dim3(256), 0, h->stream, (const cplx<T>*)h->store[h->cur],
                           (const T*)h->V, (const T*)h->phi, p.nx, p.ny, p.nz, 1.0 / (p.dx * p.dx),
                           1.0 / (p.dy * p.dy), 1.0 / (p.dz * p.dz), p.m, p.cdd, gq, gmq, 0.0, pass, part,
                           pass ? sums + 5 : nullptr);
        GPE_CHECK(hipGetLastError());
        hipLaunchKernelGGL(k_energy_sums, dim3(1), dim3(64), 0, h->stream, part, p.nx, pass, sums, out_dev, p.m,
                           p.n_atoms, dV, h->ggp11, h->gamma, p.cdd, pow(dV, 1.5), pow(p.n_atoms, 2.5));
        GPE_CHECK(hipGetLastError());
    }
    return 0;
}

// gpe_calc_energy without allocation, host copies or synchronisation: the
// partials and sums live in the device's `scratch`, the outputs go to out_dev.
template <typename T>
static int do_energy_noalloc(gpe_handle* h, double* out_dev) {
    const gpe_params& p = h->p;
    if (p.nx > kNoAllocPlanes)
        return fail(h, "gpe_calc_energy_noalloc: nx = " + std::to_string(p.nx) + " exceeds its scratch (" +
                           std::to_string(kNoAllocPlanes) + " planes)");
    double* sc = energy_scratch(h->dev);
    if (!sc) return fail(h, "gpe_calc_energy_noalloc: no scratch address for this device");
    if (materialize<T>(h)) return 1;
    return enqueue_energy<T>(h, sc, sc + (size_t)kNoAllocPlanes * 8, out_dev);
}

// Snapshot of the state and its energies, behind the launched work: the
// state is copied on the device and its energies computed on the handle's
// stream; a second stream then copies both to pinned host memory once that
// point is reached, while the handle's stream goes on with later work.
template <typename T>
static int do_snapshot_begin(gpe_handle* h) {
    if (h->snap_in_flight) return fail(h, "gpe_snapshot_begin: a snapshot is already in flight");
    const size_t bytes = h->N * csize(h);
    if (!h->copy_stream) {
        GPE_CHECK(hipStreamCreateWithFlags(&h->copy_stream, hipStreamNonBlocking));
        GPE_CHECK(hipEventCreateWithFlags(&h->snap_ev, hipEventDisableTiming));
        GPE_CHECK(hipEventCreateWithFlags(&h->copy_ev, hipEventDisableTiming));
        GPE_CHECK(hipMalloc(&h->snap, bytes));
        GPE_CHECK(hipMalloc((void**)&h->snap_dev, 16 * sizeof(double)));
        GPE_CHECK(hipHostMalloc(&h->snap_host, bytes, hipHostMallocDefault));
        GPE_CHECK(hipHostMalloc((void**)&h->snap_host_e, 8 * sizeof(double), hipHostMallocDefault));
    }
    if (materialize<T>(h)) return 1;
    if (enqueue_energy<T>(h, h->epart, h->snap_dev, h->snap_dev + 8)) return 1;
    GPE_CHECK(hipMemcpyAsync(h->snap, h->store[h->cur], bytes, hipMemcpyDeviceToDevice, h->stream));
    GPE_CHECK(hipEventRecord(h->snap_ev, h->stream));
    GPE_CHECK(hipStreamWaitEvent(h->copy_stream, h->snap_ev, 0));
    GPE_CHECK(hipMemcpyAsync(h->snap_host, h->snap, bytes, hipMemcpyDeviceToHost, h->copy_stream));
    GPE_CHECK(hipMemcpyAsync(h->snap_host_e, h->snap_dev + 8, 8 * sizeof(double), hipMemcpyDeviceToHost,
                             h->copy_stream));
    GPE_CHECK(hipEventRecord(h->copy_ev, h->copy_stream));
    h->snap_in_flight = true;
    return 0;
}

template <typename T>
static int do_snapshot_end(gpe_handle* h, double* psi, gpe_energies* e) {
    if (!h->snap_in_flight) return fail(h, "gpe_snapshot_end: no snapshot in flight");
    GPE_CHECK(hipEventSynchronize(h->copy_ev));
    h->snap_in_flight = false;
    if (psi) {
        const T* src = (const T*)h->snap_host;
        const double sc = 1.0 / h->sqdv;
        for (size_t t = 0; t < 2 * h->N; ++t) psi[t] = (double)src[t] * sc;
    }
    if (e) {
        const double* v = h->snap_host_e;
        e->e_kin = v[0];
        e->e_pot = v[1];
        e->e_int = v[2];
        e->e_ext = v[3];
        e->e_bmf = v[4];
        e->e_total = v[5];
        e->mu = v[6];
        e->residual = v[7];
    }
    return 0;
}

// ===========================================================================
// Batch: many instances (grids may differ) stepped together, one launch per
// pass over all active instances, steady-state chunks replayed as hipGraphs.
// ===========================================================================
struct gpe_batch {
    int prec = GPE_FP32;
    hipStream_t stream = nullptr;
    std::vector<gpe_handle*> inst;
    std::vector<int> active;
    std::string err;
    int cur = 0, pcur = 0;  // shared ping-pong parity of every instance
    // device argument blocks: [0] prologue, [1..4] steady by parity, [5..8] last by parity
    void* d_pargs = nullptr;
    void* d_largs = nullptr;
    void* d_pwork = nullptr;
    void* d_lwork = nullptr;
    int npwork = 0, nlwork = 0;
    // one launch per grid shape: plane groups by (ny, nz), line groups by nx
    // (kind 1: register x pass); work items of a group are contiguous
    struct Group {
        int a, b, kind, off, cnt;
        size_t lds;
    };
    std::vector<Group> pgroups, lgroups;
    // one launch per pass over every shape of the active set (k_plane_multi,
    // k_line_multi) when it has more than one and all have those bodies:
    // work items carry the shape slot
    bool multi_p = false, multi_l = false;
    void* d_pmwork = nullptr;
    void* d_lmwork = nullptr;
    int npmwork = 0, nlmwork = 0;
    size_t cap_pmwork = 0, cap_lmwork = 0, lds_pm = 0, lds_lm = 0;
    hipGraphExec_t graph[4] = {nullptr, nullptr, nullptr, nullptr};
    bool use_graphs = true;
    // Device buffers sized at the first setup for every instance and kept
    // while the active set shrinks: argument blocks and work lists of the
    // steps and of the energy check.
    size_t cap_pargs = 0, cap_largs = 0, cap_pwork = 0, cap_lwork = 0, cap_eargs = 0, cap_swork = 0, cap_ework = 0;
    void* d_eargs = nullptr;   // EArgs per active instance
    void* d_swork = nullptr;   // (instance, chunk) of the grouped scale
    void* d_ework = nullptr;   // (instance, plane) of the grouped energy
    int nswork = 0, nework = 0;
    double* d_esum = nullptr;  // per instance: sums[8], out[8]
    double* h_eout = nullptr;  // pinned: out[8] per active instance
    // Fallback states (gpe_batch_set_fallback): the state an instance
    // restarts from when its energy is non-finite or rises, or dtau mu >= 1.
    std::vector<std::vector<double>> fallback;
    std::vector<char> fell_back;
};

static int bfail(gpe_batch* b, const std::string& m) {
    if (b) b->err = m;
    g_create_err = m;
    return 1;
}

#define BCHECK(call)                                                                            \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) return bfail(b, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

static void batch_drop_graphs(gpe_batch* b) {
    for (auto& g : b->graph)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
}

// Argument blocks and work lists for the current active set.  Slots:
// [0..3] prologue, [4..7] steady step, [8..11] last step, indexed by parity
// (cur*2 + pcur).
// Slot of a plane shape in k_plane_multi (-1: no compiled body).
static int plane_slot(int ny, int nz) {
    int s = 0, r = -1;
#define GPE_PS(NY, NZ)                     \
    if (r < 0 && ny == NY && nz == NZ) r = s; \
    ++s;
    GPE_PLANE_SHAPES(GPE_PS)
#undef GPE_PS
    return r;
}
// Slot of an x-pass group in k_line_multi (-1: none).
static int line_slot(int kind, int nx) {
    if (kind == 1) return nx == 128 ? 0 : nx == 64 ? 1 : -1;
    int s = 2, r = -1;
#define GPE_LS(NX)                 \
    if (r < 0 && nx == NX) r = s; \
    ++s;
    GPE_MULTI_LINE_SHAPES(GPE_LS)
#undef GPE_LS
    return r;
}
// GPE_BATCH_SINGLE=0: a lone active instance keeps the batch's launches.
static bool batch_single_on() {
    static const bool on = [] {
        const char* e = getenv("GPE_BATCH_SINGLE");
        return !(e && e[0] == '0');
    }();
    return on;
}
// GPE_BATCH_MULTI=0: one launch per grid shape and pass.
static bool batch_multi_on() {
    static const bool on = [] {
        const char* e = getenv("GPE_BATCH_MULTI");
        return !(e && e[0] == '0');
    }();
    return on;
}

template <typename T>
static int batch_setup(gpe_batch* b) {
    batch_drop_graphs(b);
    const size_t na = b->active.size();
    std::vector<PlaneArgs<T>> pa(12 * na);
    std::vector<LineArgs<T>> la(na);
    std::vector<int2> pw, lw;
    b->pgroups.clear();
    b->lgroups.clear();
    for (size_t q = 0; q < na; ++q) {
        gpe_handle* h = b->inst[b->active[q]];
        for (int par = 0; par < 4; ++par) {
            const int c = par >> 1, pc = par & 1;
            PlaneArgs<T> a = plane_args<T>(h, M_RHO_FWD | M_NORM_OUT);
            a.in = (const cplx<T>*)h->store[c];
            a.part_out = h->part[pc];
            pa[(size_t)par * na + q] = a;
            for (int last = 0; last < 2; ++last) {
                int mode = M_PINV | M_NORM_IN | M_IT | M_NORM_OUT | M_RHO_FWD;
                if (last) mode = (mode & ~M_RHO_FWD) | M_PHI_STORE;
                a = plane_args<T>(h, mode);
                a.in = (const cplx<T>*)h->store[c];
                a.out = (cplx<T>*)h->store[1 - c];
                a.part_in = h->part[pc];
                a.part_out = h->part[1 - pc];
                pa[(size_t)(4 + 4 * last + par) * na + q] = a;
            }
        }
        la[q] = line_args<T>(h, true, false);
    }
    // work lists grouped by shape, instances in order within a group
    const bool gen = !b->inst.empty() && b->inst[0]->force_generic;
    for (size_t q = 0; q < na; ++q) {
        gpe_handle* h = b->inst[b->active[q]];
        bool seen = false;
        for (auto& g : b->pgroups) seen = seen || (g.a == h->p.ny && g.b == h->p.nz);
        if (seen) continue;
        {
            gpe_batch::Group g{h->p.ny, h->p.nz, 0, (int)pw.size(), 0, 0};
            for (size_t r = q; r < na; ++r) {
                gpe_handle* hr = b->inst[b->active[r]];
                if (hr->p.ny != g.a || hr->p.nz != g.b) continue;
                for (int i = 0; i < hr->p.nx; ++i) pw.push_back(make_int2((int)r, i));
                g.lds = std::max(g.lds, hr->plane_lds + (plane_compiled(hr) ? hr->plane_stage : 0));
            }
            g.cnt = (int)pw.size() - g.off;
            b->pgroups.push_back(g);
        }
    }
    for (size_t q = 0; q < na; ++q) {
        gpe_handle* h = b->inst[b->active[q]];
        bool seen = false;
        for (auto& g : b->lgroups) seen = seen || g.a == h->p.nx;
        if (seen) continue;
        const int nx = h->p.nx;
        const bool reg = !gen && h->line_reg && (nx == 128 || nx == 64 || nx == 256);
        gpe_batch::Group g{nx, 0, reg ? 1 : 0, (int)lw.size(), 0, 0};
        for (size_t r = q; r < na; ++r) {
            gpe_handle* hr = b->inst[b->active[r]];
            if (hr->p.nx != nx) continue;
            int nt = la[r].ntile_p;
            if (reg) {
                const int lpb = nx == 128 ? XReg<128>::LPB_P : nx == 64 ? XReg<64>::LPB_P : XReg<256>::LPB_P;
                const int exa = nx == 128 ? XReg<128>::exa<T>() : nx == 64 ? XReg<64>::exa<T>() : XReg<256>::exa<T>();
                nt = (2 * hr->p.ny * hr->p.nz + lpb - 1) / lpb;
                g.lds = (size_t)(exa + 4 * nx) * sizeof(cplx<T>);
            } else {
                g.lds = std::max(g.lds, hr->line_lds);
            }
            for (int t = 0; t < nt; ++t) lw.push_back(make_int2((int)r, t));
        }
        g.cnt = (int)lw.size() - g.off;
        b->lgroups.push_back(g);
    }
    // the single-launch work lists, when the active set spans several shapes
    std::vector<int2> pm, lm;
    b->multi_p = b->multi_l = false;
    if (sizeof(T) == 4 && !gen && batch_multi_on()) {  // FP64: the joint bodies spill
        bool ok = b->pgroups.size() > 1;
        size_t lds = 0;
        for (const auto& g : b->pgroups) {
            const int slot = plane_slot(g.a, g.b);
            if (slot < 0) ok = false;
            for (int e = 0; ok && e < g.cnt; ++e) {
                const int2 v = pw[(size_t)g.off + e];
                pm.push_back(make_int2(v.x, v.y | (slot << kSlotShift)));
            }
            lds = std::max(lds, g.lds);
        }
        if (ok) {
            b->multi_p = true;
            b->lds_pm = lds;
        } else {
            pm.clear();
        }
        ok = b->lgroups.size() > 1;
        lds = 0;
        for (const auto& g : b->lgroups) {
            const int slot = line_slot(g.kind, g.a);
            if (slot < 0) ok = false;
            for (int e = 0; ok && e < g.cnt; ++e) {
                const int2 v = lw[(size_t)g.off + e];
                lm.push_back(make_int2(v.x, v.y | (slot << kSlotShift)));
            }
            lds = std::max(lds, g.lds);
        }
        if (ok) {
            b->multi_l = true;
            b->lds_lm = lds;
        } else {
            lm.clear();
        }
    }
    // energy check: per-instance arguments, (instance, chunk) and (instance,
    // plane) work lists
    std::vector<EArgs<T>> ea(na);
    std::vector<int2> sw, ew;
    for (size_t q = 0; q < na; ++q) {
        gpe_handle* h = b->inst[b->active[q]];
        const gpe_params& p = h->p;
        const double dV = p.dx * p.dy * p.dz;
        EArgs<T> e;
        e.psi[0] = (cplx<T>*)h->store[0];
        e.psi[1] = (cplx<T>*)h->store[1];
        e.V = (const T*)h->V;
        e.phi = (const T*)h->phi;
        e.npart[0] = h->part[0];
        e.npart[1] = h->part[1];
        e.part = h->epart;
        e.sums = b->d_esum + (size_t)b->active[q] * 16;
        e.out = e.sums + 8;
        e.n = h->N;
        e.nx = p.nx;
        e.ny = p.ny;
        e.nz = p.nz;
        e.ix2 = 1.0 / (p.dx * p.dx);
        e.iy2 = 1.0 / (p.dy * p.dy);
        e.iz2 = 1.0 / (p.dz * p.dz);
        e.m = p.m;
        e.cdd = p.cdd;
        e.gq = (h->ggp11 - p.cdd / 3) * p.n_atoms / dV;
        e.gmq = h->gamma * p.w_15 / pow(dV, 1.5);
        e.N = p.n_atoms;
        e.dV = dV;
        e.g = h->ggp11;
        e.gamma = h->gamma;
        e.dv15 = pow(dV, 1.5);
        e.n25 = pow(p.n_atoms, 2.5);
        ea[q] = e;
        for (size_t c = 0; c * kScaleChunk < h->N; ++c) sw.push_back(make_int2((int)q, (int)c));
        for (int i = 0; i < p.nx; ++i) ew.push_back(make_int2((int)q, i));
    }
    // device buffers: sized by the first (largest) active set, then reused
    auto ensure = [&](void** d, size_t& cap, size_t bytes) -> int {
        if (bytes <= cap && *d) return 0;
        if (*d) BCHECK(hipFree(*d));
        *d = nullptr;
        cap = 0;
        BCHECK(hipMalloc(d, std::max<size_t>(bytes, 1)));
        cap = bytes;
        return 0;
    };
    b->npwork = (int)pw.size();
    b->nlwork = (int)lw.size();
    b->npmwork = (int)pm.size();
    b->nlmwork = (int)lm.size();
    b->nswork = (int)sw.size();
    b->nework = (int)ew.size();
    if (na == 0) return 0;
    if (ensure(&b->d_pargs, b->cap_pargs, pa.size() * sizeof(PlaneArgs<T>)) ||
        ensure(&b->d_largs, b->cap_largs, la.size() * sizeof(LineArgs<T>)) ||
        ensure(&b->d_pwork, b->cap_pwork, pw.size() * sizeof(int2)) ||
        ensure(&b->d_lwork, b->cap_lwork, lw.size() * sizeof(int2)) ||
        ensure(&b->d_eargs, b->cap_eargs, ea.size() * sizeof(EArgs<T>)) ||
        ensure(&b->d_swork, b->cap_swork, sw.size() * sizeof(int2)) ||
        ensure(&b->d_ework, b->cap_ework, ew.size() * sizeof(int2)) ||
        ensure(&b->d_pmwork, b->cap_pmwork, pm.size() * sizeof(int2)) ||
        ensure(&b->d_lmwork, b->cap_lmwork, lm.size() * sizeof(int2)))
        return 1;
    BCHECK(hipMemcpy(b->d_pargs, pa.data(), pa.size() * sizeof(PlaneArgs<T>), hipMemcpyHostToDevice));
    BCHECK(hipMemcpy(b->d_largs, la.data(), la.size() * sizeof(LineArgs<T>), hipMemcpyHostToDevice));
    BCHECK(hipMemcpy(b->d_pwork, pw.data(), pw.size() * sizeof(int2), hipMemcpyHostToDevice));
    BCHECK(hipMemcpy(b->d_lwork, lw.data(), lw.size() * sizeof(int2), hipMemcpyHostToDevice));
    BCHECK(hipMemcpy(b->d_eargs, ea.data(), ea.size() * sizeof(EArgs<T>), hipMemcpyHostToDevice));
    BCHECK(hipMemcpy(b->d_swork, sw.data(), sw.size() * sizeof(int2), hipMemcpyHostToDevice));
    BCHECK(hipMemcpy(b->d_ework, ew.data(), ew.size() * sizeof(int2), hipMemcpyHostToDevice));
    if (!pm.empty()) BCHECK(hipMemcpy(b->d_pmwork, pm.data(), pm.size() * sizeof(int2), hipMemcpyHostToDevice));
    if (!lm.empty()) BCHECK(hipMemcpy(b->d_lmwork, lm.data(), lm.size() * sizeof(int2), hipMemcpyHostToDevice));
    return 0;
}

// Energy check of every active instance (what do_energy does per handle:
// normalise, then the two k_energy passes and their sums), as five launches
// over all of them and one copy of the results: out_host[q * 8 + k] for
// active slot q, gpe_energies order.
template <typename T>
static int batch_check(gpe_batch* b, double* out_host) {
    const int na = (int)b->active.size();
    if (na == 0) return 0;
    const EArgs<T>* ea = (const EArgs<T>*)b->d_eargs;
    hipLaunchKernelGGL((k_scale_group<T>), dim3(b->nswork), dim3(256), 0, b->stream, ea, (const int2*)b->d_swork,
                       b->cur, b->pcur);
    BCHECK(hipGetLastError());
    for (int pass = 0; pass < 2; ++pass) {
        hipLaunchKernelGGL((k_energy_group<T, 256>), dim3(b->nework), dim3(256), 0, b->stream, ea,
                           (const int2*)b->d_ework, pass, b->cur);
        BCHECK(hipGetLastError());
        hipLaunchKernelGGL((k_energy_sums_group<T>), dim3(na), dim3(64), 0, b->stream, ea, pass);
        BCHECK(hipGetLastError());
    }
    for (int q = 0; q < na; ++q)
        BCHECK(hipMemcpyAsync(out_host + (size_t)q * 8, b->d_esum + (size_t)b->active[q] * 16 + 8, 8 * sizeof(double),
                              hipMemcpyDeviceToHost, b->stream));
    BCHECK(hipStreamSynchronize(b->stream));
    for (int q : b->active) b->inst[q]->norm_pending = false;
    return 0;
}

template <typename T>
static int batch_plane(gpe_batch* b, int slot) {
    const PlaneArgs<T>* args = (const PlaneArgs<T>*)b->d_pargs + (size_t)slot * b->active.size();
    if constexpr (sizeof(T) == 4) {
        if (b->multi_p) {
            constexpr int NT = PlaneThreads<T, true>::value;
            hipLaunchKernelGGL((k_plane_multi<T, NT>), dim3(b->npmwork), dim3(NT), b->lds_pm, b->stream, args,
                               (const int2*)b->d_pmwork);
            BCHECK(hipGetLastError());
            return 0;
        }
    }
    const bool gen = b->inst[0]->force_generic;
    for (const auto& g : b->pgroups) {
        const int2* w = (const int2*)b->d_pwork + g.off;
        bool done = false;
        if (!gen) {
#define GPE_PL(NY, NZ)                                                                                      \
    if (!done && g.a == NY && g.b == NZ) {                                                                  \
        constexpr int NT = PlaneThreads<T, true>::value;                                                    \
        hipLaunchKernelGGL((k_plane_group<T, NT, PlanePolCT<NY, NZ>>), dim3(g.cnt), dim3(NT), g.lds, b->stream, \
                           args, w);                                                                        \
        done = true;                                                                                        \
    }
            GPE_PLANE_SHAPES(GPE_PL)
#undef GPE_PL
        }
        if (!done) {
            constexpr int NT = PlaneThreads<T, false>::value;
            hipLaunchKernelGGL((k_plane_group<T, NT, PlanePolGeneric>), dim3(g.cnt), dim3(NT), g.lds, b->stream, args,
                               w);
        }
        BCHECK(hipGetLastError());
    }
    return 0;
}

template <typename T>
static int batch_line(gpe_batch* b) {
    const LineArgs<T>* args = (const LineArgs<T>*)b->d_largs;
    if constexpr (sizeof(T) == 4) {
        if (b->multi_l) {
            hipLaunchKernelGGL((k_line_multi<T>), dim3(b->nlmwork), dim3(kLineThreads), b->lds_lm, b->stream, args,
                               (const int2*)b->d_lmwork);
            BCHECK(hipGetLastError());
            return 0;
        }
    }
    const bool gen = b->inst[0]->force_generic;
    for (const auto& g : b->lgroups) {
        const int2* w = (const int2*)b->d_lwork + g.off;
        bool done = false;
        if (g.kind == 1 && g.a == 128) {
            hipLaunchKernelGGL((k_xpass_reg_group<T, 128>), dim3(g.cnt), dim3(256), g.lds, b->stream, args, w);
            done = true;
        } else if (g.kind == 1 && g.a == 64) {
            hipLaunchKernelGGL((k_xpass_reg_group<T, 64>), dim3(g.cnt), dim3(256), g.lds, b->stream, args, w);
            done = true;
        } else if (g.kind == 1 && g.a == 256) {
            hipLaunchKernelGGL((k_xpass_reg_group<T, 256>), dim3(g.cnt), dim3(256), g.lds, b->stream, args, w);
            done = true;
        }
        if (!done && !gen) {
#define GPE_LN(NX)                                                                                           \
    if (!done && g.a == NX) {                                                                                \
        hipLaunchKernelGGL((k_xpass_group<T, kLineThreads, LinePolCT<NX>>), dim3(g.cnt), dim3(kLineThreads), \
                           g.lds, b->stream, args, w);                                                       \
        done = true;                                                                                         \
    }
            GPE_LINE_SHAPES(GPE_LN)
#undef GPE_LN
        }
        if (!done)
            hipLaunchKernelGGL((k_xpass_group<T, kLineThreads, LinePolGeneric>), dim3(g.cnt), dim3(kLineThreads),
                               g.lds, b->stream, args, w);
        BCHECK(hipGetLastError());
    }
    return 0;
}

// n fused IT steps of every active instance (same semantics as gpe_imag_time_iters).
template <typename T>
static int batch_it(gpe_batch* b, long n) {
    if (n <= 0 || b->active.empty()) return 0;
    if (batch_plane<T>(b, b->cur * 2 + b->pcur)) return 1;
    long steady = n - 1;
    if (b->use_graphs && steady >= kGraphSteps) {
        const int gi = b->cur * 2 + b->pcur;
        if (!b->graph[gi]) {
            hipGraph_t g = nullptr;
            BCHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            int c = b->cur, pc = b->pcur, rc = 0;
            for (int t = 0; t < kGraphSteps && !rc; ++t) {
                rc = batch_line<T>(b) || batch_plane<T>(b, 4 + c * 2 + pc);
                c = 1 - c;
                pc = 1 - pc;
            }
            hipError_t e = hipStreamEndCapture(b->stream, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return 1;
            }
            if (e != hipSuccess) return bfail(b, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
            e = hipGraphInstantiate(&b->graph[gi], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) return bfail(b, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
        while (steady >= kGraphSteps) {
            BCHECK(hipGraphLaunch(b->graph[gi], b->stream));
            steady -= kGraphSteps;
        }
    }
    for (long t = 0; t < steady; ++t) {
        if (batch_line<T>(b) || batch_plane<T>(b, 4 + b->cur * 2 + b->pcur)) return 1;
        b->cur = 1 - b->cur;
        b->pcur = 1 - b->pcur;
    }
    if (batch_line<T>(b) || batch_plane<T>(b, 8 + b->cur * 2 + b->pcur)) return 1;
    b->cur = 1 - b->cur;
    b->pcur = 1 - b->pcur;
    // only the stepped instances move: a dropped one keeps the buffer its
    // state was left in (n may be odd)
    for (int q : b->active) {
        gpe_handle* h = b->inst[q];
        h->cur = b->cur;
        h->pcur = b->pcur;
        h->norm_pending = true;
    }
    return 0;
}

// Energy rise that sends a warm-started instance to its fallback state.  On
// a stable explicit map imaginary time lowers E/N from one check to the next,
// so E/N above the previous check's by more than the convergence tolerance
// (kRiseTol relative when no tolerance is given) means the state is not
// converging.  Nor is a state with dtau mu >= 1: the map's factor 1 - dtau mu
// is not positive on it, so it is a mode an unstable map (dtau beyond the
// stencil's limit) has grown, which E/N no longer changes on.  A growing
// residual ||H psi - mu psi|| alone is not taken as a failure: it grows while
// a state leaves a saddle point for a lower minimum, with E/N falling all the
// way.
constexpr double kRiseTol = 1e-9;

template <typename T>
static int batch_it_until(gpe_batch* b, long max_steps, long check_every, double tol, long* steps_out,
                          double* e_out) {
    const int ni = (int)b->inst.size();
    std::vector<double> last(ni, 0.0);
    const double rise_tol = tol > 0 ? tol : kRiseTol;
    std::vector<char> have(ni, 0);
    b->fell_back.assign(ni, 0);
    b->active.clear();
    for (int q = 0; q < ni; ++q) {
        b->active.push_back(q);
        if (steps_out) steps_out[q] = 0;
    }
    long steps = 0;
    if (check_every <= 0) check_every = 100;
    if (!b->d_esum) BCHECK(hipMalloc((void**)&b->d_esum, (size_t)ni * 16 * sizeof(double)));
    if (!b->h_eout) BCHECK(hipHostMalloc((void**)&b->h_eout, (size_t)ni * 8 * sizeof(double), hipHostMallocDefault));
    if (batch_setup<T>(b)) return 1;
    while (!b->active.empty() && steps < max_steps) {
        const long n = std::min(check_every, max_steps - steps);
        if (b->active.size() == 1 && batch_single_on()) {
            // A lone instance (the tail of a batch) runs its handle's own
            // fused loop (gpe_imag_time_iters): its tuned plane pass, two
            // blocks per plane, and graph replay, on the batch's stream and
            // state parity.  The same steps; in FP32 the last bits may round
            // as the single-instance path rounds them.
            gpe_handle* h = b->inst[b->active[0]];
            h->cur = b->cur;
            h->pcur = b->pcur;
            if (do_it_iters<T>(h, n)) return bfail(b, h->err);
            b->cur = h->cur;
            b->pcur = h->pcur;
        } else if (batch_it<T>(b, n)) {
            return 1;
        }
        steps += n;
        if (batch_check<T>(b, b->h_eout)) return 1;
        std::vector<int> keep;
        for (size_t s = 0; s < b->active.size(); ++s) {
            const int q = b->active[s];
            gpe_handle* h = b->inst[q];
            const double* e = b->h_eout + s * 8;
            const double epa = e[5] / h->p.n_atoms, res = e[7];
            if (steps_out) steps_out[q] = steps;
            // fallback: non-finite state (e.g. a NaN or zero warm state that
            // load_buffer could not normalise) or a rising energy
            const bool bad = !std::isfinite(epa) || !std::isfinite(res);
            const bool rose = have[q] && epa - last[q] > rise_tol * std::abs(epa);
            const bool flip = e[6] * h->p.imag_time_dt >= 1.0;
            if (!b->fallback[q].empty() && !b->fell_back[q] && (bad || rose || flip)) {
                if (gpe_load_buffer(h, b->fallback[q].data())) return bfail(b, h->err);
                b->fell_back[q] = 1;
                have[q] = 0;
                keep.push_back(q);
                continue;
            }
            if (e_out) e_out[q] = epa;
            const bool conv = tol > 0 && have[q] && std::abs(epa - last[q]) < tol * std::abs(epa);
            last[q] = epa;
            have[q] = 1;
            if (!conv) keep.push_back(q);
        }
        if (keep.size() != b->active.size()) {
            b->active = keep;
            if (batch_setup<T>(b)) return 1;
        }
    }
    BCHECK(hipStreamSynchronize(b->stream));
    return 0;
}

template <typename T>
static int do_it_until(gpe_handle* h, long max_steps, long check_every, double tol, long* steps_out,
                       double* e_out) {
    if (check_every <= 0) check_every = 100;
    long steps = 0;
    double last = 0.0, epa = 0.0;
    bool have = false;
    while (steps < max_steps) {
        const long n = std::min(check_every, max_steps - steps);
        if (do_it_iters<T>(h, n)) return 1;
        steps += n;
        gpe_energies e;
        if (do_energy<T>(h, &e)) return 1;
        epa = e.e_total / h->p.n_atoms;
        const bool conv = tol > 0 && have && std::abs(epa - last) < tol * std::abs(epa);
        last = epa;
        have = true;
        if (conv) break;
    }
    if (steps_out) *steps_out = steps;
    if (e_out) *e_out = epa;
    return 0;
}

static InitArgs init_args(const gpe_handle* h) {
    InitArgs a;
    memset(&a, 0, sizeof(a));
    a.nx = h->p.nx;
    a.ny = h->p.ny;
    a.nz = h->p.nz;
    a.dx = h->p.dx;
    a.dy = h->p.dy;
    a.dz = h->p.dz;
    a.m = h->p.m;
    a.scale = h->sqdv;
    return a;
}

template <typename T>
static int do_init_pote(gpe_handle* h, const InitArgs& a) {
    hipLaunchKernelGGL((k_init_pote<T>), dim3(1024), dim3(256), 0, h->stream, a, (T*)h->V);
    GPE_CHECK(hipGetLastError());
    return 0;
}

template <typename T>
static int do_init_psi(gpe_handle* h, const InitArgs& a) {
    hipLaunchKernelGGL((k_init_psi<T>), dim3(1024), dim3(256), 0, h->stream, a, (cplx<T>*)h->store[h->cur],
                       (cplx<T>*)h->psii);
    GPE_CHECK(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// Warm start from a network prediction on the device (no reference
// counterpart; the state enters as load_buffer does, src/solver/solver.cpp:
// 118-125).

// Network input of one instance, [P][cin]: channel 0 the potential over
// v_scale, channels 1 .. cin-1 the instance's constants.
struct WarmConsts {
    float c[8];
};

template <typename T>
__global__ void k_warm_features(const T* __restrict__ V, size_t n, int cin, double v_scale, WarmConsts c,
                                float* __restrict__ x) {
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        float* xe = x + e * cin;
        xe[0] = (float)((double)V[e] / v_scale);
        for (int q = 1; q < cin; ++q) xe[q] = c.c[q - 1];
    }
}

// psi0 from a predicted amplitude: interior points max(a, 0), the
// boundary zero; `mirror` first averages a with its mirror images i -> n - i
// along x, then y, then z over the interior pairs 2 <= i <= n - 2 (indices 0,
// 1 and n - 1 keep their value), in FP64 and in that order, so the state
// equals the host's warm state fed to load_buffer bit for bit.  Stored as
// (T)(psi0 sqrt(dV)) in the state and in psii, as load_buffer stores.
template <typename T>
__global__ void k_warm_state(const float* __restrict__ amp, int nx, int ny, int nz, int mirror, double sqdv,
                             cplx<T>* __restrict__ store, cplx<T>* __restrict__ psii) {
    const size_t n = (size_t)nx * ny * nz;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(e % nz), j = (int)((e / nz) % ny), i = (int)(e / ((size_t)ny * nz));
        double w = 0.0;
        if (i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2 && k >= 1 && k <= nz - 2) {
            auto A = [&](int ii, int jj, int kk) { return (double)amp[((size_t)ii * ny + jj) * nz + kk]; };
            if (!mirror) {
                w = A(i, j, k);
            } else {
                const bool mx = i >= 2 && i <= nx - 2, my = j >= 2 && j <= ny - 2, mz = k >= 2 && k <= nz - 2;
                auto tx = [&](int jj, int kk) {
                    const double a = A(i, jj, kk);
                    return mx ? 0.5 * (a + A(nx - i, jj, kk)) : a;
                };
                auto ty = [&](int kk) {
                    const double t = tx(j, kk);
                    return my ? 0.5 * (t + tx(ny - j, kk)) : t;
                };
                const double u = ty(k);
                w = mz ? 0.5 * (u + ty(nz - k)) : u;
            }
            w = isnan(w) ? w : (w > 0.0 ? w : 0.0);
        }
        const T v = (T)(w * sqdv);
        store[e] = gpe::mk(v, (T)(0.0 * sqdv));
        psii[e] = store[e];
    }
}

template <typename T>
static int do_warm_features(gpe_handle* h, float* x, int cin, double v_scale, const WarmConsts& c) {
    hipLaunchKernelGGL((k_warm_features<T>), dim3(1024), dim3(256), 0, h->stream, (const T*)h->V, h->N, cin, v_scale,
                       c, x);
    GPE_CHECK(hipGetLastError());
    GPE_CHECK(hipStreamSynchronize(h->stream));
    return 0;
}

template <typename T>
static int do_warm_state(gpe_handle* h, const float* amp, int mirror, double* norm_out) {
    if (flush_rt<T>(h)) return 1;
    hipLaunchKernelGGL((k_warm_state<T>), dim3(1024), dim3(256), 0, h->stream, amp, h->p.nx, h->p.ny, h->p.nz, mirror,
                       h->sqdv, (cplx<T>*)h->store[h->cur], (cplx<T>*)h->psii);
    GPE_CHECK(hipGetLastError());
    h->norm_pending = false;
    h->rt_pending = false;  // the loaded state replaces a pending RT step
    // calc_norm + normalize, as load_buffer (solver.cpp:118-125)
    double xn = 0.0;
    if (do_calc_norm<T>(h, &xn)) return 1;
    if (norm_out) *norm_out = xn;
    return do_normalize<T>(h);
}

#define DISPATCH(h, fn, ...) ((h)->prec == GPE_FP64 ? fn<double>(__VA_ARGS__) : fn<float>(__VA_ARGS__))

extern "C" {

const char* gpe_version(void) { return "gpesolver-amd 0.1 (gfx950)"; }

const char* gpe_last_error(const gpe_handle* h) { return h ? h->err.c_str() : g_create_err.c_str(); }

int gpe_create(const gpe_params* p, int precision, gpe_handle** out) {
    if (!p || !out) return fail(nullptr, "null argument");
    *out = nullptr;
    if (precision != GPE_FP32 && precision != GPE_FP64) return fail(nullptr, "precision must be 32 or 64");
    if (p->nx < 3 || p->ny < 3 || p->nz < 3) return fail(nullptr, "grid dimensions must be >= 3");
    if (!(p->dx > 0 && p->dy > 0 && p->dz > 0 && p->m > 0 && p->n_atoms > 0))
        return fail(nullptr, "spacings, mass and n_atoms must be positive");
    const int maxe = precision == GPE_FP64 ? Lim<double>::maxe : Lim<float>::maxe;
    const long pe = 2L * p->ny * (p->nz + 1);
    // planes beyond one block's LDS run the pencil pass: its column tiles
    // hold 2ny x 2 and its row tiles 2nz+1 complex elements at least
    // GPE_FORCE_PENCIL=1 runs the pencil pass on any grid (tests compare the
    // two plane passes on small grids)
    const char* fpen = getenv("GPE_FORCE_PENCIL");
    const bool pencil = pe > maxe || (long)p->ny * p->nz > kMaxPlanePts || (fpen && fpen[0] == '1');
    const int pmaxe = precision == GPE_FP64 ? PencilLim<double>::maxe : PencilLim<float>::maxe;
    if (pencil && (4L * p->ny > pmaxe || 2L * p->nz + 1 > pmaxe || p->nz > kPencilPPT * kPencilThreads))
        return fail(nullptr, "grid plane " + std::to_string(p->ny) + "x" + std::to_string(p->nz) +
                                 " exceeds the pencil plane path at this precision");
    gpe_handle* h = new gpe_handle();
    h->pencil = pencil;
    h->p = *p;
    h->prec = precision;
    h->N = (size_t)p->nx * p->ny * p->nz;
    h->ggp11 = p->ggp11;
    h->gamma = p->gamma;
    h->sqdv = sqrt(p->dx * p->dy * p->dz);
    {
        const char* fg = getenv("GPE_FORCE_GENERIC");
        h->force_generic = fg && fg[0] == '1';
        const char* ng = getenv("GPE_NO_GRAPH");
        h->use_graphs = !(ng && ng[0] == '1');
        const char* ll = getenv("GPE_LINE_LDS");
        h->line_reg = !(ll && ll[0] == '1');
        const char* ns = getenv("GPE_NO_SPLIT");
        h->plane_split = !(ns && ns[0] == '1') && !pencil;
    }
    auto bad = [&](const std::string& m) {
        g_create_err = m;
        gpe_destroy(h);
        return 1;
    };
    if (hipGetDevice(&h->dev) != hipSuccess) return bad("hipGetDevice failed");
    {
        const hipError_t e = init_scratch(h->dev);
        if (e != hipSuccess) return bad(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    }
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) return bad("stream create failed");
    const size_t cs = csize(h), rs = rsize(h);
    const size_t n = h->N;
    const size_t ns1 = (size_t)2 * p->ny * p->nx * p->nz;  // kz = 1..nz (the kz = 0 column is zero)
    if (hipMalloc(&h->store[0], n * cs) || hipMalloc(&h->store[1], n * cs) || hipMalloc(&h->psii, n * cs) ||
        hipMalloc(&h->V, n * rs) || hipMalloc(&h->phi, n * rs) || hipMalloc(&h->s1, ns1 * cs) ||
        hipMalloc(&h->tk, n * cs) || hipMalloc(&h->part[0], p->nx * 8) || hipMalloc(&h->part[1], p->nx * 8) ||
        hipMalloc(&h->epart, (size_t)p->nx * 8 * 8) || hipMalloc(&h->cbase, sizeof(int)))
        return bad("device allocation failed");
    if (hipMemset(h->store[0], 0, n * cs) || hipMemset(h->store[1], 0, n * cs) || hipMemset(h->psii, 0, n * cs) ||
        hipMemset(h->V, 0, n * rs) || hipMemset(h->phi, 0, n * rs) || hipMemset(h->part[0], 0, p->nx * 8) ||
        hipMemset(h->part[1], 0, p->nx * 8))
        return bad("device memset failed");
    int rc = precision == GPE_FP64 ? setup_tables<double>(h) : setup_tables<float>(h);
    if (rc) return bad(h->err);
    // launch shapes: Poisson tiles of cp kz-columns, kinetic tiles of ck k
    if (2L * p->nx > kLineMaxE) return bad("x length too large for the line pass");
    h->cp_max = std::min(p->nz, std::max(1, kLineMaxE / (2 * p->nx)));
    h->ck_max = std::min(p->nz, std::max(1, kLineMaxE / p->nx));
    h->cp = std::min(h->cp_max, 12);
    h->ck = std::min(h->ck_max, 16);
    h->plane_lds = (size_t)std::max(2L * p->ny * (p->nz + 1), (long)p->ny * (2 * p->nz + 1)) * cs;
    h->plane_stage = (size_t)(h->tw_n + 2 * (p->ny + p->nz)) * cs;
    h->line_lds = (size_t)std::max(2L * p->nx * (h->cp + 1), (long)p->nx * (h->ck + 1)) * cs;
    if (pencil) {
        h->pen_c = std::min(p->nz, pmaxe / (2 * p->ny) - 1);
        h->pen_r = std::min(std::min(p->ny, kPencilPPT * kPencilThreads / p->nz), pmaxe / (2 * p->nz + 1));
        // GPE_PENCIL_TILES="c,r" caps the tile shapes (tests: partial last tiles)
        const char* yr = getenv("GPE_PENCIL_YREG");
        h->pen_yreg = !(yr && yr[0] == '0');
        const char* zr = getenv("GPE_PENCIL_ZREG");
        h->pen_zreg = !(zr && zr[0] == '0');
        const char* pt = getenv("GPE_PENCIL_TILES");
        int tc = 0, tr = 0;
        if (pt && sscanf(pt, "%d,%d", &tc, &tr) == 2 && tc > 0 && tr > 0) {
            h->pen_c = std::min(h->pen_c, tc);
            h->pen_r = std::min(h->pen_r, tr);
        }
        // compile-time pencil shapes use their own tiles
        if (!h->force_generic) {
#define GPE_PP(NY, NZ, CC, RR)          \
    if (p->ny == NY && p->nz == NZ) { \
        h->pen_c = CC;                  \
        h->pen_r = RR;                  \
    }
            GPE_PENCIL_SHAPES(GPE_PP)
#undef GPE_PP
        }
        h->pen_lds = (size_t)std::max(2L * p->ny * (h->pen_c + 1), (long)h->pen_r * (2 * p->nz + 1)) * cs;
        h->pen_c_max = h->pen_c;
        h->pen_r_max = h->pen_r;
        const size_t nrt = p->ny;  // partials for row tiles down to one row (gpe_set_tiles)
        if (hipMalloc(&h->pen_gw, n * cs) || hipMalloc(&h->pen_h, n * cs) ||
            hipMalloc(&h->pen_part, (size_t)p->nx * nrt * sizeof(double)))
            return bad("device allocation failed");
    }
    if (hipDeviceSynchronize() != hipSuccess) return bad("device synchronize failed after setup");
    *out = h;
    return 0;
}

void gpe_destroy(gpe_handle* h) {
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    void* bufs[] = {h->store[0], h->store[1], h->psii,     h->V,       h->phi,    h->s1,  h->tk,
                    h->s1b,      h->tkb,      h->tw,       h->twl,     h->ex,     h->ey,  h->ez,  h->kx2,
                    h->ky2,      h->kz2,      h->part[0],  h->part[1], h->epart,  h->pen_gw, h->pen_h, h->pen_part};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    for (hipEvent_t e : h->ev) (void)hipEventDestroy(e);
    for (auto& kk : h->kev)
        for (auto& v : kk)
            for (hipEvent_t e : v) (void)hipEventDestroy(e);
    if (h->rt_graph) (void)hipGraphExecDestroy(h->rt_graph);
    for (hipGraphExec_t g : h->it_graph)
        if (g) (void)hipGraphExecDestroy(g);
    if (h->coef) (void)hipFree(h->coef);
    if (h->coef_host) (void)hipHostFree(h->coef_host);
    if (h->coef_ev) (void)hipEventDestroy(h->coef_ev);
    if (h->cbase) (void)hipFree(h->cbase);
    if (h->copy_stream) {
        (void)hipStreamSynchronize(h->copy_stream);
        (void)hipStreamDestroy(h->copy_stream);
    }
    if (h->snap_ev) (void)hipEventDestroy(h->snap_ev);
    if (h->copy_ev) (void)hipEventDestroy(h->copy_ev);
    if (h->snap) (void)hipFree(h->snap);
    if (h->snap_dev) (void)hipFree(h->snap_dev);
    if (h->snap_host) (void)hipHostFree(h->snap_host);
    if (h->snap_host_e) (void)hipHostFree(h->snap_host_e);
    if (h->stream && h->own_stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

int gpe_load_buffer(gpe_handle* h, const double* psi) {
    if (!h || !psi) return fail(h, "null argument");
    const size_t n2 = 2 * h->N;
    if (upload_real(h, h->store[h->cur], psi, n2, h->sqdv)) return 1;
    if (upload_real(h, h->psii, psi, n2, h->sqdv)) return 1;
    h->norm_pending = false;
    h->rt_pending = false;  // the loaded state replaces a pending RT step
    // calc_norm + normalize (solver.cpp:118-125)
    if (DISPATCH(h, do_calc_norm, h, nullptr)) return 1;
    return DISPATCH(h, do_normalize, h);
}

int gpe_warm_features(gpe_handle* h, float* x_dev, int cin, double v_scale, const float* consts) {
    if (!h || !x_dev || !consts) return fail(h, "null argument");
    if (cin < 1 || cin > 9) return fail(h, "gpe_warm_features: 1 <= cin <= 9");
    WarmConsts c{};
    for (int q = 0; q + 1 < cin; ++q) c.c[q] = consts[q];
    return DISPATCH(h, do_warm_features, h, x_dev, cin, v_scale, c);
}

int gpe_load_amplitude(gpe_handle* h, const float* amp_dev, int mirror, double* norm_out) {
    if (!h || !amp_dev) return fail(h, "null argument");
    return DISPATCH(h, do_warm_state, h, amp_dev, mirror, norm_out);
}

int gpe_load_pote(gpe_handle* h, const double* V) {
    if (!h || !V) return fail(h, "null argument");
    if (DISPATCH(h, flush_rt, h)) return 1;  // a pending half phase uses the old V
    return upload_real(h, h->V, V, h->N, 1.0);
}

int gpe_export_psi(gpe_handle* h, double* psi) {
    if (!h || !psi) return fail(h, "null argument");
    if (DISPATCH(h, materialize, h)) return 1;
    return download_real(h, psi, h->store[h->cur], 2 * h->N, 1.0 / h->sqdv);
}

int gpe_export_fi3d(gpe_handle* h, double* fi3d) {
    if (!h || !fi3d) return fail(h, "null argument");
    if (DISPATCH(h, flush_rt, h)) return 1;
    return download_real(h, fi3d, h->phi, h->N, 1.0);
}

int gpe_import_fi3d(gpe_handle* h, const double* fi3d) {
    if (!h || !fi3d) return fail(h, "null argument");
    if (DISPATCH(h, flush_rt, h)) return 1;
    return upload_real(h, h->phi, fi3d, h->N, 1.0);
}

int gpe_init_potential(gpe_handle* h, int key, const gpe_trap* t) {
    if (!h || !t) return fail(h, "null argument");
    if (key < GPE_POT_HARMONIC || key > GPE_POT_MEXICAN_ASYMETRIC_FREE) return fail(h, "unknown potential key");
    if (DISPATCH(h, flush_rt, h)) return 1;  // a pending half phase uses the old V
    InitArgs a = init_args(h);
    a.key = key;
    // parameters.cpp:37-41 (aa, b) and potentials.hpp:121-157 (the slope)
    a.omx = t->omega_x;
    a.omy = t->omega_y;
    a.omz = t->omega_z;
    a.b = 0.5 * h->p.m * pow(t->omega_x, 2);
    a.aa = h->p.m * pow(t->omega_x, 2) / 4. / pow(t->dd, 2);
    a.slope = (0.25 / (27211.4 * 11.6 * 1e9)) / t->dd;
    return DISPATCH(h, do_init_pote, h, a);
}

int gpe_init_wavefunction(gpe_handle* h, int strategy, const gpe_trap* t, int n_gauss_max, int bx, int by, int bz,
                          double sigma_cells, double phase_noise, unsigned long long seed) {
    if (!h || !t) return fail(h, "null argument");
    const gpe_params& p = h->p;
    InitArgs a = init_args(h);
    a.key = strategy;
    a.noise = phase_noise;
    a.seed = seed;
    const double Lx = p.nx * p.dx, Ly = p.ny * p.dy, Lz = p.nz * p.dz;
    auto add = [&](double x, double y, double z) -> bool {
        if (a.ncent >= kMaxCenters) return false;
        a.cx[a.ncent] = x;
        a.cy[a.ncent] = y;
        a.cz[a.ncent] = z;
        ++a.ncent;
        return true;
    };
    switch (strategy) {
    case GPE_INIT_COS:  // initializer.cpp:46-74
        a.cos_dd = t->dd;
        a.cos_r = (p.nx / 2) * p.dx;
        break;
    case GPE_INIT_GAUSS:  // initializer.cpp:76-99
        a.sx = sigma_cells > 0 ? sigma_cells * p.dx : Lx / 10.;
        a.sy = sigma_cells > 0 ? sigma_cells * p.dy : Ly / 10.;
        a.sz = sigma_cells > 0 ? sigma_cells * p.dz : Lz / 10.;
        add(0, 0, 0);
        break;
    case GPE_INIT_MULTIPLE_GAUSS:  // initializer.cpp:162-219
        if (n_gauss_max <= 0) return fail(h, "Number of Gaussian maxima must be positive.");
        a.sx = sigma_cells > 0 ? sigma_cells * p.dx : Lx / 20.;
        a.sy = sigma_cells > 0 ? sigma_cells * p.dy : Ly / 20.;
        a.sz = sigma_cells > 0 ? sigma_cells * p.dz : Lz / 20.;
        for (int idx = 0; idx < n_gauss_max; ++idx) {
            const double cx = idx % 2 ? t->dd : -t->dd;
            double cy;
            if (n_gauss_max % 2 == 1)
                cy = (idx - n_gauss_max / 2) * (p.ny * p.dy) / 2. / (n_gauss_max + 1.);
            else
                cy = (double)(0 - p.ny / 2) * p.dy + (idx / 2 + 1) * (p.ny * p.dy) / (n_gauss_max / 2 + 1);
            if (!add(cx, cy, 0)) return fail(h, "too many Gaussian maxima for the device initializer");
        }
        break;
    case GPE_INIT_SETUP_GAUSS:  // initializer.cpp:101-160
        if (bx <= 0 || by <= 0 || bz <= 0) return fail(h, "droplet counts must be positive");
        a.sx = Lx / 20.;
        a.sy = Ly / 20.;
        a.sz = Lz / 20.;
        a.zcent = 1;
        for (int ia = 0; ia < bx; ++ia)
            for (int ib = 0; ib < by; ++ib)
                for (int ic = 0; ic < bz; ++ic)
                    if (!add(-Lx / 2 + Lx / (bx + 1) * (ia + 1), -Ly / 2 + Ly / (by + 1) * (ib + 1),
                             -Lz / 2 + Lz / (bz + 1) * (ic + 1)))
                        return fail(h, "too many droplets for the device initializer");
        break;
    default:
        return fail(h, "unknown initial-state strategy");
    }
    if (DISPATCH(h, do_init_psi, h, a)) return 1;
    h->norm_pending = false;
    h->rt_pending = false;  // the new state replaces a pending RT step
    // calc_norm + normalize (solver.cpp:118-125)
    if (DISPATCH(h, do_calc_norm, h, nullptr)) return 1;
    return DISPATCH(h, do_normalize, h);
}

int gpe_export_pote(gpe_handle* h, double* V) {
    if (!h || !V) return fail(h, "null argument");
    return download_real(h, V, h->V, h->N, 1.0);
}

int gpe_set_coupling(gpe_handle* h, double ggp11, double gamma) {
    if (!h) return fail(h, "null handle");
    h->ggp11 = ggp11;
    h->gamma = gamma;
    return 0;
}

int gpe_calc_fi3d(gpe_handle* h) { return h ? DISPATCH(h, do_calc_fi3d, h) : 1; }
int gpe_calc_norm(gpe_handle* h, double* x) { return h ? DISPATCH(h, do_calc_norm, h, x) : 1; }
int gpe_normalize(gpe_handle* h) { return h ? DISPATCH(h, do_normalize, h) : 1; }
int gpe_imag_iter_step(gpe_handle* h) { return h ? DISPATCH(h, do_it_hook, h) : 1; }
int gpe_real_potential_half_step(gpe_handle* h) { return h ? DISPATCH(h, do_potential_half, h) : 1; }
int gpe_real_kinetic_step(gpe_handle* h) { return h ? DISPATCH(h, do_kinetic, h) : 1; }
int gpe_calc_energy(gpe_handle* h, gpe_energies* out) {
    if (!h || !out) return fail(h, "null argument");
    return DISPATCH(h, do_energy, h, out);
}
int gpe_snapshot_begin(gpe_handle* h) { return h ? DISPATCH(h, do_snapshot_begin, h) : 1; }
int gpe_snapshot_end(gpe_handle* h, double* psi, gpe_energies* e) {
    return h ? DISPATCH(h, do_snapshot_end, h, psi, e) : 1;
}
int gpe_calc_energy_noalloc(gpe_handle* h, double* out_dev) {
    if (!h || !out_dev) return fail(h, "null argument");
    return DISPATCH(h, do_energy_noalloc, h, out_dev);
}
int gpe_imag_time_iters(gpe_handle* h, long n) { return h ? DISPATCH(h, do_it_iters, h, n) : 1; }
int gpe_imag_time_until(gpe_handle* h, long max_steps, long check_every, double tol, long* steps_out,
                        double* energy_per_atom_out) {
    if (!h) return 1;
    if (max_steps < 0) return fail(h, "max_steps must be >= 0");
    return DISPATCH(h, do_it_until, h, max_steps, check_every, tol, steps_out, energy_per_atom_out);
}
int gpe_real_time_iters(gpe_handle* h, long n, const double* g, const double* gm) {
    return h ? DISPATCH(h, do_rt_iters, h, n, g, gm) : 1;
}
int gpe_synchronize(gpe_handle* h) {
    if (!h) return 1;
    GPE_CHECK(hipStreamSynchronize(h->stream));
    return 0;
}
int gpe_batch_create(const gpe_params* params, int n, int precision, gpe_batch** out) {
    if (!params || !out || n <= 0) return fail(nullptr, "bad batch arguments");
    *out = nullptr;
    gpe_batch* b = new gpe_batch();
    b->prec = precision;
    {
        const char* ng = getenv("GPE_NO_GRAPH");
        b->use_graphs = !(ng && ng[0] == '1');
    }
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return fail(nullptr, "stream create failed");
    }
    for (int q = 0; q < n; ++q) {
        gpe_handle* h = nullptr;
        if (gpe_create(&params[q], precision, &h) != 0) {
            gpe_batch_destroy(b);
            return 1;
        }
        if (h->pencil) {
            gpe_destroy(h);
            gpe_batch_destroy(b);
            return fail(nullptr, "batched instances need planes of the single-block plane path");
        }
        (void)hipStreamDestroy(h->stream);
        h->stream = b->stream;
        h->own_stream = false;
        b->inst.push_back(h);
    }
    b->fallback.assign(n, std::vector<double>());
    b->fell_back.assign(n, 0);
    *out = b;
    return 0;
}

void gpe_batch_destroy(gpe_batch* b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_drop_graphs(b);
    for (gpe_handle* h : b->inst) gpe_destroy(h);
    void* bufs[] = {b->d_pargs, b->d_largs, b->d_pwork, b->d_lwork,  b->d_eargs,
                    b->d_swork, b->d_ework, b->d_esum,  b->d_pmwork, b->d_lmwork};
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    if (b->h_eout) (void)hipHostFree(b->h_eout);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

gpe_handle* gpe_batch_instance(gpe_batch* b, int i) {
    if (!b || i < 0 || i >= (int)b->inst.size()) return nullptr;
    return b->inst[i];
}

const char* gpe_batch_last_error(const gpe_batch* b) { return b ? b->err.c_str() : g_create_err.c_str(); }

int gpe_batch_set_fallback(gpe_batch* b, int i, const double* psi) {
    if (!b || i < 0 || i >= (int)b->inst.size()) return bfail(b, "bad batch instance");
    if (!psi) {
        b->fallback[i].clear();
        return 0;
    }
    b->fallback[i].assign(psi, psi + 2 * b->inst[i]->N);
    return 0;
}

int gpe_batch_fallback_taken(const gpe_batch* b, int i) {
    if (!b || i < 0 || i >= (int)b->fell_back.size()) return 0;
    return b->fell_back[i] ? 1 : 0;
}

int gpe_batch_it_until(gpe_batch* b, long max_steps, long check_every, double tol, long* steps_out,
                       double* energy_per_atom_out) {
    if (!b) return fail(nullptr, "null batch");
    // bring every instance to the same parity: normalised state in buffer 0
    for (gpe_handle* h : b->inst) {
        if (DISPATCH(h, materialize, h)) return bfail(b, h->err);
        if (h->cur != 0) {
            const size_t bytes = h->N * csize(h);
            BCHECK(hipMemcpyAsync(h->store[0], h->store[1], bytes, hipMemcpyDeviceToDevice, b->stream));
            h->cur = 0;
        }
        h->pcur = 0;
    }
    b->cur = 0;
    b->pcur = 0;
    return b->prec == GPE_FP64 ? batch_it_until<double>(b, max_steps, check_every, tol, steps_out, energy_per_atom_out)
                               : batch_it_until<float>(b, max_steps, check_every, tol, steps_out, energy_per_atom_out);
}

// ---------------------------------------------------------------------------
// Autotuner.  The launch-shape variants of the fused loops are timed on a
// scratch instance of the same grid and precision (uniform field; the
// kernels have no data-dependent control flow), and the fastest is kept for
// h and cached per (grid, precision) for the process.  Variants: one or two
// blocks per plane in the real-time plane pass (bitwise identical results),
// and the register-resident or the LDS x pass (the same transform, rounded
// differently).  Untuned handles use the defaults (split, register x pass).
static void drop_graphs(gpe_handle* h) {
    if (h->rt_graph) (void)hipGraphExecDestroy(h->rt_graph);
    h->rt_graph = nullptr;
    for (auto& g : h->it_graph) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
}

// Launch tiles within their bounds (values <= 0, or pen_yreg < 0, keep the
// current one); the LDS sizes follow.
static int apply_tiles(gpe_handle* h, int cp, int ck, int pen_c, int pen_r, int pen_yreg, int pen_zreg = -1) {
    if (cp > h->cp_max || ck > h->ck_max) return fail(h, "x-pass tile beyond the built bound");
    if (h->pencil && (pen_c > h->pen_c_max || pen_r > h->pen_r_max))
        return fail(h, "pencil tile beyond the built bound");
    if (DISPATCH(h, materialize, h)) return 1;
    const gpe_params& p = h->p;
    const size_t cs = csize(h);
    if (cp > 0) h->cp = cp;
    if (ck > 0) h->ck = ck;
    h->line_lds = (size_t)std::max(2L * p.nx * (h->cp + 1), (long)p.nx * (h->ck + 1)) * cs;
    if (h->pencil) {
        if (pen_c > 0) h->pen_c = pen_c;
        if (pen_r > 0) h->pen_r = pen_r;
        if (pen_yreg >= 0) h->pen_yreg = pen_yreg != 0;
        if (pen_zreg >= 0) h->pen_zreg = pen_zreg != 0;
        h->pen_lds = (size_t)std::max(2L * p.ny * (h->pen_c + 1), (long)h->pen_r * (2 * p.nz + 1)) * cs;
    }
    drop_graphs(h);
    return 0;
}

// Tile candidates: the bound and the usual power-of-two steps below it.
static std::vector<int> tile_candidates(int bound, int first) {
    std::vector<int> c{first};
    for (int v : {bound, 16, 12, 8, 4, 2})
        if (v >= 1 && v <= bound && std::find(c.begin(), c.end(), v) == c.end()) c.push_back(v);
    return c;
}

struct TuneResult {
    bool split, line_reg, graphs;
    double rt_us, it_us;
    int cp, ck, pen_c, pen_r, pen_yreg, pen_zreg;
};
static std::mutex g_tune_mu;
static std::map<std::tuple<int, int, int, int>, TuneResult> g_tune;

static int tune_time(gpe_handle* h, long steps, double* rt_us, double* it_us) {
    gpe_handle* s = h;  // the scratch instance
    auto run = [&](bool rt, long n) -> int {
        return rt ? DISPATCH(s, do_rt_iters, s, n, nullptr, nullptr) : DISPATCH(s, do_it_iters, s, n);
    };
    for (int m = 0; m < 2; ++m) {
        const bool rt = m == 0;
        if (run(rt, kGraphSteps + 2)) return 1;  // warm-up, captures the graphs
        GPE_CHECK(hipStreamSynchronize(s->stream));
        hipEvent_t e0, e1;
        GPE_CHECK(hipEventCreate(&e0));
        GPE_CHECK(hipEventCreate(&e1));
        (void)hipEventRecord(e0, s->stream);
        const int rc = run(rt, steps);
        (void)hipEventRecord(e1, s->stream);
        (void)hipEventSynchronize(e1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (rc) return 1;
        (rt ? *rt_us : *it_us) = 1e3 * ms / steps;
    }
    return 0;
}

int gpe_autotune(gpe_handle* h, long steps, int* config, double* rt_step_us, double* it_step_us) {
    if (!h) return 1;
    if (steps < 1) return fail(h, "autotune needs steps >= 1");
    const auto key = std::make_tuple(h->p.nx, h->p.ny, h->p.nz, h->prec);
    TuneResult best{};
    bool have = false;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune.find(key);
        if (it != g_tune.end()) {
            best = it->second;
            have = true;
        }
    }
    if (!have) {
        gpe_handle* s = nullptr;
        if (gpe_create(&h->p, h->prec, &s)) return fail(h, std::string("autotune scratch instance: ") + g_create_err);
        std::vector<double> psi(2 * h->N, 1.0);
        int rc = gpe_load_buffer(s, psi.data());
        const bool reg_ok = !s->force_generic && (h->p.nx == 128 || h->p.nx == 64 || h->p.nx == 256);
        // graph replay of the steady state, or direct launches (GPE_NO_GRAPH=1
        // leaves only the latter): which one is faster depends on the grid
        const bool graphs_ok = s->use_graphs;
        double best_cost = 0;
        for (int sp = s->pencil ? 0 : 1; sp >= 0 && !rc; --sp)
            for (int lr = reg_ok ? 1 : 0; lr >= 0 && !rc; --lr)
                for (int gr = graphs_ok ? 1 : 0; gr >= 0 && !rc; --gr) {
                    rc = DISPATCH(s, materialize, s);
                    if (rc) break;
                    s->plane_split = sp != 0;
                    s->line_reg = lr != 0;
                    s->use_graphs = gr != 0;
                    drop_graphs(s);
                    double rt = 0, itu = 0;
                    rc = tune_time(s, steps, &rt, &itu);
                    const double cost = rt + itu;
                    if (!rc && (!have || cost < best_cost)) {
                        best = TuneResult{sp != 0, lr != 0, gr != 0, rt, itu, s->cp, s->ck, s->pen_c, s->pen_r,
                                          s->pen_yreg ? 1 : 0, s->pen_zreg ? 1 : 0};
                        best_cost = cost;
                        have = true;
                    }
                }
        // then the tiles of the chosen variant, one parameter at a time: the
        // LDS x pass's kz / k columns per block, the pencil pass's register y
        // passes, row tile and (LDS y passes) column tile
        auto try_tiles = [&](int cp, int ck, int pc, int pr, int yr, int zr = -1) -> int {
            if (apply_tiles(s, cp, ck, pc, pr, yr, zr)) return 1;
            double rt = 0, itu = 0;
            if (tune_time(s, steps, &rt, &itu)) return 1;
            if (rt + itu < best_cost) {
                best = TuneResult{best.split, best.line_reg, best.graphs, rt, itu, s->cp, s->ck, s->pen_c, s->pen_r,
                                  s->pen_yreg ? 1 : 0, s->pen_zreg ? 1 : 0};
                best_cost = rt + itu;
            }
            return apply_tiles(s, best.cp, best.ck, best.pen_c, best.pen_r, best.pen_yreg, best.pen_zreg);
        };
        if (!rc) {
            rc = DISPATCH(s, materialize, s);
            s->plane_split = best.split;
            s->line_reg = best.line_reg;
            s->use_graphs = best.graphs;
            drop_graphs(s);
        }
        if (!rc && !best.line_reg) {
            for (int v : tile_candidates(s->cp_max, s->cp))
                if (!rc && v != best.cp) rc = try_tiles(v, 0, 0, 0, -1);
            for (int v : tile_candidates(s->ck_max, s->ck))
                if (!rc && v != best.ck) rc = try_tiles(0, v, 0, 0, -1);
        }
        if (!rc && s->pencil) {
            if (h->p.ny == 256 && !s->force_generic) rc = try_tiles(0, 0, 0, 0, best.pen_yreg ? 0 : 1);
            if (!rc && h->p.ny == 256 && h->p.nz == 64 && !s->force_generic)
                rc = try_tiles(0, 0, 0, 0, -1, best.pen_zreg ? 0 : 1);
            for (int v : tile_candidates(s->pen_r_max, s->pen_r))
                if (!rc && v != best.pen_r) rc = try_tiles(0, 0, 0, v, -1);
            if (!best.pen_yreg || h->p.ny != 256)
                for (int v : tile_candidates(s->pen_c_max, s->pen_c))
                    if (!rc && v != best.pen_c) rc = try_tiles(0, 0, v, 0, -1);
        }
        if (rc) {
            const std::string m = s->err;
            gpe_destroy(s);
            return fail(h, "autotune: " + m);
        }
        gpe_destroy(s);
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune[key] = best;
    }
    // a pending RT state lives in the buffers of the current variant
    if (DISPATCH(h, materialize, h)) return 1;
    if (h->plane_split != best.split || h->line_reg != best.line_reg) drop_graphs(h);
    h->plane_split = best.split;
    h->line_reg = best.line_reg;
    if (apply_tiles(h, best.cp, best.ck, h->pencil ? best.pen_c : 0, h->pencil ? best.pen_r : 0,
                    h->pencil ? best.pen_yreg : -1, h->pencil ? best.pen_zreg : -1))
        return 1;
    const char* ng = getenv("GPE_NO_GRAPH");
    h->use_graphs = best.graphs && !(ng && ng[0] == '1');
    if (config) *config = (best.split ? 1 : 0) | (best.line_reg ? 2 : 0) | (h->use_graphs ? 4 : 0);
    if (rt_step_us) *rt_step_us = best.rt_us;
    if (it_step_us) *it_step_us = best.it_us;
    return 0;
}

int gpe_set_tiles(gpe_handle* h, int cp, int ck, int pen_c, int pen_r, int pen_yreg, int pen_zreg) {
    if (!h) return 1;
    return apply_tiles(h, cp, ck, pen_c, pen_r, pen_yreg, pen_zreg);
}

int gpe_get_tiles(const gpe_handle* h, int* cur, int* bound) {
    if (!h) return 1;
    if (cur) {
        cur[0] = h->cp;
        cur[1] = h->ck;
        cur[2] = h->pencil ? h->pen_c : 0;
        cur[3] = h->pencil ? h->pen_r : 0;
        cur[4] = h->pencil ? (h->pen_yreg ? 1 : 0) : 0;
        cur[5] = h->pencil ? (h->pen_zreg ? 1 : 0) : 0;
    }
    if (bound) {
        bound[0] = h->cp_max;
        bound[1] = h->ck_max;
        bound[2] = h->pencil ? h->pen_c_max : 0;
        bound[3] = h->pencil ? h->pen_r_max : 0;
        bound[4] = h->pencil ? 1 : 0;
        bound[5] = h->pencil ? 1 : 0;
    }
    return 0;
}

// Batch autotuner: per x length (one x-pass launch per length in the batch),
// the register x pass where it exists and the LDS x pass at each kz-column
// tile, timed as whole batched IT steps on a scratch batch of the same grids
// (uniform fields), one length at a time with the others at their current or
// already chosen variant.  The plane pass has one form in the batch.
int gpe_batch_autotune(gpe_batch* b, long steps, int* config) {
    if (!b) return fail(nullptr, "null batch");
    if (steps < 1) return bfail(b, "autotune needs steps >= 1");
    const int n = (int)b->inst.size();
    std::vector<gpe_params> ps;
    for (gpe_handle* h : b->inst) ps.push_back(h->p);
    gpe_batch* s = nullptr;
    if (gpe_batch_create(ps.data(), n, b->prec, &s)) return bfail(b, "autotune scratch batch: " + g_create_err);
    int rc = 0;
    for (int q = 0; q < n && !rc; ++q) {
        gpe_handle* hs = s->inst[q];
        hs->line_reg = b->inst[q]->line_reg;
        std::vector<double> psi(2 * hs->N, 1.0);
        rc = gpe_load_buffer(hs, psi.data()) || apply_tiles(hs, b->inst[q]->cp, b->inst[q]->ck, 0, 0, -1);
        if (rc) (void)bfail(b, hs->err);
    }
    if (!rc && hipMalloc((void**)&s->d_esum, (size_t)n * 16 * sizeof(double)) != hipSuccess)
        rc = bfail(b, "autotune scratch batch: device allocation failed");
    auto time_batch = [&](double* us) -> int {
        for (gpe_handle* h : s->inst) {
            if (DISPATCH(h, materialize, h)) return 1;
            if (h->cur != 0) {
                BCHECK(hipMemcpyAsync(h->store[0], h->store[1], h->N * csize(h), hipMemcpyDeviceToDevice, s->stream));
                h->cur = 0;
            }
            h->pcur = 0;
        }
        s->cur = s->pcur = 0;
        s->active.clear();
        for (int q = 0; q < n; ++q) s->active.push_back(q);
        const int r = s->prec == GPE_FP64 ? batch_setup<double>(s) : batch_setup<float>(s);
        if (r) return bfail(b, s->err);
        auto run = [&](long k) { return s->prec == GPE_FP64 ? batch_it<double>(s, k) : batch_it<float>(s, k); };
        if (run(kGraphSteps + 2)) return bfail(b, s->err);  // warm-up, captures the graphs
        BCHECK(hipStreamSynchronize(s->stream));
        hipEvent_t e0, e1;
        BCHECK(hipEventCreate(&e0));
        BCHECK(hipEventCreate(&e1));
        (void)hipEventRecord(e0, s->stream);
        const int rr = run(steps);
        (void)hipEventRecord(e1, s->stream);
        (void)hipEventSynchronize(e1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (rr) return bfail(b, s->err);
        *us = 1e3 * ms / steps;
        return 0;
    };
    std::vector<int> lengths;
    for (gpe_handle* h : s->inst)
        if (std::find(lengths.begin(), lengths.end(), h->p.nx) == lengths.end()) lengths.push_back(h->p.nx);
    double best_us = 0;
    if (!rc) rc = time_batch(&best_us);
    for (int nx : lengths) {
        if (rc) break;
        gpe_handle* h0 = nullptr;
        for (gpe_handle* h : s->inst)
            if (h->p.nx == nx) {
                h0 = h;
                break;
            }
        const bool reg_ok = !h0->force_generic && (nx == 128 || nx == 64 || nx == 256);
        bool best_reg = h0->line_reg && reg_ok;
        int best_cp = h0->cp;
        std::vector<std::pair<bool, int>> cands;
        if (reg_ok) cands.push_back({true, best_cp});
        for (int v : tile_candidates(h0->cp_max, h0->cp)) cands.push_back({false, v});
        for (const auto& c : cands) {
            if (rc) break;
            if (c.first == best_reg && (c.first || c.second == best_cp)) continue;  // timed already
            for (gpe_handle* h : s->inst)
                if (h->p.nx == nx) {
                    h->line_reg = c.first;
                    if (apply_tiles(h, c.second, 0, 0, 0, -1)) rc = bfail(b, h->err);
                }
            double us = 0;
            if (!rc) rc = time_batch(&us);
            if (!rc && us < best_us) {
                best_us = us;
                best_reg = c.first;
                best_cp = c.second;
            }
        }
        for (gpe_handle* h : s->inst)
            if (h->p.nx == nx) {
                h->line_reg = best_reg;
                if (!rc && apply_tiles(h, best_cp, 0, 0, 0, -1)) rc = bfail(b, h->err);
            }
    }
    if (!rc) {
        for (int q = 0; q < n && !rc; ++q) {
            gpe_handle* h = b->inst[q];
            h->line_reg = s->inst[q]->line_reg;
            if (apply_tiles(h, s->inst[q]->cp, 0, 0, 0, -1)) rc = bfail(b, h->err);
            if (config) config[q] = (h->line_reg ? 2 : 0) | (h->cp << 8);
        }
        batch_drop_graphs(b);
    }
    gpe_batch_destroy(s);
    return rc;
}

int gpe_set_timing(gpe_handle* h, int on) {
    if (!h) return 1;
    h->timing = on != 0;
    return 0;
}
int gpe_get_timing(gpe_handle* h, double* plane_us, double* line_us, double* total_ms) {
    if (!h) return 1;
    if (plane_us) *plane_us = h->t_plane_us;
    if (line_us) *line_us = h->t_line_us;
    if (total_ms) *total_ms = h->t_total_ms;
    return 0;
}

}  // extern "C"
