// gpe_fft_ct.hpp -- compile-time-specialised batched in-LDS FFT (gfx950).
//
// Same self-sorting Stockham formulation as gpe_fft.hpp, but the line length
// is a template constant and each step uses a large radix (16, 12, 8, 6, 4, 3
// or 2) whose DFT runs in registers.  Every length of the solver (2^a 3^b up
// to 256) then takes at most two LDS round trips per transform, and every
// index division is by a compile-time constant.
#pragma once
#include "gpe_fft.hpp"

namespace gpe {
namespace ct {

// cos / sin of 2 pi m / 16 and 2 pi m / 12.
struct TwC {
    static constexpr double c16[16] = {1.0,
                                       0.92387953251128675613,
                                       0.70710678118654752440,
                                       0.38268343236508977173,
                                       0.0,
                                       -0.38268343236508977173,
                                       -0.70710678118654752440,
                                       -0.92387953251128675613,
                                       -1.0,
                                       -0.92387953251128675613,
                                       -0.70710678118654752440,
                                       -0.38268343236508977173,
                                       0.0,
                                       0.38268343236508977173,
                                       0.70710678118654752440,
                                       0.92387953251128675613};
    static constexpr double s16[16] = {0.0,
                                       0.38268343236508977173,
                                       0.70710678118654752440,
                                       0.92387953251128675613,
                                       1.0,
                                       0.92387953251128675613,
                                       0.70710678118654752440,
                                       0.38268343236508977173,
                                       0.0,
                                       -0.38268343236508977173,
                                       -0.70710678118654752440,
                                       -0.92387953251128675613,
                                       -1.0,
                                       -0.92387953251128675613,
                                       -0.70710678118654752440,
                                       -0.38268343236508977173};
    static constexpr double c12[12] = {1.0,  0.86602540378443864676,  0.5,  0.0, -0.5, -0.86602540378443864676,
                                       -1.0, -0.86602540378443864676, -0.5, 0.0, 0.5,  0.86602540378443864676};
    static constexpr double s12[12] = {0.0, 0.5,  0.86602540378443864676,  1.0,  0.86602540378443864676,  0.5,
                                       0.0, -0.5, -0.86602540378443864676, -1.0, -0.86602540378443864676, -0.5};
};

// W_N^m with sign (-1: exp(-2 pi i m/N), +1: exp(+2 pi i m/N)), N | 16 or N | 12.
template <typename T, int N, int SIGN>
__device__ __forceinline__ cplx<T> wconst(int m) {
    m %= N;
    if constexpr (16 % N == 0) {
        const int t = m * (16 / N);
        return mk<T>((T)TwC::c16[t], (T)(SIGN * TwC::s16[t]));
    } else {
        static_assert(12 % N == 0, "twiddle base");
        const int t = m * (12 / N);
        return mk<T>((T)TwC::c12[t], (T)(SIGN * TwC::s12[t]));
    }
}

// multiply by SIGN * i
template <typename T, int SIGN>
__device__ __forceinline__ cplx<T> mul_si(cplx<T> a) {
    return SIGN < 0 ? mk(a.y, -a.x) : mk(-a.y, a.x);
}

template <typename T, int R, int SIGN>
struct Dft;

template <typename T, int SIGN>
struct Dft<T, 1, SIGN> {
    __device__ __forceinline__ static void run(cplx<T>*) {}
};

template <typename T, int SIGN>
struct Dft<T, 2, SIGN> {
    __device__ __forceinline__ static void run(cplx<T>* v) {
        cplx<T> t = v[0];
        v[0] = cadd(t, v[1]);
        v[1] = csub(t, v[1]);
    }
};

template <typename T, int SIGN>
struct Dft<T, 3, SIGN> {
    __device__ __forceinline__ static void run(cplx<T>* v) {
        const T h = T(0.86602540378443864676);
        cplx<T> s = cadd(v[1], v[2]);
        cplx<T> d = csub(v[1], v[2]);
        cplx<T> m = mk(v[0].x - T(0.5) * s.x, v[0].y - T(0.5) * s.y);
        cplx<T> e = mul_si<T, SIGN>(mk(h * d.x, h * d.y));
        v[0] = cadd(v[0], s);
        v[1] = cadd(m, e);
        v[2] = csub(m, e);
    }
};

template <typename T, int SIGN>
struct Dft<T, 4, SIGN> {
    __device__ __forceinline__ static void run(cplx<T>* v) {
        cplx<T> t0 = cadd(v[0], v[2]);
        cplx<T> t1 = csub(v[0], v[2]);
        cplx<T> t2 = cadd(v[1], v[3]);
        cplx<T> d = mul_si<T, SIGN>(csub(v[1], v[3]));
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, d);
        v[3] = csub(t1, d);
    }
};

// Composite DFT of length A*B in registers: n = B a + b, k = k1 + A k2.
template <typename T, int A, int B, int SIGN>
__device__ __forceinline__ void dft_ab(cplx<T>* v) {
    cplx<T> y[A * B];
#pragma unroll
    for (int b = 0; b < B; ++b) {
        cplx<T> t[A];
#pragma unroll
        for (int a = 0; a < A; ++a) t[a] = v[B * a + b];
        Dft<T, A, SIGN>::run(t);
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) {
            if (b * k1 == 0)
                y[b * A + k1] = t[k1];
            else
                y[b * A + k1] = cmul(t[k1], wconst<T, A * B, SIGN>(b * k1));
        }
    }
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) {
        cplx<T> t[B];
#pragma unroll
        for (int b = 0; b < B; ++b) t[b] = y[b * A + k1];
        Dft<T, B, SIGN>::run(t);
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) v[k1 + A * k2] = t[k2];
    }
}

template <typename T, int SIGN>
struct Dft<T, 6, SIGN> {
    __device__ __forceinline__ static void run(cplx<T>* v) { dft_ab<T, 2, 3, SIGN>(v); }
};
template <typename T, int SIGN>
struct Dft<T, 8, SIGN> {
    __device__ __forceinline__ static void run(cplx<T>* v) { dft_ab<T, 2, 4, SIGN>(v); }
};
template <typename T, int SIGN>
struct Dft<T, 12, SIGN> {
    __device__ __forceinline__ static void run(cplx<T>* v) { dft_ab<T, 4, 3, SIGN>(v); }
};
template <typename T, int SIGN>
struct Dft<T, 16, SIGN> {
    __device__ __forceinline__ static void run(cplx<T>* v) { dft_ab<T, 4, 4, SIGN>(v); }
};

constexpr int pick_radix(int L) {
    return (L % 16 == 0) ? 16
           : (L % 12 == 0) ? 12
           : (L % 8 == 0)  ? 8
           : (L % 6 == 0)  ? 6
           : (L % 4 == 0)  ? 4
           : (L % 3 == 0)  ? 3
           : (L % 2 == 0)  ? 2
                           : 1;
}

constexpr bool supported_length(int L) {
    while (L > 1) {
        const int r = pick_radix(L);
        if (r == 1) return false;
        L /= r;
    }
    return true;
}

// One Stockham step: radix R, sub-transform size NS, on `nlines` lines of
// length L in LDS (element e of line l at buf[l*ls + e*es]).  MAXL bounds the
// number of lines (compile time) so each thread's butterflies fit in registers.
// Butterfly w works on line w % nlines, so the lanes of a wave touch
// neighbouring lines: with ls == 1 or ls odd their LDS addresses fall in
// distinct banks for every read and write.  (The butterfly-major order puts
// the lanes' writes R*es apart, which is a multiple of the bank count for
// the padded plane layouts: a 16-way conflict on the first step.)
template <typename T, int L, int R, int NS, int NT, int MAXL, int SIGN>
__device__ __forceinline__ void step(cplx<T>* buf, int nlines, int es, int ls, const cplx<T>* __restrict__ tw) {
    constexpr int NB = L / R;
    constexpr int MAXQ = (NB * MAXL + NT - 1) / NT;
    constexpr int TSTR = L / (NS * R);
    const int total = NB * nlines;
    cplx<T> v[MAXQ][R];
#pragma unroll
    for (int q = 0; q < MAXQ; ++q) {
        const int w = threadIdx.x + q * NT;
        if (w < total) {
            const int j = w / nlines;
            const int line = w - j * nlines;
            const cplx<T>* base = buf + line * ls + j * es;
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = base[r * NB * es];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXQ; ++q) {
        const int w = threadIdx.x + q * NT;
        if (w < total) {
            const int j = w / nlines;
            const int line = w - j * nlines;
            const int k = j % NS;
            if constexpr (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    const cplx<T> t = tw[r * k * TSTR];
                    v[q][r] = (SIGN < 0) ? cmul(v[q][r], t) : cmulc(v[q][r], t);
                }
            }
            Dft<T, R, SIGN>::run(v[q]);
            cplx<T>* base = buf + line * ls + ((j / NS) * NS * R + k) * es;
#pragma unroll
            for (int r = 0; r < R; ++r) base[r * NS * es] = v[q][r];
        }
    }
    __syncthreads();
}

template <typename T, int L, int NS, int NT, int MAXL, int SIGN>
__device__ __forceinline__ void run_steps(cplx<T>* buf, int nlines, int es, int ls, const cplx<T>* __restrict__ tw) {
    if constexpr (NS < L) {
        constexpr int R = pick_radix(L / NS);
        static_assert(R > 1, "unsupported FFT length");
        step<T, L, R, NS, NT, MAXL, SIGN>(buf, nlines, es, ls, tw);
        run_steps<T, L, NS * R, NT, MAXL, SIGN>(buf, nlines, es, ls, tw);
    }
}

// Unnormalised transform of `nlines` (<= MAXL) lines of length L held in LDS.
// tw: the table W_L^m (m in [0, L)).  Callers place a barrier between the
// writes that filled the tile and this call.
template <typename T, int L, int NT, int MAXL, int SIGN>
__device__ __forceinline__ void fft(cplx<T>* buf, int nlines, int es, int ls, const cplx<T>* __restrict__ tw) {
    static_assert(supported_length(L), "FFT length must be 2^a 3^b");
    run_steps<T, L, 1, NT, MAXL, SIGN>(buf, nlines, es, ls, tw);
}

}  // namespace ct
}  // namespace gpe
