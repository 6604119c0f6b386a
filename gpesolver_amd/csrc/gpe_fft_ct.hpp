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

// cos / sin of 2 pi m / 32 and 2 pi m / 24.
struct TwC {
    static constexpr double c32[32] = {
        1.0, 0.980785280403230449126, 0.923879532511286756128, 0.831469612302545237079,
        0.707106781186547524401, 0.555570233019602224743, 0.382683432365089771728, 0.195090322016128267848,
        0.0, -0.195090322016128267848, -0.382683432365089771728, -0.555570233019602224743,
        -0.707106781186547524401, -0.831469612302545237079, -0.923879532511286756128, -0.980785280403230449126,
        -1.0, -0.980785280403230449126, -0.923879532511286756128, -0.831469612302545237079,
        -0.707106781186547524401, -0.555570233019602224743, -0.382683432365089771728, -0.195090322016128267848,
        0.0, 0.195090322016128267848, 0.382683432365089771728, 0.555570233019602224743,
        0.707106781186547524401, 0.831469612302545237079, 0.923879532511286756128, 0.980785280403230449126};
    static constexpr double s32[32] = {
        0.0, 0.195090322016128267848, 0.382683432365089771728, 0.555570233019602224743,
        0.707106781186547524401, 0.831469612302545237079, 0.923879532511286756128, 0.980785280403230449126,
        1.0, 0.980785280403230449126, 0.923879532511286756128, 0.831469612302545237079,
        0.707106781186547524401, 0.555570233019602224743, 0.382683432365089771728, 0.195090322016128267848,
        0.0, -0.195090322016128267848, -0.382683432365089771728, -0.555570233019602224743,
        -0.707106781186547524401, -0.831469612302545237079, -0.923879532511286756128, -0.980785280403230449126,
        -1.0, -0.980785280403230449126, -0.923879532511286756128, -0.831469612302545237079,
        -0.707106781186547524401, -0.555570233019602224743, -0.382683432365089771728, -0.195090322016128267848};
    static constexpr double c24[24] = {
        1.0, 0.96592582628906828675, 0.866025403784438646764, 0.707106781186547524401,
        0.5, 0.258819045102520762349, 0.0, -0.258819045102520762349,
        -0.5, -0.707106781186547524401, -0.866025403784438646764, -0.96592582628906828675,
        -1.0, -0.96592582628906828675, -0.866025403784438646764, -0.707106781186547524401,
        -0.5, -0.258819045102520762349, 0.0, 0.258819045102520762349,
        0.5, 0.707106781186547524401, 0.866025403784438646764, 0.96592582628906828675};
    static constexpr double s24[24] = {
        0.0, 0.258819045102520762349, 0.5, 0.707106781186547524401,
        0.866025403784438646764, 0.96592582628906828675, 1.0, 0.96592582628906828675,
        0.866025403784438646764, 0.707106781186547524401, 0.5, 0.258819045102520762349,
        0.0, -0.258819045102520762349, -0.5, -0.707106781186547524401,
        -0.866025403784438646764, -0.96592582628906828675, -1.0, -0.96592582628906828675,
        -0.866025403784438646764, -0.707106781186547524401, -0.5, -0.258819045102520762349};
};

// multiply by SIGN * i
template <typename T, int SIGN>
__device__ __forceinline__ cplx<T> mul_si(cplx<T> a) {
    return SIGN < 0 ? mk(a.y, -a.x) : mk(-a.y, a.x);
}

// cos / sin of 2 pi m / N from the tables (any integer m).
template <int N>
__device__ __forceinline__ void wtab(int m, double* c, double* s) {
    m = ((m % N) + N) % N;
    if constexpr (32 % N == 0) {
        *c = TwC::c32[m * (32 / N)];
        *s = TwC::s32[m * (32 / N)];
    } else {
        static_assert(24 % N == 0, "twiddle base");
        *c = TwC::c24[m * (24 / N)];
        *s = TwC::s24[m * (24 / N)];
    }
}

// a * i^q (SIGN -1: a * (-i)^q)
template <typename T, int SIGN>
__device__ __forceinline__ cplx<T> quarter(cplx<T> a, int q) {
    q &= 3;
    return q == 0 ? a : q == 1 ? mul_si<T, SIGN>(a) : q == 2 ? mk(-a.x, -a.y) : mul_si<T, -SIGN>(a);
}

// a * W_N^m (W = exp(SIGN 2 pi i / N), N | 32 or N | 24).  Quarter turns are
// exact; odd multiples of pi/4 are a sum and a difference scaled by 1/sqrt 2.
// EX (the transforms of psi, whose norm every real-time step must keep):
// the constants enter as hi + lo pairs (mulk, cmulk), one rounding of the
// exact product.  The rounded cos / sin pairs of pi/8, pi/6 ... fall short
// of |w| = 1 by 2-6e-8, and even the increment form 1 + (cos r - 1, sin r)
// of the residual angle leaves |w| - 1 at 1-5e-9: over thousands of steps
// either shows as a norm drift.  Otherwise the constants are rounded to T.
template <typename T, int N, int SIGN, bool EX>
__device__ __forceinline__ cplx<T> mul_w(cplx<T> a, int m) {
    m %= N;
    if ((4 * m) % N == 0) return quarter<T, SIGN>(a, 4 * m / N);
    double c, s;
    wtab<N>(m, &c, &s);
    s *= SIGN;
    constexpr double r = 0.70710678118654752440;
    if ((8 * m) % N == 0) {
        // c = +-r, s = +-r
        const T re = (c > 0 ? a.x : -a.x) - (s > 0 ? a.y : -a.y);
        const T im = (s > 0 ? a.x : -a.x) + (c > 0 ? a.y : -a.y);
        if constexpr (EX) return mk(mulk(re, r), mulk(im, r));
        return mk(re * (T)r, im * (T)r);
    }
    if constexpr (EX) return cmulk(a, c, s);
    return cmul(a, mk((T)c, (T)s));
}

template <typename T, int R, int SIGN, bool EX = false>
struct Dft;

template <typename T, int SIGN, bool EX>
struct Dft<T, 1, SIGN, EX> {
    __device__ __forceinline__ static void run(cplx<T>*) {}
};

template <typename T, int SIGN, bool EX>
struct Dft<T, 2, SIGN, EX> {
    __device__ __forceinline__ static void run(cplx<T>* v) {
        cplx<T> t = v[0];
        v[0] = cadd(t, v[1]);
        v[1] = csub(t, v[1]);
    }
};

template <typename T, int SIGN, bool EX>
struct Dft<T, 3, SIGN, EX> {
    __device__ __forceinline__ static void run(cplx<T>* v) {
        constexpr double h = 0.86602540378443864676;
        cplx<T> s = cadd(v[1], v[2]);
        cplx<T> d = csub(v[1], v[2]);
        cplx<T> m = mk(v[0].x - T(0.5) * s.x, v[0].y - T(0.5) * s.y);
        cplx<T> e;
        if constexpr (EX)
            e = mul_si<T, SIGN>(mk(mulk(d.x, h), mulk(d.y, h)));
        else
            e = mul_si<T, SIGN>(mk((T)h * d.x, (T)h * d.y));
        v[0] = cadd(v[0], s);
        v[1] = cadd(m, e);
        v[2] = csub(m, e);
    }
};

template <typename T, int SIGN, bool EX>
struct Dft<T, 4, SIGN, EX> {
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
template <typename T, int A, int B, int SIGN, bool EX>
__device__ __forceinline__ void dft_ab(cplx<T>* v) {
    cplx<T> y[A * B];
#pragma unroll
    for (int b = 0; b < B; ++b) {
        cplx<T> t[A];
#pragma unroll
        for (int a = 0; a < A; ++a) t[a] = v[B * a + b];
        Dft<T, A, SIGN, EX>::run(t);
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) y[b * A + k1] = mul_w<T, A * B, SIGN, EX>(t[k1], b * k1);
    }
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) {
        cplx<T> t[B];
#pragma unroll
        for (int b = 0; b < B; ++b) t[b] = y[b * A + k1];
        Dft<T, B, SIGN, EX>::run(t);
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) v[k1 + A * k2] = t[k2];
    }
}

template <typename T, int SIGN, bool EX>
struct Dft<T, 6, SIGN, EX> {
    __device__ __forceinline__ static void run(cplx<T>* v) { dft_ab<T, 2, 3, SIGN, EX>(v); }
};
template <typename T, int SIGN, bool EX>
struct Dft<T, 8, SIGN, EX> {
    __device__ __forceinline__ static void run(cplx<T>* v) { dft_ab<T, 2, 4, SIGN, EX>(v); }
};
template <typename T, int SIGN, bool EX>
struct Dft<T, 12, SIGN, EX> {
    __device__ __forceinline__ static void run(cplx<T>* v) { dft_ab<T, 4, 3, SIGN, EX>(v); }
};
template <typename T, int SIGN, bool EX>
struct Dft<T, 16, SIGN, EX> {
    __device__ __forceinline__ static void run(cplx<T>* v) { dft_ab<T, 4, 4, SIGN, EX>(v); }
};
template <typename T, int SIGN, bool EX>
struct Dft<T, 32, SIGN, EX> {
    __device__ __forceinline__ static void run(cplx<T>* v) { dft_ab<T, 4, 8, SIGN, EX>(v); }
};

// 64 = 8 x 8 rather than 16 x 4: on a tile of 64-point lines both passes then
// run twice the butterflies (threads) of a radix-16 pass.
constexpr int pick_radix(int L) {
    return (L == 64) ? 8
           : (L % 16 == 0) ? 16
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
// EX: twiddles enter as the table value plus the lo table twl (tw + twl =
// the FP64 twiddle to ~1e-15) and the register DFTs take the EX constants.
template <typename T, int L, int R, int NS, int NT, int MAXL, int SIGN, bool EX>
__device__ __forceinline__ void step(cplx<T>* buf, int nlines, int es, int ls, const cplx<T>* __restrict__ tw,
                                     const cplx<T>* __restrict__ twl) {
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
                    if constexpr (EX) {
                        const cplx<T> tl = twl[r * k * TSTR];
                        v[q][r] = (SIGN < 0) ? cmul2(v[q][r], t, tl) : cmulc2(v[q][r], t, tl);
                    } else {
                        v[q][r] = (SIGN < 0) ? cmul(v[q][r], t) : cmulc(v[q][r], t);
                    }
                }
            }
            Dft<T, R, SIGN, EX>::run(v[q]);
            cplx<T>* base = buf + line * ls + ((j / NS) * NS * R + k) * es;
#pragma unroll
            for (int r = 0; r < R; ++r) base[r * NS * es] = v[q][r];
        }
    }
    __syncthreads();
}

template <typename T, int L, int NS, int NT, int MAXL, int SIGN, bool EX>
__device__ __forceinline__ void run_steps(cplx<T>* buf, int nlines, int es, int ls, const cplx<T>* __restrict__ tw,
                                          const cplx<T>* __restrict__ twl) {
    if constexpr (NS < L) {
        constexpr int R = pick_radix(L / NS);
        static_assert(R > 1, "unsupported FFT length");
        step<T, L, R, NS, NT, MAXL, SIGN, EX>(buf, nlines, es, ls, tw, twl);
        run_steps<T, L, NS * R, NT, MAXL, SIGN, EX>(buf, nlines, es, ls, tw, twl);
    }
}

// Unnormalised transform of `nlines` (<= MAXL) lines of length L held in LDS.
// tw: the table W_L^m (m in [0, L)); with EX (FP32 transforms of psi) twl
// holds the lo parts W_L^m - tw[m].  Callers place a barrier between the
// writes that filled the tile and this call.
template <typename T, int L, int NT, int MAXL, int SIGN, bool EX = false>
__device__ __forceinline__ void fft(cplx<T>* buf, int nlines, int es, int ls, const cplx<T>* __restrict__ tw,
                                    const cplx<T>* __restrict__ twl = nullptr) {
    static_assert(supported_length(L), "FFT length must be 2^a 3^b");
    constexpr bool X = EX && sizeof(T) == 4;
    run_steps<T, L, 1, NT, MAXL, SIGN, X>(buf, nlines, es, ls, tw, twl);
}

}  // namespace ct
}  // namespace gpe
