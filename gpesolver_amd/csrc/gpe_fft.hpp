// gpe_fft.hpp -- batched in-LDS Stockham FFT for gfx950 (CDNA4).
//
// Lines of a tile held in LDS are transformed in place by all threads of the
// block.  Each Stockham step loads every butterfly's inputs of the whole tile
// into registers, passes a barrier, then writes the butterfly outputs to their
// self-sorted positions, so one LDS buffer suffices (no ping-pong).  Radices
// 4, 2, 3, 5 and 7 cover every length 2^a 3^b 5^c 7^d.
//
// Step with radix R over a line of length L, sub-transform size Ns:
//   butterfly j in [0, L/R):  k = j mod Ns
//   a_q = x[j + q L/R] * W_L^(sign * q k L/(Ns R))      q = 0..R-1
//   a = DFT_R(a, sign)
//   x[(j/Ns) Ns R + k + q Ns] = a_q
// starting from Ns = 1 and multiplying Ns by R after each step.
#pragma once
#include <hip/hip_runtime.h>

namespace gpe {

template <typename T>
struct cplx {
    T x, y;
};

template <typename T>
__device__ __forceinline__ cplx<T> mk(T a, T b) {
    cplx<T> r;
    r.x = a;
    r.y = b;
    return r;
}
template <typename T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
    return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
template <typename T>
__device__ __forceinline__ cplx<T> cmulc(cplx<T> a, cplx<T> b) {
    return mk(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <typename T>
__device__ __forceinline__ cplx<T> cscale(cplx<T> a, T s) { return mk(a.x * s, a.y * s); }

// Radix plan of one line length, stored in a small struct passed by value.
struct LinePlan {
    int L;        // line length
    int nst;      // number of Stockham steps
    int tw;       // offset of this length's twiddle table (W_L^m, m in [0,L))
    int radix[12];
};

constexpr int kMaxStages = 12;

template <typename T, int R>
__device__ __forceinline__ void dft_small(cplx<T> (&a)[R], int sign) {
    if constexpr (R == 2) {
        cplx<T> t = a[0];
        a[0] = cadd(t, a[1]);
        a[1] = csub(t, a[1]);
    } else if constexpr (R == 4) {
        cplx<T> t0 = cadd(a[0], a[2]);
        cplx<T> t1 = csub(a[0], a[2]);
        cplx<T> t2 = cadd(a[1], a[3]);
        cplx<T> t3 = csub(a[1], a[3]);
        // d = t3 * (sign * i)
        cplx<T> d = (sign < 0) ? mk(t3.y, -t3.x) : mk(-t3.y, t3.x);
        a[0] = cadd(t0, t2);
        a[2] = csub(t0, t2);
        a[1] = cadd(t1, d);
        a[3] = csub(t1, d);
    } else if constexpr (R == 3) {
        const T h = T(0.86602540378443864676);  // sqrt(3)/2
        cplx<T> s = cadd(a[1], a[2]);
        cplx<T> d = csub(a[1], a[2]);
        cplx<T> m = mk(a[0].x - T(0.5) * s.x, a[0].y - T(0.5) * s.y);
        // e = sign * i * h * d
        cplx<T> e = (sign < 0) ? mk(h * d.y, -h * d.x) : mk(-h * d.y, h * d.x);
        a[0] = cadd(a[0], s);
        a[1] = cadd(m, e);
        a[2] = csub(m, e);
    } else if constexpr (R == 5 || R == 7) {
        // direct DFT with the R-th roots of unity (cos / sin of 2 pi m / R)
        constexpr double c5[5] = {1.0, 0.30901699437494742410, -0.80901699437494742410, -0.80901699437494742410,
                                  0.30901699437494742410};
        constexpr double s5[5] = {0.0, 0.95105651629515357212, 0.58778525229247312917, -0.58778525229247312917,
                                  -0.95105651629515357212};
        constexpr double c7[7] = {1.0,
                                  0.62348980185873353053,
                                  -0.22252093395631440429,
                                  -0.90096886790241912624,
                                  -0.90096886790241912624,
                                  -0.22252093395631440429,
                                  0.62348980185873353053};
        constexpr double s7[7] = {0.0,
                                  0.78183148246802980871,
                                  0.97492791218182360702,
                                  0.43388373911755812048,
                                  -0.43388373911755812048,
                                  -0.97492791218182360702,
                                  -0.78183148246802980871};
        cplx<T> out[R];
#pragma unroll
        for (int q = 0; q < R; ++q) {
            cplx<T> acc = a[0];
#pragma unroll
            for (int t = 1; t < R; ++t) {
                const int m = (q * t) % R;
                const T c = (T)(R == 5 ? c5[m] : c7[m]);
                const T sn = (T)(sign * (R == 5 ? s5[m] : s7[m]));  // W^(sign m) = c + i sign s
                acc = cadd(acc, mk(a[t].x * c - a[t].y * sn, a[t].x * sn + a[t].y * c));
            }
            out[q] = acc;
        }
#pragma unroll
        for (int q = 0; q < R; ++q) a[q] = out[q];
    }
}

// One Stockham step of radix R on `nlines` lines held in LDS.
// Element e of line l lives at buf[l*ls + e*es].  tw points at W_L^m.
// MAXE bounds the number of complex elements of the tile (compile time) so
// that each thread's butterflies fit in registers.
template <typename T, int NT, int MAXE, int R>
__device__ __forceinline__ void stockham_step(cplx<T>* buf, int L, int Ns, int nlines, int es, int ls, int sign,
                                              const cplx<T>* __restrict__ tw) {
    constexpr int MAXQ = (MAXE / R + NT - 1) / NT;
    const int nb = L / R;
    const int total = nb * nlines;
    const int tstride = L / (Ns * R);
    cplx<T> v[MAXQ][R];
#pragma unroll
    for (int q = 0; q < MAXQ; ++q) {
        const int w = threadIdx.x + q * NT;
        if (w < total) {
            const int line = w / nb;
            const int j = w - line * nb;
            cplx<T>* base = buf + line * ls;
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = base[(j + r * nb) * es];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXQ; ++q) {
        const int w = threadIdx.x + q * NT;
        if (w < total) {
            const int line = w / nb;
            const int j = w - line * nb;
            const int k = j % Ns;
            if (k) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    cplx<T> t = tw[r * k * tstride];
                    v[q][r] = (sign < 0) ? cmul(v[q][r], t) : cmulc(v[q][r], t);
                }
            }
            dft_small<T, R>(v[q], sign);
            cplx<T>* base = buf + line * ls;
            const int o = (j / Ns) * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) base[(o + r * Ns) * es] = v[q][r];
        }
    }
    __syncthreads();
}

// Full transform of every line (unnormalised; sign -1 forward, +1 backward).
// Requires a __syncthreads() between the writes that filled the tile and this
// call (callers provide it).
template <typename T, int NT, int MAXE>
__device__ void lds_fft(cplx<T>* buf, const LinePlan& pl, int nlines, int es, int ls, int sign,
                        const cplx<T>* __restrict__ twbase) {
    const cplx<T>* tw = twbase + pl.tw;
    int Ns = 1;
    for (int s = 0; s < pl.nst; ++s) {
        const int R = pl.radix[s];
        if (R == 4)
            stockham_step<T, NT, MAXE, 4>(buf, pl.L, Ns, nlines, es, ls, sign, tw);
        else if (R == 2)
            stockham_step<T, NT, MAXE, 2>(buf, pl.L, Ns, nlines, es, ls, sign, tw);
        else if (R == 3)
            stockham_step<T, NT, MAXE, 3>(buf, pl.L, Ns, nlines, es, ls, sign, tw);
        else if (R == 5)
            stockham_step<T, NT, MAXE, 5>(buf, pl.L, Ns, nlines, es, ls, sign, tw);
        else
            stockham_step<T, NT, MAXE, 7>(buf, pl.L, Ns, nlines, es, ls, sign, tw);
        Ns *= R;
    }
}

}  // namespace gpe
