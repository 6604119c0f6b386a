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

// Products with FP64 constants.  In FP32 the constant is split into hi + lo
// FP32 parts, so the product carries one rounding of the exact value instead
// of the rounding of the constant: rounded to nearest, the FP32 values of
// cos / sin of pi/8, pi/6 and pi/4 all leave |w| below 1 by 2-6e-8, a
// magnitude loss that every transform pass using them would add to the norm.
template <typename T>
__device__ __forceinline__ T mulk(T x, double k) {
    if constexpr (sizeof(T) == 4) {
        const float hi = (float)k;
        const float lo = (float)(k - (double)hi);
        if (lo == 0.0f) return x * hi;
        return fmaf(x, hi, x * lo);
    } else {
        return x * (T)k;
    }
}
// a * (c + i s)
template <typename T>
__device__ __forceinline__ cplx<T> cmulk(cplx<T> a, double c, double s) {
    if constexpr (sizeof(T) == 4) {
        const float ch = (float)c, cl = (float)(c - (double)ch);
        const float sh = (float)s, sl = (float)(s - (double)sh);
        const float re_lo = fmaf(a.x, cl, -(a.y * sl));
        const float im_lo = fmaf(a.x, sl, a.y * cl);
        return mk(fmaf(a.x, ch, fmaf(-a.y, sh, re_lo)), fmaf(a.x, sh, fmaf(a.y, ch, im_lo)));
    } else {
        return mk(a.x * (T)c - a.y * (T)s, a.x * (T)s + a.y * (T)c);
    }
}

// a * (h + l) and a * conj(h + l) for a twiddle held as a table value h and
// its lo part l (FP32 transforms of psi): one rounding of the exact product.
template <typename T>
__device__ __forceinline__ cplx<T> cmul2(cplx<T> a, cplx<T> h, cplx<T> l) {
    const T re_lo = fma(a.x, l.x, -(a.y * l.y));
    const T im_lo = fma(a.x, l.y, a.y * l.x);
    return mk(fma(a.x, h.x, fma(-a.y, h.y, re_lo)), fma(a.x, h.y, fma(a.y, h.x, im_lo)));
}
template <typename T>
__device__ __forceinline__ cplx<T> cmulc2(cplx<T> a, cplx<T> h, cplx<T> l) {
    const T re_lo = fma(a.x, l.x, a.y * l.y);
    const T im_lo = fma(a.y, l.x, -(a.x * l.y));
    return mk(fma(a.x, h.x, fma(a.y, h.y, re_lo)), fma(a.y, h.x, fma(-a.x, h.y, im_lo)));
}

// a (1 + e) k for a kinetic phase entry e = (cos f - 1, sin f) and the
// normalisation k = 1/n.  The table holds the deviation from 1, so its FP32
// rounding keeps |1 + e| = 1 to ~1e-9 at the small phases of the low
// wavenumbers that hold the mass; an FP32 (cos f, sin f) pair is off by up
// to 6e-8 there, and every real-time step would add that to the norm.  The
// increment b e is formed first and b + b e rounded once: at small f the
// -f^2/2 part of the increment is below half an ulp of b, and a chain of
// roundings through b + i f b would drop it (a norm gain of f^2 per multiply).
template <typename T>
__device__ __forceinline__ cplx<T> kin_mul(cplx<T> a, cplx<T> e, double k) {
    const cplx<T> b = mk(mulk(a.x, k), mulk(a.y, k));
    const T dx = fma(b.x, e.x, -(b.y * e.y));
    const T dy = fma(b.y, e.x, b.x * e.y);
    return mk(b.x + dx, b.y + dy);
}

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
        const double h = 0.86602540378443864676;  // sqrt(3)/2
        cplx<T> s = cadd(a[1], a[2]);
        cplx<T> d = csub(a[1], a[2]);
        cplx<T> m = mk(a[0].x - T(0.5) * s.x, a[0].y - T(0.5) * s.y);
        // e = sign * i * h * d
        const T hx = mulk(d.x, h), hy = mulk(d.y, h);
        cplx<T> e = (sign < 0) ? mk(hy, -hx) : mk(-hy, hx);
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
                const double c = R == 5 ? c5[m] : c7[m];
                const double sn = sign * (R == 5 ? s5[m] : s7[m]);  // W^(sign m) = c + i sign s
                acc = cadd(acc, cmulk(a[t], c, sn));
            }
            out[q] = acc;
        }
#pragma unroll
        for (int q = 0; q < R; ++q) a[q] = out[q];
    }
}

// One Stockham step of radix R on `nlines` lines held in LDS.
// Element e of line l lives at buf[l*ls + e*es].  tw points at W_L^m; twl,
// when given (FP32 transforms of psi), at the lo parts W_L^m - tw[m].
// MAXE bounds the number of complex elements of the tile (compile time) so
// that each thread's butterflies fit in registers.
template <typename T, int NT, int MAXE, int R>
__device__ __forceinline__ void stockham_step(cplx<T>* buf, int L, int Ns, int nlines, int es, int ls, int sign,
                                              const cplx<T>* __restrict__ tw, const cplx<T>* __restrict__ twl) {
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
                    if (sizeof(T) == 4 && twl) {
                        const cplx<T> tl = twl[r * k * tstride];
                        v[q][r] = (sign < 0) ? cmul2(v[q][r], t, tl) : cmulc2(v[q][r], t, tl);
                    } else {
                        v[q][r] = (sign < 0) ? cmul(v[q][r], t) : cmulc(v[q][r], t);
                    }
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
// twl (optional): the lo parts of this length's twiddles (transforms of psi).
template <typename T, int NT, int MAXE>
__device__ void lds_fft(cplx<T>* buf, const LinePlan& pl, int nlines, int es, int ls, int sign,
                        const cplx<T>* __restrict__ twbase, const cplx<T>* __restrict__ twl = nullptr) {
    const cplx<T>* tw = twbase + pl.tw;
    int Ns = 1;
    for (int s = 0; s < pl.nst; ++s) {
        const int R = pl.radix[s];
        if (R == 4)
            stockham_step<T, NT, MAXE, 4>(buf, pl.L, Ns, nlines, es, ls, sign, tw, twl);
        else if (R == 2)
            stockham_step<T, NT, MAXE, 2>(buf, pl.L, Ns, nlines, es, ls, sign, tw, twl);
        else if (R == 3)
            stockham_step<T, NT, MAXE, 3>(buf, pl.L, Ns, nlines, es, ls, sign, tw, twl);
        else if (R == 5)
            stockham_step<T, NT, MAXE, 5>(buf, pl.L, Ns, nlines, es, ls, sign, tw, twl);
        else
            stockham_step<T, NT, MAXE, 7>(buf, pl.L, Ns, nlines, es, ls, sign, tw, twl);
        Ns *= R;
    }
}

}  // namespace gpe
