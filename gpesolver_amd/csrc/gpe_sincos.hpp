// gpe_sincos.hpp -- the real-time phase factor exp(i theta) of an FP64
// phase, applied to an FP32 or FP64 value.
//
// The phase theta is reduced in FP64 by the nearest multiple of pi/2,
// r = theta - q pi/2 with |r| <= pi/4, and sin r, cos r are even / odd
// polynomials in r^2 evaluated in FP32 (Horner):
//   sin r = r + r^3 (S0 + S1 u + S2 u^2 + S3 u^3),   u = r^2
//   cos r = 1 + u   (C0 + C1 u + C2 u^2 + C3 u^3)
// The coefficients are a weighted least-squares fit on 4000 Chebyshev nodes
// of [-pi/4, pi/4] (relative error weighting), rounded to FP32; the maximum
// absolute error of the FP32 evaluation is 4.3e-8 (sin) and 6.8e-8 (cos),
// about one FP32 ulp (checked on a dense grid by tests/test_sincos_poly.py).
// The quadrant q turns the value exactly by q quarter turns.
#pragma once
#include <hip/hip_runtime.h>

#include "gpe_fft.hpp"

namespace gpe {

struct SinCosPoly {
    static constexpr float S[4] = {-0.1666666716337204f, 0.008333329111337662f, -0.0001983929832931608f,
                                   2.7180080905964132e-06f};
    static constexpr float C[4] = {-0.5f, 0.04166661947965622f, -0.0013886665692552924f, 2.4382301489822567e-05f};
};

// c * exp(i theta) in FP32: c turned exactly by the quarter turns q, then
// b + b (cos r - 1 + i sin r) with the increment formed first and one
// rounding of the sum.  cos r - 1 = u * pc keeps its relative precision at
// small r, where a rounded cos r would be 1 and |(cos r, sin r)| > 1 (a norm
// gain of r^2 per multiply on every point with a small phase).
__device__ __forceinline__ cplx<float> phase_rot(cplx<float> c, double theta) {
    const double q = rint(theta * 0.63661977236758134308);  // 2/pi
    const float r = (float)(theta - q * 1.57079632679489661923);
    const float u = r * r;
    float ps = SinCosPoly::S[3];
    ps = ps * u + SinCosPoly::S[2];
    ps = ps * u + SinCosPoly::S[1];
    ps = ps * u + SinCosPoly::S[0];
    float pc = SinCosPoly::C[3];
    pc = pc * u + SinCosPoly::C[2];
    pc = pc * u + SinCosPoly::C[1];
    pc = pc * u + SinCosPoly::C[0];
    const float sr = r + r * u * ps;
    const float cm1 = u * pc;
    const int k = ((int)(long long)q) & 3;
    // b = c * i^k
    const float bx = (k == 0) ? c.x : (k == 1) ? -c.y : (k == 2) ? -c.x : c.y;
    const float by = (k == 0) ? c.y : (k == 1) ? c.x : (k == 2) ? -c.y : -c.x;
    const float dx = fmaf(bx, cm1, -(by * sr));
    const float dy = fmaf(by, cm1, bx * sr);
    cplx<float> o;
    o.x = bx + dx;
    o.y = by + dy;
    return o;
}

// c * exp(i r) for a small FP32 phase, |r| <= pi/4 (no reduction): the
// increment form of phase_rot.
__device__ __forceinline__ cplx<float> phase_rot_small(cplx<float> c, float r) {
    const float u = r * r;
    float ps = SinCosPoly::S[3];
    ps = ps * u + SinCosPoly::S[2];
    ps = ps * u + SinCosPoly::S[1];
    ps = ps * u + SinCosPoly::S[0];
    float pc = SinCosPoly::C[3];
    pc = pc * u + SinCosPoly::C[2];
    pc = pc * u + SinCosPoly::C[1];
    pc = pc * u + SinCosPoly::C[0];
    const float sr = r + r * u * ps;
    const float cm1 = u * pc;
    const float dx = fmaf(c.x, cm1, -(c.y * sr));
    const float dy = fmaf(c.y, cm1, c.x * sr);
    cplx<float> o;
    o.x = c.x + dx;
    o.y = c.y + dy;
    return o;
}

// c * exp(i theta) in FP64 with the library sincos of theta reduced by 2 pi:
// the run-time-size plane kernel, whose registers are already at the limit,
// keeps it (the inline polynomials below make it spill).
__device__ __forceinline__ cplx<double> phase_rot_lib(cplx<double> c, double theta) {
    double s, co;
    theta -= 6.283185307179586477 * rint(theta * 0.15915494309189533577);
    sincos(theta, &s, &co);
    cplx<double> o;
    o.x = c.x * co - c.y * s;
    o.y = c.x * s + c.y * co;
    return o;
}

// c * exp(i theta) in FP64.  theta is reduced by the nearest multiple of
// pi/2 with Cody and Waite's two-constant method (r = (theta - q P1) - q P2,
// P1 = pi/2 rounded to FP64, P2 = pi/2 - P1, each product inside an FMA), so
// |r| <= pi/4 to ~1e-16 for the phases of a step.  sin r and cos r are their
// Taylor series through r^15 and r^16 (coefficients (-1)^n / (2n+1)! and
// (-1)^n / (2n)!, Horner in u = r^2); the first omitted terms, r^17 / 17! and
// r^18 / 18!, are below 5e-17 on |r| <= pi/4.  The quadrant q turns the value
// exactly by q quarter turns.
__device__ __forceinline__ cplx<double> phase_rot(cplx<double> c, double theta) {
    const double q = rint(theta * 0.63661977236758134308);  // 2/pi
    double r = fma(-q, 1.5707963267948966, theta);
    r = fma(-q, 6.123233995736766e-17, r);
    const double u = r * r;
    double ps = -1.0 / 1307674368000.0;
    ps = fma(ps, u, 1.0 / 6227020800.0);
    ps = fma(ps, u, -1.0 / 39916800.0);
    ps = fma(ps, u, 1.0 / 362880.0);
    ps = fma(ps, u, -1.0 / 5040.0);
    ps = fma(ps, u, 1.0 / 120.0);
    ps = fma(ps, u, -1.0 / 6.0);
    const double sr = fma(r * u, ps, r);
    double pc = 1.0 / 20922789888000.0;
    pc = fma(pc, u, -1.0 / 87178291200.0);
    pc = fma(pc, u, 1.0 / 479001600.0);
    pc = fma(pc, u, -1.0 / 3628800.0);
    pc = fma(pc, u, 1.0 / 40320.0);
    pc = fma(pc, u, -1.0 / 720.0);
    pc = fma(pc, u, 1.0 / 24.0);
    pc = fma(pc, u, -0.5);
    const double cr = fma(u, pc, 1.0);
    const int k = ((int)(long long)q) & 3;
    // b = c * i^k
    const double bx = (k == 0) ? c.x : (k == 1) ? -c.y : (k == 2) ? -c.x : c.y;
    const double by = (k == 0) ? c.y : (k == 1) ? c.x : (k == 2) ? -c.y : -c.x;
    cplx<double> o;
    o.x = fma(bx, cr, -(by * sr));
    o.y = fma(bx, sr, by * cr);
    return o;
}

}  // namespace gpe
