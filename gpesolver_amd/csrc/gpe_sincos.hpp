// gpe_sincos.hpp -- sin and cos of an FP64 phase for the FP32 bulk path.
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
// The quadrant q then rotates (sin, cos) by q quarter turns.
#pragma once
#include <hip/hip_runtime.h>

namespace gpe {

struct SinCosPoly {
    static constexpr float S[4] = {-0.1666666716337204f, 0.008333329111337662f, -0.0001983929832931608f,
                                   2.7180080905964132e-06f};
    static constexpr float C[4] = {-0.5f, 0.04166661947965622f, -0.0013886665692552924f, 2.4382301489822567e-05f};
};

// (sin theta, cos theta) in FP32 for an FP64 theta of any magnitude.
__device__ __forceinline__ void sincos_phase(double theta, float* s, float* c) {
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
    const float cr = 1.0f + u * pc;
    const int k = ((int)(long long)q) & 3;
    // sin(r + k pi/2), cos(r + k pi/2)
    *s = (k == 0) ? sr : (k == 1) ? cr : (k == 2) ? -sr : -cr;
    *c = (k == 0) ? cr : (k == 1) ? -sr : (k == 2) ? -cr : sr;
}

// FP64 phase for the FP64 path: reduced to [-pi, pi], then the FP64 sincos.
__device__ __forceinline__ void sincos_phase(double theta, double* s, double* c) {
    theta -= 6.283185307179586477 * rint(theta * 0.15915494309189533577);
    sincos(theta, s, c);
}

}  // namespace gpe
