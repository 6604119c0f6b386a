"""The FP32 real-time phase factor (gpesolver_amd/csrc/gpe_sincos.hpp): its
polynomial coefficients, evaluated exactly as the kernel does (FP64
quarter-turn reduction, FP32 Horner), stay within 1e-7 of the FP64 sin / cos
for phases of any magnitude (pinned by numpy's sin / cos); and the rotation
in increment form (phase_rot, and kin_mul of gpe_fft.hpp) keeps |psi|^2 on
average where the rounded (cos, sin) pair gains r^2 at small phases.  FP32
arithmetic is emulated with numpy float32 (an FMA as one rounding of the
exact FP64 value)."""
import os
import re

import numpy as np

HDR = os.path.join(os.path.dirname(__file__), "..", "gpesolver_amd", "csrc", "gpe_sincos.hpp")


def coeffs(name):
    txt = open(HDR).read()
    m = re.search(r"static constexpr float %s\[4\] = \{([^}]*)\}" % name, txt)
    return np.array([float(x.strip().rstrip("f")) for x in m.group(1).split(",")], dtype=np.float32)


def sincos_phase(theta):
    S, C = coeffs("S"), coeffs("C")
    theta = np.asarray(theta, dtype=np.float64)
    q = np.rint(theta * (2 / np.pi))
    r = (theta - q * (np.pi / 2)).astype(np.float32)
    u = r * r
    ps = S[3]
    for k in (2, 1, 0):
        ps = (ps * u + S[k]).astype(np.float32)
    pc = C[3]
    for k in (2, 1, 0):
        pc = (pc * u + C[k]).astype(np.float32)
    sr = (r + r * u * ps).astype(np.float32)
    cr = (np.float32(1) + u * pc).astype(np.float32)
    k = q.astype(np.int64) & 3
    s = np.select([k == 0, k == 1, k == 2], [sr, cr, -sr], -cr)
    c = np.select([k == 0, k == 1, k == 2], [cr, -sr, -cr], sr)
    return s.astype(np.float64), c.astype(np.float64)


def test_quarter_turn_polynomial_accuracy():
    r = np.linspace(-np.pi / 4, np.pi / 4, 400001)
    s, c = sincos_phase(r)
    assert np.abs(s - np.sin(r)).max() < 1e-7
    assert np.abs(c - np.cos(r)).max() < 1e-7


def test_large_and_negative_phases():
    rng = np.random.default_rng(3)
    th = np.concatenate([rng.uniform(-1e4, 1e4, 200000), rng.uniform(-10, 10, 200000),
                         np.array([0.0, np.pi / 4, -np.pi / 4, np.pi / 2, np.pi, -np.pi, 3 * np.pi / 4])])
    s, c = sincos_phase(th)
    assert np.abs(s - np.sin(th)).max() < 1.5e-7
    assert np.abs(c - np.cos(th)).max() < 1.5e-7
    assert np.abs(s * s + c * c - 1).max() < 5e-7


def f32(x):
    return np.asarray(x, dtype=np.float64).astype(np.float32).astype(np.float64)


def fma32(a, b, c):
    # a*b is exact in FP64 for FP32 operands; one rounding of the sum
    return f32(a * b + c)


def poly_parts(r):
    S, C = coeffs("S"), coeffs("C")
    r = f32(r)
    u = f32(r * r)
    ps = np.float64(S[3])
    for k in (2, 1, 0):
        ps = fma32(ps, u, np.float64(S[k]))
    pc = np.float64(C[3])
    for k in (2, 1, 0):
        pc = fma32(pc, u, np.float64(C[k]))
    sr = fma32(f32(r * u), ps, r)
    cm1 = f32(u * pc)
    return sr, cm1


def rot_increment(cx, cy, sr, cm1):
    dx = fma32(cx, cm1, -f32(cy * sr))
    dy = fma32(cy, cm1, f32(cx * sr))
    return f32(cx + dx), f32(cy + dy)


def rot_pair(cx, cy, s, c):
    return fma32(cx, c, -f32(cy * s)), fma32(cx, s, f32(cy * c))


def test_small_phase_rotation_keeps_norm_on_average():
    rng = np.random.default_rng(11)
    n = 400000
    cx, cy = f32(rng.standard_normal(n)), f32(rng.standard_normal(n))
    n0 = cx * cx + cy * cy
    for r0 in (1e-5, 1e-4, 2e-4, 1e-2):
        r = f32(rng.uniform(-r0, r0, n))
        sr, cm1 = poly_parts(r)
        ox, oy = rot_increment(cx, cy, sr, cm1)
        gain_inc = np.mean((ox * ox + oy * oy - n0) / n0)
        # the same rotation through the rounded pair (1 + cm1, sr)
        px, py = rot_pair(cx, cy, sr, f32(1.0 + cm1))
        gain_pair = np.mean((px * px + py * py - n0) / n0)
        mean_r2 = np.mean(r * r)
        # one rounding per component: the mean gain is second order (~1e-15)
        assert abs(gain_inc) < 2e-9, (r0, gain_inc)
        if r0 in (1e-4, 2e-4):
            # the pair form keeps cos r = 1 (r^2 / 2 below half an ulp of 1)
            # and gains r^2, well above the sampling noise of the mean (~1e-10)
            assert gain_pair > 0.5 * mean_r2, (r0, gain_pair, mean_r2)


def test_kinetic_factor_deviation_form():
    """kin_mul: the table entry (cos f - 1, sin f) and one rounding of
    b + b e give |b (1 + e)| = |b| on average at the small phases of the low
    wavenumbers."""
    rng = np.random.default_rng(12)
    n = 400000
    bx, by = f32(rng.standard_normal(n)), f32(rng.standard_normal(n))
    n0 = bx * bx + by * by
    f = rng.uniform(-2e-4, 2e-4, n)
    ex = f32(-2.0 * np.sin(0.5 * f) ** 2)
    ey = f32(np.sin(f))
    ox, oy = rot_increment(bx, by, ey, ex)
    assert abs(np.mean((ox * ox + oy * oy - n0) / n0)) < 2e-9
    # the nested form b + i f b (rounded) + b (cos f - 1) drops the sub-ulp
    # second-order part and gains f^2
    nx_ = fma32(bx, ex, fma32(-by, ey, bx))
    ny_ = fma32(by, ex, fma32(bx, ey, by))
    assert np.mean((nx_ * nx_ + ny_ * ny_ - n0) / n0) > 0.5 * np.mean(f * f)


def _fp64_phase_method():
    """The FP64 phase_rot's constants, read from the header: the pi/2 split
    (P1, P2) and the Horner coefficients of sin and cos in kernel order."""
    txt = open(HDR).read()
    body = txt[txt.index("cplx<double> phase_rot(cplx<double> c, double theta)"):]
    body = body[:body.index("\n}\n")]

    def const(expr):
        expr = expr.strip()
        if "/" in expr:
            a, b = expr.split("/")
            return float(a) / float(b)
        return float(expr)

    split = [float(x) for x in re.findall(r"fma\(-q, ([-0-9.e]+), ", body)]
    ps = [const(m) for m in re.findall(r"ps = (?:fma\(ps, u, )?([-0-9.e /]+)\)?;", body)]
    pc = [const(m) for m in re.findall(r"pc = (?:fma\(pc, u, )?([-0-9.e /]+)\)?;", body)]
    return split, ps, pc


def test_fp64_phase_method_accuracy():
    """The FP64 phase factor: Cody-Waite quarter-turn reduction with the
    two-part pi/2 and the Taylor series through r^15 / r^16, every FMA
    emulated as one rounding of the exact value (fractions), is within
    4e-16 of numpy's FP64 sin / cos for phases up to +-100 rad, and its
    coefficients are (-1)^n / (2n+1)! and (-1)^n / (2n)!."""
    import math
    from fractions import Fraction as F

    (p1, p2), ps, pc = _fp64_phase_method()
    half_pi = F("3.14159265358979323846264338327950288419716939937510") / 2
    assert p1 == math.pi / 2 and abs(F(p1) + F(p2) - half_pi) < F(1, 10**32)
    assert ps == [(-1) ** n / math.factorial(2 * n + 1) for n in range(7, 0, -1)]
    assert pc == [(-1) ** n / math.factorial(2 * n) for n in range(8, 0, -1)]

    def fma(a, b, c):
        return float(F(a) * F(b) + F(c))

    rng = np.random.default_rng(5)
    thetas = np.concatenate([rng.uniform(-100, 100, 1500), np.arange(-40, 41) * (math.pi / 4),
                             np.arange(-40, 41) * (math.pi / 4) + 1e-9, [0.0, 1e-300, -2.5e-8]])
    err = 0.0
    for th in thetas:
        q = float(np.rint(th * 0.63661977236758134308))
        r = fma(-q, p1, th)
        r = fma(-q, p2, r)
        u = r * r
        s = ps[0]
        for cf in ps[1:]:
            s = fma(s, u, cf)
        sr = fma(r * u, s, r)
        c = pc[0]
        for cf in pc[1:]:
            c = fma(c, u, cf)
        cr = fma(u, c, 1.0)
        k = int(q) & 3
        cos_t, sin_t = [(cr, sr), (-sr, cr), (-cr, -sr), (sr, -cr)][k]
        err = max(err, abs(cos_t - math.cos(th)), abs(sin_t - math.sin(th)))
    assert err < 4e-16, err
