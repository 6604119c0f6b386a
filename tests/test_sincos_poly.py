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
