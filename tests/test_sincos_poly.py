"""The FP32 sin / cos of the real-time phase (gpesolver_amd/csrc/gpe_sincos.hpp):
its coefficients, evaluated exactly as the kernel does (FP64 quarter-turn
reduction, FP32 Horner), stay within 1e-7 of the FP64 sin / cos for phases of
any magnitude.  Pinned by numpy's sin / cos."""
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
