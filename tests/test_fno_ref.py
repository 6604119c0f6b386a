"""Pins the numpy FNO restatement (oracle/fno_ref.py) that checks the GPU
network: the spectral layer with identity mode weights is the orthogonal
projection onto the retained Fourier modes (idempotent, exact on a retained
mode, zero on a dropped one), and its finite-difference gradient agrees with
the analytic derivative of a quadratic in one weight."""
import numpy as np

from oracle.fno_ref import FnoRef


def ident_R(ref):
    R = np.zeros((ref.nm, ref.W, ref.W))
    R[:] = np.eye(ref.W)
    return R, np.zeros_like(R)


def test_spectral_identity_is_projection():
    # (kz = 0 and the Nyquist plane keep only the real part, so exactness is
    # checked on modes with 0 < kz < nz/2)
    ref = FnoRef(16, 12, 8, 3, 4, 1, 8, 6)
    Rre, Rim = ident_R(ref)
    x, y, z = np.meshgrid(np.arange(16), np.arange(12), np.arange(8), indexing="ij")
    rng = np.random.default_rng(0)
    h = np.zeros((2, 16, 12, 8, 4))
    for kx in (0, 1, 2, -1, -3):
        for ky in (0, 2, -2):
            for kz in (1, 2, 3):
                ph = rng.uniform(0, 2 * np.pi)
                h += np.cos(2 * np.pi * (kx * x / 16 + ky * y / 12 + kz * z / 8) + ph)[None, ..., None]
    p1 = ref.spectral(h, Rre, Rim)
    np.testing.assert_allclose(p1, h, atol=1e-11)
    np.testing.assert_allclose(ref.spectral(p1, Rre, Rim), p1, atol=1e-11)
    kept = np.cos(2 * np.pi * (2 * x / 16 + 1 * y / 12 + 2 * z / 8))
    dropped = np.cos(2 * np.pi * (5 * x / 16))  # kx = 5 is outside the 6 retained
    for f, expect in ((kept, kept), (dropped, 0 * dropped)):
        hh = np.repeat(f[None, ..., None], 4, axis=-1)
        out = ref.spectral(hh, Rre, Rim)
        np.testing.assert_allclose(out[0, ..., 0], expect, atol=1e-12)


def test_fd_gradient_of_output_bias():
    # d/db_q2 of mean((y - t)^2) = 2 mean(y - t), exactly
    ref = FnoRef(8, 8, 4, 2, 4, 1, 4, 4)
    rng = np.random.default_rng(1)
    w = rng.standard_normal(ref.nparam) * 0.1
    x = rng.standard_normal((1, 256, 2))
    t = rng.standard_normal((1, 256))
    o, _ = ref.off["q2_b"]
    g = ref.fd_grad(w, x, t, [o])[0]
    y = ref.forward(w, x)
    assert abs(g - 2 * np.mean(y - t)) < 1e-7
