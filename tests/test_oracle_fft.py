"""Pins the oracle's own FFT (oracle/gpe_oracle.c, written from the Stockham
autosort formulation) against a direct DFT and numpy.fft.  The reference uses
FFTW's unnormalised forward (sign -1) / backward (sign +1) transforms
(src/solver/fft_solvers/cpu/fftw_poisson_solver.cpp:115-118,
src/solver/fft_solvers/cpu/fftw_split_solver.cpp:104-107)."""
import numpy as np
import pytest


def direct_dft(x, sign):
    n = x.shape[0]
    idx = np.arange(n)
    W = np.exp(sign * 2j * np.pi * np.outer(idx, idx) / n)
    return W @ x


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 16, 24, 35])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft1d_vs_direct_dft(oracle, n, sign):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    got = oracle.fft1(x, sign)
    np.testing.assert_allclose(got, direct_dft(x, sign), rtol=0, atol=1e-12 * max(1, n))


@pytest.mark.parametrize("n", [32, 48, 64, 96, 128, 192, 256, 384, 512])
def test_fft1d_vs_numpy(oracle, n):
    rng = np.random.default_rng(100 + n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    np.testing.assert_allclose(oracle.fft1(x, -1), np.fft.fft(x), atol=1e-11 * np.sqrt(n))
    np.testing.assert_allclose(oracle.fft1(x, +1), np.fft.ifft(x) * n, atol=1e-11 * np.sqrt(n))


@pytest.mark.parametrize("shape", [(4, 6, 2), (8, 8, 4), (12, 6, 10), (32, 32, 8), (48, 24, 12)])
def test_fft3d_vs_numpy_and_roundtrip(oracle, shape):
    rng = np.random.default_rng(sum(shape))
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    X = oracle.fft3(x, -1)
    np.testing.assert_allclose(X, np.fft.fftn(x), atol=1e-10 * np.sqrt(x.size))
    # inverse o forward = N * I
    back = oracle.fft3(X, +1) / x.size
    np.testing.assert_allclose(back, x, atol=1e-12)
    # Parseval
    assert abs(np.sum(np.abs(X) ** 2) / x.size - np.sum(np.abs(x) ** 2)) < 1e-10 * np.sum(np.abs(x) ** 2)
