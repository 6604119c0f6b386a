"""How FP32 round-off accumulates in a split-step real-time run: the yardsticks
for the FP32 path's real-time error (3000 steps at 128x128x32: psi 1.8e-5
against the FP64 path on MI355X).

A weakly nonlinear Gross-Pitaevskii split step (Strang: half phase, kinetic
FFT step, half phase; the composition of src/solver/solver.cpp:107-112) on a
24x24x12 grid, run in FP64 as the reference and in two FP32 variants:

  * FP64 arithmetic, psi stored in FP32 (rounded 1 or 4 times per step):
    the floor that FP32 storage alone sets.  Rounding errors are independent
    from step to step, so the error is a random walk, ~sqrt(steps).
  * all of it in numpy's complex64 (float32 FFT and phase factors): the
    rounded factors are not of unit modulus, the amplitude drifts every step
    in the same direction and the error grows ~linearly.

The FP32 path sits between the two: its transforms of psi use exactly
rounded twiddles and unit-modulus phase factors (no systematic drift), and
what remains is the round-off of the FFT butterflies (~3e-7 per step over
the ~38 FP32 butterfly stages of one step), ~7x the storage floor.
"""
import numpy as np
import pytest

SHAPE = (24, 24, 12)
DT = 0.05
G = 0.02


def _setup():
    xs = [np.arange(s) - s / 2 for s in SHAPE]
    X, Y, Z = np.meshgrid(*xs, indexing="ij")
    V = 0.01 * (X**2 + Y**2 + 2 * Z**2)
    ks = [2 * np.pi * np.fft.fftfreq(s) for s in SHAPE]
    KX, KY, KZ = np.meshgrid(*ks, indexing="ij")
    kin = np.exp(-0.5j * DT * (KX**2 + KY**2 + KZ**2))
    psi0 = np.exp(-0.07 * (X**2 + Y**2 + 2 * Z**2)) * np.exp(0.05j * X)
    return V, kin, psi0 / np.abs(psi0).max()


def _half(psi, V):
    return psi * np.exp(-0.5j * DT * (V + G * np.abs(psi) ** 2))


def _run(checkpoints, variant):
    V, kin, psi0 = _setup()
    a = psi0.astype(np.complex128)
    r = lambda x: x.astype(np.complex64).astype(np.complex128)
    if variant == "c64":
        b = psi0.astype(np.complex64)
        V32, k32 = V.astype(np.float32), kin.astype(np.complex64)
    else:
        b = r(psi0)
    out, t = {}, 0
    for ck in checkpoints:
        while t < ck:
            a = _half(np.fft.ifftn(kin * np.fft.fftn(_half(a, V))), V)
            if variant == "c64":
                ph = lambda x: np.exp((-0.5j * DT * (V32 + np.float32(G) * np.abs(x) ** 2)).astype(np.complex64))
                b = b * ph(b)
                b = np.fft.ifftn(k32 * np.fft.fftn(b))
                b = b * ph(b)
                assert b.dtype == np.complex64
            else:
                b = _half(b, V)
                if variant == "store4":
                    b = r(b)
                b = np.fft.fftn(b)
                if variant == "store4":
                    b = r(b)
                b = np.fft.ifftn(kin * b)
                if variant == "store4":
                    b = r(b)
                b = r(_half(b, V))
            t += 1
        out[ck] = float(np.linalg.norm(b - a) / np.linalg.norm(a))
    return out


@pytest.mark.parametrize("variant", ["store1", "store4"])
def test_fp32_storage_floor_is_a_random_walk(variant):
    e = _run((300, 3000), variant)
    print(f"\nFP32 storage only ({variant}): 300 steps {e[300]:.2e}, 3000 steps {e[3000]:.2e}")
    assert 3e-7 < e[3000] < 5e-6
    assert 2.0 < e[3000] / e[300] < 5.0  # ~sqrt(10)


def test_complex64_split_step_drifts_linearly():
    e = _run((300, 3000), "c64")
    print(f"\ncomplex64 split step: 300 steps {e[300]:.2e}, 3000 steps {e[3000]:.2e}")
    assert e[3000] > 5e-5
    assert e[3000] / e[300] > 7.0  # ~10: linear growth
