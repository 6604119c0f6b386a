"""Grid lengths outside the transforms' radix set are refused when the
handle (or a batch instance) is created, with the reason, and leave the
library usable.  The reference's FFTW plans any length
(src/solver/fft_solvers/cpu/fftw_poisson_solver.cpp:105-118); here the
lengths nx, ny, nz and their doubles must be 2^a 3^b 5^c 7^d."""
import numpy as np
import pytest

from gpesolver_amd import hip
from gpesolver_amd import params as P

pytestmark = pytest.mark.gpu

MSG = r"is not of the form 2\^a 3\^b 5\^c 7\^d"


def make(shape):
    return P.PhysicalParameters.from_physical(*shape, 120.0, 120.0, 600.0, n_atoms=1e4, edd=1.4)


@pytest.mark.parametrize("shape", [(22, 16, 8), (16, 26, 8), (16, 16, 11)])
def test_grid_length_outside_the_radix_set_is_refused(shape):
    p = make(shape)
    for prec in (hip.GPE_FP64, hip.GPE_FP32):
        with pytest.raises(hip.GpeError, match=MSG):
            hip.Device(p.to_gpe(), prec)
    with pytest.raises(hip.GpeError, match=MSG):
        hip.Batch([make((16, 16, 8)).to_gpe(), p.to_gpe()], hip.GPE_FP32)
    # a supported grid still runs after the refusals
    ok = make((16, 16, 8))
    d = hip.Device(ok.to_gpe(), hip.GPE_FP64)
    d.load_pote(P.potential("HARMONIC", ok))
    d.load_buffer(P.wavefunction("GAUSS", ok))
    d.it_steps(5)
    assert np.isfinite(d.energy()["e_total"])
    d.close()


@pytest.mark.parametrize("shape", [(3, 3, 3), (4, 3, 5), (5, 7, 3)])
@pytest.mark.parametrize("prec", [hip.GPE_FP64, hip.GPE_FP32], ids=["fp64", "fp32"])
def test_smallest_grids_against_oracle(oracle, shape, prec):
    """The smallest grids create accepts (every length >= 3; one interior
    point at 3x3x3) run IT and RT as the oracle does.  50 atoms: a density
    per point like the other run-time-size tests' (a dense 3x3x3 state
    amplifies the FP32 rounding as the dense small double wells do)."""
    nx, ny, nz = shape
    p = P.PhysicalParameters.from_physical(nx, ny, nz, 120.0, 120.0, 600.0, n_atoms=50.0, edd=1.4)
    op = oracle.make_params(nx, ny, nz, 120.0, 120.0, 600.0, n_atoms=50.0, edd=1.4)
    V = P.potential("HARMONIC", p)
    psi = P.add_phase_noise(P.wavefunction("GAUSS", p), 1e-2, 9)
    print(f"smallest grid {shape} prec {prec}", flush=True)
    d = hip.Device(p.to_gpe(), prec)
    d.load_pote(V)
    d.load_buffer(psi)
    o = oracle.Oracle(op)
    o.load_pote(V)
    o.load_buffer(psi)
    tol = 1e-10 if prec == hip.GPE_FP64 else 1e-5
    d.it_steps(10)
    o.it_steps(10)
    e_it = np.linalg.norm(d.export_psi() - o.psi) / np.linalg.norm(o.psi)
    d.rt_steps(5)
    o.rt_steps(0, 5)
    e_rt = np.linalg.norm(d.export_psi() - o.psi) / np.linalg.norm(o.psi)
    print(f"  IT {e_it:.2e}  RT {e_rt:.2e}", flush=True)
    assert e_it < tol and e_rt < tol
    e = d.energy()
    assert np.isfinite(e["e_total"]) and np.isfinite(e["mu"])
    d.close()
