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
