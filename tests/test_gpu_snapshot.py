"""gpe_snapshot_begin / _end (the checkpoint and energy cadence of
src/solver/solver.cpp:54-59,81-86 without stopping the loop): the snapshot
equals what gpe_export_psi and gpe_calc_energy give at the same point, bit for
bit, while later steps are already enqueued; and those later steps are not
disturbed."""
import numpy as np
import pytest

from gpesolver_amd import hip
from gpesolver_amd import params as P

pytestmark = pytest.mark.gpu


def twin(prec):
    p = P.PhysicalParameters.from_physical(32, 24, 12, 120.0, 120.0, 600.0, n_atoms=2e4, edd=1.5, dd_nm=800.0)
    devs = []
    for _ in range(2):
        d = hip.Device(p.to_gpe(), prec)
        d.init_potential("MEXICAN", p)
        d.init_wavefunction("MULTIPLE_GAUSS", p, n_gauss_max=2, sigma_cells=3.0, phase_noise=1e-3, seed=5)
        devs.append(d)
    return devs


@pytest.mark.parametrize("prec", [hip.GPE_FP64, hip.GPE_FP32])
@pytest.mark.parametrize("kind", ["it", "rt"])
def test_snapshot_equals_synchronous_output(prec, kind):
    a, b = twin(prec)
    step = (lambda d, n: d.it_steps(n)) if kind == "it" else (lambda d, n: d.rt_steps(n))
    for d in (a, b):
        step(d, 40)  # RT leaves a pending half phase: the snapshot applies it
    a.snapshot_begin()
    step(a, 30)  # enqueued behind the snapshot
    psi_s, e_s = a.snapshot_end()
    psi_b, e_b = b.export_psi(), b.energy()
    assert np.array_equal(psi_s, psi_b)
    assert e_s == e_b
    step(b, 30)
    assert np.array_equal(a.export_psi(), b.export_psi())
    with pytest.raises(hip.GpeError):
        a.snapshot_end()  # nothing in flight
    a.close()
    b.close()
