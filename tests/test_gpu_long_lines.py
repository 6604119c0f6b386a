"""Long transform lines against the oracle: x lines of 512 and 1024 points
(padded Poisson lines of 1024 and 2048 in the LDS x pass), y lines of 512
and z lines of 256 on the run-time-size plane kernels, each with the other
two axes small so the oracle finishes in seconds.  10 fused IT steps, then
3 RT steps from the oracle's state (so only the RT error is measured)."""
import numpy as np
import pytest

from gpesolver_amd import hip
from gpesolver_amd import params as P

pytestmark = pytest.mark.gpu

TOL_IT = {hip.GPE_FP64: 1e-10, hip.GPE_FP32: 1e-5}
TOL_RT = {hip.GPE_FP64: 1e-10, hip.GPE_FP32: 3e-6}


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("shape", [(512, 8, 6), (1024, 6, 4), (8, 512, 8), (8, 16, 256)])
@pytest.mark.parametrize("prec", [hip.GPE_FP64, hip.GPE_FP32], ids=["fp64", "fp32"])
def test_long_lines_against_oracle(oracle, shape, prec):
    nx, ny, nz = shape
    d = (120.0, 120.0, 600.0)
    p = P.PhysicalParameters.from_physical(nx, ny, nz, *d, n_atoms=500.0, edd=1.5)
    op = oracle.make_params(nx, ny, nz, *d, n_atoms=500.0, edd=1.5)
    V = P.potential("HARMONIC", p)
    psi = P.add_phase_noise(P.wavefunction("GAUSS", p), 1e-3, 5)
    print(f"long lines {shape} prec {prec}", flush=True)
    dev = hip.Device(p.to_gpe(), prec)
    dev.load_pote(V)
    dev.load_buffer(psi)
    orc = oracle.Oracle(op)
    orc.load_pote(V)
    orc.load_buffer(psi)
    dev.it_steps(10)
    orc.it_steps(10)
    assert rel(dev.export_psi(), orc.psi) < TOL_IT[prec]
    orc.calc_fi3d()
    dev.load_buffer(orc.psi)
    dev.import_fi3d(orc.fi3d.copy())
    dev.rt_steps(3)
    orc.rt_steps(0, 3)
    assert rel(dev.export_psi(), orc.psi) < TOL_RT[prec]
    dev.close()
