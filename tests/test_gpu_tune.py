"""Autotuner (gpe_autotune): picks a variant, caches it per grid, and a tuned
handle computes exactly what an untuned handle forced to the same variant
computes."""
import numpy as np
import pytest

from gpesolver_amd import hip
from gpesolver_amd import params as P

pytestmark = pytest.mark.gpu


def case(shape):
    nx, ny, nz = shape
    p = P.PhysicalParameters.from_physical(nx, ny, nz, 120.0, 120.0, 600.0, n_atoms=500.0, edd=1.5)
    return p, P.potential("HARMONIC", p), P.add_phase_noise(P.wavefunction("GAUSS", p), 1e-3, 5)


def run(d, V, psi):
    d.load_pote(V)
    d.load_buffer(psi)
    d.it_steps(70)
    a = d.export_psi()
    d.calc_fi3d()
    d.rt_steps(69)
    return a, d.export_psi(), d.export_fi3d()


@pytest.mark.parametrize("prec", [hip.GPE_FP32, hip.GPE_FP64])
def test_autotune_matches_forced_variant(monkeypatch, prec):
    shape = (64, 64, 16)
    p, V, psi = case(shape)
    d = hip.Device(p.to_gpe(), prec)
    cfg = d.autotune(40)
    print("tuned", shape, prec, cfg, flush=True)
    assert cfg["rt_step_us"] > 0 and cfg["it_step_us"] > 0
    assert d.autotune(40) == cfg  # cached for this grid and precision
    tuned = run(d, V, psi)
    d.close()
    monkeypatch.setenv("GPE_NO_SPLIT", "0" if cfg["split_plane_pass"] else "1")
    monkeypatch.setenv("GPE_LINE_LDS", "0" if cfg["register_x_pass"] else "1")
    monkeypatch.setenv("GPE_NO_GRAPH", "0" if cfg["graph_replay"] else "1")
    f = hip.Device(p.to_gpe(), prec)
    forced = run(f, V, psi)
    f.close()
    for u, v in zip(tuned, forced):
        assert np.array_equal(u, v)
