"""hipGraph-replayed steady-state steps give the same results as direct
launches (GPE_NO_GRAPH=1), for the single-instance IT and RT loops and the
batched IT loop; graph replays with eager work between them (energies,
exports) stay correct."""
import numpy as np
import pytest

from gpesolver_amd import hip
from gpesolver_amd import params as P

pytestmark = pytest.mark.gpu


def case(shape=(32, 24, 12)):
    nx, ny, nz = shape
    p = P.PhysicalParameters.from_physical(nx, ny, nz, 120.0, 120.0, 600.0, n_atoms=500.0, edd=1.5)
    return p, P.potential("HARMONIC", p), P.add_phase_noise(P.wavefunction("GAUSS", p), 1e-3, 2)


def dev_for(p, V, psi, prec):
    d = hip.Device(p.to_gpe(), prec)
    d.load_pote(V)
    d.load_buffer(psi)
    return d


@pytest.mark.parametrize("prec", [hip.GPE_FP64, hip.GPE_FP32])
def test_it_and_rt_graph_equals_direct(monkeypatch, prec):
    p, V, psi = case()
    outs = {}
    for flag in ("0", "1"):
        monkeypatch.setenv("GPE_NO_GRAPH", flag)
        d = dev_for(p, V, psi, prec)
        d.it_steps(150)  # 149 steady steps: two graph replays + 21 direct
        e1 = d.energy()["e_total"]
        d.it_steps(150)
        a = d.export_psi()
        d.calc_fi3d()
        d.rt_steps(200)
        d.rt_steps(130)
        outs[flag] = (e1, a, d.export_psi(), d.export_fi3d())
        d.close()
    g, n = outs["0"], outs["1"]
    assert g[0] == n[0]
    for u, v in zip(g[1:], n[1:]):
        assert np.array_equal(u, v)


def test_batch_graph_equals_direct(monkeypatch):
    p, V, psi = case()
    p2, V2, psi2 = case((16, 16, 8))
    res = {}
    for flag in ("0", "1"):
        monkeypatch.setenv("GPE_NO_GRAPH", flag)
        b = hip.Batch([p.to_gpe(), p2.to_gpe()], hip.GPE_FP32)
        for d, (VV, ps) in zip(b.instances, ((V, psi), (V2, psi2))):
            d.load_pote(VV)
            d.load_buffer(ps)
        steps, e = b.it_until(300, check_every=100, tol=0.0)
        res[flag] = (steps, e, [d.export_psi() for d in b.instances])
        b.close()
    assert np.array_equal(res["0"][0], res["1"][0])
    assert np.array_equal(res["0"][1], res["1"][1])
    for u, v in zip(res["0"][2], res["1"][2]):
        assert np.array_equal(u, v)


@pytest.mark.parametrize("prec", [hip.GPE_FP64, hip.GPE_FP32])
@pytest.mark.parametrize("shape", [(32, 24, 12), (64, 64, 16)])
def test_split_plane_pass_equals_single(monkeypatch, prec, shape):
    """Two blocks per plane with ping-pong S1 / Tk (the default real-time and
    imaginary-time plane pass) is bitwise the one-block-per-plane pass
    (GPE_NO_SPLIT=1), for step counts of both parities, with and without
    graph replay, and through the energy-checked IT loop."""
    p, V, psi = case(shape)
    outs = {}
    for split in ("0", "1"):
        monkeypatch.setenv("GPE_NO_SPLIT", split)
        d = dev_for(p, V, psi, prec)
        r = []
        for n in (1, 2, 67, 130):
            d.it_steps(n)
            r.append(d.export_psi())
        steps, epa = d.it_until(450, 100, 1e-12)
        r += [d.export_psi(), np.array([steps, epa])]
        d.calc_fi3d()
        for n in (1, 2, 67, 130):
            d.rt_steps(n)
            r.append(d.export_psi())
        r.append(d.export_fi3d())
        outs[split] = r
        d.close()
    for u, v in zip(outs["0"], outs["1"]):
        assert np.array_equal(u, v)
