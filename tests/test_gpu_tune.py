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


@pytest.mark.parametrize("prec", [hip.GPE_FP32, hip.GPE_FP64])
def test_lds_x_pass_tiles_are_bitwise_equal(monkeypatch, prec):
    """The LDS x pass groups kz / k columns into blocks of cp / ck (every
    candidate of the tuner): each line's transform is the same, so the fused
    IT and RT loops give the same bits at every tile."""
    monkeypatch.setenv("GPE_LINE_LDS", "1")
    shape = (96, 24, 12)
    p, V, psi = case(shape)
    outs = []
    d = hip.Device(p.to_gpe(), prec)
    t = d.tiles()
    print("tiles", t, flush=True)
    assert t["bound"]["cp"] == 12 and t["bound"]["ck"] == 12 and t["cur"]["pen_r"] == 0
    for cp, ck in [(12, 12), (8, 4), (5, 7), (2, 2), (1, 12)]:
        d.set_tiles(cp=cp, ck=ck)
        assert d.tiles()["cur"]["cp"] == cp and d.tiles()["cur"]["ck"] == ck
        outs.append(run(d, V, psi))
    with pytest.raises(hip.GpeError):
        d.set_tiles(cp=13)
    d.close()
    for o in outs[1:]:
        for u, v in zip(outs[0], o):
            assert np.array_equal(u, v)


def test_pencil_tiles_agree(monkeypatch):
    """The pencil plane pass at several row / column tiles (GPE_FORCE_PENCIL
    on a small grid): the same transforms; only the summation order of the
    FP64 norm partials changes with the row tile."""
    monkeypatch.setenv("GPE_FORCE_PENCIL", "1")
    shape = (24, 20, 12)
    p, V, psi = case(shape)
    d = hip.Device(p.to_gpe(), hip.GPE_FP64)
    t = d.tiles()
    print("pencil tiles", t, flush=True)
    cb, rb = t["bound"]["pen_c"], t["bound"]["pen_r"]
    assert cb >= 2 and rb >= 2
    outs = []
    for c, r in [(cb, rb), (max(1, cb // 2), rb), (cb, max(1, rb // 2)), (1, 3)]:
        d.set_tiles(pen_c=c, pen_r=r)
        outs.append(run(d, V, psi))
    with pytest.raises(hip.GpeError):
        d.set_tiles(pen_r=rb + 1)
    d.close()
    for o in outs[1:]:
        for u, v in zip(outs[0], o):
            assert float(np.linalg.norm(u - v) / np.linalg.norm(v)) < 1e-13


def test_autotune_searches_tiles(monkeypatch):
    """Forced onto the LDS x pass, the tuner searches its cp / ck tiles and
    a tuned handle computes what an untuned one at those tiles computes."""
    monkeypatch.setenv("GPE_LINE_LDS", "1")
    shape = (96, 24, 12)
    p, V, psi = case(shape)
    d = hip.Device(p.to_gpe(), hip.GPE_FP32)
    cfg = d.autotune(30)
    print("tuned", cfg, flush=True)
    assert not cfg["register_x_pass"]
    tl = cfg["tiles"]
    assert 1 <= tl["cp"] <= 12 and 1 <= tl["ck"] <= 12
    tuned = run(d, V, psi)
    d.close()
    monkeypatch.setenv("GPE_NO_SPLIT", "0" if cfg["split_plane_pass"] else "1")
    monkeypatch.setenv("GPE_NO_GRAPH", "0" if cfg["graph_replay"] else "1")
    f = hip.Device(p.to_gpe(), hip.GPE_FP32)
    f.set_tiles(cp=tl["cp"], ck=tl["ck"])
    forced = run(f, V, psi)
    f.close()
    for u, v in zip(tuned, forced):
        assert np.array_equal(u, v)


def test_batch_autotune_per_x_length(monkeypatch):
    """gpe_batch_autotune picks one x-pass variant per x length for every
    instance of that length, and the tuned batch equals single handles set to
    the same variant, bit for bit."""
    shapes = [(64, 64, 16), (96, 24, 12), (64, 64, 16), (96, 24, 12)]
    cases = [case(s) for s in shapes]
    b = hip.Batch([c[0].to_gpe() for c in cases], hip.GPE_FP32)
    cfg = b.autotune(20)
    print("batch tuned", cfg, flush=True)
    assert cfg[0] == cfg[2] and cfg[1] == cfg[3]
    assert not cfg[1]["register_x_pass"] and 1 <= cfg[1]["cp"] <= 12  # nx = 96: LDS x pass only
    for d, (p, V, psi) in zip(b.instances, cases):
        d.load_pote(V)
        d.load_buffer(psi)
    steps, e = b.it_until(300, 100, 0.0)
    assert list(steps) == [300] * len(cases)
    got = [d.export_psi() for d in b.instances]
    b.close()
    for q in (0, 1):
        p, V, psi = cases[q]
        monkeypatch.setenv("GPE_LINE_LDS", "0" if cfg[q]["register_x_pass"] else "1")
        s = hip.Device(p.to_gpe(), hip.GPE_FP32)
        s.set_tiles(cp=cfg[q]["cp"])
        s.load_pote(V)
        s.load_buffer(psi)
        s.it_until(300, 100, 0.0)
        assert np.array_equal(s.export_psi(), got[q]), q
        s.close()
