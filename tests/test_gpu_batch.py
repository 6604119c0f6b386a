"""Batched imaginary time: several instances (different constants and grids)
stepped by one launch per pass, against the single-instance path and the
oracle, and the per-instance convergence test."""
import numpy as np
import pytest

from gpesolver_amd import hip
from gpesolver_amd import params as P

pytestmark = pytest.mark.gpu

CASES = [((32, 24, 12), 1.5, 2e4), ((16, 16, 8), 1.3, 8e3), ((32, 24, 12), 1.45, 3e4), ((24, 12, 6), 1.4, 1e4)]


def setup_instance(shape, edd, n_atoms):
    nx, ny, nz = shape
    p = P.PhysicalParameters.from_physical(nx, ny, nz, 120.0, 120.0, 600.0, n_atoms=n_atoms, edd=edd)
    V = P.potential("HARMONIC", p)
    psi = P.add_phase_noise(P.wavefunction("GAUSS", p), 1e-3, 11)
    return p, V, psi


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("prec", [hip.GPE_FP64, hip.GPE_FP32], ids=["fp64", "fp32"])
def test_batch_matches_single_and_oracle(oracle, prec):
    setups = [setup_instance(*c) for c in CASES]
    b = hip.Batch([s[0].to_gpe() for s in setups], prec)
    for d, (p, V, psi) in zip(b.instances, setups):
        d.load_pote(V)
        d.load_buffer(psi)
    steps, e = b.it_until(100, check_every=50, tol=0.0)
    assert list(steps) == [100] * len(CASES)
    for q, (p, V, psi) in enumerate(setups):
        single = hip.Device(p.to_gpe(), prec)
        single.load_pote(V)
        single.load_buffer(psi)
        single.it_steps(50)
        single.energy()
        single.it_steps(50)
        es = single.energy()
        got = b.instances[q].export_psi()
        assert rel(got, single.export_psi()) < (1e-14 if prec == hip.GPE_FP64 else 1e-6)
        assert abs(e[q] - es["e_total"] / p.n_atoms) <= 1e-12 * abs(e[q]) + (0 if prec == hip.GPE_FP64 else 1e-6 * abs(e[q]))
        op = oracle.make_params(p.nx, p.ny, p.nz, 120.0, 120.0, 600.0, n_atoms=p.n_atoms, edd=p.edd)
        o = oracle.Oracle(op)
        o.load_pote(V)
        o.load_buffer(psi)
        o.it_steps(100)
        assert rel(got, o.psi) < (1e-10 if prec == hip.GPE_FP64 else 1e-5)
        single.close()
    b.close()


def test_batch_convergence_drops_instances():
    # a small box converges in a few hundred steps, the 64x64x16 one needs more
    conv_cases = [((16, 16, 8), 1.4, 1e4), ((64, 64, 16), 1.5, 2e4), ((32, 24, 12), 1.3, 8e3)]
    setups = [setup_instance(*c) for c in conv_cases]
    b = hip.Batch([s[0].to_gpe() for s in setups], hip.GPE_FP32)
    for d, (p, V, psi) in zip(b.instances, setups):
        d.load_pote(V)
        d.load_buffer(psi)
    steps, e = b.it_until(2000, check_every=100, tol=1e-6)
    assert np.all(steps % 100 == 0) and np.all(steps <= 2000) and np.all(steps >= 200)
    assert len(set(steps.tolist())) > 1  # instances converge at different steps
    # each converged instance satisfies the criterion at its last check
    for q, d in enumerate(b.instances):
        if steps[q] < 2000:
            e1 = d.energy()["e_total"]
            d.it_steps(100)
            e2 = d.energy()["e_total"]
            assert abs(e2 - e1) < 1e-6 * abs(e2) * 1.5
    b.close()


@pytest.mark.parametrize("prec", [hip.GPE_FP64, hip.GPE_FP32], ids=["fp64", "fp32"])
def test_single_instance_until_matches_batch_of_one(prec):
    """gpe_imag_time_until on one handle stops where a batch of that one
    instance stops (same test, same steps, same E/N), and max_steps caps it."""
    p, V, psi = setup_instance((32, 24, 12), 1.4, 1e4)
    d = hip.Device(p.to_gpe(), prec)
    d.load_pote(V)
    d.load_buffer(psi)
    steps, e = d.it_until(4000, check_every=100, tol=1e-6)
    b = hip.Batch([p.to_gpe()], prec)
    b.instances[0].load_pote(V)
    b.instances[0].load_buffer(psi)
    bsteps, be = b.it_until(4000, check_every=100, tol=1e-6)
    print("until", prec, steps, e, int(bsteps[0]), flush=True)
    assert 0 < steps < 4000 and steps % 100 == 0
    if prec == hip.GPE_FP64:
        assert steps == int(bsteps[0])
        assert abs(e - be[0]) <= 1e-13 * abs(e)
        assert rel(d.export_psi(), b.instances[0].export_psi()) < 1e-13
    else:
        # the batched kernels round differently: the stop may fall one check apart
        assert abs(steps - int(bsteps[0])) <= 100
        assert abs(e - be[0]) <= 1e-6 * abs(e)
    # the cap, and a check interval that does not divide it
    d.load_buffer(psi)
    steps2, _ = d.it_until(250, check_every=100, tol=0.0)
    assert steps2 == 250
    d.close()
    b.close()


@pytest.mark.parametrize("generic", ["0", "1"])
def test_batch_groups_share_shapes_and_generic_kernels(monkeypatch, generic):
    """Instances sharing a plane shape but not nx (and the reverse) fall into
    the same launch group; with GPE_FORCE_GENERIC=1 every group runs the
    run-time-size kernels.  Each batched instance equals its single handle
    bit for bit (FP64)."""
    monkeypatch.setenv("GPE_FORCE_GENERIC", generic)
    cases = [((32, 24, 12), 1.5, 2e4), ((24, 24, 12), 1.4, 1e4), ((32, 16, 8), 1.3, 8e3), ((64, 16, 8), 1.45, 3e4)]
    setups = [setup_instance(*c) for c in cases]
    b = hip.Batch([s[0].to_gpe() for s in setups], hip.GPE_FP64)
    for d, (p, V, psi) in zip(b.instances, setups):
        d.load_pote(V)
        d.load_buffer(psi)
    steps, e = b.it_until(60, check_every=30, tol=0.0)
    assert list(steps) == [60] * len(cases)
    for q, (p, V, psi) in enumerate(setups):
        s = hip.Device(p.to_gpe(), hip.GPE_FP64)
        s.load_pote(V)
        s.load_buffer(psi)
        s.it_steps(30)
        s.energy()
        s.it_steps(30)
        assert np.array_equal(b.instances[q].export_psi(), s.export_psi()), q
        s.close()
    b.close()


def test_dropped_instances_keep_their_state_with_odd_check_interval():
    """A check interval that is odd flips the state buffer on every check:
    an instance dropped on convergence keeps the buffer its state is in while
    the others step on, and the last one left runs its handle's own loop.
    Every instance stops where its single handle stops, with its state (FP64,
    where the batched and single kernels agree to rounding)."""
    conv_cases = [((16, 16, 8), 1.4, 1e4), ((64, 64, 16), 1.5, 2e4), ((32, 24, 12), 1.3, 8e3)]
    setups = [setup_instance(*c) for c in conv_cases]
    b = hip.Batch([s[0].to_gpe() for s in setups], hip.GPE_FP64)
    for d, (p, V, psi) in zip(b.instances, setups):
        d.load_pote(V)
        d.load_buffer(psi)
    steps, e = b.it_until(3000, check_every=75, tol=1e-6)
    print("odd interval", steps.tolist(), flush=True)
    srt = sorted(steps.tolist())
    assert srt[0] < srt[-1] and srt[-2] < srt[-1]  # dropped early, then a lone instance
    for q, (p, V, psi) in enumerate(setups):
        s = hip.Device(p.to_gpe(), hip.GPE_FP64)
        s.load_pote(V)
        s.load_buffer(psi)
        st, es = s.it_until(3000, check_every=75, tol=1e-6)
        assert st == int(steps[q]), q
        assert abs(es - e[q]) <= 1e-12 * abs(es), q
        assert rel(b.instances[q].export_psi(), s.export_psi()) < 1e-12, q
        s.close()
    b.close()
