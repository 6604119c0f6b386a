"""On-device initial states and potentials (gpe_init_wavefunction /
gpe_init_potential, restating src/initializer/initializer.cpp:46-248 and
src/parameters/potentials.hpp:76-157) against the host restatement
gpesolver_amd/params.py, which tests/test_ref_pins.py pins to the
reference-built initializer and potentials (1e-13 / 1e-14 relative): every
potential and initial state, FP64 and FP32, plus the SplitMix64 phase noise
and load_buffer's normalisation."""
import numpy as np
import pytest

from gpesolver_amd import hip
from gpesolver_amd import params as P

pytestmark = pytest.mark.gpu

SHAPE = (24, 20, 12)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def params(shape=SHAPE, **kw):
    return P.PhysicalParameters.from_physical(*shape, 120.0, 120.0, 600.0, n_atoms=2e4, edd=1.5,
                                              dd_nm=0.2 * shape[0] * 120.0, **kw)


@pytest.mark.parametrize("prec", [hip.GPE_FP64, hip.GPE_FP32])
@pytest.mark.parametrize("key", sorted(hip.POT_KEYS))
def test_potential_on_device(key, prec):
    p = params(omega_hz=(60.0, 70.0, 120.0))
    dev = hip.Device(p.to_gpe(), prec)
    dev.init_potential(key, p)
    got, want = dev.export_pote(), P.potential(key, p)
    tol = 1e-13 if prec == hip.GPE_FP64 else 1e-7
    assert np.abs(got - want).max() <= tol * np.abs(want).max(), key
    dev.close()


CASES = [("COS", {}), ("GAUSS", {}), ("GAUSS", {"sigma_cells": 3.0}), ("MULTIPLE_GAUSS", {"n_gauss_max": 2}),
         ("MULTIPLE_GAUSS", {"n_gauss_max": 3}), ("MULTIPLE_GAUSS", {"n_gauss_max": 4, "sigma_cells": 2.5}),
         ("SETUP_GAUSS", {"drops": (3, 3, 2)}), ("SETUP_GAUSS", {"drops": (2, 1, 1)})]


@pytest.mark.parametrize("prec", [hip.GPE_FP64, hip.GPE_FP32])
@pytest.mark.parametrize("strategy,kw", CASES, ids=[f"{s}-{i}" for i, (s, _) in enumerate(CASES)])
def test_wavefunction_on_device(strategy, kw, prec):
    p = params()
    host = P.wavefunction(strategy, p, **kw)
    noise, seed = 1e-3, 20261203
    host = P.add_phase_noise(host, noise, seed)
    ref = hip.Device(p.to_gpe(), prec)
    ref.load_buffer(host)  # host generation + load_buffer
    dev = hip.Device(p.to_gpe(), prec)
    dev.init_wavefunction(strategy, p, phase_noise=noise, seed=seed, **kw)
    got, want = dev.export_psi(), ref.export_psi()
    r = rel(got, want)
    print(f"\n{strategy} {kw} fp{prec}: device vs host psi_0 {r:.2e}, norm {dev.calc_norm() - 1:+.1e}", flush=True)
    assert r < (1e-13 if prec == hip.GPE_FP64 else 1e-7)
    assert abs(dev.calc_norm() - 1.0) < (1e-13 if prec == hip.GPE_FP64 else 1e-6)
    dev.close()
    ref.close()


def test_config2_state_and_potential_on_device():
    """128x128x32, MEXICAN dd = 1500 nm, MULTIPLE_GAUSS(2) + 1e-3 rad noise
    (seed 20261203), FP32: the device-built run equals the host-built one
    after 20 IT steps bit for bit or to FP32 rounding."""
    p = P.PhysicalParameters.from_physical(128, 128, 32, 120.0, 120.0, 600.0, n_atoms=2e4, edd=1.5, dd_nm=1500.0)
    V = P.potential("MEXICAN", p)
    psi = P.add_phase_noise(P.wavefunction("MULTIPLE_GAUSS", p, n_gauss_max=2), 1e-3, 20261203)
    a = hip.Device(p.to_gpe(), hip.GPE_FP32)
    a.load_pote(V)
    a.load_buffer(psi)
    b = hip.Device(p.to_gpe(), hip.GPE_FP32)
    b.init_potential("MEXICAN", p)
    b.init_wavefunction("MULTIPLE_GAUSS", p, n_gauss_max=2, phase_noise=1e-3, seed=20261203)
    assert np.abs(b.export_pote() - a.export_pote()).max() <= 1e-7 * np.abs(V).max()
    for d in (a, b):
        d.it_steps(20)
    r = rel(b.export_psi(), a.export_psi())
    print(f"\nconfig 2 device-built vs host-built after 20 IT steps: {r:.2e}", flush=True)
    assert r < 1e-6
    a.close()
    b.close()
