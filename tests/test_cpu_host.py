"""CPU-only checks: the C-ABI library loads and exports what include/gpehip.h
declares, argument validation at the boundary, and the host configuration
mirror (units, derived constants, potentials, initial states) against the
oracle's restatement of parameters.cpp / potentials.hpp / initializer.cpp."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from gpesolver_amd import hip
from gpesolver_amd import params as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols(name="gpehip.h", prefix="gpe_"):
    txt = open(os.path.join(ROOT, "include", name)).read()
    return sorted(set(re.findall(r"\b(" + prefix + r"[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = hip.load()
    declared = header_symbols()
    assert declared, "no declarations found"
    assert set(declared) == set(hip.EXPORTS)
    host = header_symbols("gpesolver_host.h", "gpes_")
    assert set(host) == set(hip.HOST_EXPORTS)
    declared = declared + host
    for name in declared:
        assert hasattr(lib, name), name
    assert b"gfx950" in lib.gpe_version()


def test_create_rejects_bad_arguments_without_touching_the_device():
    lib = hip.load()
    g = P.PhysicalParameters.from_physical(8, 8, 4, 120, 120, 600).to_gpe()
    h = C.c_void_p()
    assert lib.gpe_create(C.byref(g), 16, C.byref(h)) != 0
    assert b"precision" in lib.gpe_last_error(None)
    g.nx = 2
    assert lib.gpe_create(C.byref(g), 64, C.byref(h)) != 0
    assert b">= 3" in lib.gpe_last_error(None)
    g.nx = 8
    g.n_atoms = -1.0
    assert lib.gpe_create(C.byref(g), 64, C.byref(h)) != 0
    assert b"positive" in lib.gpe_last_error(None)
    g.n_atoms = 1.0
    g.nz = 4096  # z rows beyond the row tiles of the pencil plane pass
    assert lib.gpe_create(C.byref(g), 32, C.byref(h)) != 0
    assert b"pencil plane path" in lib.gpe_last_error(None)


def test_null_handle_calls_fail_cleanly():
    lib = hip.load()
    assert lib.gpe_calc_fi3d(None) != 0
    assert lib.gpe_imag_time_iters(None, 3) != 0
    assert lib.gpe_synchronize(None) != 0


def test_derived_constants_match_oracle(oracle):
    p = P.PhysicalParameters.from_physical(128, 128, 32, 120.0, 120.0, 600.0, n_atoms=2e4, edd=1.5)
    op = oracle.make_params(128, 128, 32, 120.0, 120.0, 600.0, n_atoms=2e4, edd=1.5)
    for k in ("dx", "dy", "dz", "m", "omega_x", "omega_z", "dd", "aa", "b", "cdd", "ggp11", "gamma", "w_15"):
        a, b = getattr(p, k), getattr(op, k)
        assert abs(a - b) <= 1e-15 * abs(b), k
    # magnitudes quoted for this configuration
    assert abs(p.m - 298824) / 298824 < 1e-5
    assert abs(p.cdd - 0.01653) / 0.01653 < 1e-3
    assert abs(p.ggp11 - 0.003673) / 0.003673 < 1e-3


@pytest.mark.parametrize("key", P.POTENTIALS)
def test_potentials_match_oracle(oracle, key):
    p = P.PhysicalParameters.from_physical(20, 14, 10, 120.0, 150.0, 600.0, dd_nm=700.0)
    op = oracle.make_params(20, 14, 10, 120.0, 150.0, 600.0, dd_nm=700.0)
    np.testing.assert_allclose(P.potential(key, p), oracle.potential(key, op), rtol=1e-13, atol=0)


@pytest.mark.parametrize("strategy,kw", [("COS", {}), ("GAUSS", {}), ("GAUSS", {"sigma_cells": 4.0}),
                                         ("MULTIPLE_GAUSS", {"n_gauss_max": 2}),
                                         ("MULTIPLE_GAUSS", {"n_gauss_max": 3}),
                                         ("SETUP_GAUSS", {"drops": (3, 3, 2)})])
def test_initial_states_match_oracle(oracle, strategy, kw):
    p = P.PhysicalParameters.from_physical(24, 20, 12, 120.0, 120.0, 600.0, dd_nm=800.0)
    op = oracle.make_params(24, 20, 12, 120.0, 120.0, 600.0, dd_nm=800.0)
    got = P.wavefunction(strategy, p, **kw)
    okw = dict(kw)
    okw.setdefault("n_gauss_max", 2)
    exp = oracle.wavefunction(strategy, op, n_gauss_max=okw["n_gauss_max"], drops=okw.get("drops", (3, 3, 2)),
                              sigma_cells=okw.get("sigma_cells", 0.0))
    np.testing.assert_allclose(got, exp, rtol=1e-12, atol=1e-300)
    assert np.all(got[0] == 0) and np.all(got[:, -1] == 0) and np.all(got[:, :, 0] == 0)


def test_edd_ramp_matches_oracle(oracle):
    p = P.PhysicalParameters.from_physical(8, 8, 4, 120, 120, 600, edd=None, edd_start=1.5, edd_stop=1.3,
                                           iter_real=3000)
    op = oracle.make_params(8, 8, 4, 120, 120, 600, edd=None, edd_start=1.5, edd_stop=1.3, iter_real=3000)
    for it in (0, 1, 1499, 3000):
        oracle.lib().orc_update_edd(C.byref(op), it)
        edd, g, gm = p.coupling_at(it)
        assert edd == op.edd and g == op.ggp11 and gm == op.gamma


def test_splitmix_phase_noise_is_seeded_and_bounded():
    psi = np.ones((6, 5, 4), dtype=np.complex128)
    a = P.add_phase_noise(psi, 1e-3, 20261203)
    b = P.add_phase_noise(psi, 1e-3, 20261203)
    c = P.add_phase_noise(psi, 1e-3, 1)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    ang = np.angle(a)
    assert np.all(np.abs(ang) <= 1e-3) and np.ptp(ang) > 1.5e-3
    np.testing.assert_allclose(np.abs(a), 1.0, rtol=1e-15)
    # SplitMix64 known first output for seed 0: 0xE220A8397B1DCDAF
    u = P.splitmix64(0, 1)[0]
    assert u == (0xE220A8397B1DCDAF >> 11) / 2.0**53
