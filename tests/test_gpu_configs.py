"""Workload-level GPU checks at 128x128x32 (dx = dy = 120 nm, dz = 600 nm,
m = 163.929 Da, 2e4 atoms, edd = 1.5, MEXICAN with dd = 1500 nm,
MULTIPLE_GAUSS with 2 maxima plus seeded phase noise of 1e-3 rad).

* A full simulation (1000 IT steps, the FS switch to MEXICAN_FREE, then 3000
  RT steps with dt = 1e-4 / omega_x; src/manager/sim_manager.cpp:62-90) on
  the FP64 path and on the FP32-bulk path: the FP32 errors in psi
  (relative L2), E/N and mu against the FP64 run are printed and bounded.
* The first 20 IT and 20 RT steps of the same workload on both paths
  against the FP64 oracle, at the full 128x128x32 size.
"""
import math
import os

import numpy as np
import pytest

from gpesolver_amd import hip
from gpesolver_amd import params as P

pytestmark = pytest.mark.gpu

SHAPE = (128, 128, 32)
D = (120.0, 120.0, 600.0)
SEED = 20261203
# FP32 against FP64 after the 3000 real-time steps, about twice what MI355X
# measures (norm -2.6e-7, psi 1.8e-5, E/N 4.4e-7, mu 8.9e-7).  The psi error
# is FP32 round-off of the transforms accumulating over the steps (its
# amplitude part grows as sqrt(steps)); the dynamics do not amplify it: an
# FP64 run started 8.5e-8 away from another stays 9e-8 away after 3000 steps.
NORM_DRIFT_FP32 = 6e-7
PSI_RT_FP32 = 4e-5
EN_RT_FP32 = 1e-6
MU_RT_FP32 = 2e-6


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def case(oracle=None):
    nx, ny, nz = SHAPE
    p = P.PhysicalParameters.from_physical(nx, ny, nz, *D, n_atoms=2e4, edd=1.5, dd_nm=1500.0)
    p.real_time_dt = 1e-4 / p.omega_x
    V = P.potential("MEXICAN", p)
    Vf = P.potential("MEXICAN_FREE", p)
    psi = P.add_phase_noise(P.wavefunction("MULTIPLE_GAUSS", p, n_gauss_max=2), 1e-3, SEED)
    op = None
    if oracle is not None:
        op = oracle.make_params(nx, ny, nz, *D, n_atoms=2e4, edd=1.5, dd_nm=1500.0, real_dt=p.real_time_dt)
    return p, op, V, Vf, psi


def device(p, prec, V, psi):
    dev = hip.Device(p.to_gpe(), prec)
    dev.load_pote(V)
    dev.load_buffer(psi)
    return dev


def errors(dev, ref, n_atoms):
    e, eo = dev.energy(), ref.energy()
    return {"psi": rel(dev.export_psi(), ref.export_psi() if hasattr(ref, "export_psi") else ref.psi),
            "E/N": abs(e["e_total"] - eo["e_total"]) / abs(eo["e_total"]),
            "mu": abs(e["mu"] - eo["mu"]) / abs(eo["mu"]),
            "E/N_value": eo["e_total"] / n_atoms}


def test_full_simulation_fp32_vs_fp64_128x128x32():
    p, _, V, Vf, psi = case()
    runs = {}
    for prec in (hip.GPE_FP64, hip.GPE_FP32):
        runs[prec] = device(p, prec, V, psi)
    for prec, dev in runs.items():
        dev.it_steps(1000)
    err_it = errors(runs[hip.GPE_FP32], runs[hip.GPE_FP64], p.n_atoms)
    for dev in runs.values():
        dev.load_pote(Vf)
        dev.rt_steps(3000)
    err_rt = errors(runs[hip.GPE_FP32], runs[hip.GPE_FP64], p.n_atoms)
    print(f"\nFS 128x128x32 FP32 vs FP64 after 1000 IT: psi {err_it['psi']:.3e}  E/N {err_it['E/N']:.3e}  "
          f"mu {err_it['mu']:.3e}  (E/N = {err_it['E/N_value']:.9e} a.u.)", flush=True)
    print(f"FS 128x128x32 FP32 vs FP64 after 3000 RT: psi {err_rt['psi']:.3e}  E/N {err_rt['E/N']:.3e}  "
          f"mu {err_rt['mu']:.3e}  (E/N = {err_rt['E/N_value']:.9e} a.u.)", flush=True)
    n64, n32 = runs[hip.GPE_FP64].calc_norm(), runs[hip.GPE_FP32].calc_norm()
    print(f"norm after 3000 RT steps: FP64 {n64 - 1:+.3e}  FP32 {n32 - 1:+.3e}", flush=True)
    for d in runs.values():
        assert math.isfinite(d.energy()["e_total"])
    # real time is unitary: FP64 keeps the norm to rounding, FP32 drifts by
    # its per-step rounding only (no systematic loss in the transforms)
    assert abs(n64 - 1.0) < 1e-12
    assert abs(n32 - 1.0) < NORM_DRIFT_FP32
    assert err_it["psi"] < 1e-5 and err_it["E/N"] < 1e-6 and err_it["mu"] < 1e-6
    assert err_rt["psi"] < PSI_RT_FP32 and err_rt["E/N"] < EN_RT_FP32 and err_rt["mu"] < MU_RT_FP32
    for d in runs.values():
        d.close()


def test_first_steps_against_oracle_128x128x32(oracle):
    oracle.set_threads(min(16, os.cpu_count() or 1))
    p, op, V, Vf, psi = case(oracle)
    orc = oracle.Oracle(op)
    orc.load_pote(V)
    orc.load_buffer(psi)
    orc.it_steps(20)
    eo_it = orc.energy()
    psi_it = orc.psi.copy()
    orc.load_pote(Vf)
    orc.rt_steps(0, 20)
    eo_rt = orc.energy()
    for prec in (hip.GPE_FP64, hip.GPE_FP32):
        dev = device(p, prec, V, psi)
        dev.it_steps(20)
        e = dev.energy()
        tol = 1e-10 if prec == hip.GPE_FP64 else 1e-6
        r = rel(dev.export_psi(), psi_it)
        print(f"\n{'FP64' if prec == hip.GPE_FP64 else 'FP32'} 20 IT vs oracle: psi {r:.3e}  "
              f"E/N {abs(e['e_total'] - eo_it['e_total']) / abs(eo_it['e_total']):.3e}", flush=True)
        assert r < (1e-10 if prec == hip.GPE_FP64 else 1e-5)
        assert abs(e["e_total"] - eo_it["e_total"]) < tol * abs(eo_it["e_total"])
        assert abs(e["mu"] - eo_it["mu"]) < tol * abs(eo_it["mu"])
        dev.load_pote(Vf)
        dev.rt_steps(20)
        e = dev.energy()
        r = rel(dev.export_psi(), orc.psi)
        print(f"{'FP64' if prec == hip.GPE_FP64 else 'FP32'} +20 RT vs oracle: psi {r:.3e}  "
              f"E/N {abs(e['e_total'] - eo_rt['e_total']) / abs(eo_rt['e_total']):.3e}", flush=True)
        assert r < (1e-10 if prec == hip.GPE_FP64 else 2.5e-6)  # FP32 measured 1.0e-6
        assert abs(e["e_total"] - eo_rt["e_total"]) < tol * abs(eo_rt["e_total"])
        dev.close()


def test_config3_batched_it_sweep():
    """64 instances over 64x64x16, 96x96x24 and 128x128x32 (evenly mixed), FP32
    bulk, one launch per pass and grid shape over every active instance, x-pass
    variants from the batch autotuner, each to |dE/N| < 1e-6 between 100-step
    checks or 5000 steps; one instance per grid again on its own handle."""
    from gpesolver_amd import sweep

    setups = sweep.instances(64)
    r = sweep.run_batch(setups, 5000, 100, 1e-6, tune_steps=30)
    steps, e = r["steps"], r["e_per_atom"]
    print(f"\nconfig 3 sweep: {r['wall_s']:.2f} s (tuning {r['tune_s']:.2f} s), {r['converged']} of 64 converged, "
          f"steps mean {steps.mean():.0f} max {steps.max()}, {r['G_point_steps_per_s']:.2f} G point-steps/s, "
          f"tuned {r['tuned_per_nx']}", flush=True)
    assert np.all(np.isfinite(e)) and np.all(steps % 100 == 0) and np.all(steps <= 5000)
    assert r["converged"] >= 32
    assert set(r["tuned_per_nx"]) == {"64", "96", "128"}
    for q in range(3):
        p, V, psi = setups[q]
        d = hip.Device(p.to_gpe(), hip.GPE_FP32)
        d.load_pote(V)
        d.load_buffer(psi)
        s1, e1 = d.it_until(5000, check_every=100, tol=1e-6)
        print(f"  instance {q} {sweep.SHAPES[q]}: batch {int(steps[q])} steps E/N {e[q]:.9e}, single {s1} steps "
              f"E/N {e1:.9e}", flush=True)
        assert abs(s1 - int(steps[q])) <= 200
        assert abs(e1 - e[q]) <= 1e-5 * abs(e1)
        d.close()
