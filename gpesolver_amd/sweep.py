"""Batched imaginary-time parameter sweep (no reference counterpart: the
reference runs one parameter set per process, src/parameters/parameters.hpp:14-20).

Instances over 64x64x16, 96x96x24 and 128x128x32 (evenly mixed; dx = dy =
120 nm, dz = 600 nm), eps_dd ~ U[1.2, 1.6], n_atoms log-uniform in
[5e3, 5e4], omega_z ~ U[100, 140] Hz, MEXICAN with dd = 1500 nm and
MULTIPLE_GAUSS(2) (4-cell widths) plus 1e-3 rad phase noise, each run to
|dE/N| < tol between check_every-step energy checks or max_steps.

run_batch steps them all in one gpe_batch (one launch per pass and grid
shape, x-pass variants from gpe_batch_autotune); run_sequential runs the
same instances one after another on single handles tuned by gpe_autotune.
"""
from __future__ import annotations

import time

import numpy as np

from . import hip
from . import params as P

SHAPES = ((64, 64, 16), (96, 96, 24), (128, 128, 32))
SPACING_NM = (120.0, 120.0, 600.0)
SEED = 20261203


def instance(rng: np.random.Generator, shape):
    """(params, V, psi0) of one sweep instance on grid `shape`."""
    nx, ny, nz = shape
    edd = rng.uniform(1.2, 1.6)
    n_atoms = 10 ** rng.uniform(np.log10(5e3), np.log10(5e4))
    wz = rng.uniform(100.0, 140.0)
    p = P.PhysicalParameters.from_physical(nx, ny, nz, *SPACING_NM, n_atoms=n_atoms, edd=edd,
                                           omega_hz=(60.0, 60.0, wz), dd_nm=1500.0)
    V = P.potential("MEXICAN", p)
    psi = P.add_phase_noise(P.wavefunction("MULTIPLE_GAUSS", p, n_gauss_max=2, sigma_cells=4.0), 1e-3,
                            int(rng.integers(1 << 30)))
    return p, V, psi


def instances(n: int, seed: int = SEED, shapes=SHAPES) -> list:
    rng = np.random.default_rng(seed)
    return [instance(rng, shapes[q % len(shapes)]) for q in range(n)]


def _points(setups) -> np.ndarray:
    return np.array([s[0].nx * s[0].ny * s[0].nz for s in setups], dtype=np.float64)


def run_batch(setups, max_steps=5000, check_every=100, tol=1e-6, tune_steps=50, precision=hip.GPE_FP32) -> dict:
    b = hip.Batch([s[0].to_gpe() for s in setups], precision)
    try:
        t0 = time.perf_counter()
        tuned = b.autotune(tune_steps) if tune_steps > 0 else None
        t_tune = time.perf_counter() - t0
        for d, (p, V, psi) in zip(b.instances, setups):
            d.load_pote(V)
            d.load_buffer(psi)
        t0 = time.perf_counter()
        steps, e = b.it_until(max_steps, check_every, tol)
        wall = time.perf_counter() - t0
    finally:
        b.close()
    per_nx = {}
    if tuned is not None:
        for (p, _, _), c in zip(setups, tuned):
            per_nx[str(p.nx)] = c
    return {"wall_s": wall, "tune_s": t_tune, "steps": steps, "e_per_atom": e, "tuned_per_nx": per_nx,
            "converged": int(np.sum(steps < max_steps)),
            "G_point_steps_per_s": float(np.sum(steps * _points(setups)) / wall / 1e9)}


def run_sequential(setups, max_steps=5000, check_every=100, tol=1e-6, tune_steps=200,
                   precision=hip.GPE_FP32) -> dict:
    """The same instances one after another on autotuned single handles (the
    tuning itself, cached per grid, is outside the timed loop)."""
    steps, es = [], []
    wall = 0.0
    for p, V, psi in setups:
        d = hip.Device(p.to_gpe(), precision)
        try:
            if tune_steps > 0:
                d.autotune(tune_steps)
            d.load_pote(V)
            d.load_buffer(psi)
            d.synchronize()
            t0 = time.perf_counter()
            s, e = d.it_until(max_steps, check_every, tol)
            wall += time.perf_counter() - t0
        finally:
            d.close()
        steps.append(s)
        es.append(e)
    steps = np.array(steps, dtype=np.int64)
    return {"wall_s": wall, "steps": steps, "e_per_atom": np.array(es), "converged": int(np.sum(steps < max_steps)),
            "G_point_steps_per_s": float(np.sum(steps * _points(setups)) / wall / 1e9)}
