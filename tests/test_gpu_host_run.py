"""The host front end end-to-end on the GPU (gpes_run: gpe_params.json ->
SimulationManager::initialize + run_simulation, src/manager/sim_manager.cpp:16-27):
a full (FS) simulation's energy records equal the oracle's composition of
the same steps (src/solver/solver.cpp:25-97), and the reference's output
files are written."""
import json

import numpy as np
import pytest

from gpesolver_amd import hip
from gpesolver_amd import params as P

pytestmark = pytest.mark.gpu


def test_full_simulation_matches_oracle(tmp_path, oracle):
    nx, ny, nz = 24, 20, 12
    cfg = {"n_atoms": 20000, "m": 163.929, "dx": 120, "dy": 120, "dz": 600, "nx": nx, "ny": ny, "nz": nz,
           "omega_x": 60, "omega_y": 60, "omega_z": 120, "calc_strategy": "FS", "init_strategy": "GAUSS",
           "gauss_sigma_cells": 4, "pote_strategy": "MEXICAN", "dd": 800, "iter_imag": 200, "iter_real": 1000,
           "edd": 1.5, "fftw_n_threads": 1, "precision": "fp64"}
    (tmp_path / "gpe_params.json").write_text(json.dumps(cfg))
    print("host run FS", (nx, ny, nz), flush=True)
    res = hip.run_directory(str(tmp_path))
    assert res["it_steps"] == 201  # iterations 0..iter_imag inclusive
    rows = np.array([[float(x) for x in line.split()] for line in (tmp_path / "energy.gpe.dat").read_text().splitlines()])
    # energies after IT steps 1, 101, 201 and RT steps 1, 1001 (checks at iter % 100 / % 1000 == 0)
    p = P.PhysicalParameters.from_physical(nx, ny, nz, 120, 120, 600, n_atoms=2e4, edd=1.5, dd_nm=800)
    op = oracle.make_params(nx, ny, nz, 120, 120, 600, n_atoms=2e4, edd=1.5, dd_nm=800)
    orc = oracle.Oracle(op)
    orc.load_pote(P.potential("MEXICAN", p))
    orc.load_buffer(P.wavefunction("GAUSS", p, sigma_cells=4.0))
    ref = []
    for n in (1, 100, 100):
        orc.it_steps(n)
        ref.append(orc.energy())
    orc.load_pote(P.potential("MEXICAN_FREE", p))  # the FS switch (sim_manager.cpp:78-80)
    orc.rt_steps(0, 1)
    ref.append(orc.energy())
    orc.rt_steps(1, 1000)
    ref.append(orc.energy())
    assert rows.shape == (5, 7)
    exp = np.array([[e["e_kin"], e["e_pot"], e["e_int"], e["e_ext"], e["e_bmf"], e["e_total"]] for e in ref])
    # energy.gpe.dat holds 6 significant digits (default ostream precision, as the reference writes it)
    np.testing.assert_allclose(rows[:, 1:], exp, rtol=1e-5, atol=1e-12 * np.abs(exp).max())
    # the last energy in full precision (FP64 path)
    assert abs(res["energy_per_atom"] - ref[-1]["e_total"] / 2e4) <= 1e-10 * abs(ref[-1]["e_total"] / 2e4)
    for f in ("initial_potential.gpe.dat", "changed_potential.gpe.dat", "wavefunction_0.gpe.bin",
              "wavefunction_2.gpe.bin", "energy.gpe.dat"):
        assert (tmp_path / f).exists(), f


def _read_bin(path):
    raw = open(path, "rb").read()
    n = np.frombuffer(raw[:12], dtype=np.int32)
    return np.frombuffer(raw[36:], dtype=np.complex128).reshape(tuple(n))


def test_binary_resume_continues_the_run(tmp_path, oracle):
    """Checkpoint / resume (src/manager/sim_manager.cpp:92-116,
    src/filemanager/file_manager.cpp:290-385): a real-time run writes
    wavefunction_1 after 1001 steps; a second run started from that file
    (BINARY_FILE) numbers its checkpoints on from 2 and equals the oracle
    replaying the reference's resume: the loaded state normalised, fi3d zero
    for the first half step (fresh containers), iterations counted from 0.
    Against the uninterrupted run the resumed one differs only by that first
    half step's missing dipolar phase."""
    nx, ny, nz = 24, 20, 12
    base = {"n_atoms": 20000, "m": 163.929, "dx": 120, "dy": 120, "dz": 600, "nx": nx, "ny": ny, "nz": nz,
            "omega_x": 60, "omega_y": 60, "omega_z": 120, "calc_strategy": "RT", "pote_strategy": "HARMONIC",
            "dd": 800, "iter_imag": 0, "edd": 1.5, "fftw_n_threads": 1, "precision": "fp64"}
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    (a / "gpe_params.json").write_text(json.dumps(dict(base, init_strategy="GAUSS", gauss_sigma_cells=3,
                                                       iter_real=1001)))
    print("host run RT, then resume", (nx, ny, nz), flush=True)
    hip.run_directory(str(a))
    cp1 = _read_bin(a / "wavefunction_1.gpe.bin")
    (b / "wavefunction_1.gpe.bin").write_bytes((a / "wavefunction_1.gpe.bin").read_bytes())
    (b / "gpe_params.json").write_text(json.dumps(dict(base, init_strategy="BINARY_FILE",
                                                       load_filename="wavefunction_1", iter_real=1000)))
    hip.run_directory(str(b))
    got2, got3 = _read_bin(b / "wavefunction_2.gpe.bin"), _read_bin(b / "wavefunction_3.gpe.bin")
    p = P.PhysicalParameters.from_physical(nx, ny, nz, 120, 120, 600, n_atoms=2e4, edd=1.5, dd_nm=800)
    op = oracle.make_params(nx, ny, nz, 120, 120, 600, n_atoms=2e4, edd=1.5, dd_nm=800)
    orc = oracle.Oracle(op)
    orc.load_pote(P.potential("HARMONIC", p))
    orc.load_buffer(cp1)
    orc.rt_steps(0, 1)
    r2 = np.linalg.norm(got2 - orc.psi) / np.linalg.norm(orc.psi)
    orc.rt_steps(1, 1000)
    r3 = np.linalg.norm(got3 - orc.psi) / np.linalg.norm(orc.psi)
    # the uninterrupted trajectory over the same 2002 steps
    full = oracle.Oracle(op)
    full.load_pote(P.potential("HARMONIC", p))
    full.load_buffer(P.wavefunction("GAUSS", p, sigma_cells=3.0))
    full.rt_steps(0, 2002)
    dev = np.linalg.norm(got3 - full.psi) / np.linalg.norm(full.psi)
    print(f"resume vs oracle: after 1 step {r2:.2e}, after 1001 steps {r3:.2e}; vs uninterrupted run {dev:.2e}",
          flush=True)
    assert r2 < 1e-10 and r3 < 1e-10
    assert dev < 1e-2
