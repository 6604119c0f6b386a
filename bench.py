#!/usr/bin/env python3
"""Benchmark: real-time split-step throughput at 128x128x32 on one MI355X.

Workload (one "step" = one real_time_iter of the reference,
src/solver/solver.cpp:107-112): grid 128x128x32, dx=dy=120 nm, dz=600 nm,
m=163.929 Da, n_atoms=2e4, eps_dd=1.5, MEXICAN potential with dd=1500 nm,
MULTIPLE_GAUSS(2) initial state with SplitMix64 phase noise (1e-3 rad, seed
20261203), dt = 1e-4/omega_x.  FP32 bulk arithmetic, FP64 reductions.  The
field, potential and all tables are resident in HBM before the timed region.

Timing: after the warm-up, exactly --steps steps are timed between two
device synchronisations, 5 times over; `value` is the median.  Between
consecutive calls the solver keeps the last step's closing half phase pending
and merges it into the next call's first step, so each timed call runs
exactly --steps whole steps of work (one half phase carried in, one out).

Prints ONE JSON line.  Besides steps/s it reports:
  roofline      the dominant kernel's achieved bandwidth (its algorithmic
                bytes per launch / its average launch time, from start / stop
                HIP events attached to each steady-state dispatch on the
                solver's stream) against the 8 TB/s HBM peak; `traffic` is the
                HBM bytes per launch that the FETCH_SIZE / WRITE_SIZE counter
                runs recorded for that kernel (profiles/traffic_*.json, each
                counter scaled by its reading for a copy of known size);
  step_roofline the model bytes of one RT step (112N + 20H, FP32) x steps/s
                against the same peak;
  cpu_baseline  the FP64 CPU oracle (the project's restatement of
                cpu_solver + FFTW solvers) timed on this host's cores on a
                bounded sample of the same workload;
  rt_fp64       the same RT workload on the FP64 path (the reference's
                precision), median steps/s and its step roofline;
  it_steps_per_s, it_time_to_converge
                fused imaginary-time throughput on the same grid, and the
                wall time of plain IT from MULTIPLE_GAUSS to |dE/N| < 1e-6
                between 100-step checks (energies included);
  it_warm_vs_plain
                the learned warm start on 64x64x16 (FNO width 32, 4 spectral
                layers, 16 modes), trained in this run by its own kernels on
                ground states the batched IT solver generates in this run
                (the recipe of gpesolver_amd.warmstart_experiment, fewer
                epochs), then batched IT to |dE/N| < 1e-6 on held-out
                instances from the FNO's prediction against plain
                MULTIPLE_GAUSS: wall times (inference included) and steps;
                `converged_in_both` times both runs again over the instances
                that reach the tolerance within the 20000-step cap in both
                (an instance at the cap in both runs ends both runs);
  it_batch_sweep
                the batched IT sweep (64 instances over 64x64x16, 96x96x24 and
                128x128x32, each to |dE/N| < 1e-6 or 5000 steps) in one batch
                with the batch autotuner's x-pass variants, against the same
                instances one after another on autotuned single handles;
                `batch_step_roofline` is the IT step model summed over every
                instance's steps, per second of the batched run.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12
GRID = (128, 128, 32)
REPS = 5
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_rt_128x128x32_fp32.json")


def measured_traffic(kernel: str):
    """Counter-measured HBM bytes per launch of `kernel` (None if absent)."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        return t["kernels"][kernel]["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def build_case(precision, sigma_cells=0.0):
    from gpesolver_amd import hip
    from gpesolver_amd import params as P

    nx, ny, nz = GRID
    p = P.PhysicalParameters.from_physical(nx, ny, nz, 120.0, 120.0, 600.0, m_da=163.929, n_atoms=2e4, edd=1.5,
                                           omega_hz=(60.0, 60.0, 120.0), dd_nm=1500.0)
    p.real_time_dt = 1e-4 / p.omega_x
    V = P.potential("MEXICAN", p)
    psi = P.add_phase_noise(P.wavefunction("MULTIPLE_GAUSS", p, n_gauss_max=2, sigma_cells=sigma_cells), 1e-3,
                            20261203)
    dev = hip.Device(p.to_gpe(), precision)
    dev.load_pote(V)
    dev.load_buffer(psi)
    dev.calc_fi3d()
    return p, dev, V, psi


def calib_copy(nbytes: int):
    """One hipMemcpy device-to-device of nbytes (a plain copy kernel that
    reads and writes nbytes), for scaling the memory counters."""
    import ctypes as C
    hiprt = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    if hiprt.hipMalloc(C.byref(a), C.c_size_t(nbytes)) or hiprt.hipMalloc(C.byref(b), C.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    try:
        hiprt.hipMemset(a, 1, C.c_size_t(nbytes))
        hiprt.hipDeviceSynchronize()
        if hiprt.hipMemcpy(b, a, C.c_size_t(nbytes), 3):  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
        hiprt.hipDeviceSynchronize()
    finally:
        hiprt.hipFree(a)
        hiprt.hipFree(b)


def model_bytes(nx, ny, nz, fp64=False):
    """Algorithmic bytes of one RT step: 112N + 20H (FP32), H = 2nx*2ny*(nz+1)."""
    N = nx * ny * nz
    H = 2 * nx * 2 * ny * (nz + 1)
    b = 112 * N + 20 * H
    return 2 * b if fp64 else b


def kernel_bytes(nx, ny, nz, cs=8, rs=4):
    """Algorithmic bytes of one steady-state real-time launch (its own inputs
    and outputs, each once).  S1 holds 2ny x nx x nz complex values (kz =
    1..nz), Tk N.  Plane pass, per x-plane: read S1 + Tk + V, write S1 + Tk
    = 52 ny nz bytes in FP32, so 52 N per launch; line pass: read and write
    S1 + Tk = 48 N per launch."""
    N = nx * ny * nz
    s1 = cs * 2 * ny * nx * nz
    tk = cs * N
    return {"plane": 2 * s1 + 2 * tk + rs * N, "line": 2 * s1 + 2 * tk}


def cpu_baseline(threads: int, budget_s: float = 12.0):
    """FP64 oracle RT steps at the same grid, on `threads` host threads."""
    from gpesolver_amd import params as P
    from oracle import pyoracle

    nx, ny, nz = GRID
    pyoracle.set_threads(threads)
    op = pyoracle.make_params(nx, ny, nz, 120.0, 120.0, 600.0, n_atoms=2e4, edd=1.5, dd_nm=1500.0)
    op.real_time_dt = 1e-4 / op.omega_x
    o = pyoracle.Oracle(op)
    p = P.PhysicalParameters.from_physical(nx, ny, nz, 120.0, 120.0, 600.0, n_atoms=2e4, edd=1.5, dd_nm=1500.0)
    o.load_pote(P.potential("MEXICAN", p))
    o.load_buffer(P.add_phase_noise(P.wavefunction("MULTIPLE_GAUSS", p, n_gauss_max=2), 1e-3, 20261203))
    o.calc_fi3d()
    o.rt_step()  # warm-up (plans, first touch)
    n = 0
    t0 = time.perf_counter()
    while True:
        o.rt_step()
        n += 1
        el = time.perf_counter() - t0
        if el > budget_s or n >= 200:
            break
    return {"value": n / el, "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": f"{n} FP64 RT steps of the same 128x128x32 workload ({el:.1f} s), FFT threaded with OpenMP"}


def large_rt(steps: int, warmup: int, fp64: bool) -> dict:
    """256x256x64 (dx = dy = 60 nm, dz = 300 nm), 1e5 atoms, MEXICAN_ASYMETRIC,
    SETUP_GAUSS 3x3x2, eps_dd ramped 1.5 -> 1.3 over 20000 steps: RT steps/s
    of the pencil plane pass, FP32 (and FP64 with --fp64)."""
    from gpesolver_amd import hip
    from gpesolver_amd import params as P

    nx, ny, nz = 256, 256, 64
    p = P.PhysicalParameters.from_physical(nx, ny, nz, 60.0, 60.0, 300.0, n_atoms=1e5, edd=None, edd_start=1.5,
                                           edd_stop=1.3, iter_real=20000)
    V = P.potential("MEXICAN_ASYMETRIC", p)
    psi = P.wavefunction("SETUP_GAUSS", p, drops=(3, 3, 2))
    n = warmup + steps
    g = np.array([p.coupling_at(t)[1] for t in range(n)])
    gm = np.array([p.coupling_at(t)[2] for t in range(n)])
    out = {"metric": "RT split-step steps/s @256x256x64 (pencil plane pass)", "unit": "steps/s", "steps": steps,
           "warmup": warmup, "config": {"workload": "RT 256x256x64, MEXICAN_ASYMETRIC, SETUP_GAUSS 3x3x2, "
                                                    "edd ramp 1.5->1.3 over 20000 steps"}}
    for prec, name in ((hip.GPE_FP32, "fp32"), (hip.GPE_FP64, "fp64")):
        if prec == hip.GPE_FP64 and not fp64:
            continue
        d = hip.Device(p.to_gpe(), prec)
        d.autotune(50)
        d.load_pote(V)
        d.load_buffer(psi)
        d.rt_steps(warmup, g[:warmup], gm[:warmup])
        d.synchronize()
        t0 = time.perf_counter()
        d.rt_steps(steps, g[warmup:], gm[warmup:])
        d.synchronize()
        el = time.perf_counter() - t0
        d.set_timing(True)
        d.rt_steps(min(steps, 200), g[:min(steps, 200)], gm[:min(steps, 200)])
        tm = d.timing()
        d.set_timing(False)
        mb = model_bytes(nx, ny, nz, fp64=prec == hip.GPE_FP64)
        out[name] = {"steps_per_s": steps / el, "ms_per_step": 1e3 * el / steps, "plane_pass_us": tm["plane_us"],
                     "x_pass_us": tm["line_us"],
                     "step_roofline": {"model_bytes_per_step": mb, "achieved_GBps": mb * steps / el / 1e9,
                                       "frac": mb * steps / el / HBM_PEAK}}
        d.close()
    return out


def timed_rt(dev, steps: int, reps: int = REPS) -> list:
    """Seconds of `reps` timed calls of exactly `steps` RT steps each."""
    out = []
    for _ in range(reps):
        dev.synchronize()
        t0 = time.perf_counter()
        dev.rt_steps(steps)
        dev.synchronize()
        out.append(time.perf_counter() - t0)
    return out


def host_cores() -> dict:
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    return {"cpu_count": os.cpu_count(), "affinity": aff, "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def batch_sweep(n: int) -> dict:
    """The batched IT sweep: n instances over 64x64x16, 96x96x24 and
    128x128x32, FP32, each to |dE/N| < 1e-6 between 100-step checks or 5000
    steps, in one gpe_batch with the batch autotuner's x-pass variants,
    against the same instances one after another on autotuned single handles."""
    from gpesolver_amd import sweep

    setups = sweep.instances(n)
    rb = sweep.run_batch(setups, 5000, 100, 1e-6, tune_steps=50)
    rs = sweep.run_sequential(setups, 5000, 100, 1e-6, tune_steps=200)
    de = np.abs(rb["e_per_atom"] - rs["e_per_atom"]) / np.abs(rs["e_per_atom"])
    # the IT step model (60N + 20H bytes, FP32) summed over every instance's
    # steps, per second of the batched run
    model = 0.0
    for (p, *_), st in zip(setups, rb["steps"]):
        N, H = p.nx * p.ny * p.nz, 2 * p.nx * 2 * p.ny * (p.nz + 1)
        model += float(st) * (60 * N + 20 * H)
    return {"instances": n, "grids": [list(s) for s in sweep.SHAPES], "tol_dE_per_N": 1e-6, "max_steps": 5000,
            "batch_s": rb["wall_s"], "batch_tune_s": rb["tune_s"], "sequential_s": rs["wall_s"],
            "speedup": rs["wall_s"] / rb["wall_s"], "batch_converged": rb["converged"],
            "sequential_converged": rs["converged"], "batch_G_point_steps_per_s": rb["G_point_steps_per_s"],
            "sequential_G_point_steps_per_s": rs["G_point_steps_per_s"],
            "steps_mean": float(np.mean(rb["steps"])), "energy_rel_diff_max": float(de.max()),
            "batch_tuned_per_nx": rb["tuned_per_nx"],
            "batch_step_roofline": {"model_bytes": model, "achieved_GBps": model / rb["wall_s"] / 1e9,
                                    "frac": model / rb["wall_s"] / HBM_PEAK}}


def warm_vs_plain(n_train: int, n_test: int, epochs: int, seed: int = 20261203) -> dict:
    """Train the warm-start FNO in this run and time batched IT, warm vs plain."""
    from gpesolver_amd import fno as F
    from gpesolver_amd import warmstart as W

    rng = np.random.default_rng(seed)
    insts = W.sample_instances(rng, n_train + n_test)
    t0 = time.perf_counter()
    X, T, S = W.generate(insts[:n_train], 1e-6, 5000, log=lambda m: None)
    t_gen = time.perf_counter() - t0
    model = F.Fno(*W.GRID, width=32, layers=4, hidden=64, modes=16, max_batch=8, seed=seed % 1000)
    t0 = time.perf_counter()
    hist = W.train(model, X, T, epochs, 1e-3, 8, log=lambda m: None)
    t_train = time.perf_counter() - t0
    del X, T
    res = W.compare(model, insts[n_train:], 1e-6, 20000)
    nparam = model.num_params
    model.close()
    ps, ws = np.array(res["plain_steps"]), np.array(res["warm_steps"])
    return {
        "grid": list(W.GRID), "instances": n_test, "tol_dE_per_N": 1e-6, "check_every": 100, "max_steps": 20000,
        "plain_wall_s": res["plain_wall_s"], "warm_wall_s": res["warm_wall_s"],
        "warm_inference_s": res["warm_inference_s"], "speedup_wall": res["speedup_wall"],
        "plain_steps_mean": res["plain_steps_mean"], "warm_steps_mean": res["warm_steps_mean"],
        "plain_steps_max": res["plain_steps_max"], "warm_steps_max": res["warm_steps_max"],
        "step_ratio_median": float(np.median(ps / ws)),
        "converged_in_both": res["converged_in_both"],
        "fallbacks_before_run": res["fallbacks"], "fallbacks_in_run": res["fallbacks_in_run"],
        "energy_rel_diff_median": res["energy_rel_diff_median"],
        "weights": f"trained in this run: {n_train} generated ground states, {epochs} epochs, batch 8, Adam lr 1e-3 "
                   f"(cosine), {nparam} parameters, BF16 GEMM operands / FP32 accumulation",
        "train_mse_last": hist[-1], "generate_s": t_gen, "train_s": t_train,
        "train_steps_median": float(np.median(S)),
    }


def it_only(dev, args):
    """IT leg alone on the default launch shapes: exactly --steps fused
    imaginary-time steps, REPS times (median), then per-kernel event times.
    Algorithmic bytes per IT launch (FP32): plane pass reads S1 + psi + V and
    writes S1 + psi (52 N, as in real time with psi in place of Tk), x pass
    reads and writes S1 (32 N)."""
    nx, ny, nz = GRID
    N = nx * ny * nz
    dev.it_steps(args.warmup)
    dev.synchronize()
    reps = []
    for _ in range(REPS):
        dev.synchronize()
        t0 = time.perf_counter()
        dev.it_steps(args.steps)
        dev.synchronize()
        reps.append(time.perf_counter() - t0)
    el = float(np.median(reps))
    dev.set_timing(True)
    dev.it_steps(min(args.steps, 1000))
    tm = dev.timing()
    dev.set_timing(False)
    dev.close()
    kb = {"plane": 52 * N, "line": 32 * N}
    kern = {"plane": tm["plane_us"], "line": tm["line_us"]}
    if args.calib_copy_mib > 0:
        calib_copy(args.calib_copy_mib << 20)
    H = 2 * nx * 2 * ny * (nz + 1)
    print(json.dumps({"metric": "it_steps_per_s", "value": args.steps / el, "unit": "steps/s", "steps": args.steps,
                      "warmup": args.warmup, "ms_per_step": 1e3 * el / args.steps, "it_only": True,
                      "avg_launch_us": kern, "algorithmic_bytes_per_launch": kb,
                      "kernel_frac": {k: kb[k] / (kern[k] * 1e-6) / HBM_PEAK for k in kb},
                      "step_model_bytes": 60 * N + 20 * H,
                      "step_frac": (60 * N + 20 * H) * args.steps / el / HBM_PEAK,
                      "calib_copy_mib": args.calib_copy_mib}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--fp64", action="store_true", help="--large: also time the FP64 path")
    ap.add_argument("--it-max-steps", type=int, default=20000, help="cap of the IT time-to-convergence leg")
    ap.add_argument("--no-warm", action="store_true", help="skip the warm-start leg (trains the FNO, ~2 min)")
    ap.add_argument("--warm-train", type=int, default=2000, help="warm-start leg: generated training states")
    ap.add_argument("--warm-epochs", type=int, default=60, help="warm-start leg: training epochs")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched IT sweep leg (~20 s)")
    ap.add_argument("--batch-n", type=int, default=64, help="batched IT sweep leg: instances")
    ap.add_argument("--rt-only", action="store_true",
                    help="profiling runs: the RT leg only, on the default launch shapes (split plane pass, "
                         "register x pass: what the autotuner picks at this grid), no autotune and no IT legs, "
                         "so every steady-state launch in a trace or counter run is an RT one")
    ap.add_argument("--it-only", action="store_true",
                    help="profiling runs: the IT leg only (fused steps, default launch shapes), timed and "
                         "per-kernel event times, so every steady-state launch in a trace or counter run is "
                         "an IT one")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="end with one device-to-device copy of this many MiB (the known-size "
                         "reference dispatch for FETCH_SIZE / WRITE_SIZE counter runs)")
    ap.add_argument("--large", action="store_true",
                    help="time the 256x256x64 ramped real-time workload (pencil plane pass) instead")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs one solver instance on one GPU; --gpus must be 1 "
              "(multi-GPU slab decomposition is design-only)", file=sys.stderr)
        sys.exit(2)
    if args.large:
        out = large_rt(args.steps, args.warmup, args.fp64)
        if args.calib_copy_mib > 0:
            calib_copy(args.calib_copy_mib << 20)
        print(json.dumps(out))
        return

    from gpesolver_amd import hip

    nx, ny, nz = GRID
    p, dev, V, psi = build_case(hip.GPE_FP32)
    if args.it_only:
        it_only(dev, args)
        return
    # launch-shape variant for this grid (scratch instance)
    tuned = None if args.rt_only else dev.autotune(200)
    # warm-up
    dev.rt_steps(args.warmup)
    dev.synchronize()
    # timed region: exactly K steps, REPS times; the median
    reps = timed_rt(dev, args.steps)
    el = float(np.median(reps))
    steps_per_s = args.steps / el

    # per-kernel durations from HIP events on the solver stream (separate pass)
    dev.set_timing(True)
    dev.rt_steps(min(args.steps, 1000))
    tm = dev.timing()
    dev.set_timing(False)
    kb = kernel_bytes(nx, ny, nz)
    kern = {"plane": tm["plane_us"], "line": tm["line_us"]}
    dom = max(kern, key=kern.get)
    achieved = kb[dom] / (kern[dom] * 1e-6)

    if args.rt_only:
        dev.close()
        if args.calib_copy_mib > 0:
            calib_copy(args.calib_copy_mib << 20)
        print(json.dumps({"metric": "rt_steps_per_s", "value": steps_per_s, "unit": "steps/s", "steps": args.steps,
                          "warmup": args.warmup, "ms_per_step": 1e3 / steps_per_s, "rt_only": True,
                          "avg_launch_us": kern, "calib_copy_mib": args.calib_copy_mib}), flush=True)
        return

    # IT throughput on the same grid (FP32), fused loop
    dev.it_steps(args.warmup)
    dev.synchronize()
    nit = max(100, args.steps // 3)
    t1 = time.perf_counter()
    dev.it_steps(nit)
    dev.synchronize()
    it_el = time.perf_counter() - t1
    e = dev.energy()
    dev.close()

    # IT time to |dE/N| < 1e-6 between 100-step checks from the plain
    # MULTIPLE_GAUSS start (4-cell widths), FP32, this grid
    pc, dc, _, _ = build_case(hip.GPE_FP32, sigma_cells=4.0)
    dc.autotune(200)
    dc.synchronize()
    t1 = time.perf_counter()
    it_conv_steps, it_conv_e = dc.it_until(args.it_max_steps, check_every=100, tol=1e-6)
    it_conv_s = time.perf_counter() - t1
    dc.close()

    # the same RT workload on the FP64 path
    _, d64, _, _ = build_case(hip.GPE_FP64)
    tuned64 = d64.autotune(100)
    d64.rt_steps(args.warmup)
    d64.synchronize()
    reps64 = timed_rt(d64, args.steps)
    el64 = float(np.median(reps64))
    d64.set_timing(True)
    d64.rt_steps(min(args.steps * 10, 500))
    tm64 = d64.timing()
    d64.set_timing(False)
    d64.close()
    mb64 = model_bytes(nx, ny, nz, fp64=True)
    extra = {"rt_fp64": {"steps_per_s": args.steps / el64, "ms_per_step": 1e3 * el64 / args.steps,
                         "rep_ms": [1e3 * r for r in reps64],
                         "avg_launch_us": {"plane": tm64["plane_us"], "line": tm64["line_us"]},
                         "tuned": {k: tuned64[k] for k in ("split_plane_pass", "register_x_pass", "graph_replay")},
                         "step_roofline": {"model_bytes_per_step": mb64,
                                           "achieved_GBps": mb64 * args.steps / el64 / 1e9,
                                           "frac": mb64 * args.steps / el64 / HBM_PEAK}}}

    if not args.no_warm:
        extra["it_warm_vs_plain"] = warm_vs_plain(args.warm_train, 64, args.warm_epochs)
    if not args.no_batch:
        extra["it_batch_sweep"] = batch_sweep(args.batch_n)

    if args.calib_copy_mib > 0:
        calib_copy(args.calib_copy_mib << 20)

    mb = model_bytes(nx, ny, nz)
    out = {
        "metric": "RT split-step steps/s @128²×32 (%HBM roof); IT time-to-ΔE<1e-6, warm vs plain",
        "value": steps_per_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * el / args.steps,
        "rep_ms": [1e3 * r for r in reps],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "RT split-step 128x128x32, MEXICAN dd=1500nm, MULTIPLE_GAUSS(2)+phase noise, "
                        "dt=1e-4/omega_x, FP32 bulk / FP64 reductions",
            "grid": list(GRID),
            "spacing_nm": [120.0, 120.0, 600.0],
            "n_atoms": 2e4,
            "edd": 1.5,
            "parallelism": "single GPU",
            "tuned": {k: tuned[k] for k in ("split_plane_pass", "register_x_pass", "graph_replay")},
        },
        "roofline": {
            "bound": "hbm",
            "kernel": f"k_{'plane' if dom == 'plane' else 'xpass'} (RT steady state)",
            "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "traffic": measured_traffic(dom),
            "bytes_per_launch": kb[dom],
            "avg_launch_us": kern[dom],
            "other_kernel_us": {k: v for k, v in kern.items() if k != dom},
        },
        "step_roofline": {
            "model_bytes_per_step": mb,
            "achieved_GBps": mb * steps_per_s / 1e9,
            "frac": mb * steps_per_s / HBM_PEAK,
        },
        "it_steps_per_s": nit / it_el,
        "it_energy_per_atom": e["e_total"] / p.n_atoms,
        "it_time_to_converge": {"tol_dE_per_N": 1e-6, "check_every": 100, "steps": it_conv_steps,
                                "seconds": it_conv_s, "energy_per_atom": it_conv_e,
                                "converged": it_conv_steps < args.it_max_steps,
                                "start": "MULTIPLE_GAUSS(2), 4-cell widths, 1e-3 rad phase noise"},
        **extra,
    }
    hc = host_cores()
    out["host"] = hc
    if not args.no_cpu:
        # the GPU box allots 16 host CPUs to a one-GPU job (OMP_NUM_THREADS=16
        # there); use what the process may run on, up to that share
        threads = max(1, min(16, hc["affinity"]))
        out["cpu_baseline"] = cpu_baseline(threads)
        out["cpu_baseline"]["host_cpu_count"] = hc["cpu_count"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
