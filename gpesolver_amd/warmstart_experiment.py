"""Learned warm start experiment on 64x64x16 (FNO width 32, 4 spectral
layers, 16 modes per axis capped at nz/2+1 along z, BF16 GEMM operands with
FP32 accumulation): generate ground states with the batched IT solver, train
the network on the GPU, then time batched IT to |dE/N| < tol from the
network's prediction against plain MULTIPLE_GAUSS, on held-out instances,
and the same per instance on single handles (time to convergence including
that instance's inference).

    python -m gpesolver_amd.warmstart_experiment --n-train 2000 --n-test 64 --out result.json
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import fno as F
from . import warmstart as W


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n-train", type=int, default=2000)
    ap.add_argument("--n-test", type=int, default=64)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--max-steps", type=int, default=5000)
    ap.add_argument("--width", type=int, default=32)
    ap.add_argument("--layers", type=int, default=4)
    ap.add_argument("--modes", type=int, default=16)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--grid", type=int, nargs=3, default=list(W.GRID))
    ap.add_argument("--out", default="")
    ap.add_argument("--save-weights", default="")
    args = ap.parse_args(argv)
    grid = tuple(args.grid)
    rng = np.random.default_rng(args.seed)
    insts = W.sample_instances(rng, args.n_train + args.n_test)
    train_i, test_i = insts[:args.n_train], insts[args.n_train:]
    t0 = time.perf_counter()
    X, T, S = W.generate(train_i, args.tol, args.max_steps, grid=grid, log=lambda m: print(m, flush=True))
    t_gen = time.perf_counter() - t0
    model = F.Fno(*grid, width=args.width, layers=args.layers, hidden=64, modes=args.modes,
                  max_batch=args.batch, seed=args.seed % 1000)
    print(f"[warmstart] FNO parameters: {model.num_params}", flush=True)
    t0 = time.perf_counter()
    hist = W.train(model, X, T, args.epochs, args.lr, args.batch, log=lambda m: print(m, flush=True))
    t_train = time.perf_counter() - t0
    res = W.compare(model, test_i, args.tol, args.max_steps, grid=grid)
    res["single"] = W.compare_single(model, test_i, args.tol, args.max_steps, grid=grid)
    res.update({"grid": list(grid), "n_train": args.n_train, "epochs": args.epochs, "train_mse": hist,
                "generate_s": t_gen, "train_s": t_train, "train_steps_median": float(np.median(S)),
                "fno": {"width": args.width, "layers": args.layers, "modes": args.modes,
                        "params": model.num_params}})
    print(json.dumps(res), flush=True)
    if args.out:
        with open(args.out, "w") as f:
            json.dump(res, f, indent=1)
    if args.save_weights:
        w = model.get_params()
        np.save(args.save_weights, (w.view(np.uint32) >> 16).astype(np.uint16))  # BF16 bits
    model.close()


if __name__ == "__main__":
    main()
