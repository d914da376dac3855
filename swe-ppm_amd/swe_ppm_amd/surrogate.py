"""Training pipeline of the learned middle state (include/swe_hip_mlp.h).

The network is trained on middle states that the library's own exact AugRie
logs while it steps the plain workloads:
  * the 256^2 radial dam break (f-wave stepping, 200 steps),
  * the 2048^2 dam break over 16 bumps with 3 dry islands (AugRie, 500 steps),
  * the 28-tile still-water tiling of the LTS workload as one 2048^2 block
    (AugRie, 300 steps: the depth steps between tiles radiate waves),
every `every`-th step.  The solver settles most Newton edges with the
linearised middle state; the network is for the rest, so only edges whose
linearised guess fails the residual check are logged (hard_only), each kept
with a fixed probability so that about `target` samples are collected.
Training runs the library's FP32 forward / backward / Adam kernels on the GPU.

  python -m swe_ppm_amd.surrogate --out swe-ppm_amd/swe_ppm_amd/data/middle_state_mlp.npy
"""
from __future__ import annotations

import argparse
import json
import time
from pathlib import Path

import numpy as np

from . import AUGRIE, FWAVE, MiddleStateMLP, SweBlock, scenarios

DEFAULT_WEIGHTS = Path(__file__).resolve().parent / "data" / "middle_state_mlp.npy"


def _log_run(mlp, state, solver, steps, every, keep, seed, hard_only=True):
    h, hu, hv, b, dx, dy, bnd = state
    nx, ny = h.shape[0] - 2, h.shape[1] - 2
    added = 0
    with SweBlock(nx, ny, dx, dy, solver=solver) as blk:
        blk.set_boundaries(bnd)
        blk.upload(h, hu, hv, b)
        for k in range(0, steps, every):
            added += mlp.log_block(blk, keep=keep, seed=seed + k, hard_only=hard_only)
            blk.run(every, want_dts=False)
    return added


def log_workloads(mlp, keep: float = 1.0, seed: int = 1, scale: int = 1, every: int = 2):
    """Log the hard Newton edges of the three workloads (sizes divided by
    `scale` for quick runs) every `every` steps.  Returns per-workload
    sample counts."""
    n2 = 2048 // scale
    plan = [("dam_break_256", scenarios.dam_break_c1(256 // scale if scale > 1 else 256), FWAVE, 200),
            ("islands_2048", scenarios.dam_break_islands(n=n2, seed=2048), AUGRIE, 500),
            ("lts_tiles_2048", scenarios.lts_tiling_as_block(seed=28, n=n2), AUGRIE, 300)]
    counts = {}
    for name, state, solver, steps in plan:
        counts[name] = _log_run(mlp, state, solver, steps, every, keep, seed)
    return counts


def train(mlp, steps: int = 6000, batch: int = 8192, lr: float = 2e-3, seed: int = 7):
    mlp.sort_samples()  # the logging order varies between runs; training by index should not
    return mlp.train(steps, batch=batch, lr=lr, seed=seed)


def load_default(mlp=None, path: Path = DEFAULT_WEIGHTS):
    """A network with the committed trained weights (float32 .npy, no pickle)."""
    p = np.load(path, allow_pickle=False)
    m = mlp if mlp is not None else MiddleStateMLP(capacity=0)
    m.params = p
    return m


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", type=Path, default=DEFAULT_WEIGHTS)
    ap.add_argument("--capacity", type=int, default=60_000_000)
    ap.add_argument("--every", type=int, default=2)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--scale", type=int, default=1)
    ap.add_argument("--report", type=Path, default=None)
    args = ap.parse_args()
    t0 = time.perf_counter()
    mlp = MiddleStateMLP(capacity=args.capacity, seed=11)
    counts = log_workloads(mlp, scale=args.scale, every=args.every)
    t1 = time.perf_counter()
    print(f"logged {mlp.count} samples {counts} in {t1 - t0:.1f} s", flush=True)
    loss = train(mlp, steps=args.steps, batch=args.batch, lr=args.lr)
    t2 = time.perf_counter()
    print(f"trained {args.steps} steps in {t2 - t1:.1f} s: loss {loss[0]:.3e} -> {np.mean(loss[-50:]):.3e}",
          flush=True)
    args.out.parent.mkdir(parents=True, exist_ok=True)
    np.save(args.out, mlp.params)
    if args.report:
        args.report.write_text(json.dumps({
            "samples": int(mlp.count), "per_workload": counts, "steps": args.steps, "batch": args.batch,
            "lr": args.lr, "loss_first": float(loss[0]), "loss_last50": float(np.mean(loss[-50:])),
            "loss_curve_every_100": [float(v) for v in loss[::100]],
            "log_s": round(t1 - t0, 2), "train_s": round(t2 - t1, 2)}, indent=1))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
