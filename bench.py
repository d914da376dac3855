"""Throughput bench of the MI355X SWE step.

A "step" is one global time step of the plain flavour-U path (boundary
conditions, every edge solved, CFL dt reduced on the device, update with dry
clamp) on the 2048^2 AugRie workload: radial dam break plus 16 seeded
Gaussian bumps of which 3 emerge as dry islands, FP32, outflow boundaries.
The state is resident in HBM before timing starts.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps 500] [--warmup 20]

cpu_baseline: the CPU oracle (the same scheme restated in C++, OpenMP over
the host's cores) timed on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "swe-ppm_amd"))

METRIC = "MCell-updates/s, 2048² AugRie FP32 + % HBM peak; batched/surrogate vs plain"
HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL = 28.0  # read h, hu, hv, b + write h, hu, hv (FP32)


def cpu_threads() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except Exception:
        return os.cpu_count() or 1


def timed_run(swe, blk, nsteps, per_kernel):
    tt = C.c_float(0.0)
    total = C.c_float(0.0)
    kms = (C.c_float * 3)()
    kl = (C.c_int * 3)()
    rc = swe.lib.swe_hip_run_timed(blk._h, nsteps, C.byref(tt), C.byref(total), kms if per_kernel else None,
                                   kl if per_kernel else None)
    if rc != 0:
        raise RuntimeError(swe.lib.swe_hip_last_error().decode())
    return tt.value, total.value, list(kms), list(kl)


def cpu_baseline(h, hu, hv, b, dx, dy, bnd, solver, budget_s: float):
    """Oracle steps on all host threads until the time budget is used."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle_py

    threads = min(cpu_threads(), 64)
    nx, ny = h.shape[0] - 2, h.shape[1] - 2
    steps, elapsed = 0, 0.0
    state = (h, hu, hv)
    # first step alone to size the sample
    t0 = time.perf_counter()
    state = oracle_py.run_u(*state, b, dx, dy, bnd, solver, 1, nthreads=threads)[:3]
    elapsed += time.perf_counter() - t0
    steps += 1
    per = elapsed
    more = max(0, int(budget_s / max(per, 1e-6)) - 1)
    more = min(more, 200)
    if more:
        t0 = time.perf_counter()
        oracle_py.run_u(*state, b, dx, dy, bnd, solver, more, nthreads=threads)
        elapsed += time.perf_counter() - t0
        steps += more
    mcells = nx * ny * steps / elapsed / 1e6
    return {"value": mcells, "unit": "MCell-updates/s", "cores": threads, "kind": "port",
            "sample": f"{steps} flavour-U AugRie steps of the same {nx}x{ny} workload from t=0 "
                      f"({elapsed:.1f} s, OpenMP {threads} threads, C++ oracle -O2)"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--seed", type=int, default=2048)
    ap.add_argument("--strategy", type=int, default=-1)
    ap.add_argument("--xseg", type=int, default=-1)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--math", type=int, default=0, help="0 IEEE (bit-compatible), 1 fast")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch

        n = args.calibrate_copy * (1 << 20) // 4
        a = torch.ones(n, device="cuda", dtype=torch.float32)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes_read": 4 * n, "calibrate_copy_bytes_written": 4 * n,
                          "launches": 3}))
        return 0
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design in DESIGN.md, not a measured path)", file=sys.stderr)
        return 2

    import swe_ppm_amd as swe
    from swe_ppm_amd import scenarios

    n = args.n
    h, hu, hv, b, dx, dy, bnd = scenarios.dam_break_islands(n=n, seed=args.seed)
    blk = swe.SweBlock(n, n, dx, dy, solver=swe.AUGRIE)
    blk.set_boundaries(bnd)
    blk.set_tuning(args.strategy, args.xseg)
    blk.set_math(args.math)
    strategy, xseg = blk.get_tuning()

    # warmup from the initial state, then the timed K steps from the initial state again
    blk.upload(h, hu, hv, b)
    if args.warmup:
        blk.run(args.warmup, want_dts=False)
    blk.upload(h, hu, hv, b)
    blk.synchronize()
    t_wall0 = time.perf_counter()
    t_sim, total_ms, _, _ = timed_run(swe, blk, args.steps, per_kernel=False)
    wall_s = time.perf_counter() - t_wall0
    ms_per_step = total_ms / args.steps
    cells = n * n
    mcells = cells * args.steps / (total_ms * 1e-3) / 1e6

    # per-kernel attribution (separate run with events between launches)
    blk.upload(h, hu, hv, b)
    _, total2, kms, kl = timed_run(swe, blk, args.steps, per_kernel=True)
    pass1_ms = kms[1] / max(kl[1], 1)
    pass2_ms = kms[2] / max(kl[2], 1)
    # kernel time per step: every launch of the loop (boundary, passes) / steps
    step_ms_k = (kms[0] + kms[1] + kms[2]) / args.steps
    achieved = BYTES_PER_CELL * cells / (step_ms_k * 1e-3) / 1e9
    gh, ghu, ghv = blk.download()
    blk.close()
    if not np.all(np.isfinite(gh)):
        print("bench.py: non-finite state after the timed run", file=sys.stderr)
        return 3

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(h, hu, hv, b, dx, dy, bnd, swe.AUGRIE, args.cpu_budget)

    names = {0: ("sweep_u_kernel<AUGRIE,MODE_MAX>", "sweep_u_kernel<AUGRIE,MODE_UPDATE>"),
             1: ("sweep_u_kernel<AUGRIE,MODE_STORE>", "apply_u_kernel"),
             2: ("fused_u_kernel<AUGRIE> (first launch: sweep_u_kernel<AUGRIE,MODE_STORE>)",
                 "apply_u_kernel (once, after the last step)")}[strategy]
    line = {
        "metric": METRIC,
        "value": round(mcells, 3),
        "unit": "MCell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{n}x{n} flavour-U AugRie FP32 global-time-stepping step: radial dam break "
                        f"(15/10 m) + 16 seeded Gaussian bumps with 3 dry islands (seed {args.seed}), "
                        f"outflow boundaries, {args.steps} steps from t=0",
            "cells": cells,
            "strategy": ["recompute", "store", "fused"][strategy],
            "xseg": xseg,
            "simulated_s": round(t_sim, 6),
            "wall_s_incl_sync": round(wall_s, 4),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": None,
            "kernel": f"{names[0]}: {kl[1]} launches, avg {pass1_ms * 1e3:.2f} us; {names[1]}: {kl[2]} launches, "
                      f"avg {pass2_ms * 1e3:.2f} us; boundary {kms[0] * 1e3:.1f} us total; "
                      f"{step_ms_k * 1e3:.2f} us of kernel time per step",
            "algorithmic_bytes_per_cell_update": BYTES_PER_CELL,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
