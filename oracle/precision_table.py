"""FP32 against FP64 for the five benchmark configurations, on the CPU oracle.

TEST INFRASTRUCTURE ONLY (it runs the oracle, never the product).  For each
configuration the same FP32 initial state is advanced by the oracle in FP32
and in FP64 (FP64 arithmetic throughout, the state widened once at the
start), and the relative L-infinity difference of h, hu, hv after the run is
reported (max |a - b| over the interior / max |b|), with the FP32 run's
relative change of total mass.  The GPU's plain FP32 path is bit-identical to
the FP32 oracle on these configurations (tests/test_gpu_*.py), so the FP32
column is also the GPU's error against FP64.

    python3 oracle/precision_table.py [--threads N] [--quick] > table.md
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))
sys.path.insert(0, str(HERE.parent / "swe-ppm_amd"))

import oracle_py  # noqa: E402
from swe_ppm_amd import scenarios  # noqa: E402

FWAVE, AUGRIE = 1, 2


def rel_linf(a, b):
    a = np.asarray(a, np.float64)[1:-1, 1:-1]
    b = np.asarray(b, np.float64)[1:-1, 1:-1]
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / den) if den > 0 else float(np.max(np.abs(a - b)))


def single(state, solver, steps, threads):
    h, hu, hv, b, dx, dy, bnd = state
    r32 = oracle_py.run_u(h, hu, hv, b, dx, dy, bnd, solver, steps, nthreads=threads)
    r64 = oracle_py.run_u(h, hu, hv, b, dx, dy, bnd, solver, steps, nthreads=threads, dtype=np.float64)
    m0 = float(np.sum(h[1:-1, 1:-1], dtype=np.float64))
    m1 = float(np.sum(r32[0][1:-1, 1:-1], dtype=np.float64))
    return [rel_linf(a, c) for a, c in zip(r32[:3], r64[:3])], (m1 - m0) / m0


def row(name, config, errs, drift, seconds):
    return {"config": name, "run": config, "rel_linf_h": errs[0], "rel_linf_hu": errs[1], "rel_linf_hv": errs[2],
            "mass_drift_fp32": drift, "cpu_s": round(seconds, 1)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--quick", action="store_true", help="smaller step counts (a check of the script)")
    ap.add_argument("--json", type=Path, default=None)
    args = ap.parse_args()
    th = args.threads
    q = args.quick
    rows = []

    t0 = time.time()
    n = 200 if not q else 20
    e, d = single(scenarios.dam_break_c1(256), FWAVE, n, th)
    rows.append(row("C1", f"256^2 dam break, f-wave, {n} steps", e, d, time.time() - t0))

    t0 = time.time()
    n = 500 if not q else 10
    e, d = single(scenarios.dam_break_islands(n=2048 if not q else 256, seed=2048), AUGRIE, n, th)
    rows.append(row("C2", f"{2048 if not q else 256}^2 islands, AugRie, {n} steps", e, d, time.time() - t0))

    t0 = time.time()
    blocks, dx, dy = scenarios.lts_tiling(seed=28, n=2048 if not q else 256)
    T, dur = 1.0, (10.0 if not q else 1.0)
    nmacro = int(round(dur / T))
    r32 = oracle_py.multiblock_run(blocks, dx, dy, AUGRIE, True, T, dur, nmacro, nthreads=th)
    r64 = oracle_py.multiblock_run(blocks, dx, dy, AUGRIE, True, T, dur, nmacro, nthreads=th, dtype=np.float64)
    errs = []
    for f in range(3):
        num = max(float(np.max(np.abs(np.float64(a[f][1:-1, 1:-1]) - c[f][1:-1, 1:-1]))) for a, c in zip(r32[0], r64[0]))
        den = max(float(np.max(np.abs(c[f][1:-1, 1:-1]))) for c in r64[0])
        errs.append(num / den if den > 0 else num)
    m0 = sum(float(np.sum(bl["h"][1:-1, 1:-1], dtype=np.float64)) for bl in blocks)
    m1 = sum(float(np.sum(a[0][1:-1, 1:-1], dtype=np.float64)) for a in r32[0])
    rows.append(row("C3", f"28-block LTS tiling of {2048 if not q else 256}^2, AugRie, macro step {T} s, "
                          f"{dur} s simulated", errs, (m1 - m0) / m0, time.time() - t0))

    t0 = time.time()
    count, n = (512, 300) if not q else (8, 10)
    inst = scenarios.ensemble_instances(count=count, seed=512)
    worst = [0.0, 0.0, 0.0]
    drift = 0.0
    for i in inst:
        e, d = single((i["h"], i["hu"], i["hv"], i["b"], i["dx"], i["dy"], i["bnd"]), AUGRIE, n, th)
        worst = [max(a, b) for a, b in zip(worst, e)]
        drift = max(drift, abs(d), key=abs)
    rows.append(row("C4", f"{count}-instance ensemble, AugRie, {n} steps each (worst instance)", worst, drift,
                    time.time() - t0))

    t0 = time.time()
    n = 1000 if not q else 10
    size = 4096 if not q else 256
    e, d = single(scenarios.trench_hump(n=size, seed=4096), AUGRIE, n, th)
    rows.append(row("C5", f"{size}^2 trench, AugRie (exact middle state), {n} steps", e, d, time.time() - t0))

    print("| config | run | rel L-inf h | rel L-inf hu | rel L-inf hv | FP32 mass change (outflow included) | CPU s |")
    print("|---|---|---|---|---|---|---|")
    for r in rows:
        print(f"| {r['config']} | {r['run']} | {r['rel_linf_h']:.2e} | {r['rel_linf_hu']:.2e} | "
              f"{r['rel_linf_hv']:.2e} | {r['mass_drift_fp32']:.1e} | {r['cpu_s']} |")
    if args.json:
        args.json.write_text(json.dumps(rows, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
