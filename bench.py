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
CLOCK_GHZ = 2.4  # MI355X peak engine clock
# one wave64 VALU instruction per 4 cycles per SIMD, 4 SIMDs per CU
VALU_WAVE_INSTR_PER_CU_CYCLE = 1.0
BYTES_PER_CELL = 28.0  # read h, hu, hv, b + write h, hu, hv (FP32)


def profile_field(key: str, workload_file: str = "fused_u_augrie_2048_islands.md"):
    """An integer `key: value` line of a committed rocprofv3 summary, or None."""
    import re

    path = ROOT / "profiles" / workload_file
    if not path.exists():
        return None
    m = re.search(r"^" + key + r":\s*(\d+)", path.read_text(), re.M)
    return int(m.group(1)) if m else None


def measured_traffic(xseg: int):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3
    counter summary (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, calibrated on a plain
    copy), or None if the summary is absent or profiled another strip width."""
    if profile_field("profiled_xseg") != xseg:
        return None
    return profile_field("traffic_bytes_per_launch")


def measured_valu(xseg: int):
    """VALU wave instructions per launch of the dominant kernel (rocprofv3
    SQ_INSTS_VALU) from the committed counter summary, or None if absent or
    profiled at another strip width."""
    if profile_field("profiled_xseg") != xseg:
        return None
    return profile_field("valu_wave_instructions_per_launch")


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


def timed_run_from(swe, blk, nsteps, t0):
    """swe_hip_run_timed continuing from simulated time t0."""
    tt = C.c_float(t0)
    total = C.c_float(0.0)
    rc = swe.lib.swe_hip_run_timed(blk._h, nsteps, C.byref(tt), C.byref(total), None, None)
    if rc != 0:
        raise RuntimeError(swe.lib.swe_hip_last_error().decode())
    return tt.value, total.value, None, None


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform

    return platform.processor() or "unknown"


def _tiles(state, bnd, cols, rows):
    """The state cut into cols x rows connected blocks (ghosts from the
    neighbours' interiors or the domain's ghosts) for the overdecomposed
    driver: block k = c * rows + r."""
    h, hu, hv, b = state
    nx, ny = h.shape[0] - 2, h.shape[1] - 2
    bw, bh = nx // cols, ny // rows
    out = []
    for c in range(cols):
        for r in range(rows):
            x0, x1 = c * bw, (c + 1) * bw if c < cols - 1 else nx
            y0, y1 = r * bh, (r + 1) * bh if r < rows - 1 else ny
            k = c * rows + r
            nbr = [k - rows if c > 0 else -1, k + rows if c < cols - 1 else -1, k - 1 if r > 0 else -1,
                   k + 1 if r < rows - 1 else -1]
            out.append(dict(nx=x1 - x0, ny=y1 - y0, nbr=nbr, bnd=[4 if m >= 0 else t for m, t in zip(nbr, bnd)],
                            **{key: np.ascontiguousarray(a[x0:x1 + 2, y0:y1 + 2]) for key, a in
                               zip(("h", "hu", "hv", "b"), (h, hu, hv, b))}))
    return out


def cpu_baseline(h, hu, hv, b, dx, dy, bnd, solver, budget_s: float, reps: int = 5):
    """The CPU oracle on the host's cores, as the baseline protocol asks: 5
    warm-up steps from t=0, then `reps` timed repetitions of the same K steps
    (monotonic clock), median; two modes -- one block with OpenMP over rows
    (swe_simple's structure) and overdecomposed, one connected block per
    thread with global dt (swe_mpi_overdecomp's) -- and two builds of the same
    source: the IEEE parity build (-O2, no contraction) and the reference's
    Release flags (-O3 -ffast-math, timing only).  K is sized so that the
    whole baseline takes about budget_s.  value: the fastest mode of the
    Release build."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle_py

    threads = min(cpu_threads(), 64)
    nx, ny = h.shape[0] - 2, h.shape[1] - 2
    rows = max(d for d in range(1, int(threads ** 0.5) + 1) if threads % d == 0)
    cols = threads // rows
    # 5 warm-up steps from t=0 (the state every repetition starts from)
    w = oracle_py.run_u(h, hu, hv, b, dx, dy, bnd, solver, 5, nthreads=threads)
    state = (w[0], w[1], w[2], b)
    tiles = _tiles(state, bnd, cols, rows)

    def single(release, k):
        t0 = time.perf_counter()
        oracle_py.run_u(*state, dx, dy, bnd, solver, k, nthreads=threads, release=release)
        return time.perf_counter() - t0

    def overdecomposed(release, k):
        t0 = time.perf_counter()
        oracle_py.multiblock_run(tiles, dx, dy, solver, False, 1.0, 1e30, k, nthreads=threads, release=release)
        return time.perf_counter() - t0

    modes = {}
    per_sample = budget_s / (4 * reps + 4)
    for mode, fn in (("single_block", single), ("overdecomposed", overdecomposed)):
        modes[mode] = {}
        for build, release in (("release", True), ("ieee", False)):
            # a 1-step call is mostly the call's fixed cost (thread start, array
            # copies): size K from it, time K steps once, then rescale K
            one = fn(release, 1)
            k = int(min(400, max(2, per_sample / max(one, 1e-6))))
            k = int(min(400, max(2, k * per_sample / max(fn(release, k), 1e-6))))
            times = sorted(fn(release, k) for _ in range(reps))
            med = times[len(times) // 2]
            modes[mode][build] = {"value": round(nx * ny * k / med / 1e6, 2), "steps": k,
                                  "median_s": round(med, 4), "spread": round(times[-1] / times[0] - 1.0, 4)}
    best_mode = max(modes, key=lambda m: modes[m]["release"]["value"])
    best = modes[best_mode]["release"]
    return {"value": best["value"], "unit": "MCell-updates/s", "cores": threads, "kind": "port",
            "sample": f"median of {reps} repetitions of {best['steps']} flavour-U AugRie steps of the same {nx}x{ny} "
                      f"workload after 5 warm-up steps from t=0; C++ oracle built with the reference's Release flags, "
                      f"OpenMP {threads} threads, {best_mode.replace('_', ' ')} mode (fastest of the two)",
            "cpu_model": cpu_model(),
            "modes": modes,
            "overdecomposed_blocks": f"{cols} x {rows}"}


def batched_lts(swe, scenarios, autotune: bool, seed: int = 28, n: int = 2048, T: float = 1.0,
                duration: float = 80.0):
    """28 uneven blocks tiling n x n, per-block depth U(2,40) m, AugRie, local
    time stepping with macro step T up to `duration` (one unit = one macro step)."""
    blocks, dx, dy = scenarios.lts_tiling(seed=seed, n=n)
    nmacro = int(round(duration / T))
    fields = [[bl[f] for bl in blocks] for f in ("h", "hu", "hv", "b")]
    with swe.SweBatch(blocks, solver=swe.AUGRIE, mode=swe.BATCH_LTS, dx=dx, dy=dy, max_local=T,
                      duration=duration) as bt:
        bt.upload(*fields)
        cand = None
        if autotune:
            # tuned on the state a tenth into the run: at t = 0 the blocks are
            # still water except at their interfaces, where narrow strips win,
            # while the run is spent on the loaded state
            bt.run(max(1, nmacro // 10))
            cand = bt.autotune(iterations=32)
        bt.upload(*fields)
        bt.run(1)  # warm-up macro step
        bt.upload(*fields)
        _, its, ms = bt.run(nmacro)
        upd, _ = bt.stats()
        xs, groups, chunk, graph, fused = bt.get_tuning()
        h = bt.download(packed=True)[0]
    if not np.all(np.isfinite(h)):
        raise RuntimeError("non-finite LTS state")
    cells = (bt.nx.astype(np.int64) * bt.ny).astype(np.int64)
    updates = int(np.sum(cells * upd))
    plain = plain_lts_domain(swe, scenarios, duration, seed=seed, n=n)
    return {
        "workload": f"{len(blocks)} blocks tiling {n}x{n} (sizes {int(bt.nx.min())}..{int(bt.nx.max())} x "
                    f"{int(bt.ny.min())}..{int(bt.ny.max())}), still water U(2,40) m per block (seed {seed}), AugRie "
                    f"FP32, LTS macro step {T} s, {duration} s simulated ({nmacro} macro steps)",
        "value": round(updates / (ms * 1e-3) / 1e6, 3), "unit": "MCell-updates/s",
        "elapsed_ms": round(ms, 3), "cell_updates": updates, "driver_iterations": int(np.sum(its)),
        "sim_s_per_wall_s": round(duration / (ms * 1e-3), 3),
        "plain_global_dt_same_domain": plain,
        "sim_speed_vs_plain": round((duration / (ms * 1e-3)) / plain["sim_s_per_wall_s"], 4),
        "hbm_frac_algorithmic": round(BYTES_PER_CELL * updates / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
        "tuning": {"xs": xs, "groups": groups, "chunk": chunk, "graph": graph, "fused": fused,
                   "autotune_ms_per_iteration": None if cand is None else [round(float(c), 4) for c in cand]},
    }


def plain_lts_domain(swe, scenarios, duration: float, seed: int = 28, n: int = 2048):
    """The LTS workload's domain as one plain block with the global CFL dt:
    simulated seconds per wall second, for comparison with local time stepping."""
    h, hu, hv, b, dx, dy, bnd = scenarios.lts_tiling_as_block(seed=seed, n=n)
    with swe.SweBlock(n, n, dx, dy, solver=swe.AUGRIE) as blk:
        blk.set_boundaries(bnd)
        blk.upload(h, hu, hv, b)
        blk.run(10, want_dts=False)
        blk.upload(h, hu, hv, b)
        # advance in chunks of 200 steps until the simulated time is reached
        t, total_ms, steps = 0.0, 0.0, 0
        while t < duration:
            t, ms, _, _ = timed_run_from(swe, blk, 200, t)
            total_ms += ms
            steps += 200
    return {"sim_s_per_wall_s": round(t / (total_ms * 1e-3), 3), "steps": steps, "simulated_s": round(t, 3),
            "value": round(n * n * steps / (total_ms * 1e-3) / 1e6, 3), "unit": "MCell-updates/s",
            "cell_updates": n * n * steps}


def plain_ensemble(swe, inst, steps: int):
    """The ensemble's instances one after another through the plain block path."""
    total_ms, updates = 0.0, 0
    for i in inst:
        with swe.SweBlock(i["nx"], i["ny"], i["dx"], i["dy"], solver=swe.AUGRIE) as blk:
            blk.set_boundaries(i["bnd"])
            blk.upload(i["h"], i["hu"], i["hv"], i["b"])
            _, ms, _, _ = timed_run(swe, blk, steps, per_kernel=False)
            total_ms += ms
            updates += i["nx"] * i["ny"] * steps
    return {"value": round(updates / (total_ms * 1e-3) / 1e6, 3), "unit": "MCell-updates/s", "steps": steps,
            "elapsed_ms": round(total_ms, 3)}


def batched_ensemble(swe, scenarios, autotune: bool, count: int = 512, seed: int = 512, steps: int = 300):
    """512 independent dam-break instances of uneven size (64..384)^2, each with its own dt."""
    inst = scenarios.ensemble_instances(count=count, seed=seed)
    fields = [[i[f] for i in inst] for f in ("h", "hu", "hv", "b")]
    with swe.SweBatch(inst, solver=swe.AUGRIE, mode=swe.BATCH_ENSEMBLE) as bt:
        bt.upload(*fields)
        cand = bt.autotune(iterations=16) if autotune else None
        bt.upload(*fields)
        bt.run(5)  # warm-up
        bt.upload(*fields)
        _, _, ms = bt.run(steps)
        upd, t = bt.stats()
        xs, groups, chunk, graph, fused = bt.get_tuning()
        h = bt.download(packed=True)[0]
    if not np.all(np.isfinite(h)):
        raise RuntimeError("non-finite ensemble state")
    updates = int(bt.cells) * steps
    plain = plain_ensemble(swe, inst, steps=min(steps, 50))
    return {
        "workload": f"{count} independent dam-break instances, NX,NY ~ U{{64..384}} ({bt.cells} cells, seed {seed}), "
                    f"AugRie FP32, per-instance CFL dt, {steps} steps each",
        "value": round(updates / (ms * 1e-3) / 1e6, 3), "unit": "MCell-updates/s",
        "elapsed_ms": round(ms, 3), "cell_updates": updates,
        "instance_time_range_s": [round(float(t.min()), 4), round(float(t.max()), 4)],
        "plain_one_instance_after_another": plain,
        "vs_plain_same_instances": round((updates / (ms * 1e-3) / 1e6) / plain["value"], 4),
        "hbm_frac_algorithmic": round(BYTES_PER_CELL * updates / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
        "tuning": {"xs": xs, "groups": groups, "chunk": chunk, "graph": graph, "fused": fused,
                   "autotune_ms_per_iteration": None if cand is None else [round(float(c), 4) for c in cand]},
    }


def surrogate_compare(swe, state, weights, steps):
    """Plain AugRie against AugRie with the learned middle state on one
    workload, `steps` global steps each: MCell-updates/s of both."""
    from swe_ppm_amd import surrogate

    h, hu, hv, b, dx, dy, bnd = state
    n = h.shape[0] - 2
    res = {}
    mlp = surrogate.load_default(path=weights)
    for name, net in (("plain", None), ("surrogate", mlp)):
        with swe.SweBlock(n, n, dx, dy, solver=swe.AUGRIE) as blk:
            blk.set_boundaries(bnd)
            blk.set_surrogate(net)
            blk.upload(h, hu, hv, b)
            blk.run(10, want_dts=False)  # warm-up
            blk.upload(h, hu, hv, b)
            t_sim, total_ms, _, _ = timed_run(swe, blk, steps, per_kernel=False)
            gh = blk.download()[0]
        if not np.all(np.isfinite(gh)):
            raise RuntimeError(f"non-finite trench state ({name})")
        res[name] = {"value": round(n * n * steps / (total_ms * 1e-3) / 1e6, 3), "ms_per_step": round(total_ms / steps, 6),
                     "simulated_s": round(t_sim, 4)}
    mlp.close()
    return {"unit": "MCell-updates/s", "plain": res["plain"], "surrogate": res["surrogate"],
            "surrogate_vs_plain": round(res["surrogate"]["value"] / res["plain"]["value"], 4),
            "weights": str(Path(weights).name)}


def solver_lines(swe, state, steps: int = 200, autotune: bool = True, names=("hlle", "fwave", "hybrid", "augrie")):
    """The headline workload with each solver of the plain path (IEEE):
    MCell-updates/s and the algorithmic HBM fraction, to show where the same
    step stands when the Riemann solver costs less.  Each block's strip width
    (and, for HLLE and f-wave, the strategy) is autotuned on the workload's
    first 20 steps (the state is restored before the timed run)."""
    h, hu, hv, b, dx, dy, bnd = state
    n = h.shape[0] - 2
    out = {}
    for name in names:
        sid = swe.SOLVER_NAMES[name]
        with swe.SweBlock(n, n, dx, dy, solver=sid) as blk:
            blk.set_boundaries(bnd)
            blk.upload(h, hu, hv, b)
            default = blk.get_tuning()[1]
            tune = blk.autotune(20) if autotune else None
            blk.run(10, want_dts=False)  # warm-up
            blk.upload(h, hu, hv, b)
            _, total_ms, _, _ = timed_run(swe, blk, steps, per_kernel=False)
            strategy, xseg = blk.get_tuning()
        v = n * n * steps / (total_ms * 1e-3) / 1e6
        out[name] = {"value": round(v, 1), "ms_per_step": round(total_ms / steps, 5),
                     "hbm_frac_algorithmic": round(BYTES_PER_CELL * v * 1e6 / 1e9 / HBM_PEAK_GBS, 4),
                     "strategy": ["recompute", "store", "fused"][strategy], "xseg": xseg, "default_xseg": default,
                     "autotune_ms_per_step": None if tune is None else {k: round(t, 5) for k, t in tune.items()}}
    return {"workload": f"{n}x{n} islands dam break, {steps} steps per solver, strip width "
                        f"{'and strategy autotuned' if autotune else 'default'}",
            "unit": "MCell-updates/s", "solvers": out}


def seeded_init_line(swe, scenarios, n: int, seed: int, host_state, dx: float, bnd, host_gen_s: float):
    """The headline workload sampled on the device (swe_hip_init_seeded)
    against the host generator: time and bit identity."""
    spec = scenarios.dam_break_islands_spec(seed)
    with swe.SweBlock(n, n, dx, dx, solver=swe.AUGRIE) as blk:
        blk.set_boundaries(bnd)
        blk.init_seeded(spec)  # first call: module load
        blk.synchronize()
        t0 = time.perf_counter()
        blk.init_seeded(spec)
        blk.synchronize()
        dev_s = time.perf_counter() - t0
        dev = blk.download(with_b=True)
    same = all(np.array_equal(np.asarray(a).view(np.uint32), np.asarray(e, np.float32).view(np.uint32))
               for a, e in zip(dev, host_state))
    return {"workload": f"the headline workload's initial state ({n}x{n} cells, ghosts and boundary conditions "
                        f"included), sampled on the device from its parameter set and on the host by numpy",
            "device_ms": round(dev_s * 1e3, 3), "host_s": round(host_gen_s, 3), "bit_identical": bool(same)}


def every_tile_line(swe, state, steps: int, xseg: int):
    """The headline workload with quiet-tile skipping off: every strip tile is
    computed every step (same results, bit for bit)."""
    h, hu, hv, b, dx, dy, bnd = state
    n = h.shape[0] - 2
    with swe.SweBlock(n, n, dx, dy, solver=swe.AUGRIE) as blk:
        blk.set_boundaries(bnd)
        blk.set_tuning(-1, xseg)
        blk.set_quiet_skip(False)
        blk.upload(h, hu, hv, b)
        blk.run(10, want_dts=False)  # warm-up
        blk.upload(h, hu, hv, b)
        _, total_ms, kms, kl = timed_run(swe, blk, steps, per_kernel=True)
    v = n * n * steps / (total_ms * 1e-3) / 1e6
    return {"workload": "the headline workload with every strip tile computed every step (quiet-tile skipping off)",
            "value": round(v, 3), "unit": "MCell-updates/s", "ms_per_step": round(total_ms / steps, 6),
            "fused_kernel_avg_us": round(kms[1] / max(kl[1], 1) * 1e3, 3), "xseg": xseg}


def per_call_line(state, steps: int, warmup: int):
    """The headline workload driven through the drop-in block one call at a
    time, as the reference's single-block driver does (swe_simple.cpp:225-239:
    setGhostLayer, computeNumericalFluxes, getMaxTimestep, updateUnknowns per
    step): swe_simple_hip from the same initial state, its own wall clock
    around the loop (host round trip for dt included), with and without the
    look-ahead launches."""
    import re
    import subprocess
    import tempfile

    h, hu, hv, b, dx, dy, bnd = state
    n = h.shape[0] - 2
    exe = ROOT / "swe-ppm_amd" / "host" / "build" / "swe_simple_hip"
    if not exe.exists():
        return {"skipped": f"{exe} not built"}
    out = {"workload": f"the headline workload ({n}x{n} islands, AugRie, {steps} steps from t=0) through "
                       "SWE_DimensionalSplittingHip's setGhostLayer / computeNumericalFluxes / getMaxTimestep / "
                       "updateUnknowns, one call each per step (swe_simple_hip --input)",
           "unit": "MCell-updates/s"}
    with tempfile.TemporaryDirectory() as tmp:
        snap = Path(tmp) / "islands.bin"
        with open(snap, "wb") as f:
            np.array([1, n, n], np.int32).tofile(f)
            for a in (h, hu, hv, b):
                np.ascontiguousarray(a, np.float32).tofile(f)
        for key, extra in (("lookahead", []), ("no_lookahead", ["--lookahead", "0"])):
            cmd = [str(exe), "--input", str(snap), "-t", "1e9", "-n", "1", "--solver", "augrie", "--max-steps",
                   str(steps), "--warmup", str(warmup)] + extra
            p = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
            m = re.search(r"iterations (\d+) simulated ([\d.]+) s wall ([\d.]+) s\s+([\d.]+) MCell", p.stdout)
            if p.returncode != 0 or not m:
                out[key] = {"error": (p.stderr or p.stdout)[-300:]}
                continue
            it, wall = int(m.group(1)), float(m.group(3))
            out[key] = {"value": round(float(m.group(4)), 3), "ms_per_step": round(wall / it * 1e3, 6),
                        "simulated_s": float(m.group(2))}
    if "value" in out.get("lookahead", {}):
        out["value"] = out["lookahead"]["value"]
        out["ms_per_step"] = out["lookahead"]["ms_per_step"]
    return out


def disturbed_line(swe, scenarios, n: int, seed: int, steps: int, noise: float = 0.01):
    """The headline workload with seeded surface noise of +-noise m on every
    cell: no edge starts still, so every wet edge runs the full augmented
    decomposition (the still-water shortcut saves nothing here)."""
    h, hu, hv, b, dx, dy, bnd = scenarios.dam_break_islands(n=n, seed=seed, surface_noise=noise)
    with swe.SweBlock(n, n, dx, dy, solver=swe.AUGRIE) as blk:
        blk.set_boundaries(bnd)
        blk.upload(h, hu, hv, b)
        blk.run(10, want_dts=False)  # warm-up
        blk.upload(h, hu, hv, b)
        _, total_ms, kms, kl = timed_run(swe, blk, steps, per_kernel=True)
        xseg = blk.get_tuning()[1]
    v = n * n * steps / (total_ms * 1e-3) / 1e6
    fused_us = kms[1] / max(kl[1], 1) * 1e3
    return {"workload": f"{n}x{n} islands dam break (seed {seed}) with uniform surface noise +-{noise} m on every "
                        f"cell (seed {seed + 1}), AugRie FP32, {steps} steps from t=0",
            "value": round(v, 3), "unit": "MCell-updates/s", "ms_per_step": round(total_ms / steps, 6),
            "fused_kernel_avg_us": round(fused_us, 3), "xseg": xseg,
            "hbm_frac_algorithmic": round(BYTES_PER_CELL * n * n / (fused_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 5)}


def batched_surrogate_ensemble(swe, scenarios, weights, count: int = 512, seed: int = 512, steps: int = 100,
                               autotune: bool = True):
    """The 512-instance ensemble through the batched engine, exact and with
    the learned middle state attached: the same instances, steps and tuning
    procedure, so the ratio compares like with like."""
    from swe_ppm_amd import surrogate

    inst = scenarios.ensemble_instances(count=count, seed=seed)
    fields = [[i[f] for i in inst] for f in ("h", "hu", "hv", "b")]
    updates = int(sum(i["nx"] * i["ny"] for i in inst)) * steps
    out = {"workload": f"the {count}-instance ensemble (seed {seed}), AugRie FP32, {steps} steps each, exact and "
                       "with the learned middle state", "unit": "MCell-updates/s", "weights": str(Path(weights).name)}
    for key in ("exact", "learned"):
        mlp = surrogate.load_default(path=weights) if key == "learned" else None
        with swe.SweBatch(inst, solver=swe.AUGRIE, mode=swe.BATCH_ENSEMBLE) as bt:
            if mlp is not None:
                bt.set_surrogate(mlp)
            bt.upload(*fields)
            if autotune:
                bt.autotune(iterations=16)
            bt.upload(*fields)
            bt.run(5)  # warm-up
            bt.upload(*fields)
            _, _, ms = bt.run(steps)
            h = bt.download(packed=True)[0]
            tuning = bt.get_tuning()
        if mlp is not None:
            mlp.close()
        if not np.all(np.isfinite(h)):
            raise RuntimeError(f"non-finite ensemble state ({key})")
        out[key] = {"value": round(updates / (ms * 1e-3) / 1e6, 3), "elapsed_ms": round(ms, 3),
                    "tuning": dict(zip(("xs", "groups", "chunk", "graph", "fused"), tuning))}
    out["learned_vs_exact"] = round(out["learned"]["value"] / out["exact"]["value"], 4)
    return out


def surrogate_trench(swe, scenarios, weights, n: int = 4096, steps: int = 1000, seed: int = 4096):
    """The n x n trench (depth 10..4000 m, 2 m hump), plain and learned."""
    out = {"workload": f"{n}x{n} trench (depth 10 m far field, 4000 m on the axis, Gaussian profile) with a 2 m "
                       f"Gaussian sea-surface hump (seed {seed}), AugRie FP32, outflow, {steps} steps"}
    out.update(surrogate_compare(swe, scenarios.trench_hump(n=n, seed=seed), weights, steps))
    return out


def surrogate_islands(swe, scenarios, weights, n: int = 2048, steps: int = 500, seed: int = 2048):
    """The plain line's dam break over islands, plain and learned."""
    out = {"workload": f"{n}x{n} dam break over 16 bumps with 3 dry islands (seed {seed}), AugRie FP32, {steps} steps"}
    out.update(surrogate_compare(swe, scenarios.dam_break_islands(n=n, seed=seed), weights, steps))
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--seed", type=int, default=2048)
    ap.add_argument("--strategy", type=int, default=-1)
    ap.add_argument("--xseg", type=int, default=-1)
    ap.add_argument("--cpu-budget", type=float, default=24.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--math", type=int, default=0, help="0 IEEE (bit-compatible), 1 fast")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched LTS / ensemble lines")
    ap.add_argument("--batched-only", choices=["lts", "ensemble", "surrogate"], default=None,
                    help="only run one batched workload (for profiling) and print its line")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--lts-duration", type=float, default=80.0, help="simulated seconds of the LTS workload")
    ap.add_argument("--ens-steps", type=int, default=300, help="steps of the ensemble workload")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the learned-middle-state line")
    ap.add_argument("--no-solvers", action="store_true", help="skip the per-solver lines")
    ap.add_argument("--surrogate-weights", type=Path, default=None,
                    help="trained network (.npy); default: the packaged weights")
    ap.add_argument("--trench-n", type=int, default=4096)
    ap.add_argument("--trench-steps", type=int, default=1000)
    ap.add_argument("--only-headline", action="store_true",
                    help="print the main line only (no sub-lines): the command the profiles are taken with")
    ap.add_argument("--solver", choices=["augrie", "hlle", "fwave", "hybrid"], default="augrie",
                    help="solver of the main line (profiling of the light solvers; the metric is AugRie's)")
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
        print(f"bench.py: --gpus {args.gpus} is not supported: the step runs and is measured on one GPU "
              "(sharding blocks across GPUs is not implemented)", file=sys.stderr)
        return 2

    import swe_ppm_amd as swe
    from swe_ppm_amd import scenarios

    weights = args.surrogate_weights
    if weights is None:
        from swe_ppm_amd import surrogate as _sur

        weights = _sur.DEFAULT_WEIGHTS
    if args.batched_only == "surrogate":
        trench = surrogate_trench(swe, scenarios, weights, n=args.trench_n, steps=args.trench_steps)
        islands = surrogate_islands(swe, scenarios, weights, steps=args.steps)
        ens = batched_surrogate_ensemble(swe, scenarios, weights, autotune=not args.no_autotune)
        print(json.dumps({"trench": trench, "islands": islands, "batched_ensemble_512": ens}))
        return 0
    if args.batched_only:
        if args.batched_only == "lts":
            out = batched_lts(swe, scenarios, not args.no_autotune, duration=args.lts_duration)
        else:
            out = batched_ensemble(swe, scenarios, not args.no_autotune, steps=args.ens_steps)
        print(json.dumps(out))
        return 0

    n = args.n
    t_gen = time.perf_counter()
    h, hu, hv, b, dx, dy, bnd = scenarios.dam_break_islands(n=n, seed=args.seed)
    host_gen_s = time.perf_counter() - t_gen
    solver = {"augrie": swe.AUGRIE, "hlle": swe.HLLE, "fwave": swe.FWAVE, "hybrid": swe.HYBRID}[args.solver]
    blk = swe.SweBlock(n, n, dx, dy, solver=solver)
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

    # per-kernel attribution (separate run with events between launches):
    # class 0 boundary + quiet-tile plans, 1 the fused (or first) pass, 2 the
    # final apply (or second pass)
    blk.upload(h, hu, hv, b)
    _, total2, kms, kl = timed_run(swe, blk, args.steps, per_kernel=True)
    pass1_ms = kms[1] / max(kl[1], 1)
    pass2_ms = kms[2] / max(kl[2], 1)
    plans = max(kl[0] - 1, 0)  # one boundary launch, the rest are plans
    # kernel time per step: every launch of the loop (boundary, plans, passes) / steps
    step_ms_k = (kms[0] + kms[1] + kms[2]) / args.steps
    # the dominant kernel's own average launch time (the fused launch for the fused strategy)
    achieved = BYTES_PER_CELL * cells / (pass1_ms * 1e-3) / 1e9
    gh, ghu, ghv = blk.download()
    blk.close()
    if not np.all(np.isfinite(gh)):
        print("bench.py: non-finite state after the timed run", file=sys.stderr)
        return 3

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(h, hu, hv, b, dx, dy, bnd, swe.AUGRIE, args.cpu_budget)

    profiled = strategy == 2 and n == 2048 and args.math == 0 and args.solver == "augrie"
    traffic = measured_traffic(xseg) if profiled else None
    valu = measured_valu(xseg) if profiled else None
    cus = 256
    try:
        import torch

        if torch.cuda.is_available():
            cus = torch.cuda.get_device_properties(0).multi_processor_count
    except Exception:
        pass
    valu_line = None
    if valu:
        rate = valu / (pass1_ms * 1e-3) / 1e9
        peak = cus * VALU_WAVE_INSTR_PER_CU_CYCLE * CLOCK_GHZ
        valu_line = {"achieved": round(rate, 1), "peak": round(peak, 1), "unit": "G VALU wave-instructions/s",
                     "frac": round(rate / peak, 4), "instructions_per_launch": valu,
                     "source": "rocprofv3 SQ_INSTS_VALU of fused_u_kernel on this workload "
                               "(profiles/fused_u_augrie_2048_islands.md); peak = CUs x 4 SIMDs x 1/4 "
                               "wave64 instruction per cycle x 2.4 GHz"}
    # what binds the dominant kernel, from the committed counters: VALU issue
    # or HBM traffic near their peaks, else latency (neither saturated)
    binding = None
    if valu_line and traffic:
        hbm_frac = traffic / (pass1_ms * 1e-3) / 1e9 / HBM_PEAK_GBS
        binding = "valu" if valu_line["frac"] >= 0.8 else "hbm" if hbm_frac >= 0.7 else "latency"
    sname = args.solver.upper() if args.solver != "augrie" else "AUGRIE"
    names = {0: (f"sweep_u_kernel<{sname},MODE_MAX>", f"sweep_u_kernel<{sname},MODE_UPDATE>"),
             1: (f"sweep_u_kernel<{sname},MODE_STORE>", "apply_u_kernel"),
             2: (f"fused_u_kernel<{sname}> (first launch: sweep_u_kernel<{sname},MODE_STORE>)",
                 "apply_u_kernel (once, after the last step)")}[strategy]
    solver_label = {"augrie": "AugRie", "hlle": "HLLE", "fwave": "f-wave", "hybrid": "Hybrid"}[args.solver]
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
            "workload": f"{n}x{n} flavour-U {solver_label} FP32 global-time-stepping step: radial dam break "
                        f"(15/10 m) + 16 seeded Gaussian bumps with 3 dry islands (seed {args.seed}), "
                        f"outflow boundaries, {args.steps} steps from t=0",
            "cells": cells,
            "strategy": ["recompute", "store", "fused"][strategy],
            "xseg": xseg,
            "simulated_s": round(t_sim, 6),
            "wall_s_incl_sync": round(wall_s, 4),
        },
        "roofline": {
            # algorithmic bytes over the dominant kernel's time, against HBM
            # peak; "binding" names what limits that kernel (valu, hbm or
            # latency, from the committed counters)
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": traffic,
            "traffic_source": "rocprofv3 FETCH_SIZE/WRITE_SIZE of fused_u_kernel on this workload "
                              "(profiles/fused_u_augrie_2048_islands.md)" if traffic else None,
            "kernel": f"{names[0]}: {kl[1]} launches, avg {pass1_ms * 1e3:.2f} us; {names[1]}: {kl[2]} launches, "
                      f"avg {pass2_ms * 1e3:.2f} us; boundary launch, {plans} quiet_plan_kernel launches and the gaps "
                      f"between launches: {kms[0] * 1e3:.1f} us; {step_ms_k * 1e3:.2f} us per step in this "
                      f"attribution run (fused launches timed on their own dispatches)",
            "algorithmic_bytes_per_cell_update": BYTES_PER_CELL,
            "binding": binding,
            "valu": valu_line,
        },
        "cpu_baseline": cpu,
    }
    if args.only_headline:
        print(json.dumps(line))
        return 0
    line["init"] = seeded_init_line(swe, scenarios, n, args.seed, (h, hu, hv, b), dx, bnd, host_gen_s)
    line["every_tile"] = every_tile_line(swe, (h, hu, hv, b, dx, dy, bnd), args.steps, xseg)
    line["every_tile"]["vs_headline"] = round(line["every_tile"]["value"] / mcells, 4)
    line["per_call"] = per_call_line((h, hu, hv, b, dx, dy, bnd), args.steps, args.warmup)
    if "value" in line["per_call"]:
        line["per_call"]["vs_headline"] = round(line["per_call"]["value"] / mcells, 4)
    line["disturbed"] = disturbed_line(swe, scenarios, n, args.seed, args.steps)
    line["disturbed"]["vs_headline"] = round(line["disturbed"]["value"] / mcells, 4)
    if not args.no_solvers:
        line["solvers"] = solver_lines(swe, (h, hu, hv, b, dx, dy, bnd), autotune=not args.no_autotune)
        # the light solvers where the state does not fit the GPU's caches
        big = scenarios.dam_break_islands(n=2 * n, seed=args.seed)
        line["solvers_hbm_resident"] = solver_lines(swe, big, autotune=not args.no_autotune, names=("hlle", "fwave"))
        del big
    if not args.no_batched:
        lts = batched_lts(swe, scenarios, not args.no_autotune, duration=args.lts_duration)
        ens = batched_ensemble(swe, scenarios, not args.no_autotune, steps=args.ens_steps)
        line["batched"] = {"lts_28_blocks": lts, "ensemble_512": ens}
    if not args.no_surrogate:
        if Path(weights).exists():
            line["surrogate"] = {
                "trench": surrogate_trench(swe, scenarios, weights, n=args.trench_n, steps=args.trench_steps),
                "islands": surrogate_islands(swe, scenarios, weights, n=n, steps=args.steps, seed=args.seed),
                "batched_ensemble_512": batched_surrogate_ensemble(swe, scenarios, weights,
                                                                   autotune=not args.no_autotune)}
        else:
            line["surrogate"] = {"skipped": f"no trained network at {weights}"}
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
