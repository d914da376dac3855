"""Parity at the full sizes of the measured workloads, through the C ABI.

- 2048^2 islands (AugRie): bit-identical to the oracle over 100 steps; over
  500 steps the interior mass is conserved (the dam-break front stays far
  from the outflow boundaries).
- 4096^2 trench with a sea-surface hump (AugRie): bit-identical over 40
  steps; with the learned middle state the fused and in-sweep forms agree
  bit for bit and stay within 1.5x the exact path's distance to FP64.  The
  same trench at rest (no hump) stays bit for bit at rest.
- the 28-block LTS tiling of 2048^2: one macro step bit-identical to the
  oracle's overdecomposed LTS loop.
- all 512 ensemble instances: 5 steps bit-identical to 512 oracle runs.
- the learned middle state (packaged network) on the 2048^2 islands and on
  128 ensemble instances through the batched engine: within 1.5x the exact
  path's distance to FP64, in the maximum and in the relative L2 norm.
"""
import os

import numpy as np
import pytest

from helpers import same_bits

from test_gpu_batch import _assert_blocks_equal

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


def _run(swe, state, steps, solver):
    """steps of the default (fused) loop; returns the state and the dts"""
    h, hu, hv, b, dx, dy, bnd = state
    n, m = h.shape[0] - 2, h.shape[1] - 2
    with swe.SweBlock(n, m, dx, dy, solver=solver) as blk:
        blk.set_boundaries(bnd)
        blk.upload(h, hu, hv, b)
        _, dts = blk.run(steps)
        out = blk.download()
    return out, dts


def test_islands_2048_bitwise_and_mass(swe, oracle):
    from swe_ppm_amd import scenarios

    st = scenarios.dam_break_islands(n=2048, seed=2048)
    h, hu, hv, b, dx, dy, bnd = st
    print("2048^2 islands, 100 steps GPU and oracle", flush=True)
    (gh, ghu, ghv), gdts = _run(swe, st, 100, swe.AUGRIE)
    oh, ohu, ohv, odts = oracle.run_u(h, hu, hv, b, dx, dy, bnd, swe.AUGRIE, 100, nthreads=THREADS)
    assert same_bits(np.asarray(gdts, np.float32), odts)
    for name, a, c in (("h", gh, oh), ("hu", ghu, ohu), ("hv", ghv, ohv)):
        assert same_bits(a[1:-1, 1:-1], c[1:-1, 1:-1]), name
    print("2048^2 islands, 500 steps GPU", flush=True)
    (gh5, _, _), _ = _run(swe, st, 500, swe.AUGRIE)
    m0 = scenarios.interior_mass(h, dx, dy)
    m1 = scenarios.interior_mass(gh5, dx, dy)
    assert abs(m1 - m0) / m0 < 1e-5
    assert np.any(gh5[1:-1, 1:-1] == 0.0), "the islands stay dry"


def test_trench_4096_bitwise(swe, oracle):
    from swe_ppm_amd import scenarios

    st = scenarios.trench_hump(n=4096, seed=4096)
    h, hu, hv, b, dx, dy, bnd = st
    print("4096^2 trench, 40 steps GPU and oracle", flush=True)
    (gh, ghu, ghv), gdts = _run(swe, st, 40, swe.AUGRIE)
    oh, ohu, ohv, odts = oracle.run_u(h, hu, hv, b, dx, dy, bnd, swe.AUGRIE, 40, nthreads=THREADS)
    assert same_bits(np.asarray(gdts, np.float32), odts)
    for name, a, c in (("h", gh, oh), ("hu", ghu, ohu), ("hv", ghv, ohv)):
        assert same_bits(a[1:-1, 1:-1], c[1:-1, 1:-1]), name


def test_trench_4096_lake_at_rest(swe):
    """The trench without its hump is a lake at rest (h + b = 0 exactly, the
    depth changing by far less than a factor 2 between neighbours): every
    edge is still, and the state stays bit for bit at rest."""
    from swe_ppm_amd import scenarios

    _, _, _, b, dx, dy, bnd = scenarios.trench_hump(n=4096, seed=4096)
    h0 = np.maximum(-b, 0.0).astype(np.float32)
    z = np.zeros_like(h0)
    print("4096^2 trench at rest, 100 steps", flush=True)
    (gh, ghu, ghv), _ = _run(swe, (h0, z, z.copy(), b, dx, dy, bnd), 100, swe.AUGRIE)
    inner = (slice(1, -1), slice(1, -1))
    assert same_bits(gh[inner], h0[inner])
    assert np.all(ghu[inner] == 0.0) and np.all(ghv[inner] == 0.0)


def test_trench_4096_learned_middle_state(swe, oracle):
    """Config 5 with the packaged network: 40 steps of the 4096^2 trench with
    the learned middle state.  The fused loop (hard edges deferred to a
    second launch) equals the store strategy (the network inside the sweep)
    bit for bit, and its distance to the FP64 oracle run of the same scheme
    is within 1.5x that of the exact solver's FP32 run."""
    from swe_ppm_amd import scenarios, surrogate

    st = scenarios.trench_hump(n=4096, seed=4096)
    h, hu, hv, b, dx, dy, bnd = st
    steps = 40
    net = surrogate.load_default()
    out = {}
    for name, strategy, attach in (("exact", 2, False), ("learned", 2, True), ("learned_in_sweep", 1, True)):
        print(f"4096^2 trench, {steps} steps, {name}", flush=True)
        with swe.SweBlock(4096, 4096, dx, dy, solver=swe.AUGRIE) as blk:
            blk.set_boundaries(bnd)
            blk.set_tuning(strategy, -1)
            if attach:
                blk.set_surrogate(net)
            blk.upload(h, hu, hv, b)
            blk.run(steps, want_dts=False)
            out[name] = blk.download()
    net.close()
    for f in range(3):
        assert same_bits(out["learned"][f], out["learned_in_sweep"][f]), f
    print("4096^2 trench, FP64 oracle", flush=True)
    ref = oracle.run_u(h, hu, hv, b, dx, dy, bnd, swe.AUGRIE, steps, nthreads=THREADS, dtype=np.float64)
    err = {}
    for name in ("exact", "learned"):
        err[name] = max(float(np.max(np.abs(out[name][f][1:-1, 1:-1] - ref[f][1:-1, 1:-1])) /
                              np.max(np.abs(ref[f][1:-1, 1:-1]))) for f in range(3))
    print(f"rel Linf vs FP64: exact {err['exact']:.3e}, learned {err['learned']:.3e}", flush=True)
    assert err["learned"] <= 1.5 * err["exact"] + 1e-6


def _rel_linf(a, ref):
    return max(float(np.max(np.abs(a[f][1:-1, 1:-1] - ref[f][1:-1, 1:-1])) / np.max(np.abs(ref[f][1:-1, 1:-1])))
               for f in range(3))


def _rel_l2(a, ref):
    return max(float(np.sqrt(np.sum((a[f][1:-1, 1:-1].astype(np.float64) - ref[f][1:-1, 1:-1]) ** 2) /
                             np.sum(np.asarray(ref[f][1:-1, 1:-1], np.float64) ** 2))) for f in range(3))


def _differing(a, c):
    return sum(int(np.count_nonzero(a[f][1:-1, 1:-1] != c[f][1:-1, 1:-1])) for f in range(3))


def test_islands_2048_learned_middle_state(swe, oracle):
    """Config 2 with the packaged network: 100 steps of the 2048^2 islands
    with the learned middle state attached stay within 1.5x the exact FP32
    run's distance to the FP64 oracle run of the same scheme."""
    from swe_ppm_amd import scenarios, surrogate

    st = scenarios.dam_break_islands(n=2048, seed=2048)
    h, hu, hv, b, dx, dy, bnd = st
    steps = 100
    net = surrogate.load_default()
    out = {}
    for name, attach in (("exact", False), ("learned", True)):
        print(f"2048^2 islands, {steps} steps, {name}", flush=True)
        with swe.SweBlock(2048, 2048, dx, dy, solver=swe.AUGRIE) as blk:
            blk.set_boundaries(bnd)
            if attach:
                blk.set_surrogate(net)
            blk.upload(h, hu, hv, b)
            blk.run(steps, want_dts=False)
            out[name] = blk.download()
    net.close()
    print("2048^2 islands, FP64 oracle", flush=True)
    ref = oracle.run_u(h, hu, hv, b, dx, dy, bnd, swe.AUGRIE, steps, nthreads=THREADS, dtype=np.float64)
    linf = {k: _rel_linf(v, ref) for k, v in out.items()}
    l2 = {k: _rel_l2(v, ref) for k, v in out.items()}
    print(f"rel Linf vs FP64: exact {linf['exact']:.3e}, learned {linf['learned']:.3e}; rel L2: exact "
          f"{l2['exact']:.3e}, learned {l2['learned']:.3e}; values differing from exact: "
          f"{_differing(out['learned'], out['exact'])}", flush=True)
    assert linf["learned"] <= 1.5 * linf["exact"] + 1e-6
    assert l2["learned"] <= 1.5 * l2["exact"] + 1e-7


def test_ensemble_learned_middle_state(swe, oracle):
    """Config 4 with the packaged network: the first 128 ensemble instances,
    100 steps through the batched engine with and without the learned middle
    state; against per-instance FP64 oracle runs the learned form's largest
    and median errors stay within 1.5x the exact form's."""
    from swe_ppm_amd import scenarios, surrogate

    inst = scenarios.ensemble_instances(count=512, seed=512)[:128]
    steps = 100
    net = surrogate.load_default()
    out = {}
    for name, attach in (("exact", False), ("learned", True)):
        print(f"128-instance ensemble, {steps} steps, {name}", flush=True)
        with swe.SweBatch(inst, solver=swe.AUGRIE, mode=swe.BATCH_ENSEMBLE) as bt:
            if attach:
                bt.set_surrogate(net)
            bt.upload(*[[i[f] for i in inst] for f in ("h", "hu", "hv", "b")])
            bt.run(steps)
            out[name] = bt.download()
    net.close()
    print("128 instances, FP64 oracle", flush=True)
    err = {"exact": [], "learned": []}
    differing = 0
    for k, i in enumerate(inst):
        ref = oracle.run_u(i["h"], i["hu"], i["hv"], i["b"], i["dx"], i["dy"], i["bnd"], swe.AUGRIE, steps,
                           nthreads=THREADS, dtype=np.float64)
        got = {name: [out[name][f][k] for f in range(3)] for name in out}
        for name in err:
            err[name].append(_rel_linf(got[name], ref))
        differing += _differing(got["learned"], got["exact"])
    mx = {k: max(v) for k, v in err.items()}
    med = {k: float(np.median(v)) for k, v in err.items()}
    print(f"rel Linf vs FP64, max / median over instances: exact {mx['exact']:.3e} / {med['exact']:.3e}, "
          f"learned {mx['learned']:.3e} / {med['learned']:.3e}; values differing from exact: {differing}",
          flush=True)
    assert mx["learned"] <= 1.5 * mx["exact"] + 1e-6
    assert med["learned"] <= 1.5 * med["exact"] + 1e-7


def test_lts_tiling_2048_one_macro_step(swe, oracle):
    from swe_ppm_amd import scenarios

    blocks, dx, dy = scenarios.lts_tiling(seed=28, n=2048)
    print("28-block LTS tiling of 2048^2, 1 macro step", flush=True)
    with swe.SweBatch(blocks, solver=swe.AUGRIE, mode=swe.BATCH_LTS, dx=dx, dy=dy, max_local=1.0,
                      duration=80.0) as bt:
        bt.upload(*[[bl[f] for bl in blocks] for f in ("h", "hu", "hv", "b")])
        _, its, _ = bt.run(1)
        got = bt.download()
        upd, _ = bt.stats()
    states, oupd, oits, _ = oracle.multiblock_run(blocks, dx, dy, swe.AUGRIE, True, 1.0, 80.0, 1, nthreads=THREADS)
    assert same_bits(its, oits)
    assert same_bits(upd, oupd)
    _assert_blocks_equal(oracle, got, states, blocks, "LTS 2048^2 tiling")


def test_ensemble_512_instances(swe, oracle):
    from swe_ppm_amd import scenarios

    inst = scenarios.ensemble_instances(count=512, seed=512)
    nsteps = 5
    print("512-instance ensemble, 5 steps", flush=True)
    with swe.SweBatch(inst, solver=swe.AUGRIE, mode=swe.BATCH_ENSEMBLE) as bt:
        bt.upload(*[[i[f] for i in inst] for f in ("h", "hu", "hv", "b")])
        bt.run(nsteps)
        got = bt.download()
    want = [oracle.run_u(i["h"], i["hu"], i["hv"], i["b"], i["dx"], i["dy"], i["bnd"], swe.AUGRIE, nsteps)[:3]
            for i in inst]
    _assert_blocks_equal(oracle, got, want, inst, "512-instance ensemble")
