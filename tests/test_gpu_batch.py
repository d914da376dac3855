"""GPU parity of the batched engine (include/swe_hip_batch.h) against the CPU
oracle, through the C ABI.

Bar: bit-identical interiors, time steps, iteration counts and update
counts.  The batched engine leaves the OUTFLOW / WALL ghosts and corners of
the current state (what the next setGhostLayer writes), the oracle's loops
leave those of the state before the last update; the oracle's outputs get
one more applyBoundaryConditions before the ghost lines are compared, which
leaves the CONNECT_WITHIN_RANK ghost lines as the oracle left them.
"""
import numpy as np
import pytest

from helpers import same_bits

pytestmark = pytest.mark.gpu


def _with_bc(oracle, state, bnd):
    h, hu, hv = (a.copy() for a in state[:3])
    oracle.apply_bc(bnd, h, hu, hv)
    return h, hu, hv


def _assert_blocks_equal(oracle, got, want, blocks, what):
    gh, ghu, ghv = got
    for k, bl in enumerate(blocks):
        oh, ohu, ohv = _with_bc(oracle, want[k], bl["bnd"])
        for name, a, c in (("h", gh[k], oh), ("hu", ghu[k], ohu), ("hv", ghv[k], ohv)):
            assert same_bits(a[1:-1, 1:-1], c[1:-1, 1:-1]), (
                f"{what}: block {k} {name} interior differs (max {np.max(np.abs(a - c)):.3e})")
            assert same_bits(a, c), f"{what}: block {k} {name} ghost cells differ"


# groups: -1 the default (the direct form), 0 direct, > 0 a persistent grid of that many workgroups
@pytest.mark.parametrize("xs,chunk,graph,fused,groups", [(8, 16, 1, 1, -1), (4, 7, 1, 1, 512), (16, 5, 0, 1, 2048),
                                                         (8, 16, 1, 0, -1), (16, 5, 0, 0, 300), (8, 16, 1, 1, 0),
                                                         (4, 5, 0, 0, 0)])
def test_ensemble_matches_independent_oracle_runs(swe, oracle, xs, chunk, graph, fused, groups):
    from swe_ppm_amd import scenarios

    inst = scenarios.ensemble_instances(count=24, seed=7, nmin=8, nmax=140)
    nsteps = 37
    with swe.SweBatch(inst, solver=swe.AUGRIE, mode=swe.BATCH_ENSEMBLE) as bt:
        bt.set_tuning(xs=xs, groups=groups, chunk=chunk, graph=graph, fused=fused)
        bt.upload([i["h"] for i in inst], [i["hu"] for i in inst], [i["hv"] for i in inst], [i["b"] for i in inst])
        dts, its, ms = bt.run(nsteps)
        got = bt.download()
        upd, t = bt.stats()
    assert np.all(upd == nsteps)
    assert np.all(its == 1)
    want = []
    for k, i in enumerate(inst):
        oh, ohu, ohv, odts = oracle.run_u(i["h"], i["hu"], i["hv"], i["b"], i["dx"], i["dy"], i["bnd"],
                                          swe.AUGRIE, nsteps)
        want.append((oh, ohu, ohv))
        # the per-instance time is the float sum of its own dts, in order
        tt = np.float32(0.0)
        for d in odts:
            tt = np.float32(tt + d)
        assert t[k] == tt, f"instance {k}: t {t[k]} != {tt}"
    _assert_blocks_equal(oracle, got, want, inst, f"ensemble xs={xs}")


@pytest.mark.parametrize("fused,extreme", [(1, False), (0, False), (1, True)])
def test_ensemble_random_states_bitwise(swe, oracle, fused, extreme):
    """Instances of uncorrelated random cell states (every Riemann case on
    many edges, wet/dry included; extreme: depths log-uniform over 1e-5 ..
    5000 m, velocities to 50 m/s) through the batched sweeps: two steps
    bit-identical to independent oracle runs."""
    rng = np.random.default_rng(91)
    inst = []
    for _ in range(6):
        nx, ny = (int(v) for v in rng.integers(70, 260, size=2))
        shape = (nx + 2, ny + 2)
        if extreme:
            h = np.exp(rng.uniform(np.log(1e-5), np.log(5000.0), shape)).astype(np.float32)
            vel = 50.0
        else:
            h = rng.uniform(0.2, 20.0, shape).astype(np.float32)
            vel = 3.0
        h[rng.random(h.shape) < 0.03] = 0.0
        hu = (h * rng.uniform(-vel, vel, h.shape)).astype(np.float32)
        hv = (h * rng.uniform(-vel, vel, h.shape)).astype(np.float32)
        b = rng.uniform(-2.0, 2.0, h.shape).astype(np.float32)
        inst.append(dict(nx=nx, ny=ny, dx=1.0, dy=1.0, bnd=[0, 0, 0, 0], h=h, hu=hu, hv=hv, b=b))
    nsteps = 2
    with swe.SweBatch(inst, solver=swe.AUGRIE, mode=swe.BATCH_ENSEMBLE) as bt:
        bt.set_tuning(xs=8, groups=-1, chunk=16, graph=1, fused=fused)
        bt.upload([i["h"] for i in inst], [i["hu"] for i in inst], [i["hv"] for i in inst], [i["b"] for i in inst])
        bt.run(nsteps)
        got = bt.download()
    want = [oracle.run_u(i["h"], i["hu"], i["hv"], i["b"], i["dx"], i["dy"], i["bnd"], swe.AUGRIE, nsteps)[:3]
            for i in inst]
    _assert_blocks_equal(oracle, got, want, inst, f"random-state ensemble fused={fused} extreme={extreme}")


def test_ensemble_runs_continue_across_calls(swe, oracle):
    from swe_ppm_amd import scenarios

    inst = scenarios.ensemble_instances(count=6, seed=3, nmin=20, nmax=70)
    with swe.SweBatch(inst, solver=swe.FWAVE, mode=swe.BATCH_ENSEMBLE) as bt:
        bt.upload([i["h"] for i in inst], [i["hu"] for i in inst], [i["hv"] for i in inst], [i["b"] for i in inst])
        for n in (1, 4, 9):
            bt.run(n)
        got = bt.download()
    want = [oracle.run_u(i["h"], i["hu"], i["hv"], i["b"], i["dx"], i["dy"], i["bnd"], swe.FWAVE, 14)[:3]
            for i in inst]
    _assert_blocks_equal(oracle, got, want, inst, "ensemble 1+4+9 steps")


@pytest.mark.parametrize("solver,fused,groups", [(2, 1, -1), (3, 1, 600), (2, 0, -1)])
def test_global_dt_blocks_match_oracle(swe, oracle, solver, fused, groups):
    from swe_ppm_amd import scenarios

    blocks, dx, dy = scenarios.lts_tiling(n=256)
    nsteps = 25
    with swe.SweBatch(blocks, solver=solver, mode=swe.BATCH_GLOBAL, dx=dx, dy=dy) as bt:
        bt.set_tuning(groups=groups, chunk=6, fused=fused)
        bt.upload(*[[bl[f] for bl in blocks] for f in ("h", "hu", "hv", "b")])
        dts, its, _ = bt.run(nsteps)
        got = bt.download()
        upd, _ = bt.stats()
    states, oupd, oits, odts = oracle.multiblock_run(blocks, dx, dy, solver, False, 1.0, 0.0, nsteps, nthreads=8)
    assert same_bits(dts, odts)
    assert same_bits(upd, oupd)
    _assert_blocks_equal(oracle, got, states, blocks, "global dt")


@pytest.mark.parametrize("T,duration,nmacro,fused,groups", [(1.0, 3.0, 3, 1, -1), (1.0, 2.5, 3, 1, 1024),
                                                            (0.5, 1.2, 3, 1, -1), (1.0, 3.0, 3, 0, -1),
                                                            (0.5, 1.2, 3, 0, 700), (1.0, 2.5, 3, 1, 0)])
def test_lts_blocks_match_oracle(swe, oracle, T, duration, nmacro, fused, groups):
    """Local time stepping over 28 uneven blocks: copies, interpolated ghost
    layers, power-of-two sub-steps, blocks reaching the end time mid-step."""
    from swe_ppm_amd import scenarios

    blocks, dx, dy = scenarios.lts_tiling(n=256)
    with swe.SweBatch(blocks, solver=swe.AUGRIE, mode=swe.BATCH_LTS, dx=dx, dy=dy, max_local=T,
                      duration=duration) as bt:
        bt.set_tuning(groups=groups, chunk=5, fused=fused)
        bt.upload(*[[bl[f] for bl in blocks] for f in ("h", "hu", "hv", "b")])
        _, its, _ = bt.run(nmacro)
        got = bt.download()
        upd, _ = bt.stats()
    states, oupd, oits, _ = oracle.multiblock_run(blocks, dx, dy, swe.AUGRIE, True, T, duration, nmacro,
                                                  nthreads=8)
    assert same_bits(its, oits), (its, oits)
    assert same_bits(upd, oupd)
    assert len(set(upd.tolist())) > 1, "blocks must take different numbers of sub-steps"
    _assert_blocks_equal(oracle, got, states, blocks, f"LTS T={T}")


@pytest.mark.parametrize("fused", [1, 0])
def test_lts_random_states_match_oracle(swe, oracle, fused):
    """Local time stepping over the 28-block tiling with uncorrelated random
    cell states, each block at its own depth scale (so that blocks take
    different power-of-two sub-steps): one macro step bit-identical to the
    oracle's LTS driver, ghost copies and interpolations included."""
    from swe_ppm_amd import scenarios

    blocks, dx, dy = scenarios.lts_tiling(n=256)
    rng = np.random.default_rng(95)
    for bl in blocks:
        shape = (bl["nx"] + 2, bl["ny"] + 2)
        scale = rng.uniform(0.5, 20.0)
        h = (scale * rng.uniform(0.5, 1.5, shape)).astype(np.float32)
        h[rng.random(shape) < 0.02] = 0.0
        bl["h"] = h
        bl["hu"] = (h * rng.uniform(-1.0, 1.0, shape)).astype(np.float32)
        bl["hv"] = (h * rng.uniform(-1.0, 1.0, shape)).astype(np.float32)
        bl["b"] = rng.uniform(-1.0, 1.0, shape).astype(np.float32)
    with swe.SweBatch(blocks, solver=swe.AUGRIE, mode=swe.BATCH_LTS, dx=dx, dy=dy, max_local=1.0,
                      duration=1.0) as bt:
        bt.set_tuning(chunk=5, fused=fused)
        bt.upload(*[[bl[f] for bl in blocks] for f in ("h", "hu", "hv", "b")])
        _, its, _ = bt.run(1)
        got = bt.download()
        upd, _ = bt.stats()
    states, oupd, oits, _ = oracle.multiblock_run(blocks, dx, dy, swe.AUGRIE, True, 1.0, 1.0, 1, nthreads=8)
    assert same_bits(its, oits), (its, oits)
    assert same_bits(upd, oupd)
    assert len(set(upd.tolist())) > 1, "blocks must take different numbers of sub-steps"
    _assert_blocks_equal(oracle, got, states, blocks, f"LTS random states fused={fused}")


def test_lts_macro_steps_split_over_calls_and_autotune(swe, oracle):
    """Two runs of 1 and 2 macro steps, with the autotuner in between (it
    restores the state it times on), equal one oracle run of 3."""
    from swe_ppm_amd import scenarios

    blocks, dx, dy = scenarios.lts_tiling(seed=11, n=512)
    with swe.SweBatch(blocks, solver=swe.AUGRIE, mode=swe.BATCH_LTS, dx=dx, dy=dy, max_local=0.5,
                      duration=10.0) as bt:
        bt.upload(*[[bl[f] for bl in blocks] for f in ("h", "hu", "hv", "b")])
        _, its1, _ = bt.run(1)
        cand = bt.autotune(iterations=8)
        assert len(cand) == 12 and np.all(cand > 0)
        _, its2, _ = bt.run(2)
        got = bt.download()
    states, _, oits, _ = oracle.multiblock_run(blocks, dx, dy, swe.AUGRIE, True, 0.5, 10.0, 3, nthreads=8)
    assert same_bits(np.concatenate([its1, its2]), oits)
    _assert_blocks_equal(oracle, got, states, blocks, "LTS split runs")


def test_batch_rejects_inconsistent_neighbours(swe):
    blocks = [dict(nx=16, ny=16, bnd=[0, 4, 0, 0], nbr=[-1, 1, -1, -1]),
              dict(nx=16, ny=12, bnd=[4, 0, 0, 0], nbr=[0, -1, -1, -1])]
    with pytest.raises(swe.SweHipError):
        swe.SweBatch(blocks, mode=swe.BATCH_GLOBAL, dx=1.0, dy=1.0)
    ok = [dict(nx=16, ny=16, bnd=[0, 4, 0, 0], nbr=[-1, 1, -1, -1]),
          dict(nx=16, ny=16, bnd=[4, 0, 0, 0], nbr=[0, -1, -1, -1])]
    with pytest.raises(swe.SweHipError):
        swe.SweBatch(ok, mode=swe.BATCH_ENSEMBLE, dx=1.0, dy=1.0)
    with swe.SweBatch(ok, mode=swe.BATCH_GLOBAL, dx=1.0, dy=1.0) as bt:
        with pytest.raises(swe.SweHipError):
            bt.run(1)  # nothing uploaded yet
        with pytest.raises(swe.SweHipError):
            bt.set_tuning(xs=5)


EDGE_SIZES = [(1, 1), (1, 70), (70, 1), (200, 1), (200, 2), (33, 63), (64, 124), (9, 126), (300, 30), (37, 62),
              (8, 64), (100, 125)]


def _edge_instance(rng, nx, ny, bnd):
    """A wet state with dry patches, uneven bed and moving water."""
    h = rng.uniform(0.5, 6.0, size=(nx + 2, ny + 2)).astype(np.float32)
    h[rng.uniform(size=h.shape) < 0.1] = 0.0
    b = rng.uniform(-1.0, 1.0, size=h.shape).astype(np.float32)
    hu = (h * rng.uniform(-1.0, 1.0, size=h.shape)).astype(np.float32)
    hv = (h * rng.uniform(-1.0, 1.0, size=h.shape)).astype(np.float32)
    return dict(nx=nx, ny=ny, dx=float(np.float32(10.0)), dy=float(np.float32(7.5)), bnd=bnd, h=h, hu=hu, hv=hv, b=b)


@pytest.mark.parametrize("xs", [4, 8, 16])
def test_ensemble_edge_shapes_match_oracle(swe, oracle, xs):
    """Single-row and single-column blocks, heights around the 62-row strip
    (tails of 1..62 rows, packed several segments to a wave), widths that
    are not a multiple of xs, WALL sides, dry cells."""
    rng = np.random.default_rng(17)
    inst = []
    for k, (nx, ny) in enumerate(EDGE_SIZES):
        bnd = [int(v) for v in rng.integers(0, 2, size=4)]  # OUTFLOW or WALL per side
        inst.append(_edge_instance(rng, nx, ny, bnd))
    nsteps = 12
    with swe.SweBatch(inst, solver=swe.AUGRIE, mode=swe.BATCH_ENSEMBLE) as bt:
        bt.set_tuning(xs=xs, chunk=5)
        bt.upload(*[[i[f] for i in inst] for f in ("h", "hu", "hv", "b")])
        bt.run(nsteps)
        got = bt.download()
    want = [oracle.run_u(i["h"], i["hu"], i["hv"], i["b"], np.float32(i["dx"]), np.float32(i["dy"]), i["bnd"],
                         swe.AUGRIE, nsteps)[:3] for i in inst]
    _assert_blocks_equal(oracle, got, want, inst, f"edge shapes xs={xs}")


@pytest.mark.parametrize("lts", [False, True])
def test_connected_edge_shapes_match_oracle(swe, oracle, lts):
    """3 x 3 connected blocks with column widths (1, 45, 200) and row heights
    (63, 1, 124): thin blocks between wide ones, global dt and LTS."""
    widths, heights = [1, 45, 200], [63, 1, 124]
    rng = np.random.default_rng(23)
    rows, cols = len(heights), len(widths)
    blocks = []
    for c in range(cols):
        for r in range(rows):
            k = c * rows + r
            nbr = [k - rows if c > 0 else -1, k + rows if c < cols - 1 else -1, k - 1 if r > 0 else -1,
                   k + 1 if r < rows - 1 else -1]
            bnd = [4 if m >= 0 else 0 for m in nbr]
            nx, ny = widths[c], heights[r]
            depth = np.float32(rng.uniform(2.0, 20.0))
            h = np.full((nx + 2, ny + 2), depth, np.float32)
            h += rng.uniform(0.0, 0.5, size=h.shape).astype(np.float32)
            z = np.zeros_like(h)
            blocks.append(dict(nx=nx, ny=ny, bnd=bnd, nbr=nbr, h=h, hu=z.copy(), hv=z.copy(), b=z.copy()))
    dx = dy = 5.0
    T, duration, units = 0.5, 2.0, 3
    mode = swe.BATCH_LTS if lts else swe.BATCH_GLOBAL
    with swe.SweBatch(blocks, solver=swe.AUGRIE, mode=mode, dx=dx, dy=dy, max_local=T, duration=duration) as bt:
        bt.upload(*[[bl[f] for bl in blocks] for f in ("h", "hu", "hv", "b")])
        _, its, _ = bt.run(units if lts else 20)
        got = bt.download()
        upd, _ = bt.stats()
    states, oupd, oits, _ = oracle.multiblock_run(blocks, dx, dy, swe.AUGRIE, lts, T, duration if lts else 0.0,
                                                  units if lts else 20, nthreads=8)
    assert same_bits(upd, oupd)
    if lts:
        assert same_bits(its, oits)
    _assert_blocks_equal(oracle, got, states, blocks, f"connected edge shapes lts={lts}")


def test_ensemble_at_the_block_limit(swe, oracle):
    """2047 instances (the limit): the round keeps two blocks per thread."""
    rng = np.random.default_rng(2047)
    inst = []
    for k in range(2047):
        nx, ny = (int(v) for v in rng.integers(2, 14, size=2))
        inst.append(_edge_instance(rng, nx, ny, [int(v) for v in rng.integers(0, 2, size=4)]))
    nsteps = 4
    with swe.SweBatch(inst, solver=swe.AUGRIE, mode=swe.BATCH_ENSEMBLE) as bt:
        bt.upload(*[[i[f] for i in inst] for f in ("h", "hu", "hv", "b")])
        bt.run(nsteps)
        got = bt.download()
        upd, _ = bt.stats()
    assert np.all(upd == nsteps)
    want = [oracle.run_u(i["h"], i["hu"], i["hv"], i["b"], np.float32(i["dx"]), np.float32(i["dy"]), i["bnd"],
                         swe.AUGRIE, nsteps)[:3] for i in inst]
    _assert_blocks_equal(oracle, got, want, inst, "2047 instances")
    with pytest.raises(swe.SweHipError):
        swe.SweBatch(inst + inst[:1], solver=swe.AUGRIE, mode=swe.BATCH_ENSEMBLE)


def test_lts_with_more_blocks_than_round_threads(swe, oracle):
    """32 x 40 connected blocks (1280 > 1024 round threads) of uneven depth:
    neighbour times cross between the round's two register slots."""
    cols, rows = 32, 40
    rng = np.random.default_rng(1280)
    widths = [int(v) for v in rng.integers(4, 12, size=cols)]
    heights = [int(v) for v in rng.integers(4, 12, size=rows)]
    blocks = []
    for c in range(cols):
        for r in range(rows):
            k = c * rows + r
            nbr = [k - rows if c > 0 else -1, k + rows if c < cols - 1 else -1, k - 1 if r > 0 else -1,
                   k + 1 if r < rows - 1 else -1]
            bnd = [4 if m >= 0 else 0 for m in nbr]
            nx, ny = widths[c], heights[r]
            h = np.full((nx + 2, ny + 2), np.float32(rng.uniform(2.0, 40.0)), np.float32)
            z = np.zeros_like(h)
            blocks.append(dict(nx=nx, ny=ny, bnd=bnd, nbr=nbr, h=h, hu=z.copy(), hv=z.copy(), b=z.copy()))
    dx = dy = 2.0
    with swe.SweBatch(blocks, solver=swe.AUGRIE, mode=swe.BATCH_LTS, dx=dx, dy=dy, max_local=0.5, duration=1.0) as bt:
        bt.upload(*[[bl[f] for bl in blocks] for f in ("h", "hu", "hv", "b")])
        _, its, _ = bt.run(2)
        got = bt.download()
        upd, _ = bt.stats()
    states, oupd, oits, _ = oracle.multiblock_run(blocks, dx, dy, swe.AUGRIE, True, 0.5, 1.0, 2, nthreads=8)
    assert same_bits(its, oits)
    assert same_bits(upd, oupd)
    assert len(set(upd.tolist())) > 1
    _assert_blocks_equal(oracle, got, states, blocks, "1280 LTS blocks")
