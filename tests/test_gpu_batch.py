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
            assert np.array_equal(a[1:-1, 1:-1], c[1:-1, 1:-1]), (
                f"{what}: block {k} {name} interior differs (max {np.max(np.abs(a - c)):.3e})")
            assert np.array_equal(a, c), f"{what}: block {k} {name} ghost cells differ"


@pytest.mark.parametrize("xs,chunk,graph,fused", [(8, 16, 1, 1), (4, 7, 1, 1), (16, 5, 0, 1), (8, 16, 1, 0),
                                                  (16, 5, 0, 0)])
def test_ensemble_matches_independent_oracle_runs(swe, oracle, xs, chunk, graph, fused):
    from swe_ppm_amd import scenarios

    inst = scenarios.ensemble_instances(count=24, seed=7, nmin=8, nmax=140)
    nsteps = 37
    with swe.SweBatch(inst, solver=swe.AUGRIE, mode=swe.BATCH_ENSEMBLE) as bt:
        bt.set_tuning(xs=xs, chunk=chunk, graph=graph, fused=fused)
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


@pytest.mark.parametrize("solver,fused", [(2, 1), (3, 1), (2, 0)])
def test_global_dt_blocks_match_oracle(swe, oracle, solver, fused):
    from swe_ppm_amd import scenarios

    blocks, dx, dy = scenarios.lts_tiling(n=256)
    nsteps = 25
    with swe.SweBatch(blocks, solver=solver, mode=swe.BATCH_GLOBAL, dx=dx, dy=dy) as bt:
        bt.set_tuning(chunk=6, fused=fused)
        bt.upload(*[[bl[f] for bl in blocks] for f in ("h", "hu", "hv", "b")])
        dts, its, _ = bt.run(nsteps)
        got = bt.download()
        upd, _ = bt.stats()
    states, oupd, oits, odts = oracle.multiblock_run(blocks, dx, dy, solver, False, 1.0, 0.0, nsteps, nthreads=8)
    assert np.array_equal(dts, odts)
    assert np.array_equal(upd, oupd)
    _assert_blocks_equal(oracle, got, states, blocks, "global dt")


@pytest.mark.parametrize("T,duration,nmacro,fused", [(1.0, 3.0, 3, 1), (1.0, 2.5, 3, 1), (0.5, 1.2, 3, 1),
                                                     (1.0, 3.0, 3, 0), (0.5, 1.2, 3, 0)])
def test_lts_blocks_match_oracle(swe, oracle, T, duration, nmacro, fused):
    """Local time stepping over 28 uneven blocks: copies, interpolated ghost
    layers, power-of-two sub-steps, blocks reaching the end time mid-step."""
    from swe_ppm_amd import scenarios

    blocks, dx, dy = scenarios.lts_tiling(n=256)
    with swe.SweBatch(blocks, solver=swe.AUGRIE, mode=swe.BATCH_LTS, dx=dx, dy=dy, max_local=T,
                      duration=duration) as bt:
        bt.set_tuning(chunk=5, fused=fused)
        bt.upload(*[[bl[f] for bl in blocks] for f in ("h", "hu", "hv", "b")])
        _, its, _ = bt.run(nmacro)
        got = bt.download()
        upd, _ = bt.stats()
    states, oupd, oits, _ = oracle.multiblock_run(blocks, dx, dy, swe.AUGRIE, True, T, duration, nmacro,
                                                  nthreads=8)
    assert np.array_equal(its, oits), (its, oits)
    assert np.array_equal(upd, oupd)
    assert len(set(upd.tolist())) > 1, "blocks must take different numbers of sub-steps"
    _assert_blocks_equal(oracle, got, states, blocks, f"LTS T={T}")


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
        assert len(cand) == 9 and np.all(cand > 0)
        _, its2, _ = bt.run(2)
        got = bt.download()
    states, _, oits, _ = oracle.multiblock_run(blocks, dx, dy, swe.AUGRIE, True, 0.5, 10.0, 3, nthreads=8)
    assert np.array_equal(np.concatenate([its1, its2]), oits)
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
