"""GPU parity of the HIP step against the CPU oracle, through the C ABI.

Bar: the kernels keep the oracle's single-precision operation order with
no contraction and correctly rounded division/square root, so per-edge
results, time steps and states after N steps are required to be identical
(relative L-infinity 0); the 1e-5 relative L-infinity bound is checked beside.
"""
import numpy as np
import pytest

from helpers import same_bits, random_edges, rel_linf

pytestmark = pytest.mark.gpu

SOLVERS = [(0, "hlle"), (1, "fwave"), (2, "augrie"), (3, "hybrid")]


@pytest.mark.parametrize("solver,name", SOLVERS)
def test_solve_edges_bitwise(swe, oracle, solver, name):
    ins = random_edges(200_000, seed=11 + solver)
    g = swe.solve_edges(solver, *ins)
    o = oracle.solve_edges(solver, *ins)
    for k, (a, b) in enumerate(zip(g, o)):
        assert same_bits(a, b), f"{name} output {k}: max diff {np.nanmax(np.abs(a - b))}"


def _random_states(n, seed):
    rng = np.random.default_rng(seed)
    h = rng.uniform(0.2, 20.0, (n + 2, n + 2)).astype(np.float32)
    h[rng.random(h.shape) < 0.03] = 0.0
    hu = (h * rng.uniform(-3.0, 3.0, h.shape)).astype(np.float32)
    hv = (h * rng.uniform(-3.0, 3.0, h.shape)).astype(np.float32)
    b = rng.uniform(-2.0, 2.0, h.shape).astype(np.float32)
    return h, hu, hv, b


@pytest.mark.parametrize("solver", [0, 1, 2, 3])
@pytest.mark.parametrize("strategy,xseg,bnd", [(1, 16, (0, 0, 0, 0)), (2, 12, (0, 0, 0, 0)), (2, 8, (1, 0, 1, 1))])
def test_random_states_bitwise(swe, oracle, solver, strategy, xseg, bnd):
    """Uncorrelated random cell states, so that every Riemann case occurs on
    many edges (two shocks, shock and rarefaction, two rarefactions, dry and
    wet/dry edges): one and two steps bit-identical to the oracle, through the
    sweeps and the fused step, whose Newton loops start from the bracket's
    quantities."""
    h, hu, hv, b = _random_states(768, 77)
    bnd = list(bnd)
    for steps in (1, 2):
        g, _, gdts = _run_gpu(swe, h, hu, hv, b, 1.0, 1.0, bnd, solver, steps, strategy=strategy, xseg=xseg)
        oh, ohu, ohv, odts = oracle.run_u(h, hu, hv, b, 1.0, 1.0, bnd, solver, steps)
        for name, a, c in zip(("h", "hu", "hv"), g, (oh, ohu, ohv)):
            nd = int(np.count_nonzero(a[1:-1, 1:-1] != c[1:-1, 1:-1]))
            assert nd == 0, f"{steps} step(s): {name} differs in {nd} cells"
        assert same_bits(gdts, odts)


def _run_gpu(swe, h, hu, hv, b, dx, dy, bnd, solver, nsteps, strategy=1, xseg=16, flavour=0):
    nx, ny = h.shape[0] - 2, h.shape[1] - 2
    with swe.SweBlock(nx, ny, dx, dy, solver=solver, flavour=flavour) as blk:
        blk.set_boundaries(bnd)
        blk.set_tuning(strategy, xseg)
        blk.upload(h, hu, hv, b)
        t, dts = blk.run(nsteps)
        out = blk.download()
    return out, t, dts


def _assert_parity(g, o, gdts, odts, tol=1e-5):
    for name, a, b in zip(("h", "hu", "hv"), g, o):
        e = rel_linf(a, b)
        assert e <= tol, f"{name}: rel Linf {e:.3e}"
        assert same_bits(a[1:-1, 1:-1], b[1:-1, 1:-1]), f"{name}: not bit-identical (rel Linf {e:.3e})"
    assert same_bits(gdts, odts)


def test_c1_fwave_256_200_steps(swe, oracle):
    from swe_ppm_amd import scenarios

    h, hu, hv, b, dx, dy, bnd = scenarios.dam_break_c1(256)
    (gh, ghu, ghv), t, gdts = _run_gpu(swe, h, hu, hv, b, dx, dy, bnd, swe.FWAVE, 200)
    oh, ohu, ohv, odts = oracle.run_u(h, hu, hv, b, dx, dy, bnd, swe.FWAVE, 200)
    _assert_parity((gh, ghu, ghv), (oh, ohu, ohv), gdts, odts)
    # the front stays inside the domain: mass conserved to round-off
    m0 = scenarios.interior_mass(h, dx, dy)
    m1 = scenarios.interior_mass(gh, dx, dy)
    assert abs(m1 - m0) / m0 < 1e-6
    assert abs(t - float(np.sum(gdts, dtype=np.float32))) <= 1e-3 * t


@pytest.mark.parametrize("strategy,xseg", [(1, 16), (0, 16), (1, 8), (0, 32), (2, 8), (2, 16), (2, 32), (2, 10),
                                           (2, 12), (1, 12), (2, 4), (1, 4)])
def test_islands_augrie_parity(swe, oracle, strategy, xseg):
    from swe_ppm_amd import scenarios

    h, hu, hv, b, dx, dy, bnd = scenarios.dam_break_islands(n=192, seed=5)
    (gh, ghu, ghv), t, gdts = _run_gpu(swe, h, hu, hv, b, dx, dy, bnd, swe.AUGRIE, 60, strategy, xseg)
    oh, ohu, ohv, odts = oracle.run_u(h, hu, hv, b, dx, dy, bnd, swe.AUGRIE, 60)
    _assert_parity((gh, ghu, ghv), (oh, ohu, ohv), gdts, odts)
    assert np.any(gh[1:-1, 1:-1] == 0.0), "the islands must stay dry"


@pytest.mark.parametrize("nx,ny", [(1, 1), (1, 70), (70, 1), (63, 65), (130, 17), (61, 62), (17, 125)])
@pytest.mark.parametrize("solver", [1, 3])
def test_ragged_sizes_and_walls(swe, oracle, nx, ny, solver):
    rng = np.random.default_rng(nx * 1000 + ny)
    h = rng.uniform(0.0, 8.0, (nx + 2, ny + 2)).astype(np.float32)
    h[rng.random(h.shape) < 0.1] = 0.0
    hu = (h * rng.normal(0, 1, h.shape)).astype(np.float32)
    hv = (h * rng.normal(0, 1, h.shape)).astype(np.float32)
    b = rng.normal(0, 1, h.shape).astype(np.float32)
    bnd = [1, 0, 0, 1]  # WALL, OUTFLOW, OUTFLOW, WALL
    from swe_ppm_amd import scenarios

    scenarios.apply_bc(h, hu, hv, bnd)
    scenarios.apply_bc_bathymetry(b, bnd)
    for strategy in (0, 1, 2):
        (gh, ghu, ghv), t, gdts = _run_gpu(swe, h, hu, hv, b, 1.0, 1.5, bnd, solver, 12, strategy)
        oh, ohu, ohv, odts = oracle.run_u(h, hu, hv, b, 1.0, 1.5, bnd, solver, 12)
        _assert_parity((gh, ghu, ghv), (oh, ohu, ohv), gdts, odts)


def test_per_call_api_matches_oracle_step_by_step(swe, oracle):
    """setGhostLayer / computeNumericalFluxes / getMaxTimestep / updateUnknowns
    one call at a time, ghosts included, against the oracle's step."""
    from swe_ppm_amd import scenarios

    h, hu, hv, b, dx, dy, bnd = scenarios.dam_break_islands(n=96, seed=9)
    oh, ohu, ohv = h.copy(), hu.copy(), hv.copy()
    with swe.SweBlock(96, 96, dx, dy, solver=swe.AUGRIE) as blk:
        blk.set_boundaries(bnd)
        blk.upload(h, hu, hv, b)
        for step in range(15):
            blk.apply_bc()
            dt = blk.fluxes()
            blk.update(dt)
            oh, ohu, ohv, odts = oracle.run_u(oh, ohu, ohv, b, dx, dy, bnd, swe.AUGRIE, 1)
            assert dt == odts[0]
            gh, ghu, ghv = blk.download()
            for a, c in ((gh, oh), (ghu, ohu), (ghv, ohv)):
                assert same_bits(a, c), f"step {step}"


@pytest.mark.parametrize("solver,xseg", [(2, 16), (2, 4), (1, 12), (0, 8)])
def test_per_call_steps_without_downloads_match_oracle(swe, oracle, solver, xseg):
    """The per-call loop with no host transfer between steps: the update is
    left pending and folded into the next flux pass, ghosts included."""
    from swe_ppm_amd import scenarios

    h, hu, hv, b, dx, dy, bnd = scenarios.dam_break_islands(n=150, seed=4)
    bnd = [1, 0, 0, 1]
    scenarios.apply_bc(h, hu, hv, bnd)
    nsteps = 23
    with swe.SweBlock(150, 150, dx, dy, solver=solver) as blk:
        blk.set_boundaries(bnd)
        blk.set_tuning(-1, xseg)
        blk.upload(h, hu, hv, b)
        dts = []
        for _ in range(nsteps):
            blk.apply_bc()
            dt = blk.fluxes()
            blk.update(dt)
            dts.append(dt)
        gh, ghu, ghv = blk.download()
    oh, ohu, ohv, odts = oracle.run_u(h, hu, hv, b, dx, dy, bnd, solver, nsteps)
    assert same_bits(np.array(dts, np.float32), odts)
    for a, c in ((gh, oh), (ghu, ohu), (ghv, ohv)):
        assert same_bits(a, c)


@pytest.mark.parametrize("solver", [0, 1, 2, 3])
@pytest.mark.parametrize("lookahead", [True, False])
def test_per_call_random_states_match_oracle(swe, oracle, solver, lookahead):
    """The per-call loop (setGhostLayer, computeNumericalFluxes,
    getMaxTimestep, updateUnknowns) on uncorrelated random cell states, with
    and without the look-ahead that folds each update into the next flux
    pass: three steps bit-identical to the oracle."""
    h, hu, hv, b = _random_states(300, 84 + solver)
    bnd = [1, 0, 0, 1]
    nsteps = 3
    with swe.SweBlock(300, 300, 1.0, 1.0, solver=solver) as blk:
        blk.set_boundaries(bnd)
        blk.set_lookahead(lookahead)
        blk.upload(h, hu, hv, b)
        dts = []
        for _ in range(nsteps):
            blk.apply_bc()
            dt = blk.fluxes()
            blk.update(dt)
            dts.append(dt)
        g = blk.download()
    oh, ohu, ohv, odts = oracle.run_u(h, hu, hv, b, 1.0, 1.0, bnd, solver, nsteps)
    assert same_bits(np.array(dts, np.float32), odts)
    for name, a, c in zip(("h", "hu", "hv"), g, (oh, ohu, ohv)):
        assert same_bits(a[1:-1, 1:-1], c[1:-1, 1:-1]), name


def test_per_call_orders_equal_eager_updates(swe):
    """Runs of the drivers' order (setGhostLayer -> fluxes -> update) broken by
    other orders (fluxes without a boundary refresh, repeated setGhostLayer,
    bathymetry ghosts, a boundary type changed mid-run, a run() between
    per-call steps, an update with a smaller dt than returned, a download)
    give the same bits whether each update is forced out at once, left
    pending, or left pending with the next step queued ahead."""
    from swe_ppm_amd import scenarios

    h, hu, hv, b, dx, dy, bnd = scenarios.dam_break_islands(n=100, seed=8)
    in_order = ["bc", "f", "u"] * 8
    script = (in_order + ["f", "u", "f", "u", "bc", "bc", "f", "u", "bcb", "bc", "f", "u", "wall", "bc", "f", "u"]
              + in_order + ["run", "f", "u", "bc", "f", "half", "bc", "f", "u", "bc", "f", "u", "bc", "f", "u",
                            "bc", "f", "u", "get"] + in_order + ["bc", "f", "u", "f", "bc", "u", "bc", "f", "u"])
    outs = []
    for mode in ("lookahead", "lazy", "eager"):
        with swe.SweBlock(100, 100, dx, dy, solver=swe.AUGRIE) as blk:
            blk.set_boundaries(bnd)
            blk.set_tuning(-1, 8)
            blk.set_lookahead(mode == "lookahead")
            blk.upload(h, hu, hv, b)
            dts, mid = [], None
            for op in script:
                if op == "bc":
                    blk.apply_bc()
                elif op == "bcb":
                    blk.apply_bc_bathymetry()
                elif op == "f":
                    dts.append(blk.fluxes())
                elif op in ("u", "half"):
                    blk.update(dts[-1] if op == "u" else dts[-1] * 0.5)
                    if mode == "eager":
                        blk.download()
                elif op == "wall":
                    blk.set_boundaries([1, 1, 0, 0])
                elif op == "run":
                    _, rd = blk.run(3)
                    dts.extend(rd.tolist())
                elif op == "get":
                    mid = blk.download()
            outs.append((blk.download(with_b=True), mid, np.array(dts, np.float32)))
    a, ma, da = outs[2]
    for c, mc, dc in outs[:2]:
        assert same_bits(da, dc)
        for x, y in zip(a + ma, c + mc):
            assert same_bits(x, y)


def test_snapshots_hold_the_state_of_their_call(swe):
    """swe_hip_snapshot during stepping: each slot holds the state at its
    call (as a download there), whatever ran after it; slots alternate, and
    wait/release on a slot without a snapshot are refused."""
    from swe_ppm_amd import scenarios

    h, hu, hv, b, dx, dy, bnd = scenarios.dam_break_islands(n=120, seed=6)
    with swe.SweBlock(120, 120, dx, dy, solver=swe.AUGRIE) as ref, swe.SweBlock(120, 120, dx, dy) as blk:
        for k in (ref, blk):
            k.set_boundaries(bnd)
            k.upload(h, hu, hv, b)
        with pytest.raises(swe.SweHipError):
            blk.snapshot_wait(0)
        expected, slots = [], []
        for rnd in range(5):
            for k in (ref, blk):
                k.run(3)
            for _ in range(2):  # per-call steps left pending
                for k in (ref, blk):
                    k.apply_bc()
                    k.update(k.fluxes())
            expected.append(ref.download())
            slots.append(blk.snapshot())
            blk.run(4)  # overwrites both state buffers
            ref.run(4)
            if rnd % 2 == 1:  # release the two in flight
                for s in slots[-2:]:
                    got = blk.snapshot_wait(s)
                    for a, c in zip(got, expected[slots.index(s, len(slots) - 2)]):
                        assert same_bits(a, c)
                    blk.snapshot_release(s)
        assert slots[:4] == [0, 1, 0, 1]
        got = blk.snapshot_wait(slots[4])
        for a, c in zip(got, expected[4]):
            assert same_bits(a, c)
        blk.snapshot_release(slots[4])
        with pytest.raises(swe.SweHipError):
            blk.snapshot_release(slots[4])


@pytest.mark.parametrize("bnd_outer,states", [(0, "islands"), (1, "islands"), (1, "random")])
def test_connected_blocks_per_call_match_the_oracle(swe, oracle, bnd_outer, states):
    """Two blocks sharing an edge, driven per call as the overdecomposed
    driver does (setGhostLayer with the in-rank copy, fluxes, the global dt
    minimum, update): the copy reads the neighbour's copy layer with its
    pending update applied on the fly, so both blocks keep the fused step.
    Bitwise against the oracle's connected global-dt run, on the islands
    dam break and on uncorrelated random cell states."""
    from swe_ppm_amd import scenarios

    if states == "islands":
        h, hu, hv, b, dx, dy, bnd = scenarios.dam_break_islands(n=128, seed=12)
    else:
        h, hu, hv, b = _random_states(128, 86)
        dx = dy = 1.0
    # left block x in [1, 70], right block x in [71, 128]
    cut = 70
    parts = [(0, cut), (cut, 128)]
    blocks, ob = [], []
    for k, (x0, x1) in enumerate(parts):
        nxk = x1 - x0
        sl = slice(x0, x1 + 2)
        bk = [bnd_outer] * 4
        bk[1 if k == 0 else 0] = 4  # CONNECT_WITHIN_RANK towards the other block
        nbr = [-1, -1, -1, -1]
        nbr[1 if k == 0 else 0] = 1 - k
        blk = swe.SweBlock(nxk, 128, dx, dy, solver=swe.AUGRIE)
        blk.set_boundaries(bk)
        blk.upload(h[sl], hu[sl], hv[sl], b[sl])
        blocks.append(blk)
        ob.append(dict(nx=nxk, ny=128, bnd=bk, nbr=nbr, h=h[sl].copy(), hu=hu[sl].copy(), hv=hv[sl].copy(),
                       b=b[sl].copy()))
    nsteps = 25
    try:
        for step in range(nsteps):
            for k, blk in enumerate(blocks):
                blk.apply_bc()
                blk.copy_layer(1 if k == 0 else 0, blocks[1 - k])
            dt = min(blk.fluxes() for blk in blocks)
            for blk in blocks:
                blk.update(dt)
        got = [blk.download() for blk in blocks]
    finally:
        for blk in blocks:
            blk.close()
    res, _, _, _ = oracle.multiblock_run(ob, np.float32(dx), np.float32(dy), swe.AUGRIE, False, 1.0, 1e30, nsteps)
    for k in range(2):
        for a, c in zip(got[k], res[k][:3]):
            assert same_bits(a[1:-1, 1:-1], c[1:-1, 1:-1]), f"block {k}"


def test_update_needs_a_fresh_flux_pass(swe):
    from swe_ppm_amd import scenarios

    h, hu, hv, b, dx, dy, bnd = scenarios.dam_break_c1(32)
    with swe.SweBlock(32, 32, dx, dy, solver=swe.FWAVE) as blk:
        blk.set_boundaries(bnd)
        blk.upload(h, hu, hv, b)
        with pytest.raises(RuntimeError):
            blk.update(0.1)
        blk.apply_bc()
        dt = blk.fluxes()
        blk.update(dt)
        with pytest.raises(RuntimeError):
            blk.update(dt)


def test_run_ghosts_match_reference_sequence(swe, oracle):
    """After run(), ghost cells hold the boundary values set at the start of
    the last step, as after the reference's loop."""
    from swe_ppm_amd import scenarios

    h, hu, hv, b, dx, dy, bnd = scenarios.dam_break_c1(64)
    for bnd in ([0, 0, 0, 0], [1, 0, 1, 5], [5, 1, 0, 1]):
        for strategy in (0, 1, 2):
            for nsteps in (1, 2, 7):
                (gh, ghu, ghv), t, _ = _run_gpu(swe, h, hu, hv, b, dx, dy, bnd, swe.FWAVE, nsteps, strategy)
                oh, ohu, ohv, _ = oracle.run_u(h, hu, hv, b, dx, dy, bnd, swe.FWAVE, nsteps)
                for a, c in ((gh, oh), (ghu, ohu), (ghv, ohv)):
                    assert same_bits(a, c), (bnd, strategy, nsteps)


@pytest.mark.parametrize("solver", [1, 3])
def test_flavour_s_parity(swe, oracle, solver):
    from swe_ppm_amd import scenarios

    h, hu, hv, b, dx, dy, bnd = scenarios.dam_break_c1(96)
    nsteps = 25
    with swe.SweBlock(96, 96, dx, dy, solver=solver, flavour=swe.FLAVOUR_S) as blk:
        blk.set_boundaries(bnd)
        blk.upload(h, hu, hv, b)
        t, gdts = blk.run(nsteps)
        gh, ghu, ghv = blk.download()
    oh, ohu, ohv, odts = oracle.run_s(h, hu, hv, b, dx, bnd, solver, nsteps)
    _assert_parity((gh, ghu, ghv), (oh, ohu, ohv), gdts, odts)


def _threshold_states(n, seed):
    """Random cells crowded at the solvers' branch points: depths about the
    dry tolerance (0.01 m), the still-water minimum (0.1 m) and the hybrid
    switch (1 m); zero momentum on many cells; free surfaces from a small set,
    so that neighbours often share one and still edges (and near-still ones,
    off by a depth ratio over 2) occur next to moving and dry ones."""
    rng = np.random.default_rng(seed)
    shape = (n + 2, n + 2)
    pick = rng.random(shape)
    h = np.where(pick < 0.25, rng.uniform(0.0, 0.02, shape),
                 np.where(pick < 0.45, rng.uniform(0.09, 0.11, shape),
                          np.where(pick < 0.65, rng.uniform(0.9, 1.1, shape), rng.uniform(0.2, 20.0, shape))))
    h = h.astype(np.float32)
    h[rng.random(shape) < 0.1] = 0.0
    eta = rng.choice(np.array([0.0, 0.5, 2.0], np.float32), size=shape)
    b = (eta - h).astype(np.float32)
    still = rng.random(shape) < 0.5
    hu = np.where(still, 0.0, h * rng.uniform(-2.0, 2.0, shape)).astype(np.float32)
    hv = np.where(still, 0.0, h * rng.uniform(-2.0, 2.0, shape)).astype(np.float32)
    return h, hu, hv, b


@pytest.mark.parametrize("solver", [0, 1, 2, 3])
@pytest.mark.parametrize("strategy,xseg", [(1, 16), (2, 12)])
def test_threshold_states_bitwise(swe, oracle, solver, strategy, xseg):
    """States crowded at the branch points of the solvers (dry, still, hybrid
    switch): one and two steps bit-identical to the oracle."""
    h, hu, hv, b = _threshold_states(512, 80 + solver)
    bnd = [0, 1, 0, 1]
    for steps in (1, 2):
        g, _, gdts = _run_gpu(swe, h, hu, hv, b, 1.0, 1.0, bnd, solver, steps, strategy=strategy, xseg=xseg)
        oh, ohu, ohv, odts = oracle.run_u(h, hu, hv, b, 1.0, 1.0, bnd, solver, steps)
        for name, a, c in zip(("h", "hu", "hv"), g, (oh, ohu, ohv)):
            nd = int(np.count_nonzero(a[1:-1, 1:-1].view(np.uint32) != c[1:-1, 1:-1].view(np.uint32)))
            assert nd == 0, f"{steps} step(s): {name} differs in {nd} cells"
        assert same_bits(gdts, odts)


@pytest.mark.parametrize("nx,ny", [(1, 1), (1, 70), (70, 1), (2, 3), (63, 65), (130, 17), (61, 62), (17, 125),
                                   (5, 200), (200, 5)])
def test_ragged_random_states_bitwise(swe, oracle, nx, ny):
    """Ragged sizes with random cell states and a random mix of wall and
    outflow sides, every solver and strategy: three steps bit-identical to
    the oracle."""
    rng = np.random.default_rng(nx * 1000 + ny)
    shape = (nx + 2, ny + 2)
    h = rng.uniform(0.2, 20.0, shape).astype(np.float32)
    h[rng.random(shape) < 0.05] = 0.0
    hu = (h * rng.uniform(-3.0, 3.0, shape)).astype(np.float32)
    hv = (h * rng.uniform(-3.0, 3.0, shape)).astype(np.float32)
    b = rng.uniform(-2.0, 2.0, shape).astype(np.float32)
    bnd = [int(v) for v in rng.integers(0, 2, size=4)]
    for solver in (0, 1, 2, 3):
        for strategy, xseg in ((0, 8), (1, 16), (2, 4), (2, 12)):
            g, _, gdts = _run_gpu(swe, h, hu, hv, b, 1.0, 1.0, bnd, solver, 3, strategy=strategy, xseg=xseg)
            oh, ohu, ohv, odts = oracle.run_u(h, hu, hv, b, 1.0, 1.0, bnd, solver, 3)
            for name, a, c in zip(("h", "hu", "hv"), g, (oh, ohu, ohv)):
                assert same_bits(a[1:-1, 1:-1], c[1:-1, 1:-1]), (solver, strategy, xseg, bnd, name)
            assert same_bits(gdts, odts), (solver, strategy, xseg, bnd)


@pytest.mark.parametrize("solver", [0, 1, 2, 3])
@pytest.mark.parametrize("strategy,xseg", [(1, 16), (2, 12)])
def test_extreme_states_bitwise(swe, oracle, solver, strategy, xseg):
    """Depths log-uniform over 1e-5 .. 5000 m (5% dry), velocities up to
    50 m/s, bed levels of +-100 m, walls on two sides: two steps bit-identical
    to the oracle, with no NaN."""
    n = 384
    rng = np.random.default_rng(200 + solver)
    shape = (n + 2, n + 2)
    h = np.exp(rng.uniform(np.log(1e-5), np.log(5000.0), shape)).astype(np.float32)
    h[rng.random(shape) < 0.05] = 0.0
    hu = (h * rng.uniform(-50.0, 50.0, shape).astype(np.float32)).astype(np.float32)
    hv = (h * rng.uniform(-50.0, 50.0, shape).astype(np.float32)).astype(np.float32)
    b = rng.uniform(-100.0, 100.0, shape).astype(np.float32)
    bnd = [0, 1, 0, 1]
    g, _, gdts = _run_gpu(swe, h, hu, hv, b, 10.0, 10.0, bnd, solver, 2, strategy=strategy, xseg=xseg)
    oh, ohu, ohv, odts = oracle.run_u(h, hu, hv, b, 10.0, 10.0, bnd, solver, 2)
    for name, a, c in zip(("h", "hu", "hv"), g, (oh, ohu, ohv)):
        assert not np.any(np.isnan(a)), name
        assert same_bits(a[1:-1, 1:-1], c[1:-1, 1:-1]), name
    assert same_bits(gdts, odts)


@pytest.mark.parametrize("solver", [0, 1, 2, 3])
def test_flavour_s_random_states_bitwise(swe, oracle, solver):
    """swe_simple's split step on uncorrelated random cell states: two steps
    bit-identical to the oracle."""
    h, hu, hv, b = _random_states(384, 78)
    bnd = [0, 0, 0, 0]
    with swe.SweBlock(384, 384, 1.0, 1.0, solver=solver, flavour=swe.FLAVOUR_S) as blk:
        blk.set_boundaries(bnd)
        blk.upload(h, hu, hv, b)
        _, gdts = blk.run(2)
        g = blk.download()
    oh, ohu, ohv, odts = oracle.run_s(h, hu, hv, b, 1.0, bnd, solver, 2)
    for name, a, c in zip(("h", "hu", "hv"), g, (oh, ohu, ohv)):
        assert same_bits(a[1:-1, 1:-1], c[1:-1, 1:-1]), name
    assert same_bits(gdts, odts)


def test_apply_bc_matches_oracle(swe, oracle):
    rng = np.random.default_rng(3)
    nx, ny = 37, 29
    h, hu, hv, b = (rng.normal(5, 1, (nx + 2, ny + 2)).astype(np.float32) for _ in range(4))
    for bnd in ([0, 0, 0, 0], [1, 1, 1, 1], [1, 0, 5, 3], [5, 5, 5, 5]):
        with swe.SweBlock(nx, ny, 1.0, 1.0) as blk:
            blk.set_boundaries(bnd)
            blk.upload(h, hu, hv, b)
            blk.apply_bc()
            blk.apply_bc_bathymetry()
            gh, ghu, ghv, gb = blk.download(with_b=True)
        oh, ohu, ohv, ob = h.copy(), hu.copy(), hv.copy(), b.copy()
        oracle.apply_bc(bnd, oh, ohu, ohv)
        oracle.apply_bc_bathymetry(bnd, ob)
        for a, c in ((gh, oh), (ghu, ohu), (ghv, ohv), (gb, ob)):
            assert same_bits(a, c), bnd


def test_invalid_arguments_fail_loudly(swe):
    with pytest.raises(swe.SweHipError):
        swe.SweBlock(0, 10, 1.0, 1.0)
    with pytest.raises(swe.SweHipError):
        swe.SweBlock(10, 10, -1.0, 1.0)
    with swe.SweBlock(8, 8, 1.0, 1.0) as blk:
        with pytest.raises(swe.SweHipError):
            blk.set_boundaries([2, 0, 0, 0])  # INFLOW: the reference asserts on it
        with pytest.raises(swe.SweHipError):
            blk.set_tuning(7, -1)
        with pytest.raises(swe.SweHipError):
            blk.set_tuning(-1, 11)  # strip widths are 4, 8, 10, 12, 16 or 32


def test_autotune_restores_the_state_and_keeps_the_results(swe):
    """swe_hip_autotune times every strip width from the current state, puts
    the state back bit for bit and keeps the fastest width; a run afterwards
    gives the same bits as a run on an untuned block."""
    from swe_ppm_amd import scenarios

    h, hu, hv, b, dx, dy, bnd = scenarios.dam_break_islands(n=300, seed=7)
    with swe.SweBlock(300, 300, dx, dy, solver=swe.AUGRIE) as blk:
        blk.set_boundaries(bnd)
        blk.upload(h, hu, hv, b)
        t_ref, _ = blk.run(30)
        ref = blk.download()
    with swe.SweBlock(300, 300, dx, dy, solver=swe.AUGRIE) as blk:
        blk.set_boundaries(bnd)
        blk.upload(h, hu, hv, b)
        before = blk.download()
        ms = blk.autotune(4)
        assert sorted(ms) == [4, 8, 10, 12, 16, 32] and all(v > 0.0 for v in ms.values())
        assert blk.get_tuning()[1] == min(ms, key=ms.get)
        for a, c in zip(blk.download(), before):
            assert same_bits(a, c)
        t, _ = blk.run(30)
        assert t == t_ref
        for a, c in zip(blk.download(), ref):
            assert same_bits(a, c)
        with pytest.raises(swe.SweHipError):
            blk.autotune(1)


@pytest.mark.parametrize("solver", [0, 1])
def test_autotune_may_keep_recompute_for_light_solvers(swe, solver):
    """For HLLE and f-wave the autotuner also times the recompute strategy at
    16 and 32 columns; whatever it keeps, a run afterwards gives the bits of
    an untuned block."""
    from swe_ppm_amd import scenarios

    h, hu, hv, b, dx, dy, bnd = scenarios.dam_break_islands(n=700, seed=3)
    with swe.SweBlock(700, 700, dx, dy, solver=solver) as blk:
        blk.set_boundaries(bnd)
        blk.upload(h, hu, hv, b)
        t_ref, d_ref = blk.run(25)
        ref = blk.download()
    with swe.SweBlock(700, 700, dx, dy, solver=solver) as blk:
        blk.set_boundaries(bnd)
        blk.upload(h, hu, hv, b)
        blk.autotune(5)
        strategy, xseg = blk.get_tuning()
        assert strategy == 2 or (strategy == 0 and xseg in (16, 32))
        t, d = blk.run(25)
        assert t == t_ref and same_bits(d, d_ref)
        for a, c in zip(blk.download(), ref):
            assert same_bits(a, c)


def test_default_strip_width_follows_the_wave_count(swe):
    """The widest of 16 and 12 columns per strip whose busiest SIMD still
    holds the solver's threshold of waves (256 CUs x 4 SIMDs; AugRie 12, 4
    waves; learned 5, 3; HLLE, f-wave, Hybrid 2, 1), else 8 (widths are whole
    4-column quiet-record tiles): 4096^2 has 67 strips x 256 segments at 16
    (16.75 per SIMD), 2048^2 34 x 171 at 12 (5.68), 1024^2 17 x 86 at 12
    (1.43, below 4).  A width the caller set is kept."""
    from swe_ppm_amd import surrogate

    with swe.SweBlock(2048, 2048, 1.0, 1.0, solver=swe.HLLE) as blk:
        assert blk.get_tuning()[1] == 16
    with swe.SweBlock(1024, 1024, 1.0, 1.0, solver=swe.FWAVE) as blk:
        assert blk.get_tuning()[1] == 12
    net = surrogate.load_default()
    for nx, ny, exact, learned in ((2048, 2048, 12, 12), (4096, 4096, 16, 16), (4096, 2048, 12, 16),
                                   (2048, 4096, 12, 16), (1024, 1024, 8, 8), (64, 64, 8, 8)):
        with swe.SweBlock(nx, ny, 1.0, 1.0) as blk:
            assert blk.get_tuning()[1] == exact, (nx, ny)
            blk.set_surrogate(net)
            assert blk.get_tuning()[1] == learned, (nx, ny)
            blk.set_surrogate(None)
            assert blk.get_tuning()[1] == exact, (nx, ny)
            blk.set_tuning(-1, 32)
            blk.set_surrogate(net)
            assert blk.get_tuning()[1] == 32
    net.close()


def test_analytic_scenario_known_answer_and_errors(swe):
    """swe_hip_init_scenario on the radial dam break over [0, 1000]^2 (15 m
    within 100 m of (250, 250), 10 m elsewhere, b = 10, OUTFLOW): cell
    centres decide, ghosts copy their neighbours; unknown ids fail loudly."""
    with swe.SweBlock(10, 10, 100.0, 100.0) as blk:
        blk.set_boundaries([0, 0, 0, 0])
        blk.init_scenario(swe.SCENARIO_RADIAL_DAM_BREAK)
        h, hu, hv, b = blk.download(with_b=True)
        want = np.full((12, 12), 10.0, np.float32)
        want[3, 3] = 15.0  # the only centre within 100 m: (250, 250) is the centre of cell (3, 3)
        assert same_bits(h, want)
        assert np.all(b == 10.0) and np.all(hu == 0.0) and np.all(hv == 0.0)
        with pytest.raises(swe.SweHipError):
            blk.init_scenario(8)


def test_grid_scenario_known_answer_and_errors(swe):
    """swe_hip_init_grid_scenario on a 2 x 2 bathymetry grid of 50 m cells
    over [0, 100]^2, sampled by a 4 x 4 block of 25 m cells: each quarter of
    the block takes one grid value (interior) and the OUTFLOW ghosts copy
    their neighbours.  A displacement grid over [25, 75]^2 adds 1 m to the
    four central cells only."""
    bathy = (np.array([[-10.0, -20.0], [-30.0, 5.0]], np.float32), 25.0, 25.0, 50.0, 50.0)  # rows are y
    displ = (np.ones((1, 1), np.float32), 50.0, 50.0, 50.0, 50.0)
    with swe.SweBlock(4, 4, 25.0, 25.0) as blk:
        blk.set_boundaries([0, 0, 0, 0])
        blk.init_grid_scenario(bathy, displ)
        h, hu, hv, b = blk.download(with_b=True)
    want_b = np.zeros((4, 4), np.float32)
    want_h = np.zeros((4, 4), np.float32)
    for x in range(4):  # interior cell x + 1, centre at 12.5 + 25 x
        for y in range(4):
            bv = bathy[0][y // 2, x // 2]
            inside = 1 <= x <= 2 and 1 <= y <= 2  # centres 37.5 and 62.5 lie strictly inside (25, 75)
            want_b[x, y] = bv + (1.0 if inside else 0.0)
            want_h[x, y] = 0.0 if bv > 0 else -bv
    assert same_bits(b[1:5, 1:5], want_b)
    assert same_bits(h[1:5, 1:5], want_h)
    assert not np.any(hu) and not np.any(hv)
    assert same_bits(b[0, 1:5], b[1, 1:5]) and same_bits(h[5, 1:5], h[4, 1:5])  # OUTFLOW ghosts
    with swe.SweBlock(4, 4, 25.0, 25.0) as blk:
        for bad in [(np.ones((2, 2), np.float32), 0.0, 0.0, 0.0, 1.0), (np.ones((2, 2), np.float32), 0.0, 0.0, 1.0, -1.0),
                    (np.ones((0, 2), np.float32), 0.0, 0.0, 1.0, 1.0)]:
            with pytest.raises(swe.SweHipError):
                blk.init_grid_scenario(bad)


@pytest.mark.parametrize("solver", [2, 3])
def test_solve_edges_still_and_almost_still_bitwise(swe, oracle, solver):
    """Still edges (a lake at rest: zero net updates and the full solver's
    speed) and edges one condition short of still, per edge against the
    oracle, bit for bit."""
    from test_oracle_solvers import still_edges

    hL, hR, huL, huR, bL, bR = still_edges(20000, 51)
    f = np.float32
    n = hL.size
    sets = [(hL, hR, huL, huR, bL, bR),
            (hL, hR, np.full(n, f(1e-30)), huR, bL, bR),
            (hL, (np.nextafter(hR + bR, f(np.inf)) - bR).astype(f), huL, huR, bL, bR),
            (np.full(n, f(0.09)), np.full(n, f(0.09)), huL, huR, np.zeros(n, f), np.zeros(n, f))]
    ins = [np.concatenate([s[k] for s in sets]).astype(f) for k in range(6)]
    g = swe.solve_edges(solver, *ins)
    o = oracle.solve_edges(solver, *ins)
    for k, (a, b) in enumerate(zip(g, o)):
        assert same_bits(a, b), f"output {k}"
    if solver == 2:  # Hybrid sends edges with both depths >= 1 m to the f-wave solver
        assert np.all(g[0][:n] == 0.0) and np.all(g[3][:n] == 0.0)


@pytest.mark.parametrize("solver", [2, 3])
def test_flavour_s_parity_dx_ne_dy(swe, oracle, solver):
    """Flavour S on a rectangular block with dx != dy (the reference's split
    step uses dx for the y-part and only the x-speeds for dt): bit-identical
    to the oracle, on 2-D data with currents both ways and dry cells."""
    rng = np.random.default_rng(52)
    nx, ny, dx, dy = 70, 45, 2.0, 7.0
    x = np.arange(nx + 2)[:, None]
    y = np.arange(ny + 2)[None, :]
    h = (4.0 + 2.0 * np.sin(x / 5.0) + 1.5 * np.cos(y / 4.0) + rng.uniform(0, 0.3, (nx + 2, ny + 2)))
    h[30:34, 10:14] = 0.0
    hu = h * rng.uniform(-1.0, 1.0, h.shape)
    hv = h * rng.uniform(-1.0, 1.0, h.shape)
    b = rng.uniform(-0.3, 0.3, h.shape)
    h, hu, hv, b = (a.astype(np.float32) for a in (h, hu, hv, b))
    bnd = [0, 1, 0, 1]
    nsteps = 20
    with swe.SweBlock(nx, ny, dx, dy, solver=solver, flavour=swe.FLAVOUR_S) as blk:
        blk.set_boundaries(bnd)
        blk.upload(h, hu, hv, b)
        t, gdts = blk.run(nsteps)
        gh, ghu, ghv = blk.download()
    oh, ohu, ohv, odts = oracle.run_s(h, hu, hv, b, dx, bnd, solver, nsteps)
    _assert_parity((gh, ghu, ghv), (oh, ohu, ohv), gdts, odts)


def still_lake(n: int, seed: int, eta: float = 10.0):
    """A lake at rest over smooth random bathymetry, built so that every wet
    edge is still in single precision: depths in [3, 5.9] m (neighbours
    within a factor 2), h + b == eta exactly in float32, plus square dry
    islands (h = 0, bed above the surface)."""
    from swe_ppm_amd import scenarios

    f = np.float32
    rng = np.random.default_rng(seed)
    k = np.arange(n + 2, dtype=np.float64)
    ph = rng.uniform(0, 2 * np.pi, 4)
    field = (np.sin(k[:, None] / 97.0 + ph[0]) * np.cos(k[None, :] / 61.0 + ph[1]) +
             0.5 * np.sin(k[:, None] / 23.0 + ph[2]) * np.sin(k[None, :] / 41.0 + ph[3]))
    h = (3.0 + 2.9 * (field - field.min()) / (field.max() - field.min())).astype(f)
    b = (f(eta) - h).astype(f)
    for _ in range(4):  # make h + b == eta exact in float32
        bad = (h + b) != f(eta)
        h[bad] = (f(eta) - b[bad]).astype(f)
    bad = (h + b) != f(eta)
    h[bad], b[bad] = f(0.0), f(eta + 2.0)
    for cx, cy in rng.integers(n // 8, n - n // 8, size=(3, 2)):
        h[cx:cx + n // 40, cy:cy + n // 30] = f(0.0)
        b[cx:cx + n // 40, cy:cy + n // 30] = f(eta + 2.0)
    bnd = [0, 0, 0, 0]
    z = np.zeros_like(h)
    h, hu, hv, b = scenarios.finish(h, b, bnd)
    return h, hu, hv, b, bnd


def test_lake_at_rest_2048_stays_exactly_still(swe, oracle):
    """A 2048^2 lake at rest over smooth bathymetry with dry islands: AugRie
    keeps every cell bit for bit at rest over 60 steps of the fused loop
    (every wet edge is still, every wet/dry edge a wall at rest), with the
    oracle's dts."""
    n = 2048
    h, hu, hv, b, bnd = still_lake(n, 53)
    dx = 1000.0 / n
    (gh, ghu, ghv), t, dts = _run_gpu(swe, h, hu, hv, b, dx, dx, bnd, swe.AUGRIE, 60, strategy=2, xseg=-1)
    assert same_bits(gh[1:-1, 1:-1], h[1:-1, 1:-1])
    assert np.all(ghu[1:-1, 1:-1] == 0.0) and np.all(ghv[1:-1, 1:-1] == 0.0)
    odts = oracle.run_u(h, hu, hv, b, dx, dx, bnd, swe.AUGRIE, 2, nthreads=16)[3]
    assert same_bits(dts[:2], odts) and np.all(dts == dts[0])


@pytest.mark.parametrize("solver,n,steps,xseg", [(2, 512, 300, 12), (3, 384, 200, 16), (2, 200, 150, 8), (2, 300, 200, 4),
                                                 (2, 260, 120, 32), (2, 130, 60, 16),
                                                 (0, 256, 100, 16), (1, 256, 100, 32)])
def test_quiet_skip_is_bit_identical(swe, oracle, solver, n, steps, xseg):
    """The fused loop with quiet tiles skipped against the same loop computing
    every tile and against the oracle: identical states and dts, on dam breaks
    whose still far field makes most tiles quiet for most of the run (walls
    on two sides, dry islands)."""
    from swe_ppm_amd import scenarios

    h, hu, hv, b, dx, dy, _ = scenarios.dam_break_islands(n=n, seed=9)
    bnd = [0, 1, 1, 0]
    scenarios.apply_bc(h, hu, hv, bnd)
    scenarios.apply_bc_bathymetry(b, bnd)
    got = {}
    for skip in (1, 0):
        with swe.SweBlock(n, n, dx, dy, solver=solver) as blk:
            blk.set_boundaries(bnd)
            blk.set_tuning(2, xseg)
            blk.set_quiet_skip(bool(skip))
            blk.upload(h, hu, hv, b)
            t1, d1 = blk.run(steps // 2)
            t2, d2 = blk.run(steps - steps // 2)  # a second run on the same block starts a new record
            got[skip] = (blk.download(), np.concatenate([d1, d2]))
    for f in range(3):
        assert same_bits(got[1][0][f], got[0][0][f]), f
    assert same_bits(got[1][1], got[0][1])
    if n <= 256:
        o = oracle.run_u(h, hu, hv, b, dx, dy, bnd, solver, steps, nthreads=16)
        for f in range(3):
            assert same_bits(got[1][0][f][1:-1, 1:-1], o[f][1:-1, 1:-1]), f


@pytest.mark.parametrize("xseg", [12, 4])
def test_quiet_skip_random_patches_bit_identical(swe, oracle, xseg):
    """A lake at rest over random bathymetry (depths 5-9 m, so that every
    edge of the far field is a still edge and its tiles go quiet, and dry
    islands) with patches of uncorrelated random states: the patches'
    disturbances spread into quiet tiles through every Riemann case.
    Skipping on and off and the oracle: identical states and dts."""
    n, steps = 320, 80
    rng = np.random.default_rng(83)
    shape = (n + 2, n + 2)
    b = -rng.uniform(5.0, 9.0, shape).astype(np.float32)
    for _ in range(10):
        x0, y0 = (int(v) for v in rng.integers(1, n - 8, size=2))
        b[x0:x0 + 8, y0:y0 + 8] = 1.0
    h = np.maximum(np.float32(0.0) - b, np.float32(0.0)).astype(np.float32)
    hu = np.zeros(shape, np.float32)
    hv = np.zeros(shape, np.float32)
    for _ in range(6):
        x0, y0 = (int(v) for v in rng.integers(1, n - 24, size=2))
        sl = (slice(x0, x0 + 24), slice(y0, y0 + 24))
        h[sl] = rng.uniform(0.0, 15.0, (24, 24))
        hu[sl] = h[sl] * rng.uniform(-2.0, 2.0, (24, 24))
        hv[sl] = h[sl] * rng.uniform(-2.0, 2.0, (24, 24))
    bnd = [1, 0, 0, 1]
    got = {}
    for skip in (1, 0):
        with swe.SweBlock(n, n, 1.0, 1.0, solver=swe.AUGRIE) as blk:
            blk.set_boundaries(bnd)
            blk.set_tuning(2, xseg)
            blk.set_quiet_skip(bool(skip))
            blk.upload(h, hu, hv, b)
            _, d = blk.run(steps)
            got[skip] = (blk.download(), np.asarray(d))
    for f in range(3):
        assert same_bits(got[1][0][f], got[0][0][f]), f
    assert same_bits(got[1][1], got[0][1])
    o = oracle.run_u(h, hu, hv, b, 1.0, 1.0, bnd, swe.AUGRIE, steps, nthreads=16)
    for f in range(3):
        assert same_bits(got[1][0][f][1:-1, 1:-1], o[f][1:-1, 1:-1]), f
    assert same_bits(got[1][1], o[3])


def test_quiet_skip_learned_middle_state_bit_identical(swe):
    """The learned middle state's deferring fused loop with quiet tiles
    skipped against the same loop computing every tile."""
    from swe_ppm_amd import scenarios, surrogate

    net = surrogate.load_default()
    h, hu, hv, b, dx, dy, bnd = scenarios.dam_break_islands(n=384, seed=10)
    got = {}
    for skip in (1, 0):
        with swe.SweBlock(384, 384, dx, dy, solver=swe.AUGRIE) as blk:
            blk.set_boundaries(bnd)
            blk.set_surrogate(net)
            blk.set_quiet_skip(bool(skip))
            blk.upload(h, hu, hv, b)
            _, d = blk.run(120)
            got[skip] = (blk.download(), d)
    net.close()
    for f in range(3):
        assert same_bits(got[1][0][f], got[0][0][f]), f
    assert same_bits(got[1][1], got[0][1])
