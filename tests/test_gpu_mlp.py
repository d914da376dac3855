"""GPU tests of the learned middle state (include/swe_hip_mlp.h).

Training kernels: gradients against the float64 oracle (oracle/mlp_oracle.py,
itself pinned by finite differences), relative error bound 1e-4 in norm
(FP32 sums in another order).  Inference on the matrix cores: against the
oracle's BF16 emulation, bound 1e-3 of the output scale.  Logged samples:
the exact middle state of the scaled Riemann problem (FP64 oracle), 2e-5
relative.  Surrogate solver: within the exact solver's own FP32-vs-FP64
error.  And what the trained network gains: the share of Newton edges whose
corrected guess passes the residual check, trained against untrained.
"""
import numpy as np
import pytest

from helpers import same_bits

import mlp_oracle as M

pytestmark = pytest.mark.gpu


def _random_params(seed, scale):
    return np.random.default_rng(seed).normal(0, scale, M.NPARAMS).astype(np.float32)


def test_training_gradient_matches_oracle(swe):
    rng = np.random.default_rng(3)
    p = _random_params(4, 0.25)
    samples = np.concatenate([rng.normal(0, 1, (3000, 6)), rng.normal(0, 1, (3000, 2))], axis=1).astype(np.float32)
    with swe.MiddleStateMLP(capacity=0) as mlp:
        mlp.params = p
        g, loss = mlp.gradient(samples)
    lref, gref = M.loss_and_grad(p, samples)
    assert abs(loss - lref) <= 1e-5 * lref
    assert np.linalg.norm(g - gref) <= 1e-4 * np.linalg.norm(gref)
    for name, a, b in (("W1", M.OFF_W1, M.OFF_B1), ("W2", M.OFF_W2, M.OFF_B2), ("W3", M.OFF_W3, M.OFF_B3)):
        assert np.linalg.norm(g[a:b] - gref[a:b]) <= 1e-4 * np.linalg.norm(gref[a:b]), name


def test_matrix_core_inference_matches_bf16_emulation(swe):
    rng = np.random.default_rng(8)
    p = _random_params(9, 0.2)
    x = rng.normal(0, 1, (1000, 6)).astype(np.float32)  # 15 full waves and a ragged one
    with swe.MiddleStateMLP(capacity=0) as mlp:
        mlp.params = p
        y = mlp.infer(x)
    ye = M.infer_bf16(p, x)
    yf, _ = M.forward(p, x)
    scale = np.max(np.abs(ye))
    assert np.max(np.abs(y - ye)) <= 1e-3 * scale
    assert np.max(np.abs(y - yf)) <= 3e-2 * scale  # and BF16 stays close to the FP32 network


def test_feature_function_bitwise(swe):
    rng = np.random.default_rng(2)
    n = 5000
    hL, hR = rng.uniform(0.02, 50, n), rng.uniform(0.02, 50, n)
    uL, uR = rng.normal(0, 3, n), rng.normal(0, 3, n)
    bL, bR = rng.normal(0, 5, n), rng.normal(0, 5, n)
    got = swe.mlp_features(hL, hR, uL, uR, bL, bR)
    want, _, _ = M.features(hL, hR, uL, uR, bL, bR)
    assert same_bits(got, want)


def test_logged_samples_are_exact_middle_states(swe, oracle):
    from swe_ppm_amd import scenarios

    h, hu, hv, b, dx, dy, bnd = scenarios.dam_break_islands(n=96, seed=5)
    with swe.SweBlock(96, 96, dx, dy, solver=swe.AUGRIE) as blk, swe.MiddleStateMLP(capacity=100_000) as mlp:
        blk.set_boundaries(bnd)
        blk.upload(h, hu, hv, b)
        blk.run(15, want_dts=False)
        n_all = mlp.log_block(blk, keep=1.0, seed=1)
        n_half = mlp.log_block(blk, keep=0.5, seed=2)
        s = mlp.read_samples()
    edges = 2 * 96 * 97
    assert 0.5 * edges < n_all < edges  # most edges of the wet dam break need the iteration
    assert abs(n_half - 0.5 * n_all) < 4 * np.sqrt(n_all)
    # the middle state of the scaled problem (hs = 1, cs = sqrt(g)) equals h*/hs;
    # targets are stored as corrections to the linearised middle state
    x, t = s[:, :6].astype(np.float64), (s[:, 6:] + M.linear_part(s[:, :6])).astype(np.float64)
    sg = np.sqrt(9.81)
    hm, s1, s2, kind, _ = oracle.middle_state(x[:, 0], x[:, 1], x[:, 2] * sg, x[:, 3] * sg)
    assert np.all(kind >= 2)  # only Newton edges are logged
    assert np.max(np.abs(t[:, 0] - hm) / hm) <= 2e-5
    assert np.max(np.abs(t[:, 1] - 0.5 * (s1 + s2) / sg)) <= 2e-5 * (1 + np.max(np.abs(t[:, 1])))


def _trained_network(swe, seed=0):
    """Log a few small runs, train briefly."""
    from swe_ppm_amd import scenarios

    mlp = swe.MiddleStateMLP(capacity=2_000_000, seed=seed)
    for st, solver in ((scenarios.dam_break_islands(n=256, seed=2048), swe.AUGRIE),
                       (scenarios.lts_tiling_as_block(seed=28, n=256), swe.AUGRIE),
                       (scenarios.dam_break_c1(128), swe.FWAVE)):
        h, hu, hv, b, dx, dy, bnd = st
        with swe.SweBlock(h.shape[0] - 2, h.shape[1] - 2, dx, dy, solver=solver) as blk:
            blk.set_boundaries(bnd)
            blk.upload(h, hu, hv, b)
            for k in range(24):
                mlp.log_block(blk, keep=1.0, seed=100 * seed + k, hard_only=True)
                blk.run(4, want_dts=False)
    mlp.sort_samples()  # logging order varies between runs; fixed order makes the training reproducible
    loss = mlp.train(1500, batch=4096, lr=3e-3, seed=seed)
    return mlp, loss


def _held_out_edges(oracle):
    """Newton edges of another scenario (islands seed 9, 192^2, after 40 steps)."""
    from swe_ppm_amd import scenarios

    h, hu, hv, b, dx, dy, bnd = scenarios.dam_break_islands(n=192, seed=9)
    h, hu, hv, _ = oracle.run_u(h, hu, hv, b, dx, dy, bnd, 2, 40, nthreads=8)
    hL, hR = h[:-1, 1:-1].ravel(), h[1:, 1:-1].ravel()
    uL = np.where(hL > 0, hu[:-1, 1:-1].ravel() / np.maximum(hL, 1e-30), 0)
    uR = np.where(hR > 0, hu[1:, 1:-1].ravel() / np.maximum(hR, 1e-30), 0)
    bL, bR = b[:-1, 1:-1].ravel(), b[1:, 1:-1].ravel()
    hm, s1, s2, kind, iters = oracle.middle_state(hL.astype(np.float64), hR, uL, uR)
    sel = kind >= 2
    return [a[sel].astype(np.float32) for a in (hL, hR, uL, uR, bL, bR)], iters[sel]


@pytest.mark.parametrize("nx,ny,xseg,walls", [(256, 256, 8, False), (130, 77, 16, False), (200, 150, 32, False),
                                              (190, 140, 16, True), (203, 131, 10, False), (161, 190, 12, True)])
def test_deferred_hard_edges_match_the_learned_middle_state_in_the_sweep(swe, nx, ny, xseg, walls):
    """With a network attached, the fused loop settles the edges whose
    linearised guess passes the residual check inside the sweep and queues
    the others for a second launch (network, check, exact fallback); the store
    strategy solves every edge inside its sweep.  Both give the same bits."""
    from swe_ppm_amd import scenarios, surrogate

    net = surrogate.load_default()
    h, hu, hv, b, dx, dy, bnd = scenarios.dam_break_islands(n=max(nx, ny), seed=7)
    h, hu, hv, b = (a[:nx + 2, :ny + 2].copy() for a in (h, hu, hv, b))
    if walls:  # reflecting sides: the deferred edges at the walls read the mirrored ghosts
        bnd = [swe.WALL] * 4
    out = {}
    for strategy in (1, 2):
        with swe.SweBlock(nx, ny, dx, dy, solver=swe.AUGRIE) as blk:
            blk.set_boundaries(bnd)
            blk.set_tuning(strategy, xseg)
            blk.set_surrogate(net)
            blk.upload(h, hu, hv, b)
            _, dts = blk.run(40)
            out[strategy] = (blk.download(), np.asarray(dts))
    for name, a, c in zip(("h", "hu", "hv"), out[1][0], out[2][0]):
        assert same_bits(a.view(np.uint32), c.view(np.uint32)), name  # bit patterns, signed zeros included
    assert same_bits(out[1][1], out[2][1])
    net.close()


@pytest.mark.parametrize("extreme", [False, True])
def test_deferred_hard_edges_match_the_sweep_on_random_states(swe, extreme):
    """The same agreement on uncorrelated random cell states, where every
    Riemann case and many hard edges occur (extreme: depths log-uniform over
    1e-5 .. 5000 m, velocities to 50 m/s, bed levels of +-100 m): two steps,
    the deferring fused loop against the network inside the store sweep, bit
    for bit, with no NaN."""
    from swe_ppm_amd import surrogate

    rng = np.random.default_rng(79)
    n = 320
    shape = (n + 2, n + 2)
    if extreme:
        h = np.exp(rng.uniform(np.log(1e-5), np.log(5000.0), shape)).astype(np.float32)
        vel, bed = 50.0, 100.0
    else:
        h = rng.uniform(0.2, 20.0, shape).astype(np.float32)
        vel, bed = 3.0, 2.0
    h[rng.random(h.shape) < 0.03] = 0.0
    hu = (h * rng.uniform(-vel, vel, h.shape)).astype(np.float32)
    hv = (h * rng.uniform(-vel, vel, h.shape)).astype(np.float32)
    b = rng.uniform(-bed, bed, h.shape).astype(np.float32)
    net = surrogate.load_default()
    out = {}
    for strategy in (1, 2):
        with swe.SweBlock(n, n, 1.0, 1.0, solver=swe.AUGRIE) as blk:
            blk.set_boundaries([swe.OUTFLOW] * 4)
            blk.set_tuning(strategy, 12)
            blk.set_surrogate(net)
            blk.upload(h, hu, hv, b)
            _, dts = blk.run(2)
            out[strategy] = (blk.download(), np.asarray(dts))
    net.close()
    for name, a, c in zip(("h", "hu", "hv"), out[1][0], out[2][0]):
        assert not np.any(np.isnan(a)), name
        assert same_bits(a.view(np.uint32), c.view(np.uint32)), name
    assert same_bits(out[1][1], out[2][1])


def test_batched_deferred_hard_edges_match_the_sweep_and_single_blocks(swe):
    """The batched engine with a network attached: its fused sweep defers the
    hard edges to a second launch, its store strategy solves them inside the
    sweep.  Both agree bitwise, for an ensemble of dam breaks and for 28
    connected blocks of uneven depth under local time stepping; the ensemble
    also equals its instances run one by one as single blocks."""
    from swe_ppm_amd import scenarios, surrogate

    net = surrogate.load_default()
    inst = scenarios.ensemble_instances(count=10, seed=5, nmin=40, nmax=200)
    blocks, dx, dy = scenarios.lts_tiling(n=256)
    cases = [("ensemble", inst, dict(mode=swe.BATCH_ENSEMBLE), 30),
             ("lts", blocks, dict(mode=swe.BATCH_LTS, dx=dx, dy=dy, max_local=1.0, duration=3.0), 2)]
    results = {}
    for name, bl, kw, units in cases:
        got = results[name] = {}
        for fused in (1, 0):
            with swe.SweBatch(bl, solver=swe.AUGRIE, **kw) as bt:
                bt.set_tuning(fused=fused)
                bt.set_surrogate(net)
                bt.upload(*[[b[f] for b in bl] for f in ("h", "hu", "hv", "b")])
                bt.run(units)
                got[fused] = (bt.download(), bt.stats()[0])
        for k in range(len(bl)):
            for f, field in enumerate(("h", "hu", "hv")):
                assert same_bits(got[1][0][f][k].view(np.uint32), got[0][0][f][k].view(np.uint32)), (
                    name, k, field)
        assert same_bits(got[1][1], got[0][1]), name
    for k, i in enumerate(inst):
        with swe.SweBlock(i["nx"], i["ny"], i["dx"], i["dy"], solver=swe.AUGRIE) as blk:
            blk.set_boundaries(i["bnd"])
            blk.set_tuning(1, -1)
            blk.set_surrogate(net)
            blk.upload(i["h"], i["hu"], i["hv"], i["b"])
            blk.run(30, want_dts=False)
            one = blk.download()
        for f, field in enumerate(("h", "hu", "hv")):
            assert same_bits(one[f][1:-1, 1:-1], results["ensemble"][1][0][f][k][1:-1, 1:-1]), (
                "single block", k, field)
    net.close()


def test_batched_deferred_hard_edges_on_small_and_thin_blocks(swe):
    """Deferred hard edges on instances from 1 x 1 to 70 cells per side
    (single rows and columns, tails of a few rows packed several segments to
    a wave): the fused sweep with the deferred-edge launch against the store
    strategy, bitwise."""
    from swe_ppm_amd import scenarios, surrogate

    net = surrogate.load_default()
    inst = scenarios.ensemble_instances(count=40, seed=11, nmin=1, nmax=70)
    got = {}
    for fused in (1, 0):
        with swe.SweBatch(inst, solver=swe.AUGRIE, mode=swe.BATCH_ENSEMBLE) as bt:
            bt.set_tuning(fused=fused)
            bt.set_surrogate(net)
            bt.upload(*[[i[f] for i in inst] for f in ("h", "hu", "hv", "b")])
            bt.run(25)
            got[fused] = bt.download()
    for k in range(len(inst)):
        for f, field in enumerate(("h", "hu", "hv")):
            assert same_bits(got[1][f][k], got[0][f][k]), (k, inst[k]["nx"], inst[k]["ny"], field)
    net.close()


def test_sorted_samples_do_not_depend_on_the_logging_order(swe):
    """Logging appends edges in the order the waves reach them; sort_samples
    gives every run the same rows in the same (lexicographic) order."""
    from swe_ppm_amd import scenarios

    h, hu, hv, b, dx, dy, bnd = scenarios.dam_break_islands(n=128, seed=5)
    runs = []
    for rep in range(2):
        with swe.SweBlock(128, 128, dx, dy, solver=swe.AUGRIE) as blk, \
                swe.MiddleStateMLP(capacity=100_000) as mlp:
            blk.set_boundaries(bnd)
            blk.upload(h, hu, hv, b)
            blk.run(20, want_dts=False)
            n = mlp.log_block(blk, keep=1.0, seed=3)
            raw = mlp.read_samples()
            mlp.sort_samples()
            assert mlp.count == n
            got = mlp.read_samples()
        assert n > 1000
        assert same_bits(got, raw[np.lexsort(raw.view(np.uint32).T[::-1])])  # the same rows, sorted by bits
        runs.append(got)
    assert same_bits(runs[0], runs[1])


def test_training_reduces_loss_and_trained_network_passes_the_residual_check(swe, oracle):
    edges, iters = _held_out_edges(oracle)
    trained, loss = _trained_network(swe)
    with trained, swe.MiddleStateMLP(capacity=0, seed=99) as untrained:
        assert np.mean(loss[-50:]) < 0.05 * loss[0]
        hm_t, s1_t, s2_t, acc_t = trained.middle(*edges)
        hm_u, _, _, acc_u = untrained.middle(*edges)
    exact = oracle.middle_state(*[a.astype(np.float64) for a in edges[:4]])
    # acc: 1 linearised guess kept, 2 network guess kept, 0 exact iteration
    lin = acc_t == 1
    assert same_bits(lin, acc_u == 1)  # the linearised guess does not depend on the network
    hard = ~lin
    rescue_t = float(np.mean(acc_t[hard] == 2))
    rescue_u = float(np.mean(acc_u[hard] == 2))
    rate_t = float(np.mean(acc_t > 0))
    # the gain: of the Newton edges the linearised guess does not settle, the
    # trained network settles most with one evaluation; the exact path needs
    # its full iteration for them (mean iterations stated beside)
    print(f"linearised guess kept {np.mean(lin):.4f}; of the rest the network keeps: trained {rescue_t:.4f}, "
          f"untrained {rescue_u:.4f}; overall without iteration {rate_t:.4f}; exact Newton iterations mean "
          f"{np.mean(iters):.2f} (hard edges {np.mean(iters[hard]):.2f})")
    assert rate_t >= 0.9
    assert rescue_t >= rescue_u + 0.3
    # accepted middle depths are as close to the FP64 root as the exact FP32 iteration's
    rel = np.abs(hm_t - exact[0]) / exact[0]
    assert np.max(rel) <= 2e-5


def test_surrogate_block_matches_exact_solver_to_its_fp32_error(swe, oracle):
    from swe_ppm_amd import scenarios

    trained, _ = _trained_network(swe, seed=1)
    h, hu, hv, b, dx, dy, bnd = scenarios.dam_break_islands(n=128, seed=3)
    n = 60
    out = {}
    with trained:
        for name, net in (("exact", None), ("surrogate", trained)):
            with swe.SweBlock(128, 128, dx, dy, solver=swe.AUGRIE) as blk:
                blk.set_boundaries(bnd)
                blk.upload(h, hu, hv, b)
                blk.set_surrogate(net)
                blk.run(n, want_dts=False)
                out[name] = blk.download()
        # per-edge solver with the learned middle state against the exact one
        rng = np.random.default_rng(4)
        k = 20000
        hL, hR = rng.uniform(0.5, 20, k).astype(np.float32), rng.uniform(0.5, 20, k).astype(np.float32)
        huL, huR = (hL * rng.normal(0, 2, k)).astype(np.float32), (hR * rng.normal(0, 2, k)).astype(np.float32)
        bL, bR = rng.normal(0, 0.2, k).astype(np.float32), rng.normal(0, 0.2, k).astype(np.float32)
        ge = swe.solve_edges(swe.AUGRIE, hL, hR, huL, huR, bL, bR)
        gs = trained.solve_edges(hL, hR, huL, huR, bL, bR)
    for a, c in zip(ge, gs):
        assert np.max(np.abs(a - c)) <= 1e-4 * (1 + np.max(np.abs(a)))
    # both FP32 runs against the FP64 oracle run of the same scheme
    ref = oracle.run_u(h, hu, hv, b, dx, dy, bnd, 2, n, nthreads=8, dtype=np.float64)
    err = {}
    for name, (gh, ghu, ghv) in out.items():
        err[name] = max(np.max(np.abs(gh[1:-1, 1:-1] - ref[0][1:-1, 1:-1])) / np.max(np.abs(ref[0])),
                        np.max(np.abs(ghu[1:-1, 1:-1] - ref[1][1:-1, 1:-1])) / np.max(np.abs(ref[1])))
    d = np.max(np.abs(out["exact"][0] - out["surrogate"][0])) / np.max(np.abs(out["exact"][0]))
    print(f"rel Linf vs FP64: exact {err['exact']:.3e}, surrogate {err['surrogate']:.3e}; between {d:.3e}")
    assert err["surrogate"] <= 1.5 * err["exact"] + 1e-5


def test_destroying_an_attached_network_is_refused(swe):
    """A block or batch with a network attached reads its weights on the
    device: destroying the network is refused (SWE_HIP_ESTATE) until every
    attachment is gone, by detaching or by destroying the holder.  Creates,
    attaches and frees only; no step runs."""
    from swe_ppm_amd import scenarios, surrogate

    net = surrogate.load_default()
    blk = swe.SweBlock(64, 64, 1.0, 1.0, solver=swe.AUGRIE)
    blk.set_surrogate(net)
    with pytest.raises(swe.SweHipError, match="attached"):
        net.close()
    blk.set_surrogate(None)
    inst = scenarios.ensemble_instances(count=2, seed=3, nmin=16, nmax=20)
    bt = swe.SweBatch(inst, solver=swe.AUGRIE, mode=swe.BATCH_ENSEMBLE)
    bt.set_surrogate(net)
    blk.set_surrogate(net)
    with pytest.raises(swe.SweHipError, match="2 block"):
        net.close()
    bt.close()
    with pytest.raises(swe.SweHipError, match="1 block"):
        net.close()
    blk.close()
    net.close()  # now allowed
