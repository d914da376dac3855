"""Pinning of the network oracle (oracle/mlp_oracle.py) on the CPU: its
analytic gradients against central finite differences, its BF16 rounding
against known bit patterns, its feature function against the definition."""
import numpy as np

import mlp_oracle as M


def _random_params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    return (rng.normal(0, scale, M.NPARAMS)).astype(np.float64)


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(1)
    p = _random_params(2)
    samples = np.concatenate([rng.normal(0, 1, (40, 6)), rng.normal(0, 1, (40, 2))], axis=1)
    loss, g = M.loss_and_grad(p, samples)
    eps = 1e-6
    # every bias, and a spread of weights from each layer
    idx = list(range(M.OFF_B1, M.OFF_W2)) + list(range(M.OFF_B2, M.OFF_W3)) + list(range(M.OFF_B3, M.NPARAMS))
    idx += list(rng.choice(np.arange(M.OFF_W1, M.OFF_B1), 40, replace=False))
    idx += list(rng.choice(np.arange(M.OFF_W2, M.OFF_B2), 80, replace=False))
    idx += list(rng.choice(np.arange(M.OFF_W3, M.OFF_B3), 40, replace=False))
    for i in idx:
        q = p.copy()
        q[i] += eps
        lp = M.loss_and_grad(q, samples)[0]
        q[i] -= 2 * eps
        lm = M.loss_and_grad(q, samples)[0]
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_loss_is_mean_squared_error():
    p = np.zeros(M.NPARAMS)
    p[M.OFF_B3] = 1.0
    p[M.OFF_B3 + 1] = -2.0
    samples = np.array([[0, 0, 0, 0, 0, 0, 3.0, 0.0], [0, 0, 0, 0, 0, 0, 1.0, -2.0]])
    loss, g = M.loss_and_grad(p, samples)
    # errors (-2, -2) and (0, 0): mean of 8 and 0
    assert loss == 4.0
    assert g[M.OFF_B3] == -2.0 and g[M.OFF_B3 + 1] == -2.0


def test_bf16_rounding_bit_patterns():
    vals = np.array([1.0, 1.0 + 2**-8, 1.0 + 3 * 2**-8, 1.0 + 2**-9, -2.5, 3.0e38, 1e-40], np.float32)
    got = M.bf16(vals).view(np.uint32) >> 16
    # 1+2^-8 is a tie -> even (1.0); 1+3*2^-8 tie -> even upward; 1+2^-9 rounds down
    assert list(got) == [0x3F80, 0x3F80, 0x3F82, 0x3F80, 0xC020, 0x7F62, 0x0001]


def test_features_follow_their_definition():
    hL, hR, uL, uR, bL, bR = 2.0, 6.0, 1.0, -3.0, 0.5, 9.5
    x, hs, cs = M.features([hL], [hR], [uL], [uR], [bL], [bR])
    hs_ = 4.0
    cs_ = np.sqrt(9.81 * 4.0)
    want = [hL / hs_, hR / hs_, uL / cs_, uR / cs_, 1.0, (uR - uL) / cs_]
    assert np.allclose(x[0], want, rtol=2e-7)


def test_bf16_inference_emulation_tracks_the_float_network():
    rng = np.random.default_rng(5)
    p = _random_params(6, scale=0.2)
    x = rng.normal(0, 1, (500, 6)).astype(np.float32)
    y, _ = M.forward(p, x)
    yb = M.infer_bf16(p, x)
    scale = np.max(np.abs(y))
    assert np.max(np.abs(y - yb)) <= 0.03 * scale  # BF16 weights: a few 1e-3 relative per layer
