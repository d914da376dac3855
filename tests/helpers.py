"""Shared helpers of the tests (comparison metrics, random edge states)."""
from __future__ import annotations

import numpy as np


def same_bits(a, b) -> bool:
    """Equality for the parity tests: float32 arrays compare by their bit
    patterns (so +0 and -0 differ), other arrays by value."""
    a, b = np.asarray(a), np.asarray(b)
    if a.dtype == np.float32 and b.dtype == np.float32:
        return a.shape == b.shape and np.array_equal(np.ascontiguousarray(a).view(np.uint32),
                                                     np.ascontiguousarray(b).view(np.uint32))
    return np.array_equal(a, b)


def rel_linf(a: np.ndarray, b: np.ndarray) -> float:
    """max |a - b| / max |b| over the interior (ghost cells excluded)."""
    ai = np.asarray(a, np.float64)[1:-1, 1:-1]
    bi = np.asarray(b, np.float64)[1:-1, 1:-1]
    den = float(np.max(np.abs(bi)))
    return float(np.max(np.abs(ai - bi))) / (den if den > 0 else 1.0)


def random_edges(n: int, seed: int, dry_fraction: float = 0.15, bathy: bool = True):
    """Edge states mixing wet-wet, wet-dry, dry-dry, shocks, rarefactions,
    sonic and still-water cases, and bathymetry steps."""
    rng = np.random.default_rng(seed)
    hL = rng.uniform(0.0, 20.0, n)
    hR = rng.uniform(0.0, 20.0, n)
    # some near-equal depths (weak waves) and deep cases
    k = rng.random(n)
    hR = np.where(k < 0.2, hL * rng.uniform(0.98, 1.02, n), hR)
    hL = np.where(k > 0.9, rng.uniform(100.0, 4000.0, n), hL)
    hR = np.where(k > 0.9, hL * rng.uniform(0.5, 1.5, n), hR)
    # dry cells
    d = rng.random(n)
    hL = np.where(d < dry_fraction / 2, rng.uniform(0.0, 0.02, n), hL)
    hR = np.where((d >= dry_fraction / 2) & (d < dry_fraction), rng.uniform(0.0, 0.02, n), hR)
    hR = np.where(d > 0.98, rng.uniform(0.0, 0.009, n), hR)
    hL = np.where(d > 0.98, rng.uniform(0.0, 0.009, n), hL)
    uL = rng.normal(0.0, 3.0, n)
    uR = rng.normal(0.0, 3.0, n)
    s = rng.random(n)
    uL = np.where(s < 0.1, np.sqrt(9.81 * hL) * rng.uniform(0.9, 1.1, n), uL)  # near-sonic
    uR = np.where(s > 0.9, -np.sqrt(9.81 * hR) * rng.uniform(0.9, 1.1, n), uR)
    huL = hL * uL
    huR = hR * uR
    if bathy:
        bL = rng.normal(0.0, 2.0, n)
        bR = np.where(rng.random(n) < 0.5, bL, bL + rng.normal(0.0, 3.0, n))
        big = rng.random(n) < 0.05
        bR = np.where(big, bL + rng.uniform(10.0, 30.0, n), bR)  # walls
    else:
        bL = np.zeros(n)
        bR = np.zeros(n)
    return [np.asarray(a, np.float32) for a in (hL, hR, huL, huR, bL, bR)]
