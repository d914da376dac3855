"""Scenario sampling and the seeded workloads.

initScenario (src/blocks/SWE_Block.hh:650-682) samples a scenario at cell
centres x = ox + (j - 0.5) dx, y = oy + (i - 0.5) dy over all (nx+2) x (ny+2)
cells, ghosts included, and sets hu = u h, hv = v h; afterwards the boundary
conditions and the bathymetry ghosts are applied.  The functions here return
float32 arrays in the Float2D layout (shape (nx+2, ny+2), a[x, y]).

Workloads:
  * radial_dam_break       the reference's SWE_RadialDamBreakScenario
                           (SWE_simple_scenarios.hh:77-108)
  * dam_break_c1           256^2 f-wave dam break centred in [0,1000]^2, b = 0
  * dam_break_islands      2048^2 dam break + 16 seeded Gaussian bumps, exactly
                           3 of them emerging as dry islands
  * ensemble_instances     512 seeded dam-break instances of uneven size
  * lts_tiling             28 blocks of uneven size tiling 2048^2, per-block
                           still-water depth
  * trench_hump            4096^2 trench (10 m .. 4000 m) with a 2 m hump

The seeded workloads are parameter sets (`*_spec`, a dict per block) that
sample_seeded evaluates at the cell centres; swe_hip_init_seeded evaluates
the same parameters on the device with the same single-precision operations
(each exponential in double, rounded once), so both give the same bits.
"""
from __future__ import annotations

import math

import numpy as np

OUTFLOW, WALL = 0, 1


def cell_centres(nx: int, ny: int, dx: float, dy: float, ox: float = 0.0, oy: float = 0.0):
    """Cell-centre coordinates incl. ghosts as the reference computes them:
    (float)originX + (j - 0.5) * dx evaluated in double, stored as float."""
    j = np.arange(nx + 2, dtype=np.float64)
    i = np.arange(ny + 2, dtype=np.float64)
    x = (np.float64(np.float32(ox)) + (j - 0.5) * np.float64(np.float32(dx))).astype(np.float32)
    y = (np.float64(np.float32(oy)) + (i - 0.5) * np.float64(np.float32(dy))).astype(np.float32)
    X = np.broadcast_to(x[:, None], (nx + 2, ny + 2))
    Y = np.broadcast_to(y[None, :], (nx + 2, ny + 2))
    return X, Y


def apply_bc_bathymetry(b: np.ndarray, bnd) -> None:
    """applyBoundaryBathymetry (SWE_Block.hh:724-749), in place."""
    nx, ny = b.shape[0] - 2, b.shape[1] - 2
    ow = lambda t: t in (OUTFLOW, WALL)
    if ow(bnd[0]):
        b[0, :] = b[1, :]
    if ow(bnd[1]):
        b[nx + 1, :] = b[nx, :]
    if ow(bnd[2]):
        b[:, 0] = b[:, 1]
    if ow(bnd[3]):
        b[:, ny + 1] = b[:, ny]
    b[0, 0] = b[1, 1]
    b[0, ny + 1] = b[1, ny]
    b[nx + 1, 0] = b[nx, 1]
    b[nx + 1, ny + 1] = b[nx, ny]


def apply_bc(h, hu, hv, bnd) -> None:
    """applyBoundaryConditions (SWE_Block.hh:757-911), in place."""
    nx, ny = h.shape[0] - 2, h.shape[1] - 2
    for side, (g, s) in enumerate(((0, 1), (nx + 1, nx))):
        t = bnd[side]
        if t in (OUTFLOW, WALL):
            h[g, 1:ny + 1] = h[s, 1:ny + 1]
            hu[g, 1:ny + 1] = -hu[s, 1:ny + 1] if t == WALL else hu[s, 1:ny + 1]
            hv[g, 1:ny + 1] = hv[s, 1:ny + 1]
    for k, (g, s) in enumerate(((0, 1), (ny + 1, ny))):
        t = bnd[2 + k]
        if t in (OUTFLOW, WALL):
            h[1:nx + 1, g] = h[1:nx + 1, s]
            hu[1:nx + 1, g] = hu[1:nx + 1, s]
            hv[1:nx + 1, g] = -hv[1:nx + 1, s] if t == WALL else hv[1:nx + 1, s]
    for a in (h, hu, hv):
        a[0, 0] = a[1, 1]
        a[0, ny + 1] = a[1, ny]
        a[nx + 1, 0] = a[nx, 1]
        a[nx + 1, ny + 1] = a[nx, ny]


def finish(h, b, bnd, u=None, v=None):
    """hu = u h, hv = v h, then boundary conditions (SWE_Block.hh:671-681)."""
    h = np.ascontiguousarray(h, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    hu = np.zeros_like(h) if u is None else (np.asarray(u, np.float32) * h).astype(np.float32)
    hv = np.zeros_like(h) if v is None else (np.asarray(v, np.float32) * h).astype(np.float32)
    apply_bc(h, hu, hv, bnd)
    apply_bc_bathymetry(b, bnd)
    return h, hu, hv, b


def _radius(X, Y, cx, cy):
    dx = X - np.float32(cx)
    dy = Y - np.float32(cy)
    return np.sqrt(dx * dx + dy * dy)


def radial_dam_break(nx: int, ny: int):
    """SWE_RadialDamBreakScenario on [0,1000]^2: b = 10, h = 15 within 100 m
    of (250, 250), 10 elsewhere, outflow everywhere."""
    dx, dy = np.float32(1000.0 / nx), np.float32(1000.0 / ny)
    X, Y = cell_centres(nx, ny, dx, dy)
    h = np.where(_radius(X, Y, 250.0, 250.0) < np.float32(100.0), np.float32(15.0), np.float32(10.0))
    b = np.full(h.shape, 10.0, np.float32)
    bnd = [OUTFLOW] * 4
    return (*finish(h, b, bnd), float(dx), float(dy), bnd)


def dam_break_c1(n: int = 256):
    """256^2 f-wave workload: h = 15 within 100 m of (500, 500), 10 elsewhere,
    flat b = 0, outflow boundaries, domain [0,1000]^2."""
    dx = np.float32(1000.0 / n)
    h, b, _ = sample_seeded(dam_break_c1_spec(), n, n, dx, dx)
    bnd = [OUTFLOW] * 4
    return (*finish(h, b, bnd), float(dx), float(dx), bnd)


def dam_break_c1_spec():
    return dam_spec(500.0, 500.0, 100.0, 15.0, 10.0)


SEEDED_DAM, SEEDED_TRENCH = 0, 1  # include/swe_hip.h
SEEDED_MAX_BUMPS = 32


def _f32(v) -> float:
    return float(np.float32(v))


def _exp32(q):
    """exp of a float32 array, evaluated in double and rounded once."""
    return np.exp(np.asarray(q, np.float32).astype(np.float64)).astype(np.float32)


def dam_spec(dam_x, dam_y, dam_radius, eta_in, eta_out, centres=(), amps=(), sigmas=()):
    """A dam (free surface eta_in within dam_radius of (dam_x, dam_y), eta_out
    elsewhere) over Gaussian bumps A exp(-r^2 / (2 sigma^2)); h = max(0, eta - b)."""
    bumps = [(_f32(cx), _f32(cy), _f32(a), _f32(2.0 * s * s)) for (cx, cy), a, s in zip(centres, amps, sigmas)]
    if len(bumps) > SEEDED_MAX_BUMPS:
        raise ValueError(f"at most {SEEDED_MAX_BUMPS} bumps")
    return dict(kind=SEEDED_DAM, dam_x=_f32(dam_x), dam_y=_f32(dam_y), dam_radius=_f32(dam_radius),
                eta_in=_f32(eta_in), eta_out=_f32(eta_out), bumps=bumps)


def sample_seeded(spec, nx: int, ny: int, dx: float, dy: float, ox: float = 0.0, oy: float = 0.0):
    """h, b (and, for a dam, the free surface eta) of a seeded workload at
    every cell centre, ghosts included, before boundary conditions."""
    X, Y = cell_centres(nx, ny, dx, dy, ox, oy)
    if spec["kind"] == SEEDED_TRENCH:
        ax = X - np.float32(spec["trench_x"])
        depth = np.float32(10.0) + np.float32(3990.0) * _exp32(-(ax * ax) / np.float32(spec["trench_w2"]))
        hx = X - np.float32(spec["hump_x"])
        hy = Y - np.float32(spec["hump_y"])
        eta = np.float32(2.0) * _exp32(-(hx * hx + hy * hy) / np.float32(spec["hump_r2"]))
        return (depth + eta).astype(np.float32), (-depth).astype(np.float32), None
    b = np.zeros(X.shape, np.float32)
    for bx, by, a, s2 in spec["bumps"]:
        ex = X - np.float32(bx)
        ey = Y - np.float32(by)
        b = (b + np.float32(a) * _exp32(-(ex * ex + ey * ey) / np.float32(s2))).astype(np.float32)
    eta = np.where(_radius(X, Y, spec["dam_x"], spec["dam_y"]) < np.float32(spec["dam_radius"]),
                   np.float32(spec["eta_in"]), np.float32(spec["eta_out"]))
    h = np.maximum(np.float32(0.0), eta - b).astype(np.float32)
    return h, b, eta


def island_bumps(seed: int, length: float = 1000.0, nbumps: int = 16, n_islands: int = 3,
                 dam_centre=(500.0, 500.0), dam_radius=100.0, eta_in=15.0, eta_out=10.0):
    """Draw bump parameters (A ~ U(2,12), sigma ~ U(20,80), centres ~ U(domain))
    until exactly n_islands bump peaks stand above the local free surface."""
    rng = np.random.default_rng(seed)
    for _ in range(10000):
        centres = rng.uniform(0.0, length, size=(nbumps, 2))
        amps = rng.uniform(2.0, 12.0, size=nbumps)
        sigmas = rng.uniform(20.0, 80.0, size=nbumps)
        cx = centres[:, 0][:, None]
        cy = centres[:, 1][:, None]
        # total bathymetry at each bump centre
        r2 = (cx - centres[:, 0][None, :]) ** 2 + (cy - centres[:, 1][None, :]) ** 2
        btot = (amps[None, :] * np.exp(-r2 / (2.0 * sigmas[None, :] ** 2))).sum(axis=1)
        rd = np.hypot(centres[:, 0] - dam_centre[0], centres[:, 1] - dam_centre[1])
        eta = np.where(rd < dam_radius, eta_in, eta_out)
        if int((btot > eta).sum()) == n_islands:
            return centres, amps, sigmas
    raise RuntimeError("could not draw an island configuration")


def dam_break_islands(n: int = 2048, seed: int = 2048, length: float = 1000.0, surface_noise: float = 0.0):
    """Dam break (15 m within 100 m of the centre, 10 m elsewhere) over 16
    seeded Gaussian bumps, exactly 3 of which emerge as dry islands;
    h = max(0, eta - b); outflow boundaries.  surface_noise > 0 adds seeded
    uniform noise of that amplitude (m) to eta everywhere, so that no edge
    starts at rest (the fully disturbed form of the workload)."""
    dx = np.float32(length / n)
    h, b, eta = sample_seeded(dam_break_islands_spec(seed, length), n, n, dx, dx)
    if surface_noise > 0.0:
        rng = np.random.default_rng(seed + 1)
        eta = (eta + rng.uniform(-surface_noise, surface_noise, eta.shape).astype(np.float32)).astype(np.float32)
    h = np.maximum(np.float32(0.0), eta - b).astype(np.float32)
    bnd = [OUTFLOW] * 4
    return (*finish(h, b, bnd), float(dx), float(dx), bnd)


def dam_break_islands_spec(seed: int = 2048, length: float = 1000.0):
    centres, amps, sigmas = island_bumps(seed, length)
    return dam_spec(length / 2, length / 2, 100.0, 15.0, 10.0, centres, amps, sigmas)


def ensemble_instances(count: int = 512, seed: int = 512, nmin: int = 64, nmax: int = 384, length: float = 1000.0):
    """Independent dam-break instances of uneven size: NX, NY ~ U{nmin..nmax},
    dam radius ~ U(50,200) m, inner/outer depth ~ U(5,20) / U(1,10) m, four
    random bumps (A ~ U(0.5,4) m, sigma ~ U(20,80) m).  Returns a list of
    dicts with nx, ny, dx, dy, bnd, h, hu, hv, b and the parameter set (spec)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        nx, ny = (int(v) for v in rng.integers(nmin, nmax + 1, size=2))
        radius = rng.uniform(50.0, 200.0)
        h_in, h_out = rng.uniform(5.0, 20.0), rng.uniform(1.0, 10.0)
        centres = rng.uniform(0.0, length, size=(4, 2))
        amps = rng.uniform(0.5, 4.0, size=4)
        sigmas = rng.uniform(20.0, 80.0, size=4)
        dx, dy = np.float32(length / nx), np.float32(length / ny)
        spec = dam_spec(length / 2, length / 2, radius, h_in, h_out, centres, amps, sigmas)
        h, b, _ = sample_seeded(spec, nx, ny, dx, dy)
        bnd = [OUTFLOW] * 4
        hh, hu, hv, bb = finish(h, b, bnd)
        out.append(dict(nx=nx, ny=ny, dx=float(dx), dy=float(dy), bnd=bnd, h=hh, hu=hu, hv=hv, b=bb, spec=spec))
    return out


def _split_sizes(rng, total: int, parts: int, lo: int, hi: int, quantum: int = 128):
    for _ in range(100000):
        cuts = rng.integers(lo // quantum, hi // quantum + 1, size=parts) * quantum
        if cuts.sum() == total:
            return [int(c) for c in cuts]
    raise RuntimeError("no split found")


def lts_tiling(seed: int = 28, n: int = 2048, cols: int = 4, rows: int = 7, length: float = 1000.0,
               depth_lo: float = 2.0, depth_hi: float = 40.0):
    """28 blocks (cols x rows) of uneven size tiling n x n: column widths in
    [n/16, n/2] and row heights in [n/16, n/4], multiples of n/16 (at n = 2048:
    128..1024 x 128..512), each block at its own still-water depth
    ~ U(depth_lo, depth_hi) over flat b = 0.
    Neighbours share whole edges, as the reference's in-rank ghost copy
    requires (SWE_DimensionalSplittingMPIOverdecomp.cpp:273-304).
    Block k = c*rows + r; its neighbours follow getNeighbours
    (swe_mpi_overdecomp.cpp:41-49)."""
    rng = np.random.default_rng(seed)
    q = n // 16
    widths = _split_sizes(rng, n, cols, q, n // 2, quantum=q)
    heights = _split_sizes(rng, n, rows, q, n // 4, quantum=q)
    depths = rng.uniform(depth_lo, depth_hi, size=cols * rows)
    dx = np.float32(length / n)
    blocks = []
    xo = np.cumsum([0] + widths[:-1])
    yo = np.cumsum([0] + heights[:-1])
    for c in range(cols):
        for r in range(rows):
            k = c * rows + r
            nx, ny = widths[c], heights[r]
            ox, oy = float(np.float32(xo[c] * dx)), float(np.float32(yo[r] * dx))
            nbr = [k - rows if c > 0 else -1, k + rows if c < cols - 1 else -1, k - 1 if r > 0 else -1,
                   k + 1 if r < rows - 1 else -1]
            bnd = [4 if m >= 0 else OUTFLOW for m in nbr]  # CONNECT_WITHIN_RANK = 4
            spec = dam_spec(0.0, 0.0, 0.0, depths[k], depths[k])  # still water at the block's depth, b = 0
            h, b, _ = sample_seeded(spec, nx, ny, dx, dx, ox, oy)
            hh, hu, hv, bb = finish(h, b, bnd)
            blocks.append(dict(nx=nx, ny=ny, ox=ox, oy=oy, bnd=bnd, nbr=nbr, h=hh, hu=hu, hv=hv, b=bb, spec=spec))
    return blocks, float(dx), float(dx)


def lts_tiling_as_block(seed: int = 28, n: int = 2048):
    """The state of lts_tiling as one n x n block (each tile at its own still
    water depth), outflow boundaries."""
    blocks, dx, dy = lts_tiling(seed=seed, n=n)
    h = np.zeros((n + 2, n + 2), np.float32)
    for bl in blocks:
        x0 = int(round(bl["ox"] / dx))
        y0 = int(round(bl["oy"] / dy))
        h[1 + x0:1 + x0 + bl["nx"], 1 + y0:1 + y0 + bl["ny"]] = bl["h"][1:-1, 1:-1]
    b = np.zeros_like(h)
    bnd = [OUTFLOW] * 4
    return (*finish(h, b, bnd), dx, dy, bnd)


def trench_hump(n: int = 4096, seed: int = 4096, length: float = 100000.0):
    """Trench: still-water depth 10 m far away, 4000 m on a seeded axis, with
    a Gaussian cross-section; a 2 m Gaussian sea-surface hump at a seeded
    position; sea level 0, b = -depth, outflow boundaries."""
    dx = np.float32(length / n)
    h, b, _ = sample_seeded(trench_hump_spec(seed, length), n, n, dx, dx)
    bnd = [OUTFLOW] * 4
    return (*finish(h, b, bnd), float(dx), float(dx), bnd)


def trench_hump_spec(seed: int = 4096, length: float = 100000.0):
    rng = np.random.default_rng(seed)
    axis_x = rng.uniform(0.35, 0.65) * length
    width = rng.uniform(0.04, 0.08) * length
    hx, hy = rng.uniform(0.3, 0.7, size=2) * length
    hr = rng.uniform(0.01, 0.03) * length
    return dict(kind=SEEDED_TRENCH, trench_x=_f32(axis_x), trench_w2=_f32(2.0 * width * width), hump_x=_f32(hx),
                hump_y=_f32(hy), hump_r2=_f32(2.0 * hr * hr))


def interior_mass(h: np.ndarray, dx: float, dy: float) -> float:
    return float(np.sum(h[1:-1, 1:-1], dtype=np.float64) * dx * dy)
