"""Python binding of libswe_hip.so, the MI355X SWE block library.

The library's C ABI (include/swe_hip.h) is the drop-in boundary; this module
binds it with ctypes for the tests and the bench.  There is no CPU fallback:
if the HIP library is missing or fails to load, import raises.

Host arrays use the reference's Float2D layout (src/tools/Float2D.hh:74-80):
numpy float32 arrays of shape (nx+2, ny+2) in C order, so that a[x, y] sits at
x*(ny+2) + y with ghost cells at index 0 and n+1.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_ROOT = Path(__file__).resolve().parent.parent
LIB_PATH = PKG_ROOT / "libswe_hip.so"

# solvers (WAVE_PROPAGATION_SOLVER ids of SWE_DimensionalSplittingMpi.hh:42-50, + Hybrid)
HLLE, FWAVE, AUGRIE, HYBRID = 0, 1, 2, 3
# analytic scenarios of swe_hip_init_scenario (include/swe_hip.h)
(SCENARIO_HALF_DOMAIN_DRY, SCENARIO_RADIAL_DAM_BREAK, SCENARIO_RADIAL_BATHYMETRY_DAM_BREAK,
 SCENARIO_BATHYMETRY_DAM_BREAK, SCENARIO_SEA_AT_REST, SCENARIO_SPLASHING_POOL, SCENARIO_SPLASHING_CONE,
 SCENARIO_CENTRED_DAM_BREAK) = range(8)
SOLVER_NAMES = {"hlle": HLLE, "fwave": FWAVE, "augrie": AUGRIE, "hybrid": HYBRID}
# step flavours
FLAVOUR_U, FLAVOUR_S = 0, 1
# BoundaryType / Boundary (src/types/Boundary.hh:5-28)
OUTFLOW, WALL, INFLOW, CONNECT, CONNECT_WITHIN_RANK, PASSIVE = range(6)
BND_LEFT, BND_RIGHT, BND_BOTTOM, BND_TOP = range(4)


class SweHipError(RuntimeError):
    pass


class _Grid(C.Structure):
    """swe_hip_grid (include/swe_hip.h)"""
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("x0", C.c_double), ("y0", C.c_double), ("dx", C.c_double),
                ("dy", C.c_double), ("values", C.POINTER(C.c_float))]


_MAX_BUMPS = 32  # SWE_SEEDED_MAX_BUMPS


class _Seeded(C.Structure):
    """swe_hip_seeded (include/swe_hip.h)"""
    _fields_ = [("kind", C.c_int), ("nbumps", C.c_int), ("dam_x", C.c_float), ("dam_y", C.c_float),
                ("dam_radius", C.c_float), ("eta_in", C.c_float), ("eta_out", C.c_float), ("trench_x", C.c_float),
                ("trench_w2", C.c_float), ("hump_x", C.c_float), ("hump_y", C.c_float), ("hump_r2", C.c_float),
                ("bump_x", C.c_float * _MAX_BUMPS), ("bump_y", C.c_float * _MAX_BUMPS),
                ("bump_a", C.c_float * _MAX_BUMPS), ("bump_s2", C.c_float * _MAX_BUMPS)]


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        raise SweHipError(
            f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(str(LIB_PATH))
    fp = C.POINTER(C.c_float)
    gp = C.POINTER(_Grid)
    ip = C.POINTER(C.c_int)
    vp = C.c_void_p
    sig = {
        "swe_hip_last_error": (C.c_char_p, []),
        "swe_hip_version": (C.c_char_p, []),
        "swe_hip_create": (C.c_int, [C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int,
                                     C.c_int, C.POINTER(vp)]),
        "swe_hip_destroy": (C.c_int, [vp]),
        "swe_hip_upload": (C.c_int, [vp, fp, fp, fp, fp]),
        "swe_hip_download": (C.c_int, [vp, fp, fp, fp, fp]),
        "swe_hip_set_boundary": (C.c_int, [vp, C.c_int, C.c_int]),
        "swe_hip_apply_bc": (C.c_int, [vp]),
        "swe_hip_apply_bc_bathymetry": (C.c_int, [vp]),
        "swe_hip_init_grid_scenario": (C.c_int, [vp, gp, gp]),
        "swe_hip_init_scenario": (C.c_int, [vp, C.c_int]),
        "swe_hip_init_seeded": (C.c_int, [vp, C.POINTER(_Seeded)]),
        "swe_hip_fluxes": (C.c_int, [vp, fp]),
        "swe_hip_update": (C.c_int, [vp, C.c_float]),
        "swe_hip_run": (C.c_int, [vp, C.c_int, fp, fp]),
        "swe_hip_run_timed": (C.c_int, [vp, C.c_int, fp, fp, fp, C.POINTER(C.c_int)]),
        "swe_hip_autotune": (C.c_int, [vp, C.c_int, fp]),
        "swe_hip_synchronize": (C.c_int, [vp]),
        "swe_hip_device_pointers": (C.c_int, [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), C.POINTER(vp),
                                              C.POINTER(C.c_int)]),
        "swe_hip_solve_edges": (C.c_int, [C.c_int, C.c_int, C.c_int64] + [fp] * 11),
        "swe_hip_set_math": (C.c_int, [vp, C.c_int]),
        "swe_hip_set_quiet_skip": (C.c_int, [vp, C.c_int]),
        "swe_hip_set_lookahead": (C.c_int, [vp, C.c_int]),
        "swe_hip_copy_layer": (C.c_int, [vp, C.c_int, vp, C.c_int]),
        "swe_hip_snapshot": (C.c_int, [vp, C.POINTER(C.c_int)]),
        "swe_hip_snapshot_wait": (C.c_int, [vp, C.c_int] + [C.POINTER(fp)] * 3),
        "swe_hip_snapshot_release": (C.c_int, [vp, C.c_int]),
        "swe_hip_set_tuning": (C.c_int, [vp, C.c_int, C.c_int]),
        "swe_hip_get_tuning": (C.c_int, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "swe_hip_batch_create": (C.c_int, [C.c_int, ip, ip, fp, fp, ip, ip, C.c_int, C.c_int, C.c_float, C.c_float,
                                           C.POINTER(vp)]),
        "swe_hip_batch_destroy": (C.c_int, [vp]),
        "swe_hip_batch_upload": (C.c_int, [vp, fp, fp, fp, fp]),
        "swe_hip_batch_download": (C.c_int, [vp, fp, fp, fp, fp]),
        "swe_hip_batch_run": (C.c_int, [vp, C.c_int, fp, ip, fp]),
        "swe_hip_batch_stats": (C.c_int, [vp, C.POINTER(C.c_int64), fp]),
        "swe_hip_batch_set_tuning": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
        "swe_hip_batch_get_tuning": (C.c_int, [vp, ip, ip, ip, ip, ip]),
        "swe_hip_batch_autotune": (C.c_int, [vp, C.c_int, fp, ip]),
        "swe_hip_mlp_create": (C.c_int, [C.c_int64, C.c_uint64, C.POINTER(vp)]),
        "swe_hip_mlp_destroy": (C.c_int, [vp]),
        "swe_hip_mlp_get_params": (C.c_int, [vp, fp]),
        "swe_hip_mlp_set_params": (C.c_int, [vp, fp]),
        "swe_hip_mlp_log_block": (C.c_int, [vp, vp, C.c_float, C.c_uint64, C.c_int, C.POINTER(C.c_int64)]),
        "swe_hip_mlp_add_samples": (C.c_int, [vp, C.c_int64, fp]),
        "swe_hip_mlp_read_samples": (C.c_int, [vp, C.c_int64, C.c_int64, fp]),
        "swe_hip_mlp_sample_count": (C.c_int, [vp, C.POINTER(C.c_int64)]),
        "swe_hip_mlp_clear_samples": (C.c_int, [vp]),
        "swe_hip_mlp_train": (C.c_int, [vp, C.c_int, C.c_int, C.c_float, C.c_uint64, fp]),
        "swe_hip_mlp_gradient": (C.c_int, [vp, C.c_int64, fp, fp, fp]),
        "swe_hip_mlp_infer": (C.c_int, [vp, C.c_int64, fp, fp]),
        "swe_hip_mlp_features": (C.c_int, [C.c_int64] + [fp] * 7),
        "swe_hip_mlp_middle": (C.c_int, [vp, C.c_int64] + [fp] * 9 + [ip]),
        "swe_hip_mlp_solve_edges": (C.c_int, [vp, C.c_int64] + [fp] * 11),
        "swe_hip_set_surrogate": (C.c_int, [vp, vp]),
        "swe_hip_batch_set_surrogate": (C.c_int, [vp, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise SweHipError(f"{what} failed ({rc}): {lib.swe_hip_last_error().decode()}")


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _f32(a, shape) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


class SweBlock:
    """One SWE block on the GPU (the state lives in HBM)."""

    def __init__(self, nx: int, ny: int, dx: float, dy: float, ox: float = 0.0, oy: float = 0.0,
                 solver: int = AUGRIE, flavour: int = FLAVOUR_U):
        self.nx, self.ny, self.dx, self.dy = int(nx), int(ny), float(dx), float(dy)
        self.solver, self.flavour = solver, flavour
        h = C.c_void_p()
        _check(lib.swe_hip_create(self.nx, self.ny, dx, dy, ox, oy, solver, flavour, C.byref(h)), "create")
        self._h = h

    @property
    def shape(self):
        return (self.nx + 2, self.ny + 2)

    def close(self) -> None:
        if self._h:
            lib.swe_hip_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_boundaries(self, types) -> None:
        for side, t in enumerate(types):
            _check(lib.swe_hip_set_boundary(self._h, side, int(t)), "set_boundary")

    def upload(self, h, hu, hv, b) -> None:
        arrs = [_f32(a, self.shape) for a in (h, hu, hv, b)]
        _check(lib.swe_hip_upload(self._h, *[_fptr(a) for a in arrs]), "upload")

    def init_grid_scenario(self, bathymetry, displacement=None) -> None:
        """The file-driven scenario sampled on the device
        (swe_hip_init_grid_scenario): each grid is (values[ny, nx], x0, y0,
        dx, dy) with cell-centred values, the first centre at (x0, y0).  Set
        the boundaries first; they are applied afterwards."""
        keep = []

        def grid(g):
            values, x0, y0, dx, dy = g
            v = np.ascontiguousarray(values, dtype=np.float32)
            if v.ndim != 2:
                raise ValueError("grid values must be 2-D (ny, nx)")
            keep.append(v)
            return _Grid(v.shape[1], v.shape[0], x0, y0, dx, dy, _fptr(v))

        gb = grid(bathymetry)
        gd = grid(displacement) if displacement is not None else None
        _check(lib.swe_hip_init_grid_scenario(self._h, C.byref(gb), C.byref(gd) if gd is not None else None),
               "init_grid_scenario")

    def init_scenario(self, scenario: int) -> None:
        """One of the analytic scenarios (SCENARIO_*) sampled on the device
        (swe_hip_init_scenario).  Set the boundaries first; they are applied
        afterwards."""
        _check(lib.swe_hip_init_scenario(self._h, int(scenario)), "init_scenario")

    def init_seeded(self, spec) -> None:
        """A seeded workload's parameter set (swe_ppm_amd.scenarios.*_spec)
        sampled on the device (swe_hip_init_seeded), bit for bit as
        scenarios.sample_seeded + finish sample it on the host.  Set the
        boundaries first; they are applied afterwards."""
        w = _Seeded()
        w.kind = int(spec["kind"])
        for f in ("dam_x", "dam_y", "dam_radius", "eta_in", "eta_out", "trench_x", "trench_w2", "hump_x", "hump_y",
                  "hump_r2"):
            setattr(w, f, float(spec.get(f, 0.0)))
        bumps = spec.get("bumps", [])
        if len(bumps) > _MAX_BUMPS:
            raise ValueError(f"at most {_MAX_BUMPS} bumps")
        w.nbumps = len(bumps)
        for k, (bx, by, a, s2) in enumerate(bumps):
            w.bump_x[k], w.bump_y[k], w.bump_a[k], w.bump_s2[k] = bx, by, a, s2
        _check(lib.swe_hip_init_seeded(self._h, C.byref(w)), "init_seeded")

    def download(self, with_b: bool = False):
        out = [np.empty(self.shape, np.float32) for _ in range(4 if with_b else 3)]
        ptrs = [_fptr(a) for a in out] + ([None] if not with_b else [])
        _check(lib.swe_hip_download(self._h, *ptrs), "download")
        return tuple(out)

    def apply_bc(self) -> None:
        _check(lib.swe_hip_apply_bc(self._h), "apply_bc")

    def apply_bc_bathymetry(self) -> None:
        _check(lib.swe_hip_apply_bc_bathymetry(self._h), "apply_bc_bathymetry")

    def fluxes(self) -> float:
        dt = C.c_float()
        _check(lib.swe_hip_fluxes(self._h, C.byref(dt)), "fluxes")
        return dt.value

    def update(self, dt: float) -> None:
        _check(lib.swe_hip_update(self._h, float(dt)), "update")

    def run(self, nsteps: int, t: float = 0.0, want_dts: bool = True):
        tt = C.c_float(t)
        dts = np.empty(max(nsteps, 1), np.float32) if want_dts else None
        _check(lib.swe_hip_run(self._h, int(nsteps), C.byref(tt), _fptr(dts) if want_dts else None), "run")
        return tt.value, (dts[:nsteps] if want_dts else None)

    def synchronize(self) -> None:
        _check(lib.swe_hip_synchronize(self._h), "synchronize")

    def device_pointers(self):
        ps = [C.c_void_p() for _ in range(4)]
        pitch = C.c_int()
        _check(lib.swe_hip_device_pointers(self._h, *[C.byref(p) for p in ps], C.byref(pitch)), "device_pointers")
        return [p.value for p in ps], pitch.value

    def set_tuning(self, strategy: int = -1, xseg: int = -1) -> None:
        _check(lib.swe_hip_set_tuning(self._h, strategy, xseg), "set_tuning")

    TUNE_WIDTHS = (4, 8, 10, 12, 16, 32)

    def autotune(self, steps: int = 20) -> dict:
        """Time `steps` steps at each strip width from the current state,
        restore the state and keep the fastest width; {width: ms per step}."""
        ms = (C.c_float * len(self.TUNE_WIDTHS))()
        _check(lib.swe_hip_autotune(self._h, int(steps), ms), "autotune")
        return {w: float(m) for w, m in zip(self.TUNE_WIDTHS, ms)}

    def set_math(self, math: int) -> None:
        """0 = IEEE (bit-compatible with the CPU oracle), 1 = fast."""
        _check(lib.swe_hip_set_math(self._h, int(math)), "set_math")

    def set_quiet_skip(self, enable: bool) -> None:
        """Skip quiet strip tiles in the fused device loop (default on);
        results are identical either way."""
        _check(lib.swe_hip_set_quiet_skip(self._h, int(bool(enable))), "set_quiet_skip")

    def copy_layer(self, side: int, src: "SweBlock", fields: int = 7) -> None:
        """This block's ghost line on `side` <- src's copy layer at the shared
        edge (swe_hip_copy_layer; fields: 1 h, 2 hu, 4 hv, 8 b)."""
        _check(lib.swe_hip_copy_layer(self._h, int(side), src._h, int(fields)), "copy_layer")

    def snapshot(self) -> int:
        """Queue an asynchronous copy of h, hu, hv (swe_hip_snapshot); returns its slot."""
        slot = C.c_int()
        _check(lib.swe_hip_snapshot(self._h, C.byref(slot)), "snapshot")
        return slot.value

    def snapshot_wait(self, slot: int):
        """Wait for a snapshot and return copies of its h, hu, hv."""
        ps = [C.POINTER(C.c_float)() for _ in range(3)]
        _check(lib.swe_hip_snapshot_wait(self._h, int(slot), *[C.byref(p) for p in ps]), "snapshot_wait")
        n = self.shape[0] * self.shape[1]
        return tuple(np.ctypeslib.as_array(p, shape=(n,)).reshape(self.shape).copy() for p in ps)

    def snapshot_release(self, slot: int) -> None:
        _check(lib.swe_hip_snapshot_release(self._h, int(slot)), "snapshot_release")

    def set_lookahead(self, enable: bool) -> None:
        """Queue the next per-call step while the host runs its loop (default
        on); results are identical either way."""
        _check(lib.swe_hip_set_lookahead(self._h, int(bool(enable))), "set_lookahead")

    def set_surrogate(self, mlp) -> None:
        """Attach a MiddleStateMLP (None detaches): AugRie then starts its
        Newton iteration from the network's guess."""
        _check(lib.swe_hip_set_surrogate(self._h, mlp._h if mlp is not None else None), "set_surrogate")
        self._mlp = mlp

    def get_tuning(self):
        s, x = C.c_int(), C.c_int()
        _check(lib.swe_hip_get_tuning(self._h, C.byref(s), C.byref(x)), "get_tuning")
        return s.value, x.value


BATCH_ENSEMBLE, BATCH_GLOBAL, BATCH_LTS = 0, 1, 2


class SweBatch:
    """Many flavour-U blocks advanced together on the GPU (include/swe_hip_batch.h).

    blocks: list of dicts with nx, ny, dx, dy (or the dx/dy arguments), bnd[4]
    and nbr[4] (neighbour index or -1); mode BATCH_ENSEMBLE (independent,
    per-block dt), BATCH_GLOBAL (connected, global dt) or BATCH_LTS
    (connected, local time stepping with macro step max_local)."""

    def __init__(self, blocks, solver: int = AUGRIE, mode: int = BATCH_ENSEMBLE, dx=None, dy=None,
                 max_local: float = 1.0, duration: float = 0.0):
        self.n = len(blocks)
        self.nx = np.array([bl["nx"] for bl in blocks], np.int32)
        self.ny = np.array([bl["ny"] for bl in blocks], np.int32)
        self.dx = np.array([bl.get("dx", dx) for bl in blocks], np.float32)
        self.dy = np.array([bl.get("dy", dy) for bl in blocks], np.float32)
        self.bnd = np.ascontiguousarray(np.concatenate([bl["bnd"] for bl in blocks]), np.int32)
        self.nbr = np.ascontiguousarray(
            np.concatenate([bl.get("nbr", [-1, -1, -1, -1]) for bl in blocks]), np.int32)
        self.sizes = (self.nx.astype(np.int64) + 2) * (self.ny.astype(np.int64) + 2)
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)[:-1]])
        self.cells = int(np.sum(self.nx.astype(np.int64) * self.ny))
        self.mode = mode
        h = C.c_void_p()
        _check(lib.swe_hip_batch_create(self.n, _iptr(self.nx), _iptr(self.ny), _fptr(self.dx), _fptr(self.dy),
                                        _iptr(self.bnd), _iptr(self.nbr), solver, mode, max_local, duration,
                                        C.byref(h)), "batch_create")
        self._h = h

    def close(self) -> None:
        if self._h:
            lib.swe_hip_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def pack(self, arrays):
        """Concatenate per-block Float2D arrays into the packed layout."""
        for k, a in enumerate(arrays):
            if a.shape != (self.nx[k] + 2, self.ny[k] + 2):
                raise ValueError(f"block {k}: shape {a.shape}")
        return np.ascontiguousarray(np.concatenate([np.asarray(a, np.float32).ravel() for a in arrays]))

    def unpack(self, flat):
        return [flat[o:o + s].reshape(self.nx[k] + 2, self.ny[k] + 2)
                for k, (o, s) in enumerate(zip(self.offsets, self.sizes))]

    def upload(self, h, hu, hv, b) -> None:
        """Per-block lists of (nx+2, ny+2) arrays, or packed flat arrays."""
        arrs = [a if isinstance(a, np.ndarray) and a.ndim == 1 else self.pack(a) for a in (h, hu, hv, b)]
        arrs = [np.ascontiguousarray(a, np.float32) for a in arrs]
        self._keep = arrs
        _check(lib.swe_hip_batch_upload(self._h, *[_fptr(a) for a in arrs]), "batch_upload")

    def download(self, with_b: bool = False, packed: bool = False):
        tot = int(np.sum(self.sizes))
        out = [np.empty(tot, np.float32) for _ in range(4 if with_b else 3)]
        ptrs = [_fptr(a) for a in out] + ([None] if not with_b else [])
        _check(lib.swe_hip_batch_download(self._h, *ptrs), "batch_download")
        return tuple(out) if packed else tuple(self.unpack(a) for a in out)

    def run(self, units: int):
        """Advance `units` steps (ensemble / global) or macro steps (LTS).
        Returns (dt_log, iter_log, elapsed_ms)."""
        dts = np.zeros(max(units, 1), np.float32)
        its = np.zeros(max(units, 1), np.int32)
        ms = C.c_float()
        _check(lib.swe_hip_batch_run(self._h, int(units), _fptr(dts), _iptr(its), C.byref(ms)), "batch_run")
        return dts[:units], its[:units], ms.value

    def stats(self):
        upd = np.zeros(self.n, np.int64)
        t = np.zeros(self.n, np.float32)
        _check(lib.swe_hip_batch_stats(self._h, upd.ctypes.data_as(C.POINTER(C.c_int64)), _fptr(t)), "batch_stats")
        return upd, t

    def set_tuning(self, xs: int = -1, groups: int = -1, chunk: int = -1, graph: int = -1, fused: int = -1) -> None:
        _check(lib.swe_hip_batch_set_tuning(self._h, xs, groups, chunk, graph, fused), "batch_set_tuning")

    def get_tuning(self):
        """(xs, groups, chunk, graph, fused)"""
        v = [C.c_int() for _ in range(5)]
        _check(lib.swe_hip_batch_get_tuning(self._h, *[C.byref(x) for x in v]), "batch_get_tuning")
        return tuple(x.value for x in v)

    def set_surrogate(self, mlp) -> None:
        _check(lib.swe_hip_batch_set_surrogate(self._h, mlp._h if mlp is not None else None), "batch_set_surrogate")
        self._mlp = mlp

    def autotune(self, iterations: int = 32):
        ms = np.zeros(16, np.float32)
        n = C.c_int()
        _check(lib.swe_hip_batch_autotune(self._h, int(iterations), _fptr(ms), C.byref(n)), "batch_autotune")
        return ms[:n.value]


MLP_PARAMS = 4738
MLP_SAMPLE_FLOATS = 8


class MiddleStateMLP:
    """Learned middle state of the augmented Riemann solver (include/swe_hip_mlp.h):
    a 6-64-64-2 network with a device sample store, trained by the library's
    own kernels, run by the solver on the matrix cores."""

    def __init__(self, capacity: int = 0, seed: int = 0):
        h = C.c_void_p()
        _check(lib.swe_hip_mlp_create(int(capacity), int(seed), C.byref(h)), "mlp_create")
        self._h = h

    def close(self) -> None:
        """Free the network.  Refused (SweHipError, the handle kept) while a
        block or batch still has it attached."""
        if self._h:
            _check(lib.swe_hip_mlp_destroy(self._h), "mlp_destroy")
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def params(self) -> np.ndarray:
        p = np.empty(MLP_PARAMS, np.float32)
        _check(lib.swe_hip_mlp_get_params(self._h, _fptr(p)), "mlp_get_params")
        return p

    @params.setter
    def params(self, p) -> None:
        p = np.ascontiguousarray(p, np.float32)
        if p.shape != (MLP_PARAMS,):
            raise ValueError(f"expected {MLP_PARAMS} parameters")
        _check(lib.swe_hip_mlp_set_params(self._h, _fptr(p)), "mlp_set_params")

    def log_block(self, blk: "SweBlock", keep: float = 1.0, seed: int = 0, hard_only: bool = False) -> int:
        """Append the block's Newton edges (hard_only: only those the
        linearised guess does not settle); returns the number appended."""
        n = C.c_int64()
        _check(lib.swe_hip_mlp_log_block(self._h, blk._h, float(keep), int(seed), int(bool(hard_only)), C.byref(n)),
               "mlp_log_block")
        return n.value

    def add_samples(self, samples) -> None:
        s = np.ascontiguousarray(samples, np.float32).reshape(-1, MLP_SAMPLE_FLOATS)
        _check(lib.swe_hip_mlp_add_samples(self._h, s.shape[0], _fptr(s)), "mlp_add_samples")

    @property
    def count(self) -> int:
        n = C.c_int64()
        _check(lib.swe_hip_mlp_sample_count(self._h, C.byref(n)), "mlp_sample_count")
        return n.value

    def read_samples(self, first: int = 0, n: int = -1) -> np.ndarray:
        if n < 0:
            n = self.count - first
        out = np.empty((n, MLP_SAMPLE_FLOATS), np.float32)
        _check(lib.swe_hip_mlp_read_samples(self._h, int(first), int(n), _fptr(out)), "mlp_read_samples")
        return out

    def clear_samples(self) -> None:
        _check(lib.swe_hip_mlp_clear_samples(self._h), "mlp_clear_samples")

    def sort_samples(self) -> None:
        """Put the logged samples in a canonical order: rows sorted
        lexicographically by their bit patterns (a total order, so rows equal
        in value but not in the sign of a zero cannot tie).  Logging appends
        edges in the order the device's waves reach them, which varies from run
        to run, and training draws its batches by index: sorting first makes
        training from the same samples reproducible."""
        s = self.read_samples()
        if s.shape[0] < 2:
            return
        s = s[np.lexsort(s.view(np.uint32).T[::-1])]
        self.clear_samples()
        self.add_samples(s)

    def train(self, steps: int, batch: int = 4096, lr: float = 2e-3, seed: int = 0) -> np.ndarray:
        loss = np.zeros(max(steps, 1), np.float32)
        _check(lib.swe_hip_mlp_train(self._h, int(steps), int(batch), float(lr), int(seed), _fptr(loss)), "mlp_train")
        return loss[:steps]

    def gradient(self, samples):
        s = np.ascontiguousarray(samples, np.float32).reshape(-1, MLP_SAMPLE_FLOATS)
        g = np.empty(MLP_PARAMS, np.float32)
        loss = C.c_float()
        _check(lib.swe_hip_mlp_gradient(self._h, s.shape[0], _fptr(s), _fptr(g), C.byref(loss)), "mlp_gradient")
        return g, loss.value

    def infer(self, features) -> np.ndarray:
        x = np.ascontiguousarray(features, np.float32).reshape(-1, 6)
        out = np.empty((x.shape[0], 2), np.float32)
        _check(lib.swe_hip_mlp_infer(self._h, x.shape[0], _fptr(x), _fptr(out)), "mlp_infer")
        return out

    def middle(self, hL, hR, uL, uR, bL=None, bR=None):
        ins = [np.ascontiguousarray(np.atleast_1d(a), np.float32) for a in (hL, hR, uL, uR)]
        n = ins[0].size
        bL = np.zeros(n, np.float32) if bL is None else np.ascontiguousarray(bL, np.float32)
        bR = np.zeros(n, np.float32) if bR is None else np.ascontiguousarray(bR, np.float32)
        outs = [np.empty(n, np.float32) for _ in range(3)]
        acc = np.empty(n, np.int32)
        _check(lib.swe_hip_mlp_middle(self._h, n, *[_fptr(a) for a in ins], _fptr(bL), _fptr(bR),
                                      *[_fptr(a) for a in outs], _iptr(acc)), "mlp_middle")
        return (*outs, acc)

    def solve_edges(self, hL, hR, huL, huR, bL, bR):
        ins = [np.ascontiguousarray(np.atleast_1d(a), dtype=np.float32) for a in (hL, hR, huL, huR, bL, bR)]
        n = ins[0].size
        outs = [np.empty(n, np.float32) for _ in range(5)]
        _check(lib.swe_hip_mlp_solve_edges(self._h, n, *[_fptr(a) for a in ins], *[_fptr(a) for a in outs]),
               "mlp_solve_edges")
        return tuple(outs)


def mlp_features(hL, hR, uL, uR, bL, bR) -> np.ndarray:
    """The device feature function (6 floats per edge)."""
    ins = [np.ascontiguousarray(np.atleast_1d(a), np.float32) for a in (hL, hR, uL, uR, bL, bR)]
    n = ins[0].size
    out = np.empty((n, 6), np.float32)
    _check(lib.swe_hip_mlp_features(n, *[_fptr(a) for a in ins], _fptr(out)), "mlp_features")
    return out


def _iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_int))


def solve_edges(solver: int, hL, hR, huL, huR, bL, bR, math: int = 0):
    """Per-edge net updates on the GPU: (hNetL, hNetR, huNetL, huNetR, maxSpeed)."""
    ins = [np.ascontiguousarray(np.atleast_1d(a), dtype=np.float32) for a in (hL, hR, huL, huR, bL, bR)]
    n = ins[0].size
    outs = [np.empty(n, np.float32) for _ in range(5)]
    _check(lib.swe_hip_solve_edges(solver, math, n, *[_fptr(a) for a in ins], *[_fptr(a) for a in outs]), "solve_edges")
    return tuple(outs)


def version() -> str:
    return lib.swe_hip_version().decode()
