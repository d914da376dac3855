"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker, never the thing measured or shipped.
Only tests/, the smoke check and the bench's cpu_baseline leg import this.
Arrays use the reference's Float2D layout: numpy arrays of shape
(nx+2, ny+2), C order, a[x, y] at x*(ny+2) + y (src/tools/Float2D.hh:74-80).
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "build" / "liboracle.so"


def build() -> None:
    subprocess.run(["make", "-C", str(HERE), "-j4"], check=True, capture_output=True)


def _load(path: Path = LIB_PATH) -> C.CDLL:
    if not path.exists():
        build()
    lib = C.CDLL(str(path))
    i64 = C.c_int64
    for t, ct in (("f32", C.c_float), ("f64", C.c_double)):
        p = C.POINTER(ct)
        getattr(lib, f"oracle_solver_{t}").argtypes = [C.c_int, i64] + [p] * 11
        getattr(lib, f"oracle_middle_state_{t}").argtypes = [i64] + [p] * 7 + [C.POINTER(C.c_int32)] * 2
        getattr(lib, f"oracle_run_u_{t}").argtypes = [C.c_int, C.c_int, ct, ct, C.POINTER(C.c_int), C.c_int,
                                                      C.c_int, C.c_int, p, p, p, p, p, p]
        getattr(lib, f"oracle_run_s_{t}").argtypes = [C.c_int, C.c_int, ct, C.POINTER(C.c_int), C.c_int, C.c_int,
                                                      C.c_int, p, p, p, p, p]
    fp = C.POINTER(C.c_float)
    ip = C.POINTER(C.c_int)
    lib.oracle_apply_bc_f32.argtypes = [C.c_int, C.c_int, ip, fp, fp, fp]
    lib.oracle_apply_bc_bathymetry_f32.argtypes = [C.c_int, C.c_int, ip, fp]
    lib.oracle_fluxes_u_f32.argtypes = [C.c_int, C.c_int, C.c_int] + [fp] * 13
    lib.oracle_compute_max_timestep_f32.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, fp, fp, fp, C.c_float,
                                                    C.c_float]
    lib.oracle_compute_max_timestep_f32.restype = C.c_float
    lib.oracle_timestep_u_f32.argtypes = [C.c_float, C.c_float, C.c_float]
    lib.oracle_timestep_u_f32.restype = C.c_float
    lib.oracle_multiblock_run_f32.argtypes = [C.c_int, ip, ip, C.c_float, C.c_float, ip, ip, C.c_int, C.c_int,
                                              C.c_float, C.c_float, C.c_int, C.c_int, fp, fp, fp, fp,
                                              C.POINTER(C.c_int64), C.POINTER(C.c_int32), fp]
    dp = C.POINTER(C.c_double)
    lib.oracle_multiblock_run_f64.argtypes = [C.c_int, ip, ip, C.c_double, C.c_double, ip, ip, C.c_int, C.c_int,
                                              C.c_double, C.c_double, C.c_int, C.c_int, dp, dp, dp, dp,
                                              C.POINTER(C.c_int64), C.POINTER(C.c_int32), dp]
    lib.oracle_round_timestep_f32.argtypes = [C.c_float, C.c_float, fp]
    lib.oracle_lts_trace_f32.argtypes = [C.c_int, ip, ip, C.c_float, C.c_float, ip, ip, C.c_int, C.c_float,
                                         C.c_float, C.c_int, fp, fp, fp, fp, fp, C.POINTER(C.c_int32)]
    return lib


lib = _load()
_canonical = None


def canonical_lib() -> C.CDLL:
    """The oracle built with the textbook per-cell forms (hu / h, sqrt(g h))."""
    global _canonical
    if _canonical is None:
        _canonical = _load(HERE / "build" / "liboracle_canonical.so")
    return _canonical


_release = None


def release_lib() -> C.CDLL:
    """The same source built with the reference's Release flags (fast math):
    for the CPU baseline's timing only, never as a parity reference."""
    global _release
    if _release is None:
        _release = _load(HERE / "build" / "liboracle_release.so")
    return _release


def _p(a: np.ndarray):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(np.ctypeslib.as_ctypes_type(a.dtype)))


def _ints(v):
    a = np.ascontiguousarray(v, dtype=np.int32)
    return a, a.ctypes.data_as(C.POINTER(C.c_int))


def solve_edges(solver: int, hL, hR, huL, huR, bL, bR, dtype=np.float32):
    """Per-edge net updates: returns (hNetL, hNetR, huNetL, huNetR, maxSpeed)."""
    ins = [np.ascontiguousarray(np.atleast_1d(a), dtype=dtype) for a in (hL, hR, huL, huR, bL, bR)]
    n = ins[0].size
    outs = [np.empty(n, dtype) for _ in range(5)]
    f = lib.oracle_solver_f32 if dtype == np.float32 else lib.oracle_solver_f64
    f(solver, n, *[_p(a) for a in ins], *[_p(a) for a in outs])
    return tuple(outs)


def middle_state(hL, hR, uL, uR, dtype=np.float64):
    """Middle state (hm, s1m, s2m, kind, newton_iterations)."""
    ins = [np.ascontiguousarray(np.atleast_1d(a), dtype=dtype) for a in (hL, hR, uL, uR)]
    n = ins[0].size
    outs = [np.empty(n, dtype) for _ in range(3)]
    kind = np.empty(n, np.int32)
    its = np.empty(n, np.int32)
    f = lib.oracle_middle_state_f32 if dtype == np.float32 else lib.oracle_middle_state_f64
    f(n, *[_p(a) for a in ins], *[_p(a) for a in outs], _p(kind), _p(its))
    return (*outs, kind, its)


def apply_bc(bnd, h, hu, hv) -> None:
    nx, ny = h.shape[0] - 2, h.shape[1] - 2
    b, bp = _ints(bnd)
    lib.oracle_apply_bc_f32(nx, ny, bp, _p(h), _p(hu), _p(hv))


def apply_bc_bathymetry(bnd, b) -> None:
    nx, ny = b.shape[0] - 2, b.shape[1] - 2
    bb, bp = _ints(bnd)
    lib.oracle_apply_bc_bathymetry_f32(nx, ny, bp, _p(b))


def run_u(h, hu, hv, b, dx, dy, bnd, solver, nsteps, nthreads=1, dt_in=None, dtype=np.float32, canonical=False,
          release=False):
    """nsteps of flavour U (in place on copies); returns (h, hu, hv, dts).
    canonical: the restatement with the textbook per-cell forms; release: the
    Release-flags build (timing only)."""
    h, hu, hv, b = [np.array(a, dtype=dtype, order="C", copy=True) for a in (h, hu, hv, b)]
    nx, ny = h.shape[0] - 2, h.shape[1] - 2
    dts = np.empty(max(nsteps, 1), dtype)
    dti = None if dt_in is None else np.ascontiguousarray(dt_in, dtype=dtype)
    bb, bp = _ints(bnd)
    L = canonical_lib() if canonical else release_lib() if release else lib
    f = L.oracle_run_u_f32 if dtype == np.float32 else L.oracle_run_u_f64
    rc = f(nx, ny, dx, dy, bp, solver, nsteps, nthreads, _p(h), _p(hu), _p(hv), _p(b), _p(dti), _p(dts))
    assert rc == 0
    return h, hu, hv, dts[:nsteps]


def run_s(h, hu, hv, b, dx, bnd, solver, nsteps, nthreads=1, dtype=np.float32):
    h, hu, hv, b = [np.array(a, dtype=dtype, order="C", copy=True) for a in (h, hu, hv, b)]
    nx, ny = h.shape[0] - 2, h.shape[1] - 2
    dts = np.empty(max(nsteps, 1), dtype)
    bb, bp = _ints(bnd)
    f = lib.oracle_run_s_f32 if dtype == np.float32 else lib.oracle_run_s_f64
    rc = f(nx, ny, dx, bp, solver, nsteps, nthreads, _p(h), _p(hu), _p(hv), _p(b), _p(dts))
    assert rc == 0
    return h, hu, hv, dts[:nsteps]


def fluxes_u(h, hu, hv, b, solver):
    """Net-update arrays of one flavour-U flux pass (reference member layout)."""
    nx, ny = h.shape[0] - 2, h.shape[1] - 2
    a = [np.zeros((nx + 2, ny + 2), np.float32) for _ in range(4)]
    c = [np.zeros((nx + 1, ny + 2), np.float32) for _ in range(4)]
    s = np.zeros(1, np.float32)
    ins = [np.ascontiguousarray(x, dtype=np.float32) for x in (h, hu, hv, b)]
    lib.oracle_fluxes_u_f32(nx, ny, solver, *[_p(x) for x in ins], *[_p(x) for x in a + c], _p(s))
    return a, c, float(s[0])


def compute_max_timestep(h, hu, hv, dx, dy, dry_tol=0.1, cfl=0.4) -> float:
    nx, ny = h.shape[0] - 2, h.shape[1] - 2
    ins = [np.ascontiguousarray(x, dtype=np.float32) for x in (h, hu, hv)]
    return float(lib.oracle_compute_max_timestep_f32(nx, ny, dx, dy, *[_p(x) for x in ins], dry_tol, cfl))


def timestep_u(max_speed, dx, dy) -> float:
    return float(lib.oracle_timestep_u_f32(max_speed, dx, dy))


def round_timestep(max_local: float, ts: float):
    r = C.c_float()
    k = lib.oracle_round_timestep_f32(max_local, ts, C.byref(r))
    return k, r.value


def multiblock_run(blocks, dx, dy, solver, lts, max_local, duration, nmacro, nthreads=1, release=False,
                   dtype=np.float32):
    """Overdecomposed driver.  blocks: list of dicts with nx, ny, bnd[4], nbr[4]
    (neighbour block index or -1) and h, hu, hv, b arrays.  Returns
    (states, updates, iterations, dts).  release: the Release-flags build
    (timing only)."""
    n = len(blocks)
    nx, nxp = _ints([bl["nx"] for bl in blocks])
    ny, nyp = _ints([bl["ny"] for bl in blocks])
    bnd, bndp = _ints(np.concatenate([bl["bnd"] for bl in blocks]))
    nbr, nbrp = _ints(np.concatenate([bl["nbr"] for bl in blocks]))
    cat = [np.concatenate([np.asarray(bl[k], dtype).ravel() for bl in blocks]) for k in ("h", "hu", "hv", "b")]
    upd = np.zeros(n, np.int64)
    its = np.zeros(max(nmacro, 1), np.int32)
    dts = np.zeros(max(nmacro, 1), dtype)
    L = release_lib() if release else lib
    f = L.oracle_multiblock_run_f32 if dtype == np.float32 else L.oracle_multiblock_run_f64
    rc = f(n, nxp, nyp, dx, dy, bndp, nbrp, solver, int(lts), max_local, duration, nmacro, nthreads,
           *[_p(a) for a in cat], _p(upd), _p(its), _p(dts))
    if rc != 0:
        raise RuntimeError("oracle multiblock run hit its iteration cap")
    out, off = [], 0
    for bl in blocks:
        sz = (bl["nx"] + 2) * (bl["ny"] + 2)
        shp = (bl["nx"] + 2, bl["ny"] + 2)
        out.append(tuple(a[off:off + sz].reshape(shp).copy() for a in cat))
        off += sz
    return out, upd, its[:nmacro], dts[:nmacro]


TRACE_FIELDS = ("total", "stepSize", "stepSizeCounter", "timestepCounter", "notified", "gl", "borderTs", "updates")


def lts_trace(blocks, dx, dy, solver, max_local, duration, niter):
    """niter driver iterations of the overdecomposed LTS loop (macro steps end
    when every block has its maximum local time step).  Returns (states,
    info, macro_ends): per block (h, hu, hv) incl. ghosts, a dict of the LTS
    bookkeeping after the last iteration, and the iteration count at which
    each completed macro step ended."""
    n = len(blocks)
    nx, nxp = _ints([bl["nx"] for bl in blocks])
    ny, nyp = _ints([bl["ny"] for bl in blocks])
    bnd, bndp = _ints(np.concatenate([bl["bnd"] for bl in blocks]))
    nbr, nbrp = _ints(np.concatenate([bl["nbr"] for bl in blocks]))
    cat = [np.concatenate([np.asarray(bl[k], np.float32).ravel() for bl in blocks]) for k in ("h", "hu", "hv", "b")]
    info = np.zeros((n, 14), np.float32)
    ends = np.zeros(max(niter, 1), np.int32)
    macros = lib.oracle_lts_trace_f32(n, nxp, nyp, dx, dy, bndp, nbrp, solver, max_local, duration, niter,
                                      *[_p(a) for a in cat], _p(info), ends.ctypes.data_as(C.POINTER(C.c_int32)))
    out, off = [], 0
    for bl in blocks:
        sz = (bl["nx"] + 2) * (bl["ny"] + 2)
        shp = (bl["nx"] + 2, bl["ny"] + 2)
        out.append(tuple(a[off:off + sz].reshape(shp).copy() for a in cat[:3]))
        off += sz
    rows = []
    for r in info:
        rows.append(dict(total=float(r[0]), stepSize=int(r[1]), stepSizeCounter=int(r[2]),
                         timestepCounter=int(r[3]), notified=bool(r[4]), gl=[int(v) for v in r[5:9]],
                         borderTs=[float(v) for v in r[9:13]], updates=int(r[13])))
    return out, rows, [int(e) for e in ends[:macros]]
