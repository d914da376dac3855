"""NetCdfReader and SWE_NetCdfScenario (include/swe/tools/NetCdfReader.hh,
include/swe/scenarios/SWE_NetCdfScenario.hh) on files written by scipy's
netCDF implementation.

The scenario restates the reference's file-driven scenario
(SWE_AsagiScenario.hh:42-151): h = -bathy below sea level, 0 above; b =
bathy + displacement strictly inside the displacement grid; OUTFLOW; the
domain is the bathymetry grid's extent.  The grid lookup (cell-centred
values, the cell containing the point, clamped) restates ASAGI, a library
absent from the reference tree: parity unpinned for that rule.  The checks
here are what the rule fixes: a cell centre returns its own value, a point
just inside a cell edge returns that cell, points outside return the edge
cells, and the extent is half a cell beyond the outer centres.
"""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]

PROGRAM = r"""
#include <cstdio>
#include <string>
#include "scenarios/SWE_NetCdfScenario.hh"

int main(int argc, char** argv) {
    SWE_NetCdfScenario sc(argv[1], argc > 3 ? argv[3] : "");
    const GriddedField& g = sc.bathymetryGrid();
    std::printf("grid %d %d %a %a %a %a\n", g.nx, g.ny, g.x0, g.y0, g.dx, g.dy);
    std::printf("extent %a %a %a %a\n", sc.getBoundaryPos(BND_LEFT), sc.getBoundaryPos(BND_RIGHT),
                sc.getBoundaryPos(BND_BOTTOM), sc.getBoundaryPos(BND_TOP));
    std::FILE* q = std::fopen(argv[2], "r");
    float x, y;
    while (std::fscanf(q, "%a %a", &x, &y) == 2)
        std::printf("%a %a %d\n", sc.getBathymetry(x, y), sc.getWaterHeight(x, y), (int)sc.getBoundaryType(BND_TOP));
    std::fclose(q);
    return 0;
}
"""


@pytest.fixture(scope="module")
def query_exe(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("ncs")
    (d / "q.cpp").write_text(PROGRAM)
    p = subprocess.run([gxx, "-std=c++17", "-O1", "-Wall", "-I", str(ROOT / "include" / "swe"), str(d / "q.cpp"),
                        "-o", str(d / "q")], capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    return d / "q"


def write_grid(path, xs, ys, z, version=2, ztype="f4", ctype="f4", extra=True):
    from scipy.io import netcdf_file

    with netcdf_file(str(path), "w", version=version) as f:
        f.history = "test grid"
        if extra:  # a record dimension (scipy wants it first) and variable the reader must step over
            f.createDimension("time", None)
        f.createDimension("y", len(ys))
        f.createDimension("x", len(xs))
        if extra:
            t = f.createVariable("time", "f8", ("time",))
            t[:] = [0.0, 1.0]
            t.units = "s"
        vx = f.createVariable("x", ctype, ("x",))
        vx[:] = xs
        vx.units = "m"
        vy = f.createVariable("y", ctype, ("y",))
        vy[:] = ys
        vz = f.createVariable("z", ztype, ("y", "x"))
        vz[:] = z
        vz.long_name = "elevation"


def query(exe, bathy, points, displ=None):
    qf = Path(str(bathy) + ".q")
    qf.write_text("".join(f"{float(np.float32(x)).hex()} {float(np.float32(y)).hex()}\n" for x, y in points))
    args = [str(exe), str(bathy), str(qf)] + ([str(displ)] if displ else [])
    p = subprocess.run(args, capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    lines = p.stdout.splitlines()
    grid = lines[0].split()[1:]
    extent = [np.float32(float.fromhex(v)) for v in lines[1].split()[1:]]
    vals = [(np.float32(float.fromhex(a)), np.float32(float.fromhex(b)), int(c))
            for a, b, c in (ln.split() for ln in lines[2:])]
    return grid, extent, vals


@pytest.mark.parametrize("version,ztype,ctype", [(1, "f4", "f4"), (2, "f4", "f8"), (2, "f8", "f4"), (1, "i2", "f4")])
def test_reader_and_lookup_on_scipy_files(query_exe, tmp_path, version, ztype, ctype):
    rng = np.random.default_rng(5)
    nx, ny = 7, 5
    xs = -300.0 + 100.0 * np.arange(nx)  # centres; cells 100 m wide
    ys = 1000.0 + 50.0 * np.arange(ny)
    z = rng.integers(-400, 300, size=(ny, nx)).astype(np.float64)
    bathy = tmp_path / "bathy.nc"
    write_grid(bathy, xs, ys, z, version=version, ztype=ztype, ctype=ctype)
    pts = [(xs[i], ys[j]) for j in range(ny) for i in range(nx)]  # every centre
    pts += [(xs[2] + 49.0, ys[1] - 24.0), (xs[2] + 51.0, ys[1] + 26.0)]  # just inside / past a cell edge
    pts += [(-1e5, 0.0), (1e5, 1e5)]  # outside: clamped to the corner cells
    grid, extent, vals = query(query_exe, bathy, pts)
    assert grid[:2] == [str(nx), str(ny)]
    assert float.fromhex(grid[4]) == 100.0 and float.fromhex(grid[5]) == 50.0
    assert extent == [np.float32(-350.0), np.float32(350.0), np.float32(975.0), np.float32(1225.0)]
    want = [z[j, i] for j in range(ny) for i in range(nx)] + [z[1, 2], z[2, 3], z[0, 0], z[ny - 1, nx - 1]]
    for (b, h, bt), w in zip(vals, want):
        w = np.float32(w)
        assert b == w
        assert h == (np.float32(0.0) if w > 0 else -w)
        assert bt == 0  # OUTFLOW


def test_displacement_only_strictly_inside_its_extent(query_exe, tmp_path):
    xs = np.arange(10) * 10.0 + 5.0  # bathymetry extent [0, 100]
    z = np.full((10, 10), -50.0)
    bathy = tmp_path / "b.nc"
    write_grid(bathy, xs, xs, z, extra=False)
    dx = np.array([35.0, 45.0, 55.0])  # displacement extent [30, 60]^2
    dz = np.array([[1.0, 2.0, 3.0], [4.0, 5.0, 6.0], [7.0, 8.0, 9.0]])
    displ = tmp_path / "d.nc"
    write_grid(displ, dx, dx, dz, extra=False)
    pts = [(35.0, 35.0), (55.0, 45.0), (30.0, 45.0), (60.0, 45.0), (45.0, 29.9), (30.5, 59.5), (95.0, 5.0)]
    _, _, vals = query(query_exe, bathy, pts, displ)
    want_b = [-49.0, -44.0, -50.0, -50.0, -50.0, -43.0, -50.0]
    for (b, h, _), wb in zip(vals, want_b):
        assert b == np.float32(wb)
        assert h == np.float32(50.0)  # the water height ignores the displacement


def test_reader_rejects_netcdf4_and_garbage(query_exe, tmp_path):
    bad = tmp_path / "bad.nc"
    bad.write_bytes(b"\x89HDF\r\n\x1a\n" + b"\0" * 64)
    q = tmp_path / "q"
    q.write_text("")
    p = subprocess.run([str(query_exe), str(bad), str(q)], capture_output=True, text=True)
    assert p.returncode != 0
