"""NetCdfWriter (include/swe/writer/NetCdfWriter.hh) against an independent
reader: scipy's netCDF classic implementation.

Checks the data model of the reference's writer
(src/writer/NetCdfWriter.cpp:49-240): dimensions time (unlimited), x, y;
variables time / x / y / ah / hu / hv / b with their dimension order and
attributes; cell-centre coordinates accumulated in float; ah[t][j][i] =
h[i + left][j + bottom] (interior only, ghost widths from the boundary
sizes); b written with the first step; one record per writeTimeStep.
"""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]

PROGRAM = r"""
#include <cstdio>
#include <cstdlib>
#include "tools/Float2DNative.hh"
#include "writer/NetCdfWriter.hh"

int main(int argc, char** argv) {
    const int nX = std::atoi(argv[2]), nY = std::atoi(argv[3]);
    BoundarySize bs{{1, 2, 1, 3}};  // left, right, bottom, top ghost widths
    const int cols = nX + bs[0] + bs[1], rows = nY + bs[2] + bs[3];
    Float2DNative h(cols, rows), hu(cols, rows), hv(cols, rows), b(cols, rows);
    for (int x = 0; x < cols; x++)
        for (int y = 0; y < rows; y++) b[x][y] = 100.f + 10.f * x + y;
    NetCdfWriter w(argv[1], b, bs, nX, nY, 0.5f, 0.25f, -1.f, 2.f);
    const float times[3] = {0.f, 0.5f, 1.25f};
    for (int t = 0; t < 3; t++) {
        for (int x = 0; x < cols; x++)
            for (int y = 0; y < rows; y++) {
                h[x][y] = 1000.f * t + 10.f * x + y + 0.5f;
                hu[x][y] = -h[x][y];
                hv[x][y] = 0.25f * h[x][y];
            }
        w.writeTimeStep(h, hu, hv, times[t]);
    }
    return 0;
}
"""


@pytest.fixture(scope="module")
def writer_exe(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("ncw")
    src = d / "ncw.cpp"
    src.write_text(PROGRAM)
    exe = d / "ncw"
    p = subprocess.run([gxx, "-std=c++17", "-O1", "-Wall", "-I", str(ROOT / "include" / "swe"), str(src), "-o",
                        str(exe)], capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    return exe


def _expected(nX, nY, t):
    x = np.arange(nX)[None, :] + 1  # interior column i sits at x = i + left
    y = np.arange(nY)[:, None] + 1  # interior row j sits at y = j + bottom
    return (1000.0 * t + 10.0 * x + y + 0.5).astype(np.float32)


@pytest.mark.parametrize("nX,nY", [(5, 3), (64, 17)])
def test_netcdf_file_matches_reference_data_model(writer_exe, tmp_path, nX, nY):
    from scipy.io import netcdf_file

    base = tmp_path / "out"
    p = subprocess.run([str(writer_exe), str(base), str(nX), str(nY)], capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    with netcdf_file(str(base) + ".nc", "r", mmap=False) as f:
        assert f.version_byte == 2
        assert f.dimensions == {"time": None, "x": nX, "y": nY}
        assert list(f.variables) == ["time", "x", "y", "ah", "hu", "hv", "b"]
        v = f.variables
        assert v["time"].dimensions == ("time",)
        assert v["ah"].dimensions == ("time", "y", "x")
        assert v["hu"].dimensions == ("time", "y", "x")
        assert v["hv"].dimensions == ("time", "y", "x")
        assert v["b"].dimensions == ("y", "x")
        assert v["time"].long_name == b"Time"
        assert v["time"].units == b"seconds since simulation start"
        assert f.Conventions == b"CF-1.5"
        assert f.title == b"Computed tsunami solution"
        assert f.history == b"SWE"
        for name in ("time", "x", "y", "ah", "hu", "hv", "b"):
            assert v[name].data.dtype == np.dtype(">f4")
        # coordinates: origin + 0.5 d, then += d in float (the reference's accumulation)
        xs = np.empty(nX, np.float32)
        pos = np.float32(-1.0) + np.float32(0.5) * np.float32(0.5)
        for i in range(nX):
            xs[i] = pos
            pos = np.float32(pos + np.float32(0.5))
        ys = np.empty(nY, np.float32)
        pos = np.float32(2.0) + np.float32(0.5) * np.float32(0.25)
        for j in range(nY):
            ys[j] = pos
            pos = np.float32(pos + np.float32(0.25))
        assert np.array_equal(v["x"].data, xs)
        assert np.array_equal(v["y"].data, ys)
        assert np.array_equal(v["time"].data, np.array([0.0, 0.5, 1.25], np.float32))
        for t in range(3):
            e = _expected(nX, nY, t)
            assert np.array_equal(v["ah"].data[t], e)
            assert np.array_equal(v["hu"].data[t], -e)
            assert np.array_equal(v["hv"].data[t], (np.float32(0.25) * e).astype(np.float32))
        bx = np.arange(nX)[None, :] + 1
        by = np.arange(nY)[:, None] + 1
        assert np.array_equal(v["b"].data, (100.0 + 10.0 * bx + by).astype(np.float32))


def test_netcdf_file_is_replaced_not_appended(writer_exe, tmp_path):
    from scipy.io import netcdf_file

    base = tmp_path / "again"
    for _ in range(2):
        p = subprocess.run([str(writer_exe), str(base), "4", "4"], capture_output=True, text=True)
        assert p.returncode == 0, p.stderr
    with netcdf_file(str(base) + ".nc", "r", mmap=False) as f:
        assert f.variables["ah"].shape == (3, 4, 4)
