"""The C-ABI library loads and exports every symbol include/swe_hip.h
declares (no compute calls: runs without a GPU)."""
import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def declared_functions(header: Path):
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(swe_hip_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib_path = ROOT / "swe-ppm_amd" / "libswe_hip.so"
    assert lib_path.exists(), "build the library first (__graft_entry__.build())"
    lib = ctypes.CDLL(str(lib_path))
    for header in sorted((ROOT / "include").glob("*.h")):
        names = declared_functions(header)
        assert names, header
        for name in names:
            assert hasattr(lib, name), f"{header.name}: {name} not exported"


def test_version_and_error_string_without_gpu():
    lib = ctypes.CDLL(str(ROOT / "swe-ppm_amd" / "libswe_hip.so"))
    lib.swe_hip_version.restype = ctypes.c_char_p
    lib.swe_hip_last_error.restype = ctypes.c_char_p
    assert b"gfx950" in lib.swe_hip_version()
    # argument validation happens before any device call
    out = ctypes.c_void_p()
    rc = lib.swe_hip_create(0, 5, ctypes.c_float(1.0), ctypes.c_float(1.0), ctypes.c_float(0.0),
                            ctypes.c_float(0.0), 2, 0, ctypes.byref(out))
    assert rc == -1
    assert b"invalid argument" in lib.swe_hip_last_error()


def test_batch_create_validates_before_touching_the_device():
    import numpy as np

    lib = ctypes.CDLL(str(ROOT / "swe-ppm_amd" / "libswe_hip.so"))
    lib.swe_hip_last_error.restype = ctypes.c_char_p
    ip = ctypes.POINTER(ctypes.c_int)
    fp = ctypes.POINTER(ctypes.c_float)

    def create(nx, ny, bnd, nbr, mode, T=1.0):
        n = len(nx)
        arrs = [np.asarray(nx, np.int32), np.asarray(ny, np.int32), np.ones(n, np.float32), np.ones(n, np.float32),
                np.asarray(bnd, np.int32), np.asarray(nbr, np.int32)]
        ptrs = [a.ctypes.data_as(ip if a.dtype == np.int32 else fp) for a in arrs]
        out = ctypes.c_void_p()
        return lib.swe_hip_batch_create(n, *ptrs, 2, mode, ctypes.c_float(T), ctypes.c_float(0.0),
                                        ctypes.byref(out))

    # edge lengths differ
    assert create([16, 16], [16, 12], [0, 4, 0, 0, 4, 0, 0, 0], [-1, 1, -1, -1, 0, -1, -1, -1], 1) == -1
    assert b"does not share the edge" in lib.swe_hip_last_error()
    # link not mutual
    assert create([16, 16], [16, 16], [0, 4, 0, 0, 0, 0, 0, 0], [-1, 1, -1, -1, -1, -1, -1, -1], 1) == -1
    # connected blocks in an ensemble
    assert create([16, 16], [16, 16], [0, 4, 0, 0, 4, 0, 0, 0], [-1, 1, -1, -1, 0, -1, -1, -1], 0) == -1
    assert b"ensemble" in lib.swe_hip_last_error()
    # INFLOW is not a supported boundary (the reference asserts on it)
    assert create([8], [8], [2, 0, 0, 0], [-1, -1, -1, -1], 0) == -1
    # LTS needs a positive macro step; bad mode; empty block
    assert create([8], [8], [0, 0, 0, 0], [-1, -1, -1, -1], 2, T=0.0) == -1
    assert create([8], [8], [0, 0, 0, 0], [-1, -1, -1, -1], 3) == -1
    assert create([0], [8], [0, 0, 0, 0], [-1, -1, -1, -1], 0) == -1
