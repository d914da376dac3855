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
