"""Compile-only drop-in check: SWE_DimensionalSplittingHip builds against
the reference's own SWE_Block headers (this library's mirror headers are
not used for the base class).  Runs only where the reference tree is
present; it compiles, it does not run reference code."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
REF = Path("/root/reference/src")


@pytest.mark.skipif(not (REF / "blocks" / "SWE_Block.hh").exists(), reason="reference tree not present")
def test_subclass_compiles_against_reference_headers():
    gxx = shutil.which("g++")
    assert gxx
    cmd = [gxx, "-std=c++17", "-fsyntax-only", f"-I{REF}", f"-I{ROOT / 'include' / 'swe'}", f"-I{ROOT / 'include'}",
           str(ROOT / "swe-ppm_amd" / "host" / "SWE_DimensionalSplittingHip.cpp")]
    p = subprocess.run(cmd, capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    deps = subprocess.run(cmd[:-1] + ["-M", cmd[-1]], capture_output=True, text=True).stdout
    assert str(REF / "blocks" / "SWE_Block.hh") in deps


def test_subclass_compiles_against_mirror_headers():
    gxx = shutil.which("g++")
    cmd = [gxx, "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", f"-I{ROOT / 'include' / 'swe'}",
           f"-I{ROOT / 'include'}", str(ROOT / "swe-ppm_amd" / "host" / "SWE_DimensionalSplittingHip.cpp")]
    p = subprocess.run(cmd, capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
