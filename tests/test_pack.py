"""The host side of the packed link copies (empanada_amd/csrc/emp_pack.cpp, no GPU): kept reads in
pageable memory cross PCIe as 2-bit codes, four bases to a byte, and a batch holding any byte other
than ACGTacgt must be recognised so that it is copied as it is. The codes are those of the map
kernel's read-in (blocks.cuh:33-40: A 0, C 1, T 2, G 3 from (c >> 1) & 3, lower case alike)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "empanada_amd", "csrc")


def test_pack_codes_bounds_and_stray_bytes(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    exe = str(tmp_path / "pack_check")
    subprocess.run([gxx, "-O2", "-std=c++17", "-pthread", "-I", CSRC, "-o", exe,
                    os.path.join(ROOT, "tests", "pack", "pack_check.cpp"), os.path.join(CSRC, "emp_pack.cpp")],
                   check=True, capture_output=True, timeout=300)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "all equal" in out.stdout
