"""The output.bin writer on the CPU (empanada_amd/csrc/emp_outfile.cpp, no GPU): random record
sequences written with and without its direct (O_DIRECT, Linux AIO) chunks, and with failures
injected into the direct writes, must give the bytes write_edges_binary lays out
(overlap.cu:148-161)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "outfile")


def _build(tmp_path, name, extra):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    exe = str(tmp_path / name)
    cmd = [gxx, "-O2", "-std=c++17", "-pthread"] + extra + [
        "-o", exe, os.path.join(SRC, "outfile_check.cpp"), os.path.join(ROOT, "empanada_amd", "csrc", "emp_outfile.cpp")]
    subprocess.run(cmd, check=True, capture_output=True, timeout=300)
    return exe


def _run(exe, path, trials):
    out = subprocess.run([exe, str(path), str(trials)], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "all equal" in out.stdout
    return out.stdout


def test_writer_files_equal_the_record_layout(tmp_path):
    exe = _build(tmp_path, "outfile_check", [])
    out = _run(exe, tmp_path / "output.bin", 12)
    # on ext4 or xfs a share of the larger files goes out through the direct writes
    if subprocess.run(["stat", "-f", "-c", "%T", str(tmp_path)], capture_output=True, text=True).stdout.strip() in (
            "ext2/ext3", "xfs"):
        assert any(int(ln.split(",")[2].split()[0]) > 0 for ln in out.splitlines() if ln.startswith("trial"))


def test_writer_rewrites_failed_direct_chunks(tmp_path):
    exe = _build(tmp_path, "outfile_check_faults",
                 ["-DOUTFILE_FAULTS", "-I", SRC, "-include", os.path.join(SRC, "fault_shim.h")])
    out = _run(exe, tmp_path / "output.bin", 12)
    line = [ln for ln in out.splitlines() if ln.startswith("injected")][0]
    if subprocess.run(["stat", "-f", "-c", "%T", str(tmp_path)], capture_output=True, text=True).stdout.strip() in (
            "ext2/ext3", "xfs"):
        failed_submits = int(line.split()[1])
        failed_events = int(line.split(",")[1].split()[0])
        assert failed_submits > 0 and failed_events > 0, line
