"""Host build of the length-specialised FFT passes, checked against naive DFTs.

The per-task step functions of fft_fast.h are host+device; this compiles them
for the host with hipcc (AddressSanitizer on the host side, so any LDS or
global index outside its buffer fails) and runs the check program, which
compares every split length in both precisions with long-double DFTs,
including pruned inputs/outputs and partial tiles.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_fast_passes_match_naive_dft(tmp_path):
    exe = tmp_path / "fft_fast_check"
    src = os.path.join(ROOT, "tests", "native", "fft_fast_check.hip")
    inc = os.path.join(ROOT, "topo-dda_amd", "csrc")
    cmd = [HIPCC, "-O1", "-std=c++17", "-I", inc, "-Xarch_host", "-fsanitize=address", "-o", str(exe), src]
    subprocess.run(cmd, check=True, capture_output=True, timeout=600)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout
    worst = float(r.stdout.strip().splitlines()[-1].split()[-1])
    assert worst < 1e-6


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_plane_passes_match_naive_2d_dft(tmp_path):
    """The plane transforms (fft_plane.h) of every class, forward and inverse,
    c64 and c128, against naive 2-D DFTs with pruned rows and columns; each
    output written once, no read of an unwritten LDS entry, no index outside
    the LDS region (AddressSanitizer)."""
    exe = tmp_path / "fft_plane_check"
    src = os.path.join(ROOT, "tests", "native", "fft_plane_check.hip")
    inc = os.path.join(ROOT, "topo-dda_amd", "csrc")
    cmd = [HIPCC, "-O1", "-std=c++17", "-I", inc, "-Xarch_host", "-fsanitize=address", "-o", str(exe), src]
    subprocess.run(cmd, check=True, capture_output=True, timeout=600)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout
    last = r.stdout.strip().splitlines()[-1].split()
    assert last[:2] == ["fails", "0"] and float(last[-1]) < 1e-6
