"""The library's host-side C++ under AddressSanitizer + UndefinedBehaviorSanitizer.

tomo_io.cpp (generator, reference file formats, damping tool, ASCII kernel
reader), tile_build.cpp (brick order, ray sort, mode-2 slots with 16-bit
segment indices and round cuts, mode-1 groups with 8-bit lengths) and
batch_build.cpp (batch plan) hold no HIP code; g++ builds them with the
driver tests/native/host_layouts.cpp, which runs every layout builder on
small systems (with and without a voxel grid, several row slices, a ragged
last brick, fewer rays than columns) and checks each against a plain
recomputation.  Any sanitizer report fails the run.
"""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

CSRC = os.path.join(ROOT, "plsqr3.2013.09_amd", "csrc")
SOURCES = [os.path.join(ROOT, "tests", "native", "host_layouts.cpp")] + \
    [os.path.join(CSRC, f) for f in ("tomo_io.cpp", "tile_build.cpp", "batch_build.cpp")]


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ missing")
def test_host_code_is_clean_under_asan_and_ubsan(tmp_path):
    exe = str(tmp_path / "host_layouts")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-I", os.path.join(ROOT, "include"), "-I", CSRC] + SOURCES + ["-o", exe, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    (tmp_path / "files").mkdir()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", PLSQR_HOST_THREADS="4")
    r = subprocess.run([exe, str(tmp_path / "files")], capture_output=True, text=True, timeout=600, env=env)
    print(r.stdout[-3000:])
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert "host layouts ok" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
