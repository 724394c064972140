"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

The block partition (apple_amd/csrc/layout.cpp) and the C oracle
(oracle/c/oracle.c) are compiled with g++ / gcc -fsanitize=address,undefined
into the harness tests/host/layout_check.cpp, which runs them over edge-case
meshes (empty, one tet, block windows at the slot cap, one-tet batch
instances, empty segments, rejected inputs) and checks the partition's
invariants.  The per-tet algebra (tetmath.h) runs the same way in the SVD
harness tests/host/svd_check.cpp.  A sanitizer report aborts the harness.
"""

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]


def _need(tool):
    if shutil.which(tool) is None:
        pytest.skip(f"{tool} not available")


def test_layout_and_oracle_under_asan_ubsan(tmp_path):
    _need("g++")
    _need("gcc")
    obj = str(tmp_path / "oracle.o")
    subprocess.run(["gcc", *SAN, "-std=gnu11", "-c", os.path.join(ROOT, "oracle", "c", "oracle.c"), "-o", obj],
                   check=True, capture_output=True)
    exe = str(tmp_path / "layout_check")
    subprocess.run(["g++", *SAN, "-std=c++17", os.path.join(HERE, "host", "layout_check.cpp"),
                    os.path.join(ROOT, "apple_amd", "csrc", "layout.cpp"), obj, "-o", exe, "-lm"],
                   check=True, capture_output=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert "OK (0 failures)" in r.stdout


def test_tetmath_host_harness_under_asan_ubsan(tmp_path):
    _need("g++")
    exe = str(tmp_path / "svd_check")
    # tetmath.h is host/device code: the host compiler sees its host side
    r = subprocess.run(["g++", *SAN, "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                        os.path.join(HERE, "host", "svd_check.cpp"), "-o", exe],
                       capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("tetmath.h needs the HIP headers' host build: " + r.stderr[-300:])
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    assert len(r.stdout.strip().splitlines()) == 22
