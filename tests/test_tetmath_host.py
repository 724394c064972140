"""The per-tet algebra compiled for the host (tetmath.h is host/device code):
the refined ARAP SVD of the FP32 kernels against the plain FP64 SVD.  The
reference's Warp svd3 cannot run here; its conventions are pinned by the
oracle tests (rotation-variant signs, ordering), this pins the build's own
fast path against the FP64 one."""

import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    cxx = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(cxx):
        pytest.skip("hipcc not available")
    exe = str(tmp_path_factory.mktemp("svd") / "svd_check")
    subprocess.run([cxx, "-O2", "-std=c++17", os.path.join(HERE, "host", "svd_check.cpp"), "-o", exe],
                   check=True, capture_output=True)
    return exe


def test_refined_svd_matches_fp64_svd(harness):
    out = subprocess.run([harness], check=True, capture_output=True, text=True).stdout
    rows = np.array([[float(x) for x in line.split()] for line in out.strip().splitlines()])
    assert rows.shape == (8, 6)
    p_err, e_err, s_err, orth = rows[:, 2], rows[:, 3], rows[:, 4], rows[:, 5]
    # FP32 outputs: errors at the level of FP32 rounding of the results
    assert p_err.max() <= 1e-6
    assert e_err.max() <= 1e-6
    assert s_err.max() <= 1e-6
    assert orth.max() <= 1e-6
