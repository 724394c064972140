"""Host-side parts of bench.py's contract: more than one GPU is refused with
a message on one JSON line, and the HBM traffic the roofline object reports is
read from the committed rocprofv3 summary."""

import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120, cwd=ROOT)
    assert r.returncode == 2
    lines = [ln for ln in r.stdout.splitlines() if ln.strip()]
    assert len(lines) == 1
    msg = json.loads(lines[0])
    assert "one GPU" in msg["error"]


def test_traffic_summary_is_read_from_the_profile():
    T, V = 1310720, 274625  # the 64^3 x 5 cube of the headline pass
    algorithmic = 64 * T + 36 * V
    pass_bytes = bench.load_traffic("pass_a_snh64_traffic.md")
    elem_bytes = bench.load_traffic("pass_a_snh64_traffic.md", "element_bytes_per_launch")
    assert pass_bytes is not None and elem_bytes is not None
    # the pass moves more than its compulsory bytes (slot partials, staged
    # windows), the element kernel less than the whole pass
    assert algorithmic < pass_bytes < 2 * algorithmic
    assert 64 * T + 12 * V < elem_bytes < pass_bytes
    assert bench.load_traffic("no_such_profile.md") is None
