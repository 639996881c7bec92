"""bench.py's command-line contract on the CPU: more than one GPU is refused
with a message and exit status 2 before any device work, and the measured
HBM traffic of the 2048 x 2048 workload is read from the committed profile
summary (no other workload has one)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("n", [2, 8])
def test_more_than_one_gpu_is_refused(n):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", str(n)],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2, (r.returncode, r.stderr)
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_traffic_comes_from_the_committed_profile():
    sys.path.insert(0, ROOT)
    import bench
    t = bench.measured_traffic("k_march<2>", 2048)
    assert t is not None and 2.5e8 < t < 5e8
    # the figure exceeds the kernel's contract bytes for 2048 x 2048 wet cells
    assert t > bench.KERNEL_BYTES_PER_CELL["march2"] * 2048 * 2048
    assert bench.measured_traffic("k_march<2>", 1024) is None
