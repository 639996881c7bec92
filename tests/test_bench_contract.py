"""bench.py's command-line contract on the CPU: more than one GPU is refused
with a message and exit status 2 before any device work; measured HBM
traffic is taken from a committed profile summary only when that summary
was recorded on the same state as the run."""
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


def test_traffic_only_for_the_profiled_state(tmp_path):
    sys.path.insert(0, ROOT)
    import bench
    p = tmp_path / "summary.md"
    p.write_text("# summary\nstate: coastal 64 t0=10800 spin=5 active=7\n"
                 "traffic_bytes_per_launch k_march<2> variants 4,5,12,13: 1234.5\n"
                 "traffic_bytes_per_launch k_march<2> variants 0,1,8,9: 1300\n")
    st = "coastal 64 t0=10800 spin=5 active=7"
    assert bench.measured_traffic(str(p), "k_march<2>", st, 5) == 1234.5
    assert bench.measured_traffic(str(p), "k_march<2>", st, 9) == 1300.0
    # a variant of an unprofiled class, another state (other active tile
    # count) or another kernel: no figure
    assert bench.measured_traffic(str(p), "k_march<2>", st, 2) is None
    assert bench.measured_traffic(str(p), "k_march<2>", "coastal 64 t0=10800 spin=5 active=8", 5) is None
    assert bench.measured_traffic(str(p), "k_march<1>", st, 5) is None
    assert bench.measured_traffic(str(tmp_path / "absent.md"), "k_march<2>", "x", 5) is None


def test_stage_bytes_add_up_to_the_step():
    sys.path.insert(0, ROOT)
    import bench
    for w in bench.STAGE_BYTES.values():
        assert w["march1"] + w["march2"] == w["step"]


@pytest.mark.parametrize("which,variants", [
    ("COASTAL_PROFILE", [0, 1, 4, 5, 8, 9, 12, 13]),
    ("RAIN_PROFILE", [0, 1, 2, 3, 8, 9, 10, 11, 32, 33, 34, 35, 40, 41, 42, 43])])
def test_committed_profiles_cover_the_variants_the_autotuner_picks(which, variants):
    """The launch variants the autotuner has picked on these workloads (the
    classes within about 1 % of the fastest) have a measured stage-2 traffic
    line in the committed profile, recorded on the profile's own state, of a
    plausible size: between the algorithmic bytes of the launch and 1.6x
    them."""
    sys.path.insert(0, ROOT)
    import bench
    path = getattr(bench, which)
    states = [ln.strip()[len("state: "):] for ln in open(path) if ln.startswith("state: ")]
    assert len(states) == 1
    algo = {"COASTAL_PROFILE": 6621024 * 44, "RAIN_PROFILE": 4194304 * 56}[which]
    for v in variants:
        t = bench.measured_traffic(path, "k_march<2>", states[0], v)
        assert t is not None, (which, v)
        assert algo <= t <= 1.6 * algo, (which, v, t / algo)
