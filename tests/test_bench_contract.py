"""The benchmark's command-line contract and its link to the committed profiles (CPU only)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_more_than_one_gpu_is_refused():
    """--gpus N > 1 exits with status 2 and says why, before any device work."""
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300, cwd=ROOT)
    assert r.returncode == 2
    assert "exactly one GPU" in r.stderr
    assert r.stdout.strip() == ""


def test_roofline_traffic_comes_from_the_committed_counter_summaries():
    """roofline.traffic is read from the rocprofv3 counter summaries under profiles/: every
    machine-readable line parses, the tridiagonalisation's entry is the 792 x 792 metric's,
    and a workload without a summary gives None."""
    sys.path.insert(0, ROOT)
    import bench
    hetrd = bench.measured_traffic("fp64", "k_hetrd", 792)
    assert hetrd is not None and hetrd > 2 * 792 ** 2 * 16   # at least the algorithmic bytes
    path = os.path.join(ROOT, "profiles", "k_zgemm_fp64_hbm_counters.txt")
    keys = [ln.split()[1] for ln in open(path) if ln.startswith("traffic_bytes_per_launch")]
    assert keys, "no machine-readable lines"
    for k in keys:
        v = bench.measured_traffic("fp64", "k_zgemm", k)
        assert v is not None and v > 0
    assert bench.measured_traffic("no_such_workload", "k_hetrd", 792) is None
