"""bench.py's command-line contract that does not need a GPU: it measures one GPU and refuses
another count with a message and a non-zero status, and its help lists the workloads."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = _run("--gpus", "2")
    assert r.returncode != 0
    assert "one GPU" in r.stderr or "single MI355X" in r.stderr


def test_help_lists_the_workloads():
    r = _run("--help")
    assert r.returncode == 0
    for w in ("many-light", "media", "cornell"):
        assert w in r.stdout


def test_committed_measurements_parse():
    """The bench line's traffic, VALU-busy and time-to-relMSE fields come from the committed
    profile summaries; each parser finds its figures."""
    sys.path.insert(0, ROOT)
    import bench
    t = bench.measured_traffic("many-light")
    assert t["k_shade"]["read"] > 0 and t["k_extend"]["read"] > 0
    assert 0 < bench.measured_valu_busy("many-light")["k_extend"] <= 1
    r = bench.measured_time_to_relmse()
    assert r["plain_s"] > 0 and r["guided_s"] > 0 and r["plain_spp"] > r["guided_spp"]
