"""bench.py's command-line contract on the CPU: more than one GPU is refused with a message,
the options parse, and the roofline's traffic figure is only taken from a profile summary
measured at the same launch size."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_help_lists_configs():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0
    for opt in ("--config", "--chunk", "--variant", "--mtile", "--solve", "--concurrent", "--no-cpu", "--sweep"):
        assert opt in r.stdout


def test_read_traffic_matches_launch_size(tmp_path):
    sys.path.insert(0, ROOT)
    import bench

    f = tmp_path / "traffic.txt"
    f.write_text("# material\nitems_per_launch: 6\ntraffic_bytes_per_launch: 292490902\n")
    assert bench.read_traffic(str(f), 6) == 292490902.0
    assert bench.read_traffic(str(f), 40) is None          # measured at another launch size
    assert bench.read_traffic(str(tmp_path / "missing.txt"), 6) is None
