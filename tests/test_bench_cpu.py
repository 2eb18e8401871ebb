"""bench.py's command-line contract on the CPU: more than one GPU is refused
with a JSON message and status 2 before any GPU work, and the host description
used by the CPU baseline is readable."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_host_cpu_model_is_a_string():
    sys.path.insert(0, ROOT)
    import bench

    assert isinstance(bench.host_cpu_model(), str)
