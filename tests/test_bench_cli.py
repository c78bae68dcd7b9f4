"""bench.py's command-line contract that needs no device: one GPU only."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    # The one-image step runs on one GPU; --gpus N > 1 must stop with a
    # message (as one JSON line) and a non-zero status, before any device use.
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], cwd=ROOT,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2, r.stderr[-2000:]
    lines = [ln for ln in r.stdout.splitlines() if ln.strip()]
    assert len(lines) == 1
    msg = json.loads(lines[0])
    assert "--gpus 2" in msg["error"] and "one GPU" in msg["error"]
