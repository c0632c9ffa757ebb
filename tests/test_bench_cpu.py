"""The benchmark's CPU-side plumbing (no GPU): the CPU baseline runs the
reference harness in a child process whose C stdio output must never reach
the benchmark's stdout (the one JSON line), and a multi-GPU request is
refused with a message."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_cpu_baseline_child_result_and_clean_stdout(capfd):
    from oracle import refh
    if not refh.available():
        pytest.skip("reference harness not built (oracle/_ref)")
    import bench
    r = bench.cpu_baseline(n=9, cap_s=1.0, threads=1)
    out, err = capfd.readouterr()
    assert out == "", out[:200]          # the reference's stdio went to stderr
    assert r["kind"] == "reference" and r["cores"] == 1 and r["n"] == 9
    assert r["value"] > 0 and r["rel_reached"] >= 0
    json.dumps(r)


def test_multi_gpu_request_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert p.returncode == 2
    msg = json.loads(p.stdout.strip().splitlines()[-1])
    assert msg["n_gpus"] == 2 and "one GPU" in msg["error"]
