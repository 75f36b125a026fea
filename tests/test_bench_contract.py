"""bench.py measures one GPU: a request for more exits with status 2 and a
message on stderr, before anything touches a device, and prints no result
line (no GPU needed)."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2", "--steps", "5",
                        "--warmup", "1"], capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2, (r.returncode, r.stderr[-2000:])
    assert "measures one GPU" in r.stderr and "--gpus 1" in r.stderr
    for line in r.stdout.splitlines():
        if line.startswith("{"):
            raise AssertionError("unexpected result line: %s" % json.loads(line))
