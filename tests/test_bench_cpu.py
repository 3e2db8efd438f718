"""bench.py's contract on the CPU: more than one GPU is refused with exit
status 2 and one JSON line naming the reason (the path runs on one GPU;
nothing is imported from the device stack before the refusal)."""
import json
import os
import subprocess
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=REPO)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"] and "--gpus must be 1" in line["error"]
