"""The benchmark's command-line contract that holds without a GPU: a request
for more than one GPU is refused with exit status 2 and one JSON line that
says so (the column path is benchmarked on one GPU)."""

from __future__ import annotations

import json
import subprocess
import sys

from conftest import ROOT


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120, cwd=ROOT)
    assert r.returncode == 2, r.stdout + r.stderr
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]
