"""The benchmark's command-line contract that holds without a GPU: a request
for more than one GPU is refused with exit status 2 and one JSON line that
says so (the column path is benchmarked on one GPU), and the learned
closure's training schedule is parsed and checked before any GPU work."""

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


def test_lr_schedule_parsing():
    """--lr-schedule: EPOCHS@LR stages in order; malformed or non-positive
    stages are refused before any GPU work."""
    sys.path.insert(0, str(ROOT))
    import bench
    assert bench.parse_schedule("180@1e-3,60@3e-4,60@1e-4,60@3e-5") == [(180, 1e-3), (60, 3e-4), (60, 1e-4),
                                                                         (60, 3e-5)]
    assert bench.parse_schedule("10@0.001") == [(10, 1e-3)]
    for bad in ("0@1e-3", "10@0", "10@-1e-3", "10", "x@1e-3"):
        try:
            bench.parse_schedule(bad)
        except (SystemExit, ValueError):
            continue
        raise AssertionError(f"accepted {bad!r}")
