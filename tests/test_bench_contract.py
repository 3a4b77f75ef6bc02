"""bench.py's output contract: one JSON line on stdout carrying the record's
fields, whatever the legs print on their way (the library's solvers print the
reference's end-of-run lines through C stdio); and --gpus above 1 refused with
a message."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")

RECORD_KEYS = ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
               "scaling", "vs_baseline", "dtype", "data", "config", "roofline")


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, BENCH, "--gpus", "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert r.stdout == ""
    assert "one GPU" in r.stderr


@pytest.mark.gpu
def test_short_run_prints_one_json_line():
    """A short window of the headline leg with the 1D ensemble and the 2D batch
    legs (the batch leg solves its grids one at a time through the solver that
    prints the reference's "Time is up" lines)."""
    r = subprocess.run([sys.executable, "-u", BENCH, "--steps", "2", "--warmup", "1", "--no-cpu", "--no-two-fluid",
                        "--no-unsplit"], capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stderr[-3000:]
    lines = [line for line in r.stdout.splitlines() if line.strip()]
    assert len(lines) == 1, r.stdout[:2000]
    rec = json.loads(lines[0])
    for k in RECORD_KEYS:
        assert k in rec, k
    assert rec["n_gpus"] == 1 and rec["steps"] == 2 and rec["warmup"] == 1 and rec["value"] > 0
    assert rec["higher_is_better"] is True and rec["dtype"] == "f64" and "workload" in rec["config"]
    for k in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert k in rec["roofline"], k
    assert 0 < rec["roofline"]["frac"] < 1
    assert "ensemble_1d" in rec and "batch_2d" in rec
