"""CPU checks of bench.py's host-side legs: the CPU-baseline workers (the
oracle's restatement timed in single-threaded child processes) and the
multi-GPU refusal.  Sizes are tiny; the timings themselves are not checked."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def _child(flag, n):
    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1")
    r = subprocess.run([sys.executable, BENCH, flag, str(n), "--cpu-baseline-seed", "3"],
                       capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


@pytest.mark.parametrize("flag", ["--cpu-baseline-child", "--cpu-config1-child"])
def test_cpu_worker_prints_a_rate(flag):
    """Each worker prints one JSON line whose rate is its sample over its
    time."""
    d = _child(flag, 2)
    assert d["seconds"] > 0
    assert d["value"] == pytest.approx(2 / d["seconds"])


def test_run_cpu_workers_starts_them_together():
    sys.path.insert(0, ROOT)
    import bench
    outs = bench.run_cpu_workers("--cpu-config1-child", 2, 2)
    assert len(outs) == 2 and all(o["value"] > 0 for o in outs)
    assert 1 <= bench.host_cores() <= 16
    assert isinstance(bench.cpu_model(), str)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, BENCH, "--gpus", "2"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
