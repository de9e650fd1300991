"""bench.py's command line on the CPU: the one-GPU contract (--gpus above 1
exits with a message and status 2 before touching a device), its options, and
the synthetic inputs it builds (one-cell and corner wells, the fingered front)."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_status_2():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2", "--steps", "1"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_help_lists_the_options():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"], capture_output=True, text=True,
                       timeout=120, cwd=ROOT)
    assert r.returncode == 0
    for opt in ("--gpus", "--steps", "--warmup", "--no-cpu-baseline", "--no-modes", "--no-ensemble",
                "--no-rotated", "--no-late-time", "--no-roofline"):
        assert opt in r.stdout


def test_one_cell_source_is_the_reference_quarter_five_spot():
    n = 16
    q = bench.one_cell_source(n)
    h = 1.0 / n
    assert q.shape == (n * n,)
    assert q[0] == 1.0 / (h * h) and q[n * n - 1] == -1.0 / (h * h)
    assert np.count_nonzero(q) == 2


def test_corner_wells_source_is_the_reference_parallel_grid():
    n = 16
    q = bench.corner_wells_source(n)
    h = 2 ** 0.5 / n
    v = 1.0 / (h * h)
    assert q[0] == v and q[n * n - 1] == v and q[n - 1] == -v and q[(n - 1) * n] == -v
    assert np.count_nonzero(q) == 4 and abs(q.sum()) < 1e-9 * v


def test_fingered_front():
    n = 256
    c = bench.fingered_front(n)
    assert c.shape == (n + 4, n + 4)
    inner = c[2:-2, 2:-2]
    assert inner.min() >= 0.0 and inner.max() <= 1.0
    assert inner[0, 0] > 0.99 and inner[-1, -1] < 1e-6
    # a quarter disc of radius 0.4 covers pi 0.16 / 4 of the unit square
    assert abs(np.mean(inner > 0.5) - np.pi * 0.16 / 4) < 0.01
    assert np.all(c[:2] == 0) and np.all(c[:, :2] == 0)
