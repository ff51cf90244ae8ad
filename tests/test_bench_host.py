"""Host-side checks of bench.py (no GPU): the one-GPU contract and the
serving workload's request stream (Poisson arrivals at a mean 200/s,
degree uniform in 2..6, cells per direction from {8, 16, 24, 32, 48}, a
unit velocity, a Gaussian of width in [0.05, 0.2], end time 0.5)."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"),
                        "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_serving_stream_follows_the_request_mix():
    n = 2000
    arrivals, reqs = bench.serving_requests(n, 0.5)
    assert len(reqs) == n and arrivals.shape == (n,)
    gaps = np.diff(np.concatenate([[0.0], arrivals]))
    assert np.all(gaps > 0)
    # exponential gaps of mean 1/200 s: the sample mean lies within 5
    # standard errors (1/200/sqrt(n)), the coefficient of variation near 1
    assert abs(gaps.mean() - 1 / 200) < 5 * (1 / 200) / np.sqrt(n)
    assert abs(gaps.std() / gaps.mean() - 1) < 0.1
    deg = np.array([r.degree for r in reqs])
    cells = np.array([r.cells_per_direction for r in reqs])
    assert set(deg) == {2, 3, 4, 5, 6}
    assert set(cells) == {8, 16, 24, 32, 48}
    # uniform draws: every value within 5 binomial standard deviations
    for values, k in ((deg, 5), (cells, 5)):
        counts = np.unique(values, return_counts=True)[1]
        sd = np.sqrt(n * (1 / k) * (1 - 1 / k))
        assert np.all(np.abs(counts - n / k) < 5 * sd)
    for r in reqs:
        v = np.array(r.velocity[:])
        assert abs(np.linalg.norm(v) - 1) < 1e-12
        assert 0.05 <= r.width <= 0.2
        assert all(0.2 <= c <= 0.8 for c in r.center)
        assert r.final_time == 0.5 and r.courant == 0.3
    # velocity directions cover the sphere: the mean direction is near 0
    vs = np.array([r.velocity[:] for r in reqs])
    assert np.all(np.abs(vs.mean(axis=0)) < 5 / np.sqrt(3 * n))


def test_serving_stream_is_seeded():
    a1, r1 = bench.serving_requests(50, 0.5, seed=7)
    a2, r2 = bench.serving_requests(50, 0.5, seed=7)
    a3, _ = bench.serving_requests(50, 0.5, seed=8)
    assert np.array_equal(a1, a2)
    assert [(x.degree, x.cells_per_direction, x.width) for x in r1] == \
        [(x.degree, x.cells_per_direction, x.width) for x in r2]
    assert not np.array_equal(a1, a3)
