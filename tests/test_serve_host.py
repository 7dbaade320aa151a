"""Host logic of the load generator (no device)."""
import numpy as np

from mcrt.serve import arrivals, latency_summary


def test_poisson_arrivals_have_the_rate():
    t = arrivals(200.0, 60.0, seed=5)
    assert np.all(np.diff(t) > 0) and t[-1] < 60.0
    assert abs(len(t) - 12000) < 5 * np.sqrt(12000)
    gaps = np.diff(t)
    assert abs(gaps.mean() - 1 / 200.0) < 0.05 / 200.0
    assert np.array_equal(t, arrivals(200.0, 60.0, seed=5))


def test_latency_percentiles():
    s = latency_summary(np.arange(1, 101, dtype=float))
    assert s["n"] == 100 and np.isclose(s["p50_ms"], 50.5) and np.isclose(s["p99_ms"], 99.01)
    assert latency_summary([]) == {"n": 0}
