"""Host-side pieces of the serving loop (no GPU)."""
import numpy as np

from opensubdiv_amd.serve import Arena, poisson_requests


def test_arena_first_fit_and_merge():
    a = Arena(1024, 64)
    x = a.alloc(100)
    y = a.alloc(64)
    z = a.alloc(500)
    assert x == (0, 128) and y == (128, 64) and z == (192, 512)
    assert a.alloc(400) is None
    a.release(*y)
    assert a.alloc(64) == (128, 64)
    a.release(128, 64)
    a.release(*x)
    a.release(*z)
    assert a.free == [(0, 1024)] and a.used() == 0


def test_poisson_requests_rate_and_mix():
    reqs = poisson_requests(20000, rate=500.0, ntopo=32, seed=1)
    t = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(t) >= 0)
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert abs(gaps.mean() - 1 / 500.0) < 0.05 / 500.0
    # exponential: coefficient of variation near 1
    assert 0.95 < gaps.std() / gaps.mean() < 1.05
    kinds = [r.kind for r in reqs]
    assert 0.48 < kinds.count("limit") / len(kinds) < 0.52
    assert {r.topo for r in reqs} == set(range(32))
