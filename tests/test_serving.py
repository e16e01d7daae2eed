"""The served-workload harness (python/serving.py) on a simulated engine:
arrival statistics, replay order, latency accounting."""
import math

import numpy as np
import pytest

import serving


def test_arrival_trace_is_poisson_and_seeded():
    tr = serving.arrival_trace(200.0, 4000, 3, n_chains=16, seed=7)
    assert tr == serving.arrival_trace(200.0, 4000, 3, n_chains=16, seed=7)
    t = np.array([r["t"] for r in tr])
    assert np.all(np.diff(t) > 0)
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert gaps.mean() == pytest.approx(1 / 200.0, rel=0.05)          # exponential mean
    assert np.std(gaps) == pytest.approx(1 / 200.0, rel=0.08)          # exponential: std == mean
    meshes = np.bincount([r["mesh"] for r in tr], minlength=3)
    assert meshes.min() > 4000 / 3 * 0.9
    chains = np.bincount([r["chain"] for r in tr], minlength=16)
    assert chains.min() > 4000 / 16 * 0.7
    th = np.array([r["theta"] for r in tr])
    lo = 0.8 * np.array(serving.THETA_DEFAULTS); hi = 1.2 * np.array(serving.THETA_DEFAULTS)
    assert np.all(th >= lo) and np.all(th <= hi)
    with pytest.raises(ValueError):
        serving.arrival_trace(0.0, 10, 3)


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += max(dt, 1e-4)


class FakeResult:
    def __init__(self, status, iters):
        self.status = status
        self.simple_iters = iters


class FakeModel:
    """Each solve takes `cost` seconds of simulated time after submission."""
    def __init__(self, engine, mesh, cost):
        self.engine, self.mesh, self.cost = engine, mesh, cost

    def submit(self, theta, use_cache=True):
        assert use_cache is False
        return self.engine.add(self.cost, theta)

    def complete(self, ticket, theta):
        assert self.engine.state(ticket) == 2
        return FakeResult("EvaluationStatus.Converged", int(1000 * self.cost))


class FakeEngine:
    def __init__(self, clock):
        self.clock, self.tickets = clock, {}

    def add(self, cost, theta):
        tid = len(self.tickets)
        self.tickets[tid] = self.clock() + cost
        return tid

    def state(self, tid):
        return 2 if self.clock() >= self.tickets[tid] else 1

    def states(self, tids):
        return [self.state(t) for t in tids]


def test_serve_replays_in_time_and_measures_latency():
    clk = FakeClock()
    eng = FakeEngine(clk)
    models = [FakeModel(eng, 0, 0.5), FakeModel(eng, 1, 2.0)]
    tr = serving.arrival_trace(20.0, 60, 2, seed=3)
    out = serving.serve(models, tr, eng.states, clock=clk, sleep=clk.sleep, poll_s=1e-3)
    assert out["n"] == 60 and out["converged"] == 60
    for r, q in zip(out["records"], tr):
        cost = 0.5 if q["mesh"] == 0 else 2.0
        assert r["arrival"] == q["t"]
        # submitted within one poll of arrival, retired within one poll of finishing
        assert cost - 1e-9 <= r["latency"] <= cost + 3e-3
    lat = [r["latency"] for r in out["records"]]
    assert out["p50"] == pytest.approx(np.percentile(lat, 50))
    assert out["p99"] <= 2.0 + 3e-3
    span = max(r["done"] for r in out["records"]) - tr[0]["t"]
    assert out["throughput"] == pytest.approx(60 / span)
