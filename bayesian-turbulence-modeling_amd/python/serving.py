"""Served forward solves: a request stream replayed against the device engine.

The reference serves calibration chains one blocking ForwardModel.evaluate at
a time per process (incompressible/src/ForwardModel.cpp:29-107, called from
python/UQ/datasets/channel_calibration.py:160-243).  Here every request is
submitted to the engine when it arrives and retired when its own solve
converges, so requests from many chains and on different meshes share the
device (continuous batching in librsg.so).

arrival_trace() draws the mixed workload: Poisson arrivals at a total rate
from n_chains chains, theta within +-20 % of the SST defaults for the five
calibrated coefficients (betaStar, a1, sigma_k1, sigma_w1, beta1), the mesh
drawn uniformly per request.  serve() replays a trace in real time and
reports per-request latency (arrival -> result) and throughput.
"""
import math
import time

import numpy as np

THETA_DEFAULTS = [0.09, 0.31, 0.85, 0.5, 0.075]


def arrival_trace(rate, n_requests, n_meshes, n_chains=16, seed=0):
    """List of requests sorted by arrival time: dicts with t (s from start),
    chain, mesh (index), theta (5 values).  Inter-arrival times are
    exponential with mean 1/rate (the superposition of n_chains Poisson
    chains of rate/n_chains each is Poisson at `rate`)."""
    if rate <= 0 or n_requests < 0 or n_meshes <= 0 or n_chains <= 0:
        raise ValueError("rate, n_meshes and n_chains must be positive")
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    chains = rng.integers(0, n_chains, n_requests)
    meshes = rng.integers(0, n_meshes, n_requests)
    th = rng.uniform(0.8, 1.2, (n_requests, 5)) * np.asarray(THETA_DEFAULTS)
    return [dict(t=float(t[i]), chain=int(chains[i]), mesh=int(meshes[i]), theta=[float(v) for v in th[i]])
            for i in range(n_requests)]


def percentile(values, q):
    return float(np.percentile(np.asarray(values, dtype=float), q)) if len(values) else float("nan")


def serve(models, trace, ticket_states, clock=time.perf_counter, sleep=time.sleep, poll_s=0.002,
          on_retire=None):
    """Replay `trace` against `models` (one ForwardModel per mesh index).

    models[i].submit(theta, use_cache=False) -> ticket; ticket_states(list) ->
    engine states (>= 2: finished); models[i].complete(ticket, theta) -> result.
    Each request is submitted once the clock passes its arrival time and
    completed as soon as the engine reports it finished.  Returns a dict with
    per-request records (arrival, done, latency, status, iterations) plus
    throughput and latency percentiles."""
    t0 = clock()
    pending = list(range(len(trace)))       # not yet submitted, in arrival order
    inflight = {}                           # request index -> ticket
    records = [None] * len(trace)
    nxt = 0
    while nxt < len(trace) or inflight:
        now = clock() - t0
        while nxt < len(trace) and trace[nxt]["t"] <= now:
            r = trace[nxt]
            inflight[nxt] = models[r["mesh"]].submit(r["theta"], use_cache=False)
            nxt += 1
        done_any = False
        if inflight:
            idx = list(inflight.keys())
            states = ticket_states([inflight[i] for i in idx])
            for i, st in zip(idx, states):
                if st >= 2:
                    r = trace[i]
                    res = models[r["mesh"]].complete(inflight.pop(i), r["theta"])
                    t_done = clock() - t0
                    records[i] = dict(arrival=r["t"], done=t_done, latency=t_done - r["t"], mesh=r["mesh"],
                                      status=str(res.status).split(".")[-1], iterations=res.simple_iters)
                    done_any = True
                    if on_retire is not None:
                        on_retire(i, records[i])
        if not done_any:
            waits = [poll_s]
            if nxt < len(trace):
                waits.append(max(0.0, trace[nxt]["t"] - (clock() - t0)))
            sleep(min(waits))
    lat = [r["latency"] for r in records]
    span = max(r["done"] for r in records) - min(r["arrival"] for r in records) if records else 0.0
    conv = sum(r["status"] == "Converged" for r in records)
    return dict(records=records, n=len(records), converged=conv,
                throughput=(conv / span) if span > 0 else float("nan"),
                p50=percentile(lat, 50), p99=percentile(lat, 99))
