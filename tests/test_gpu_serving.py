"""The served workload (C3 subset) against the real engine: a Poisson request
stream from 16 chains over the channel column (128 cells) and the Couette
8 x 192 mesh, replayed in real time through serving.serve(), every request
submitted as it arrives and retired as its own solve ends (continuous
batching).  Each sampled request's fields must be bitwise equal to the same
request solved alone on an idle engine."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_served_requests_equal_their_solves_alone(rs):
    import serve_c3
    import serving
    rs.engine_set_time_slice(20000)
    # a cap keeps the test short: the channel column converges (~23,600 outer
    # iterations), the Couette member stops at 30,000 unconverged
    keep, models, names = serve_c3.build_models(rs, 1e-8, 30000)
    trace = serving.arrival_trace(200.0, 24, len(models), n_chains=16, seed=3)
    sample = {i for i in range(len(trace)) if i % 8 == 0}
    captured = {}

    def grab(i, rec):
        if i in sample:
            f = models[trace[i]["mesh"]].last_fields()
            captured[i] = {k: np.array(f[k]) for k in ("U", "p", "k", "omega", "nuT")}

    out = serving.serve(models, trace, rs.ticket_states, on_retire=grab)
    assert out["n"] == len(trace)
    recs = out["records"]
    assert all(r is not None for r in recs)
    # requests overlapped on the device: some arrived before an earlier one retired
    assert any(recs[j]["arrival"] < recs[i]["done"] for i in range(len(recs)) for j in range(i + 1, len(recs)))
    conv = {names[m]: sum(r["status"] == "Converged" for r in recs if r["mesh"] == m) for m in range(len(names))}
    print("served %d requests: %s, p50 %.2f s, p99 %.2f s" % (out["n"], conv, out["p50"], out["p99"]))
    assert conv["channel_column_128"] == sum(r["mesh"] == 0 for r in recs)
    for i in sorted(sample):
        m = trace[i]["mesh"]
        models[m].clear_cache()
        res = models[m].evaluate(trace[i]["theta"])
        assert res.simple_iters == recs[i]["iterations"], (i, res.simple_iters, recs[i]["iterations"])
        alone = models[m].last_fields()
        for k, v in captured[i].items():
            assert np.array_equal(np.array(alone[k]), v), (i, names[m], k)
