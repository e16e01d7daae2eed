"""Autotuner of the engine's launch parameters, measured on the workload itself.

The engine has three shape choices that change speed but never results
(results are bit-identical across them, tested): the member workgroup size
(rsg_set_launch: 256 or 512 threads), the persistent-launch time slice
(rsg_set_time_slice), and, for single large meshes, the path and team size
(rsg_set_grid: kernel per phase; rsg_set_team: one cooperative launch of
`blocks` workgroups).  The best values depend on the device (CU count,
occupancy) and on the workload's mesh sizes, so they are measured rather than
fixed:

tune_members(rs, specs)   times every (block, slice) candidate on the
    ensemble capped at a few hundred outer iterations per member (every member
    cold, all submitted at once, as the ensemble runs) and keeps the one with
    the most member outer iterations per second.
tune_large(rs, mesh, ...) times the kernel-per-phase path and team sizes on a
    few outer iterations of one large member and keeps the fastest per outer
    iteration (the cold start is measured apart and not scored).

Each result is stored in a JSON cache keyed by the device's compute-unit count
and the workload's size signature, so a repeated run reuses it.  The tuner
calls only the public rans_sst_py surface (engine_set_*, ForwardModel).
"""
import json
import os
import time

import numpy as np

MEMBER_CANDIDATES = [(512, 100), (512, 200), (512, 400), (256, 100), (256, 200), (256, 400)]
TEAM_CANDIDATES = (16, 32, 64, 128)


def _cache_load(path):
    if path and os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def _cache_store(path, cache):
    if not path:
        return
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def _signature(kind, cells, extra=""):
    c = np.asarray(cells)
    return "%s|cells %d-%d|n %d|%s" % (kind, int(c.min()), int(c.max()), len(c), extra)


def _device_key(rs):
    info = rs.engine_device_info()
    return "cus %d" % int(info["cu_count"])


def tune_members(rs, specs, precision="fp64", sample_iters=300, candidates=MEMBER_CANDIDATES,
                 cache_path=None, tol=1e-8):
    """Best (block, slice_ms) for the channel ensemble `specs` (dicts as
    ensembles.channel_ensemble makes).  Returns (choice, table) where table
    lists every candidate's member outer iterations per second; the engine
    is left configured with the choice."""
    import ensembles
    key = _device_key(rs) + "|" + _signature("members", [4 * sp["ny"] for sp in specs], precision)
    cache = _cache_load(cache_path)
    if key in cache:
        hit = cache[key]
        rs.engine_set_launch(500, hit["block"])
        rs.engine_set_time_slice(int(hit["slice_ms"] * 1000))
        return hit, hit.get("table", [])
    meshes, fms = ensembles.build_models(rs, specs, tol=tol, max_iterations=sample_iters)
    for fm in fms:
        fm.set_precision(precision)
    thetas = [sp["theta"] for sp in specs]
    table = []
    for block, slice_ms in candidates:
        rs.engine_set_launch(500, block)
        rs.engine_set_time_slice(int(slice_ms * 1000))
        for fm in fms:
            fm.clear_cache()
        t0 = time.perf_counter()
        res = rs.evaluate_many(fms, thetas)
        wall = time.perf_counter() - t0
        its = sum(int(r.simple_iters) for r in res)
        table.append({"block": block, "slice_ms": slice_ms, "wall_s": wall, "member_iters_per_s": its / wall})
    best = max(table, key=lambda r: r["member_iters_per_s"])
    choice = {"block": best["block"], "slice_ms": best["slice_ms"], "table": table,
              "sample": "%d members capped at %d outer iterations, cold, submitted at once" % (len(specs), sample_iters)}
    rs.engine_set_launch(500, choice["block"])
    rs.engine_set_time_slice(int(choice["slice_ms"] * 1000))
    cache[key] = choice
    _cache_store(cache_path, cache)
    return choice, table


def tune_large(rs, mesh, make_model, theta, iters=(10, 30), team_candidates=TEAM_CANDIDATES, cache_path=None):
    """Fastest large-member path for `mesh`: kernel per phase, or a team of
    `blocks` workgroups.  make_model(mesh, max_iterations) returns a
    ForwardModel; it is solved at `theta`, cold, for iters[0] and iters[1] outer
    iterations, and the difference over iters[1] - iters[0] is the score.
    Returns (choice, table); the engine is left configured with the choice
    (min_cells = the mesh's cell count)."""
    n = mesh.n_cells()
    key = _device_key(rs) + "|" + _signature("large", [n])
    cache = _cache_load(cache_path)

    def apply(c):
        if c["path"] == "grid":
            rs.engine_set_team(0, 0, 20)
            rs.engine_set_grid(n, 20)
        else:
            rs.engine_set_grid(0, 20)
            rs.engine_set_team(n, c["blocks"], 20)

    if key in cache:
        apply(cache[key])
        return cache[key], cache[key].get("table", [])
    cands = [{"path": "grid", "blocks": 0}] + [{"path": "team", "blocks": b} for b in team_candidates]
    table = []
    for c in cands:
        apply(c)
        walls = []
        for k in iters:
            fm = make_model(mesh, k)
            fm.clear_cache()
            t0 = time.perf_counter()
            fm.evaluate(list(theta))
            walls.append(time.perf_counter() - t0)
        per = (walls[1] - walls[0]) / (iters[1] - iters[0])
        table.append(dict(c, ms_per_outer_iter=1e3 * per, cold_s=walls[0] - iters[0] * per))
    best = min(table, key=lambda r: r["ms_per_outer_iter"])
    choice = {"path": best["path"], "blocks": best["blocks"], "table": table,
              "sample": "one member, cold, %d and %d outer iterations" % iters}
    apply(choice)
    cache[key] = choice
    _cache_store(cache_path, cache)
    return choice, table
