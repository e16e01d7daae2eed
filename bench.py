"""Benchmark: converged RANS forward solves per second on one MI355X.

Headline workload (C2): the batched ensemble -- 512 streamwise-periodic
channel solves (4 x ny quads, ny 96-384 from Re_tau log-uniform in
[180, 2000], betaStar/a1/sigma_k1/sigma_w1/beta1 within +-20 % of their
defaults), each driven by SIMPLE + SST k-omega to the reference's convergence
criterion at tol 1e-8 (200,000-iteration cap), all members resident on the
device (one workgroup per member, persistent launches time-sliced at 200 ms,
continuous batching, largest meshes first).

One timed step = the whole ensemble solved cold (every member submitted at
once, timed to the last retirement).  A step takes minutes, so --steps and
--warmup are upper bounds: the warmup is one pass of every member capped at
100 outer iterations (engine start-up, code-object load), and timed steps
after the first run only while they fit --step-budget-s; the JSON line
reports the steps actually timed (and the requested counts beside them).

Prints ONE JSON line.  value = converged solves / wall second over the timed
steps; p50/p99 = per-member latency (submit -> result).  --precision mixed
runs the members' linear solves as FP32 inner solves under FP64 refinement
and then re-solves a sample of converged members in FP64 to measure
max_rel_l2_vs_fp64.  roofline = algorithmic bytes of the SIMPLE chunk kernel
(algorithmic_bytes) / its HIP-event time on the engine stream, against the
8 TB/s HBM peak; traffic = the kernel's measured HBM bytes per launch (two
rocprofv3 --pmc passes over a capped run, scaled by a copy kernel of known
size).  cpu_baseline = the reference solver compiled from its own sources
(oracle/_ref) on this host's cores, P = 1 and P = all, timed on a bounded
sample and scaled by this run's outer-iteration counts.

Other workloads (--workload): c1 (the stand-alone 261,600-cell
backward-facing-step pressure system, rsg_system_solve, FP64 and mixed, next
to the reference's CPU PCG), c4p (the same at 2M cells), c0 (single-solve
latency of the Re_tau 550 channel and the same mesh batched), walldist (the
device wall-distance kernel on a 2M-cell mesh), bump (team solves of one
bump-channel member at 40k, 160k and 2M cells).
"""
import argparse
import glob
import json
import math
import os
import shutil
import sqlite3
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "bayesian-turbulence-modeling_amd")
sys.path.insert(0, PKG)
sys.path.insert(0, os.path.join(PKG, "python"))

METRIC = "converged RANS solves/sec at <=1e-6 rel-L2 vs FP64; p50/p99 solve latency"
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(n, nif, nb, outer, inner, w_inner=8.0):
    """Algorithmic HBM bytes of one member's solve, by the rule: every logical
    input and output array counted once per pass at its stored width (FP64 8,
    FP32 4, int32 4), cell values gathered through face indices once per cell.

    Per pass (w = 8, planar 2-D mesh: two face-normal components):
      SpMV(LDU, value width v)  v(n + 2 f) + 8 f + 2 v n
      Green-Gauss of q fields   f(8 + 3w) + f_b(4 + 2w + q w) + n(w + 3 q w)
      face assembly of one equation, reading c cell fields
                                f(8 + 4w) + 2w f + 2w n + c w n
      SST closure               10 w n (+ the gradients it reads, counted as Green-Gauss)
    One outer iteration: gradients of Ux Uy Uz k omega + closure, grad p,
    momentum x and y (assembly reading 8 cell fields each, grad Ux refreshed
    between), p' assembly (7 fields), grad p', corrections (9 w n), k and
    omega assemblies (8 fields each), clamps (4 w n), wall pin and change
    norms (5 w n); and for each of the five linear solves one FP64 defect
    b - A x (SpMV + 2 w n; mixed precision: two, the initial defect and the
    refinement check).
    Inner iterations (value width w_inner: 8 FP64, 4 mixed):
      BiCGSTAB iteration  2 SpMV + 14 w_inner n
      AMG-PCG iteration   SpMV + 13 w_inner n + V-cycle,
      V-cycle             4/3 (4 SpMV + 14 w_inner n)   (level 0 plus geometric coarsening)
    outer: outer iterations; inner: cumulative inner iterations
    (Ux, Uy, Uz, p, k, omega)."""
    w = 8.0

    def spmv(v):
        return v * (n + 2 * nif) + 8.0 * nif + 2 * v * n

    def gg(q):
        return nif * (8 + 3 * w) + nb * (4 + 2 * w + q * w) + n * (w + 3 * q * w)

    def face_asm(c):
        return nif * (8 + 4 * w) + 2 * w * nif + 2 * w * n + c * w * n

    defects = 2 if w_inner < 8 else 1
    per_outer = (gg(5) + 10 * w * n                # turbulence update: gradients + closure
                 + gg(1)                           # grad p
                 + 2 * face_asm(8) + gg(1)         # momentum x, y; grad Ux between
                 + face_asm(7) + gg(1) + 9 * w * n  # p' assembly, grad p', corrections
                 + 2 * face_asm(8) + 9 * w * n     # k, omega assemblies; clamps, pin, change norms
                 + 5 * defects * (spmv(w) + 2 * w * n))
    wi = w_inner
    bicg = 2 * spmv(wi) + 14 * wi * n
    vcycle = (4.0 / 3.0) * (4 * spmv(wi) + 14 * wi * n)
    pcg_amg = spmv(wi) + 13 * wi * n + vcycle
    inner_bytes = (inner[0] + inner[1] + inner[2] + inner[4] + inner[5]) * bicg + inner[3] * pcg_amg
    return outer * per_outer + inner_bytes


def configure_engine(rs, args):
    rs.engine_set_launch(args.iters_per_launch, args.block)
    rs.engine_set_time_slice(int(args.slice_ms * 1000))
    rs.engine_set_schedule(args.schedule)


def run_batch(rs, specs, tol, fms_cache=None, max_iterations=200000, precision="fp64"):
    import ensembles
    if fms_cache is None:
        meshes, fms = ensembles.build_models(rs, specs, tol=tol, max_iterations=max_iterations)
        for fm in fms:
            fm.set_precision(precision)
    else:
        meshes, fms = fms_cache
    for fm in fms:
        fm.clear_cache()              # every step solves cold (order-independent)
    t0 = time.perf_counter()
    # every member is submitted before any is awaited (what evaluate_many
    # does); finished members are collected as they retire, with a progress
    # line per 32 retirements
    rs.engine_hold_admission(True)     # the whole ensemble starts in one launch
    try:
        tickets = [fm.submit(sp["theta"]) for fm, sp in zip(fms, specs)]
    finally:
        rs.engine_hold_admission(False)
    res = [None] * len(fms)
    left = set(range(len(fms)))
    done = 0
    while left:
        idx = sorted(left)
        states = rs.ticket_states([tickets[i] for i in idx])
        fin = [i for i, st in zip(idx, states) if st >= 2]
        for i in fin:
            res[i] = fms[i].complete(tickets[i], specs[i]["theta"])
            left.discard(i)
            done += 1
            if done % 32 == 0 or done == len(fms):
                print("bench: %d/%d members retired at %.1f s" % (done, len(fms), time.perf_counter() - t0),
                      file=sys.stderr, flush=True)
        if not fin:
            time.sleep(0.05)
    wall = time.perf_counter() - t0
    return (meshes, fms), res, wall


def ensemble_alg_bytes(cache, res, precision):
    meshes, fms = cache
    total = 0.0
    for fm, mesh, r in zip(fms, meshes, res):
        n, nif = mesh.n_cells(), mesh.n_internal_faces()
        total += algorithmic_bytes(n, nif, mesh.n_faces() - nif, r.simple_iters, fm.last_inner_iterations(),
                                   4.0 if precision == "mixed" else 8.0)
    return total


def field_errors(a, b):
    """Relative L2 of U, k, omega, nu_t of solve a against solve b; p against
    the dynamic-pressure scale rms(|U|^2) of b (the periodic channel's p is
    round-off, so it cannot be held against itself)."""
    rel = lambda x, y: float(np.linalg.norm(np.ravel(x) - np.ravel(y)) / max(np.linalg.norm(np.ravel(y)), 1e-300))
    out = {k: rel(a[k], b[k]) for k in ("U", "k", "omega", "nuT")}
    q = np.sqrt(np.mean(np.sum(np.asarray(b["U"]).reshape(-1, 3) ** 2, axis=1) ** 2))
    out["p"] = float(np.linalg.norm(np.asarray(a["p"]) - np.asarray(b["p"])) / np.sqrt(len(b["p"])) / q)
    return out


def fp64_check(rs, specs, cache, res, tol, max_iterations, sample=32, frac=0.6):
    """Each observable's error of the ensemble's solves against FP64 GPU
    solves of the same members: a sample of `sample` converged members spread
    evenly over the fastest `frac` of them by outer-iteration count (the
    slowest need minutes each), re-solved in FP64 in one batch of their own.
    For a mixed-precision ensemble this is its error; for an FP64 ensemble it
    measures that a member's result does not depend on the batch it ran in
    (expected 0: every reduction has a fixed order)."""
    meshes, fms = cache
    conv = sorted((i for i, r in enumerate(res) if str(r.status).split(".")[-1] == "Converged"),
                  key=lambda i: res[i].simple_iters)
    if not conv:
        return None
    pool = conv[:max(1, int(frac * len(conv)))]
    pick = pool[::max(1, len(pool) // sample)][:sample]
    sub = [specs[i] for i in pick]
    (m64, f64), r64, wall = run_batch(rs, sub, tol, None, max_iterations, "fp64")
    worst = {k: 0.0 for k in ("U", "p", "k", "omega", "nuT")}
    iters = []
    for j, i in enumerate(pick):
        e = field_errors(fms[i].last_fields(), f64[j].last_fields())
        for k in worst:
            worst[k] = max(worst[k], e[k])
        iters.append((res[i].simple_iters, r64[j].simple_iters))
    return {"max_rel_l2": max(worst.values()), "per_field": worst, "members": len(pick),
            "cells": [4 * specs[i]["ny"] for i in pick], "fp64_wall_s": wall,
            "outer_iters_mixed_vs_fp64": iters,
            "sample": "%d converged members spread over the fastest %d%% by outer iterations, re-solved in FP64; "
                      "U, k, omega, nu_t relative L2, p against rms(|U|^2)" % (len(pick), int(100 * frac))}


def _member_mesh_file(rs, sp, path):
    y = rs.Mesh.geometric_channel_nodes(sp["ny"] // 2, sp["ratio"], 1.0)
    mesh = rs.Mesh.make_channel_periodic_2d(np.linspace(0.0, 1.0, 5), np.array(y))
    mesh.save_binary(path)
    return 4 * sp["ny"]


def lscpu_info():
    """Host CPU of this machine: model, sockets, physical cores, threads, and
    the CPUs this process may run on."""
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core", "CPU(s)": "cpus"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys:
                info[keys[k.strip()]] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        info["physical_cores"] = int(info["sockets"]) * int(info["cores_per_socket"])
    except (KeyError, ValueError):
        pass
    info["affinity_cpus"] = len(os.sched_getaffinity(0))
    return info


SOLVE_DEFAULT_COEFFS = [0.85, 0.5, 0.075, 5.0 / 9.0, 1.0, 0.856, 0.0828, 0.44, 0.09, 0.31, 0.41]


def _cpu_member_run(task):
    """One reference solve of `iters` outer iterations (tolerance 1e-300, so it
    never stops early) on one pinned core; returns seconds per iteration.
    Runs in a spawned worker process."""
    import ctypes as C
    lib_path, kind, path, nu, theta, iters, cpu = task
    if cpu is not None:
        os.sched_setaffinity(0, {cpu})
    # the solver library reports mesh loads on stdout: keep the worker's
    # stdout off the benchmark's (which carries only the JSON line)
    devnull = os.open(os.devnull, os.O_WRONLY)
    os.dup2(devnull, 1)
    os.close(devnull)
    lib = C.CDLL(lib_path)
    D = lambda a: (C.c_double * len(a))(*a)
    coeffs = list(SOLVE_DEFAULT_COEFFS)
    for idx, v in zip([8, 9, 0, 1, 2], theta):
        coeffs[idx] = v
    packed = [iters, 1e-300, 1e6, 0.7, 0.3, 0.5, 0.5, 200, 1e-3, 5, 1, 500, 0, 1.0, 0, 0,
              1e-10, 1e10, 1e-6, 0, 0.41, 0, 0, 2, 2]
    info = (C.c_int * 5)()
    t0 = time.perf_counter()
    if kind == "reference":
        lib.ref_simple_solve(path.encode(), 0, D([1.0, 1e-4, 1.0]), C.c_double(nu), D(coeffs), D(packed),
                             D([1, 0, 0, 0, 1e-4, 1.0]), *([None] * 13), None, 0, info)
    else:
        lib.om_mesh_load.restype = C.c_void_p
        m = lib.om_mesh_load(path.encode())
        lib.om_simple_solve(C.c_void_p(m), 0, D([1.0, 1e-4, 1.0]), C.c_double(nu), D(coeffs), D(packed),
                            D([1, 0, 0, 0, 1e-4, 1.0]), *([None] * 13), None, 0, info)
    return (time.perf_counter() - t0) / max(info[0], 1)


def _lpt_makespan(costs, p):
    """Longest-processing-time-first schedule of independent solves on p cores."""
    import heapq
    loads = [0.0] * p
    for c in sorted(costs, reverse=True):
        heapq.heapreplace(loads, loads[0] + c)
    return max(loads)


def cpu_baseline(specs, results, sample_iters=2000, max_workers=16):
    """The reference solver compiled from its own sources (oracle/_ref; the
    oracle's C restatement where the reference library is absent) on the host
    cores, timed on a bounded sample of the same workload.

    Sample: `sample_iters` outer iterations of the smallest, median and largest
    member of the ensemble -- once one after the other on one pinned core
    (P = 1), once as P concurrent pinned processes, one per physical core of
    this process's CPU share (at most `max_workers`), the three sizes assigned
    round-robin (P = all, so memory-bandwidth sharing is in the figure).  Each
    gives the CPU seconds per outer iteration against the cell count (linear
    fit); the ensemble's CPU time is this run's per-member outer-iteration
    counts (the same SIMPLE algorithm, same counts) times that cost, scheduled
    longest-first on P cores.  value = converged solves / that time at P = all."""
    import multiprocessing as mp
    lib_path = os.path.join(ROOT, "oracle", "_ref", "libref.so")
    kind = "reference"
    if not os.path.exists(lib_path):
        lib_path = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
        kind = "port"
    if not os.path.exists(lib_path):
        return None
    import rans_sst_py as rs
    by_size = sorted(specs, key=lambda s: s["ny"])
    picks = [by_size[0], by_size[len(by_size) // 2], by_size[-1]]
    cells, paths = [], []
    for i, sp in enumerate(picks):
        path = "/tmp/bench_cpu_member_%d.msh" % i
        cells.append(_member_mesh_file(rs, sp, path))
        paths.append(path)
    cpu = lscpu_info()
    avail = sorted(os.sched_getaffinity(0))
    P = max(1, min(len(avail), max_workers, cpu.get("physical_cores", len(avail))))
    ctx = mp.get_context("spawn")
    with ctx.Pool(P) as pool:
        # P = 1: the three members one after the other on one core
        one = [pool.apply(_cpu_member_run, ((lib_path, kind, paths[i], picks[i]["nu"], picks[i]["theta"],
                                             sample_iters, avail[0]),)) for i in range(3)]
        allc = pool.map(_cpu_member_run, [(lib_path, kind, paths[w % 3], picks[w % 3]["nu"], picks[w % 3]["theta"],
                                           sample_iters, avail[w]) for w in range(P)], chunksize=1)
    loaded = [np.mean([allc[w] for w in range(P) if w % 3 == i] or [one[i]]) for i in range(3)]
    conv = sum(r["converged"] for r in results)

    def fit(per_iter):
        a, b = np.polyfit(np.array(cells, dtype=float), np.array(per_iter), 1)
        return [max(r["outer"], 1) * (a * r["cells"] + b) for r in results]

    t1 = sum(fit(one))
    tp = _lpt_makespan(fit(loaded), P)
    return {"value": conv / tp, "unit": "solves/s", "cores": P, "kind": kind,
            "p1_value": conv / t1,
            "cpu": {k: cpu.get(k) for k in ("model", "sockets", "physical_cores", "threads_per_core",
                                            "affinity_cpus")},
            "sample": "%d outer iterations of the smallest/median/largest member (%s cells): %s ms per iteration "
                      "on one pinned core, %s ms with %d pinned processes at once; ensemble CPU time from this "
                      "run's outer-iteration counts: %.0f s on one core, %.0f s on %d cores (longest first)"
                      % (sample_iters, "/".join(str(c) for c in cells), "/".join("%.3f" % (1e3 * v) for v in one),
                         "/".join("%.3f" % (1e3 * v) for v in loaded), P, t1, tp, P)}


COPY_BYTES = 256 << 20


def _rsg_lib():
    import ctypes as C
    lib = C.CDLL(os.path.join(PKG, "librsg.so"))
    lib.rsg_last_error.restype = C.c_char_p
    return lib


def traffic_child(args):
    """Child of a rocprofv3 --pmc pass: a plain copy kernel of COPY_BYTES on the
    engine's library (the counter scale), then the headline ensemble capped at
    --max-iterations outer iterations; prints the capped run's algorithmic
    bytes (JSON)."""
    import ctypes as C
    import rans_sst_py as rs
    import ensembles
    lib = _rsg_lib()
    ctx = C.c_void_p()
    if lib.rsg_context_create(0, C.c_uint64(1 << 30), C.byref(ctx)) != 0:
        raise RuntimeError("rsg_context_create failed")
    sec = C.c_double()
    if lib.rsg_copy_calibration(ctx, C.c_uint64(COPY_BYTES), C.byref(sec)) != 0:
        raise RuntimeError("copy calibration: %s" % lib.rsg_last_error(ctx))
    lib.rsg_context_destroy(ctx)
    specs = ensembles.channel_ensemble(args.members, args.seed)
    configure_engine(rs, args)
    cache, res, _ = run_batch(rs, specs, args.tol, None, args.max_iterations, args.precision)
    print(json.dumps({"alg_bytes": ensemble_alg_bytes(cache, res, args.precision)}), flush=True)


def _pmc_run(counter, argv, tag, tmp):
    """One rocprofv3 --pmc pass over `argv` (a python program); returns
    ([(kernel name, counter KiB)], stdout)."""
    rocprof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    out = os.path.join(tmp, tag)
    cmd = ["timeout", "-s", "KILL", "120", rocprof, "--pmc", counter, "-d", out, "-o", "run", "--",
           sys.executable, "-u"] + argv
    r = subprocess.run(cmd, cwd="/tmp", env=dict(os.environ, TMPDIR="/tmp"), capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("%s pass exited with %d: %s" % (counter, r.returncode, r.stderr[-400:]))
    dbs = glob.glob(os.path.join(out, "**", "*.db"), recursive=True)
    if not dbs:
        raise RuntimeError("%s pass wrote no database" % counter)
    rows = []
    for db in dbs:
        con = sqlite3.connect(db)
        rows += list(con.execute("select kernel_name, value from counters_collection where counter_name = ?",
                                 (counter,)))
        con.close()
    return rows, r.stdout


def measure_traffic(args, cap=300):
    """HBM traffic of the chunk kernel relative to its algorithmic bytes.

    Two counter passes (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch), each
    over one child process that runs the copy kernel of known size and then the
    same ensemble capped at `cap` outer iterations; each counter is scaled by
    what it reads for the copy, and the scaled bytes are divided by the capped
    run's algorithmic bytes."""
    tmp = tempfile.mkdtemp(prefix="bench_pmc_", dir="/tmp")
    try:
        me = os.path.abspath(__file__)
        work = [me, "--traffic-child", "--members", str(args.members), "--seed", str(args.seed), "--tol",
                str(args.tol), "--iters-per-launch", str(args.iters_per_launch), "--block", str(args.block),
                "--slice-ms", str(args.slice_ms), "--schedule", args.schedule, "--precision", args.precision,
                "--max-iterations", str(cap)]
        out = {}
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            rows, stdout = _pmc_run(counter, work, counter, tmp)
            cal = [v for k, v in rows if "k_copy_calibration" in k]
            kib = [v for k, v in rows if "k_simple_chunk" in k]
            if not cal or not kib:
                raise RuntimeError("%s pass: copy or chunk kernel missing from the counters" % counter)
            scale = COPY_BYTES / (1024.0 * cal[0])
            alg = json.loads(stdout.strip().splitlines()[-1])["alg_bytes"]
            out[counter] = dict(scale=scale, bytes=1024.0 * sum(kib) * scale, alg=alg, launches=len(kib))
        f, w = out["FETCH_SIZE"], out["WRITE_SIZE"]
        ratio = f["bytes"] / f["alg"] + w["bytes"] / w["alg"]
        return {"hbm_over_algorithmic": ratio, "fetch_scale": f["scale"], "write_scale": w["scale"],
                "fetch_bytes": f["bytes"], "write_bytes": w["bytes"], "alg_bytes": f["alg"],
                "launches": f["launches"],
                "sample": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE passes over the same %d-member ensemble capped "
                          "at %d outer iterations; counters scaled by a %d MiB copy kernel (rsg_copy_calibration)"
                          % (args.members, cap, COPY_BYTES >> 20)}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def bench_c0(args):
    """The single-solve latency case: the periodic Re_tau 550 channel (4 x 128,
    geometric 1.06, default SST coefficients) solved cold to tol 1e-8 through
    the rans_sst_py drop-in, one solve at a time."""
    import ctypes as C
    import rans_sst_py as rs
    import ensembles
    rs.engine_set_launch(args.iters_per_launch, args.block)
    rs.engine_set_time_slice(int(args.slice_ms * 1000))
    spec = dict(ny=128, ratio=1.06, nu=1.0 / 550.0, theta=list(ensembles.THETA_DEFAULTS))
    meshes, fms = ensembles.build_models(rs, [spec], tol=1e-8, max_iterations=200000)
    fm = fms[0]
    lat, iters, conv = [], 0, 0
    for k in range(args.warmup + args.steps):
        fm.clear_cache()
        t0 = time.perf_counter()
        r = fm.evaluate(spec["theta"])
        dt = time.perf_counter() - t0
        print("bench: c0 solve %.2f s (%d iterations)" % (dt, r.simple_iters), file=sys.stderr, flush=True)
        if k >= args.warmup:
            lat.append(dt); iters = r.simple_iters
            conv += str(r.status).split(".")[-1] == "Converged"
    batch = None
    if args.members > 1:
        # the same mesh batched: `members` solves with the ensemble's theta
        # spread (five coefficients +-20 %), all submitted at once
        specs = ensembles.channel_ensemble(args.members, args.seed)
        bspecs = [dict(spec, theta=sp["theta"]) for sp in specs]
        (_, bfms), res, wall = run_batch(rs, bspecs, 1e-8, ensembles.build_models(rs, bspecs, tol=1e-8))
        bconv = sum(str(r.status).split(".")[-1] == "Converged" for r in res)
        batch = {"members": args.members, "wall_s": wall, "converged": int(bconv),
                 "solves_per_s": bconv / wall,
                 "outer_iters_median": float(np.median([r.simple_iters for r in res])),
                 "p50_latency_s": float(np.percentile([r.latency for r in res], 50)),
                 "p99_latency_s": float(np.percentile([r.latency for r in res], 99))}
    cpu = None
    lib_ref = os.path.join(ROOT, "oracle", "_ref", "libref.so")
    if not args.no_cpu and os.path.exists(lib_ref):
        L = C.CDLL(lib_ref)
        D = lambda a: (C.c_double * len(a))(*a)
        path = "/tmp/bench_c0.msh"
        meshes[0].save_binary(path)
        coeffs = [0.85, 0.5, 0.075, 5.0 / 9.0, 1.0, 0.856, 0.0828, 0.44, 0.09, 0.31, 0.41]
        packed = [200000, 1e-8, 1e6, 0.7, 0.3, 0.5, 0.5, 200, 1e-3, 5, 1, 500, 0, 1.0, 0, 0,
                  1e-10, 1e10, 1e-6, 0, 0.41, 0, 0, 2, 2]
        info = (C.c_int * 5)()
        t0 = time.perf_counter()
        L.ref_simple_solve(path.encode(), 0, D([1.0, 1e-4, 1.0]), C.c_double(spec["nu"]), D(coeffs), D(packed),
                           D([1, 0, 0, 0, 1e-4, 1.0]), *([None] * 13), None, 0, info)
        dt = time.perf_counter() - t0
        cpu = {"value": (1.0 / dt) if info[1] else 0.0, "unit": "solves/s", "cores": 1, "kind": "reference",
               "sample": "one full C0 solve by the reference solver: %d iterations, converged %d, %.1f s"
                         % (info[0], info[1], dt)}
    t = float(np.median(lat))
    print(json.dumps({
        "metric": METRIC, "value": conv / sum(lat), "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * t, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "C0 periodic channel 4x128, geometric 1.06, Re_tau 550, SIMPLE+SST to tol 1e-8, "
                               "one solve at a time", "cells": 512, "tol": 1e-8},
        "p50_latency_s": t, "outer_iterations": int(iters), "converged": conv, "batched_same_mesh": batch,
        "roofline": None, "cpu_baseline": cpu}), flush=True)


def bench_comp(args):
    """Compressible pressure-based SIMPLE (kernel per phase, FP64) on the
    reference's Ma 0.1 regression channel (40 x 30, its own settings):
    single-solve latency and a batch of theta samples solved concurrently,
    beside the reference solver (oracle/_ref/libref.so, one core) on the
    same case; then the outer-iteration time on a finer channel of the same
    family (--comp-cells), with the reference timed over a bounded sample of
    its iterations."""
    import ctypes as C
    import rans_sst_py as rs
    import compressible_diagnostics as cd
    D = lambda a: (C.c_double * len(a))(*a)
    rs.engine_set_graphs(bool(args.graphs))
    rs.engine_set_launch(args.iters_per_launch, 512)
    rs.engine_set_time_slice(int(args.slice_ms * 1000))
    case = cd.make_validation_case("ma_0_1", 0.1, nx=40, ny=30, max_iterations=3000)
    fm = case["forward"]
    theta0 = list(case["param_set"].pack(rs.SSTCoefficients()))
    lat, iters, conv = [], 0, 0
    for k in range(args.warmup + args.steps):
        t0 = time.perf_counter()
        r = fm.evaluate(theta0)
        dt = time.perf_counter() - t0
        print("bench: comp solve %.2f s (%d iterations, %s)" % (dt, r.simple_iters, r.status), file=sys.stderr,
              flush=True)
        if k >= args.warmup:
            lat.append(dt); iters = r.simple_iters
            conv += str(r.status).split(".")[-1] == "Converged"
    summ = cd.summarise_fields(fm.last_fields(), case["eos"])
    rng = np.random.default_rng(args.seed)
    thetas = [[0.31 * (1 + 0.1 * rng.uniform(-1, 1)), 0.09 * (1 + 0.1 * rng.uniform(-1, 1))]
              for _ in range(args.comp_members)]
    t0 = time.perf_counter()
    res = fm.evaluate_batch(thetas)
    bwall = time.perf_counter() - t0
    bconv = sum(str(r.status).split(".")[-1] == "Converged" for r in res)
    print("bench: comp batch of %d: %.2f s, %d converged" % (len(thetas), bwall, bconv), file=sys.stderr, flush=True)
    batch = {"members": len(thetas), "wall_s": bwall, "converged": int(bconv), "solves_per_s": bconv / bwall,
             "outer_iters_median": float(np.median([r.simple_iters for r in res]))}

    def ref_comp(path, s_packed, maxit, Uin, kIn, omIn):
        L = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref.so"))
        info = (C.c_int * 5)()
        t0 = time.perf_counter()
        rc = L.ref_comp_solve(path.encode(), D([Uin, 300.0, 101325.0, kIn, omIn]),
                              D([1.4, 287.0, 0.72, 0.9, 1.716e-5, 273.15, 110.4]),
                              D([0.85, 0.5, 0.075, 5.0 / 9.0, 1.0, 0.856, 0.0828, 0.44, 0.09, 0.31, 0.41]),
                              D(s_packed), C.c_double(0.7), D([Uin, 0, 0, 101325.0, 300.0, kIn, omIn]),
                              *([None] * 10), None, 0, info)
        return rc, time.perf_counter() - t0, list(info)

    def packed(s, maxit):
        return [maxit, s.convergence_tol, s.divergence_limit, s.alpha_u, s.alpha_p, s.alpha_k, s.alpha_omega,
                s.inner_iterations, s.inner_tolerance, s.turb_start_iter, s.turb_update_interval, s.nut_floor_iters,
                0, 0.0, 0.0, 0.0, s.k_min, s.k_max, s.omega_min, 0, s.von_karman, 0, 0, 2, 2]

    cpu = None
    have_ref = os.path.exists(os.path.join(ROOT, "oracle", "_ref", "libref.so"))
    if not args.no_cpu and have_ref:
        path = "/tmp/bench_comp40.msh"
        case["mesh"].save_binary(path)
        kIn = 1.5 * (case["Uin"] * 0.05) ** 2
        omIn = kIn / (case["mu_in"] / case["rho_in"] * 100.0)
        rc, dt, info = ref_comp(path, packed(case["settings"], 3000), 3000, case["Uin"], kIn, omIn)
        cpu = {"value": (1.0 / dt) if info[1] else 0.0, "unit": "solves/s", "cores": 1, "kind": "reference",
               "sample": "one full solve of the 40 x 30 case by the reference solver: %d iterations, converged %d, "
                         "%.1f s" % (info[0], info[1], dt), "lscpu": lscpu_info()}
    # the finer channel: outer-iteration time, device against the reference
    if args.comp_fine_iters <= 0:
        fine_out = None
        return _comp_line(args, lat, iters, conv, summ, batch, fine_out, cpu)
    nx = int(round(math.sqrt(args.comp_cells * 4.0 / 3.0)))
    ny = max(2, int(round(args.comp_cells / nx)))
    fine = cd.make_validation_case("ma_0_1_fine", 0.1, nx=nx, ny=ny, max_iterations=args.comp_fine_iters,
                                   convergence_tol=1e-12)
    t0 = time.perf_counter()
    rf = fine["forward"].evaluate(theta0)
    fwall = time.perf_counter() - t0
    fine_out = {"cells": nx * ny, "status": str(rf.status).split(".")[-1], "outer_iterations": int(rf.simple_iters),
                "wall_s": fwall,
                "ms_per_outer_iteration": 1e3 * fwall / max(rf.simple_iters, 1), "reference_ms_per_outer_iteration": None}
    if not args.no_cpu and have_ref:
        path = "/tmp/bench_comp_fine.msh"
        fine["mesh"].save_binary(path)
        kIn = 1.5 * (fine["Uin"] * 0.05) ** 2
        omIn = kIn / (fine["mu_in"] / fine["rho_in"] * 100.0)
        sample = max(5, min(args.comp_fine_iters, 60))
        rc, dt, info = ref_comp(path, packed(fine["settings"], sample), sample, fine["Uin"], kIn, omIn)
        fine_out["reference_ms_per_outer_iteration"] = 1e3 * dt / max(info[0], 1)
        fine_out["reference_sample"] = "%d outer iterations, one core, %.1f s" % (info[0], dt)
    print("bench: comp fine %s" % fine_out, file=sys.stderr, flush=True)
    return _comp_line(args, lat, iters, conv, summ, batch, fine_out, cpu)


def _comp_line(args, lat, iters, conv, summ, batch, fine_out, cpu):
    t = float(np.median(lat))
    print(json.dumps({
        "metric": "compressible converged solves/s", "value": conv / sum(lat), "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * t, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "compressible Ma 0.1 channel 40x30 (the reference's regression case), pressure-based "
                               "SIMPLE + SST + energy to its convergence_tol 1e-3, one solve at a time",
                   "cells": 1200},
        "p50_latency_s": t, "outer_iterations": int(iters), "converged": conv, "summary": summ, "batch": batch,
        "fine_channel": fine_out, "roofline": None, "cpu_baseline": cpu}), flush=True)


def bench_walldist(args):
    """Mesh::computeWallDistance (core/src/Mesh.cpp:345-373) at the 2M-cell
    size: a 2500 x 800 wall-bounded channel (1.0 x 0.3, wall-clustered), every
    cell against the 5000 wall-face centres, as one device kernel
    (rsg_wall_distance through Mesh.compute_wall_distance(on_device=True)).
    The reference's own loop, compiled from its sources, is timed on a
    2500 x 80 sample and checked bit for bit against the device there."""
    import ctypes as C
    import rans_sst_py as rs
    nx, ny = 2500, 800
    m = rs.Mesh.make_channel_2d(nx, ny, 1.0, 0.3, 1e4 / 0.3, 1.0)
    nw = 2 * nx
    pairs = float(m.n_cells()) * nw
    times = []
    for k in range(args.warmup + args.steps):
        t = m.compute_wall_distance_timed()
        print("bench: wall distance kernel %.3f ms" % (1e3 * t), file=sys.stderr, flush=True)
        if k >= args.warmup:
            times.append(t)
    t = float(np.median(times))
    flops = 8.0 * pairs                                   # 3 sub, 3 mul, 2 add per pair
    cpu, same = None, None
    lib_ref = os.path.join(ROOT, "oracle", "_ref", "libref.so")
    if not args.no_cpu and os.path.exists(lib_ref):
        L = C.CDLL(lib_ref)
        ms = rs.Mesh.make_channel_2d(nx, 80, 1.0, 0.3, 1e4 / 0.3, 1.0)
        path = "/tmp/bench_walldist.msh"
        ms.save_binary(path)
        sizes = (C.c_int * 4)()
        t0 = time.perf_counter(); L.ref_mesh_sizes(path.encode(), sizes); t_load = time.perf_counter() - t0
        wd = np.zeros(ms.n_cells())
        t0 = time.perf_counter()
        L.ref_wall_distance(path.encode(), wd.ctypes.data_as(C.POINTER(C.c_double)))
        t_ref = time.perf_counter() - t0 - t_load
        ms.compute_wall_distance(on_device=True)
        same = bool(np.array_equal(np.asarray(ms.wall_distance()), wd))
        sp = float(ms.n_cells()) * nw
        cpu = {"value": sp / t_ref, "unit": "cell-wall pairs/s", "cores": 1, "kind": "reference",
               "sample": "Mesh::computeWallDistance of the reference on the 2500 x 80 channel (%d cells x %d wall "
                         "faces): %.2f s (mesh load, %.2f s, subtracted)" % (ms.n_cells(), nw, t_ref, t_load)}
    print(json.dumps({
        "metric": "wall-distance cell-wall pairs/s", "value": pairs / t, "unit": "cell-wall pairs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * t, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "wall distance of a 2M-cell mesh: channel 2500 x 800 (1.0 x 0.3), 5000 wall faces",
                   "cells": m.n_cells(), "wall_faces": nw},
        "bit_equal_to_reference_on_sample": same,
        "roofline": {"bound": "fp64-valu", "achieved": flops / t / 1e12, "peak": 78.6, "unit": "TFLOP/s",
                     "frac": flops / t / 1e12 / 78.6, "traffic": None, "kernel": "k_wall_distance",
                     "flops_per_launch": flops},
        "cpu_baseline": cpu}), flush=True)


C4P_GRID = dict(nx_up=415, nx_dn=1245, ny_up=553, ny_low=111)   # 2,072,510 cells


def bump_mesh(rs, nx, ny, length=8.0, height=2.0, bump_h=0.3, bump_w=1.0, re=1667.0):
    """Streamwise-periodic channel (walls y = y_b(x) and y = height) with a
    Gaussian bump of height bump_h and width bump_w at mid-length on the lower
    wall, nx x ny quads stored unstructured (the curved periodic factory,
    tanh wall clustering for y+ = 1 at Re)."""
    x = np.linspace(0.0, length, nx + 1)
    g = np.exp(-((x - 0.5 * length) / bump_w) ** 2)
    yb = bump_h * (g - g[0])                    # exactly periodic: zero at both ends
    yb[-1] = yb[0]
    return rs.Mesh.make_curved_channel_periodic_2d(x, yb, height, ny, re, 1.0)


def bench_bump(args):
    """Team solves of single large meshes: the periodic bump channel at 40k,
    160k and 2M cells (nu = 6e-4, unit body force: Re_tau ~ 1670, bump-height
    Reynolds number ~ 1e4 at the bulk velocity), SIMPLE + SST with the AMG-PCG
    pressure solve, one member on a team of workgroups (rsg_set_team: the
    engine's size unless --team-blocks names others; cooperative launches of
    20 outer iterations).  Each size is solved cold for K1 and for K1 + K outer
    iterations; the difference gives the cost of one outer iteration without
    the cold start (mesh upload, initialisation, AMG set-up on the device).  At
    40k cells the one-workgroup member engine is timed the same way."""
    import rans_sst_py as rs
    import ensembles
    nu = 6e-4
    sizes = [(200, 200, 40), (400, 400, 40), (2000, 1000, 20)]
    if args.bump_sizes:
        keep = {int(s) for s in args.bump_sizes.split(",")}
        sizes = [z for z in sizes if z[0] * z[1] in keep]
    ps = rs.InferenceParameterSet.from_indices("c2_five", ensembles.THETA_INDICES)
    obs = rs.ObservationOperator()
    rs.engine_set_launch(20, 0)
    rs.engine_set_time_slice(0)

    rs.engine_set_graphs(bool(args.graphs))

    def run(mesh, iters, mode, blocks=0):
        rs.engine_set_team(20000 if mode == "team" else 0, blocks, 20)
        rs.engine_set_grid(20000 if mode == "grid" else 0, 20)
        if mode == "grid":
            rs.engine_set_grid_levels(blocks)
        s = rs.SolverSettings()
        s.max_iterations = iters
        s.convergence_tol = 1e-12
        s.body_force = rs.Vec3(1.0, 0.0, 0.0)
        s.verbose = False
        bcs = rs.FlowBoundaryConditions.channel_defaults(mesh, 1.0, 1e-4, 1.0)
        fm = rs.ForwardModel(mesh, ps, obs, bcs, nu, s)
        fm.clear_cache()
        st0 = rs.engine_stats()
        t0 = time.perf_counter()
        r = fm.evaluate(list(ensembles.THETA_DEFAULTS))
        dt = time.perf_counter() - t0
        st1 = rs.engine_stats()
        inner = fm.last_inner_iterations()
        return dict(wall=dt, kernel=(st1[3] - st0[3]) * 1e-3, iters=int(r.simple_iters),
                    status=str(r.status).split(".")[-1], inner=inner)

    rows = []
    for nx, ny, K in sizes:
        t0 = time.perf_counter()
        mesh = bump_mesh(rs, nx, ny)
        mesh.compute_wall_distance(True)
        n, nif = mesh.n_cells(), mesh.n_internal_faces()
        nb = mesh.n_faces() - nif
        print("bench: bump %d x %d = %d cells, mesh %.1f s" % (nx, ny, n, time.perf_counter() - t0),
              file=sys.stderr, flush=True)
        paths = args.bump_paths.split(",")
        legs = []
        if "grid" in paths:
            legs += [("grid" if lv == 5 else "grid_L%d" % lv, "grid", lv) for lv in map(int, args.grid_levels.split(","))]
        if "team" in paths:
            legs += [("team" if b == 0 else "team_%d" % b, "team", b) for b in map(int, args.team_blocks.split(","))]
        if n <= 40000 and not args.no_member:
            legs.append(("member", "member", 0))
        row = {"cells": n, "internal_faces": nif, "nx": nx, "ny": ny}
        for name, mode, nb_team in legs:
            k1 = 20
            a = run(mesh, k1, mode, nb_team)
            print("bench: bump %d %s %d iterations: %.2f s wall, %.2f s kernel (%s)"
                  % (n, name, a["iters"], a["wall"], a["kernel"], a["status"]), file=sys.stderr, flush=True)
            b = run(mesh, k1 + K, mode, nb_team)
            print("bench: bump %d %s %d iterations: %.2f s wall, %.2f s kernel (%s)"
                  % (n, name, b["iters"], b["wall"], b["kernel"], b["status"]), file=sys.stderr, flush=True)
            di = b["iters"] - a["iters"]
            leg = {"outer_iters": [a["iters"], b["iters"]], "status": b["status"], "path": name}
            if di > 0:
                dk = b["kernel"] - a["kernel"]
                din = [x - y for x, y in zip(b["inner"], a["inner"])]
                byt = algorithmic_bytes(n, nif, nb, di, din)
                leg.update(ms_per_outer_iter_wall=1e3 * (b["wall"] - a["wall"]) / di,
                           ms_per_outer_iter_kernel=1e3 * dk / di,
                           inner_per_outer=[x / di for x in din],
                           algorithmic_GBps=byt / dk * 1e-9 if dk > 0 else None,
                           setup_s=a["wall"] - k1 * (b["wall"] - a["wall"]) / di)
            row[name] = leg
        rows.append(row)
    if args.bump_members > 1:
        # several members of the smallest size at once (kernel-per-phase
        # members run concurrently on their own streams)
        nx, ny, _ = sizes[0]
        mesh = bump_mesh(rs, nx, ny)
        mesh.compute_wall_distance(True)
        rs.engine_set_team(0, 0, 20)
        rs.engine_set_grid(20000, 20)
        rs.engine_set_grid_levels(5)
        K = 100
        fms, th = [], []
        rng = np.random.default_rng(3)
        for i in range(args.bump_members):
            s = rs.SolverSettings()
            s.max_iterations = K; s.convergence_tol = 1e-12; s.verbose = False
            s.body_force = rs.Vec3(1.0, 0.0, 0.0)
            bcs = rs.FlowBoundaryConditions.channel_defaults(mesh, 1.0, 1e-4, 1.0)
            fms.append(rs.ForwardModel(mesh, ps, obs, bcs, nu, s))
            th.append([v * rng.uniform(0.8, 1.2) for v in ensembles.THETA_DEFAULTS])
        t0 = time.perf_counter()
        rs.evaluate_many(fms[:1], th[:1])
        one = time.perf_counter() - t0
        t0 = time.perf_counter()
        rs.evaluate_many(fms, th)
        many = time.perf_counter() - t0
        print("bench: bump %d members x %d iterations: %.2f s (one member alone %.2f s)"
              % (len(fms), K, many, one), file=sys.stderr, flush=True)
        rows[0]["concurrent_members"] = {"members": len(fms), "outer_iters": K, "wall_s": many,
                                         "one_member_wall_s": one,
                                         "ms_per_member_outer_iter": 1e3 * many / (len(fms) * K),
                                         "ms_per_outer_iter_alone": 1e3 * one / K}
    big = rows[-1]
    cands = [v for k, v in big.items() if (k.startswith("team") or k.startswith("grid")) and isinstance(v, dict)
             and v.get("ms_per_outer_iter_kernel")]
    best = min(cands, key=lambda v: v["ms_per_outer_iter_kernel"]) if cands else {}
    t = best.get("ms_per_outer_iter_kernel")
    gbs = best.get("algorithmic_GBps")
    print(json.dumps({
        "metric": "outer SIMPLE+SST iterations/s of one large-mesh member (fastest large-member path)",
        "value": (1e3 / t) if t else None, "unit": "outer_iters/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": t, "higher_is_better": True, "scaling": "strong", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "periodic Gaussian-bump channel (bump 0.3 x 1, length 8, height 2), "
                               "%d cells, nu 6e-4, AMG-PCG pressure, %s" % (big["cells"], best.get("path", "-"))},
        "sizes": rows,
        "roofline": ({"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "frac": gbs / HBM_PEAK_GBS, "traffic": None} if gbs else None),
        "cpu_baseline": None}), flush=True)


def bench_c4(args):
    """The large-member ensemble: the periodic Gaussian-bump channel at 2M
    cells (2000 x 1000, bump 0.3 x 1, length 8, height 2), nu = 0.3 / 1e4
    (bump-height Reynolds number 1e4 at unit bulk velocity; body force 2e-3
    gives Re_tau ~ 1500 and U_b ~ 1; tanh wall clustering for y+ = 1), 65
    members: the default SST coefficients and 64 seeded +-20 % perturbations
    of betaStar, a1, sigma_k1, sigma_w1, beta1.  Every member runs on the
    kernel-per-phase path with the 5-level p' hierarchy (coarsest inverted on
    the matrix cores), once with FP32 inner solves under FP64 refinement and
    once in FP64, each for --c4-iters outer iterations from the same cold
    start; each member's relative L2 of U, p, k, omega, nu_t between the two
    is reported.  A 2M-cell member needs ~10^4 outer iterations to reach
    1e-8, so the comparison is made at a fixed outer iteration."""
    # 65 resident 2M-cell members: about 1.2 GB of engine arena each
    os.environ.setdefault("RSG_HBM_BUDGET_GB", "120")
    import rans_sst_py as rs
    import ensembles
    n_mem, K = args.c4_members, args.c4_iters
    mesh = bump_mesh(rs, 2000, 1000, re=1500.0)
    t0 = time.perf_counter()
    mesh.compute_wall_distance(True)
    print("bench: c4 mesh %d cells, wall distance %.1f s" % (mesh.n_cells(), time.perf_counter() - t0),
          file=sys.stderr, flush=True)
    specs = ensembles.channel_ensemble(n_mem, args.seed)
    thetas = [list(ensembles.THETA_DEFAULTS)] + [sp["theta"] for sp in specs[1:]]
    ps = rs.InferenceParameterSet.from_indices("c2_five", ensembles.THETA_INDICES)
    obs = rs.ObservationOperator()
    rs.engine_set_time_slice(0)
    rs.engine_set_grid(100000, 20)
    rs.engine_set_grid_levels(5)
    rs.engine_set_graphs(bool(args.graphs))

    def run(precision):
        s = rs.SolverSettings()
        s.max_iterations = K
        s.convergence_tol = 1e-12
        s.body_force = rs.Vec3(2e-3, 0.0, 0.0)
        s.verbose = False
        bcs = rs.FlowBoundaryConditions.channel_defaults(mesh, 1.0, 1e-4, 1.0)
        fms = []
        for _ in thetas:
            fm = rs.ForwardModel(mesh, ps, obs, bcs, 0.3 / 1e4, s)
            fm.set_precision(precision)
            fm.clear_cache()
            fms.append(fm)
        st0 = rs.engine_stats()
        t = time.perf_counter()
        # members in groups of 13 (each group's members share the GPU in turns;
        # a progress line per group)
        res = []
        for g in range(0, len(fms), 13):
            res += rs.evaluate_many(fms[g:g + 13], thetas[g:g + 13])
            print("bench: c4 %s: %d/%d members done at %.1f s" % (precision, len(res), len(fms),
                                                                  time.perf_counter() - t), file=sys.stderr, flush=True)
        wall = time.perf_counter() - t
        st1 = rs.engine_stats()
        print("bench: c4 %s: %d members x %d outer iterations in %.1f s" % (precision, len(fms), K, wall),
              file=sys.stderr, flush=True)
        fields = [fm.last_fields() for fm in fms]
        inner = [fm.last_inner_iterations() for fm in fms]
        return dict(wall=wall, kernel_s=(st1[3] - st0[3]) * 1e-3, res=res, fields=fields, inner=inner)

    mixed = run("mixed")
    fp64 = run("fp64")
    per = []
    for i in range(len(thetas)):
        e = field_errors(mixed["fields"][i], fp64["fields"][i])
        a, b = np.asarray(mixed["fields"][i]["p"]), np.asarray(fp64["fields"][i]["p"])
        e["p_rel_l2"] = float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
        e["status"] = [str(mixed["res"][i].status).split(".")[-1], str(fp64["res"][i].status).split(".")[-1]]
        e["outer_iters"] = [int(mixed["res"][i].simple_iters), int(fp64["res"][i].simple_iters)]
        e["inner_iters_mixed"] = list(mixed["inner"][i])
        e["inner_iters_fp64"] = list(fp64["inner"][i])
        per.append(e)
    worst = {k: max(p[k] for p in per) for k in ("U", "p", "k", "omega", "nuT", "p_rel_l2")}
    its = len(thetas) * K
    print(json.dumps({
        "metric": "member outer SIMPLE+SST iterations/s of the 65-member 2M-cell bump ensemble (mixed precision), "
                  "and each member's rel-L2 vs its FP64 GPU solve",
        "value": its / mixed["wall"], "unit": "member_outer_iters/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": 1e3 * mixed["wall"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32-inner/f64-defect", "data": "synthetic",
        "config": {"workload": "C4: periodic Gaussian-bump channel 2000x1000 = 2M cells, nu 3e-5, body force 2e-3, "
                               "%d members (defaults + seeded +-20%% on 5 SST coefficients), kernel-per-phase path, "
                               "5-level p' AMG with the coarsest inverted on the matrix cores, %d outer iterations "
                               "per member, mixed vs FP64" % (len(thetas), K),
                   "cells": mesh.n_cells(), "members": len(thetas), "outer_iters": K},
        "mixed": {"wall_s": mixed["wall"], "kernel_s": mixed["kernel_s"],
                  "ms_per_member_outer_iter": 1e3 * mixed["wall"] / its},
        "fp64": {"wall_s": fp64["wall"], "kernel_s": fp64["kernel_s"],
                 "ms_per_member_outer_iter": 1e3 * fp64["wall"] / its,
                 "member_outer_iters_per_s": its / fp64["wall"]},
        "max_rel_l2_vs_fp64": worst,
        "per_member": per,
        "roofline": None, "cpu_baseline": None}), flush=True)


def bench_c1(args, big=False):
    """The large single system: solves/s of AMG-PCG to 1e-10 (host hierarchy
    setup + device solve + host<->device copies, wall clock), FP64 and mixed.
    big: the same generator at the 2M-cell size of the 65-member bump
    ensemble's pressure-correction system."""
    import ctypes as C
    import c1_system
    lib = C.CDLL(os.path.join(PKG, "librsg.so"))
    lib.rsg_last_error.restype = C.c_char_p
    ctx = C.c_void_p()
    if lib.rsg_context_create(0, C.c_uint64(8 << 30), C.byref(ctx)) != 0:
        raise RuntimeError("rsg_context_create failed: no usable HIP device")
    lib.rsg_set_graphs(ctx, int(args.graphs))
    s = c1_system.make_system(**(C4P_GRID if big else {}))
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    M = c1_system.csr_matrix(s)

    def solve(prec):
        x = np.zeros(s["n"]); st = np.zeros(10)
        t0 = time.perf_counter()
        rc = lib.rsg_system_solve(ctx, s["n"], s["nif"], ip(s["own"]), ip(s["nbr"]), dp(s["diag"]), dp(s["upper"]),
                                  dp(s["lower"]), dp(s["b"]), dp(x), 0, prec, 20000, C.c_double(1e-10), dp(st))
        dt = time.perf_counter() - t0
        if rc != 0:
            raise RuntimeError("rsg_system_solve: %s" % lib.rsg_last_error(ctx))
        return x, st, dt

    out = {}
    for name, prec in (("fp64", 0), ("mixed", 1)):
        for _ in range(args.warmup):
            solve(prec)
        walls, st = [], None
        for _ in range(args.steps):
            x, st, dt = solve(prec)
            walls.append(dt)
            print("bench: c1 %s %.3f s (%d iterations)" % (name, dt, st[0]), file=sys.stderr, flush=True)
        out[name] = dict(x=x, st=st, walls=walls,
                         true_res=float(np.linalg.norm(s["b"] - M @ x) / np.linalg.norm(s["b"])))
    lib.rsg_context_destroy(ctx)
    f, m = out["fp64"], out["mixed"]
    rel = float(np.linalg.norm(m["x"] - f["x"]) / np.linalg.norm(f["x"]))
    cpu = None
    lib_ref = os.path.join(ROOT, "oracle", "_ref", "libref.so")
    kind = "reference"
    if not os.path.exists(lib_ref):
        lib_ref, kind = os.path.join(ROOT, "oracle", "_build", "liboracle.so"), "port"
    if not args.no_cpu and os.path.exists(lib_ref):
        L = C.CDLL(lib_ref)
        fn = L.ref_ldu_solve if kind == "reference" else L.om_ldu_solve
        x = np.zeros(s["n"]); st = np.zeros(4)
        cap = 300 if big else 20000
        t0 = time.perf_counter()
        fn(1, s["n"], s["nif"], ip(s["own"]), ip(s["nbr"]), dp(s["diag"]), dp(s["upper"]), dp(s["lower"]),
           dp(s["b"]), dp(x), cap, C.c_double(1e-10), dp(st))
        dt = time.perf_counter() - t0
        if big:
            cpu = {"value": st[0] / dt, "unit": "PCG iterations/s", "cores": 1, "kind": kind,
                   "sample": "the reference PCGSolver (Jacobi) capped at %d iterations on this system: %.1f s "
                             "(%.1f ms per iteration; a full solve to 1e-10 needs thousands)"
                             % (st[0], dt, 1e3 * dt / max(st[0], 1))}
        else:
            cpu = {"value": (1.0 / dt) if st[3] == 1 else 0.0, "unit": "solves/s", "cores": 1, "kind": kind,
                   "sample": "one full solve: the reference PCGSolver (Jacobi) to 1e-10, %d iterations, %.1f s; "
                             "its AMGSolver does not reach 1e-10 on this system (forward Gauss-Seidel V-cycle)"
                             % (st[0], dt)}
    tf = float(np.median(f["walls"]))
    print(json.dumps({
        "metric": METRIC, "value": 1.0 / tf, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * tf, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "%s pressure system: BFS expansion 1.2, %d cells (%d tri + %d quad), "
                               "log-normal nu_t sigma 1, AMG-PCG to 1e-10 relative residual"
                               % ("2M-cell (bump-ensemble size)" if big else "C1 stand-alone", s["n"], s["n_tri"],
                                  s["n_quad"]), "cells": s["n"], "tol": 1e-10},
        "fp64": {"iterations": int(f["st"][0]), "levels": int(f["st"][4]), "setup_s": f["st"][6],
                 "device_solve_s": f["st"][7], "wall_s": tf, "true_rel_res": f["true_res"],
                 "coarsest_rows_inverted": int(f["st"][8]), "coarsest_inverse_s": f["st"][9]},
        "mixed": {"iterations": int(m["st"][0]), "refinements": int(m["st"][5]), "setup_s": m["st"][6],
                  "device_solve_s": m["st"][7], "wall_s": float(np.median(m["walls"])),
                  "coarsest_rows_inverted": int(m["st"][8]), "coarsest_inverse_s": m["st"][9],
                  "true_rel_res": m["true_res"], "rel_l2_vs_fp64": rel},
        "roofline": None, "cpu_baseline": cpu}), flush=True)


def bench_c2(args):
    """The headline: the C2 ensemble, one timed step = one full ensemble."""
    t_begin = time.perf_counter()
    remaining = lambda: args.budget_s - (time.perf_counter() - t_begin)
    import rans_sst_py as rs
    import ensembles
    specs = ensembles.channel_ensemble(args.members, args.seed)
    configure_engine(rs, args)
    tuned = None
    if args.autotune:
        # workgroup size and time slice measured on this device and ensemble
        # (outside the timed region); the engine keeps the pick
        import autotune
        t0 = time.perf_counter()
        choice, _ = autotune.tune_members(rs, specs, args.precision, cache_path=args.tune_cache or None)
        args.block, args.slice_ms = choice["block"], choice["slice_ms"]
        tuned = dict(choice, seconds=time.perf_counter() - t0)
        print("bench: autotune %.1f s: block %d, slice %g ms" % (tuned["seconds"], args.block, args.slice_ms),
              file=sys.stderr, flush=True)
    # engine start-up, code-object load and one short pass of every member
    # (100 outer iterations) outside the timed region
    warm = 0
    if args.warmup > 0:
        _, _, wall = run_batch(rs, specs, args.tol, None, 100, args.precision)
        warm = 1
        print("bench: warmup pass (100 outer iterations per member) %.2f s" % wall, file=sys.stderr, flush=True)
    else:
        _, prime = ensembles.build_models(rs, specs[:1], tol=args.tol, max_iterations=2)
        rs.evaluate_many(prime, [specs[0]["theta"]])
    stats0 = rs.engine_stats()
    walls, lat, conv, outer = [], [], 0, []
    cache, res = None, None
    # each step is a whole ensemble (minutes); further steps run only while
    # they fit the step budget
    while len(walls) < max(args.steps, 1):
        cache, res, wall = run_batch(rs, specs, args.tol, cache, args.max_iterations, args.precision)
        walls.append(wall)
        print("bench: step %d %.2f s" % (len(walls), wall), file=sys.stderr, flush=True)
        for r in res:
            lat.append(r.latency)
            conv += str(r.status).split(".")[-1] == "Converged"
            outer.append(r.simple_iters)
        if sum(walls) + wall > args.step_budget_s:
            break
    stats1 = rs.engine_stats()
    steps = len(walls)
    total = sum(walls)
    launches = stats1[0] - stats0[0]
    kernel_ms = stats1[3] - stats0[3]
    bytes_alg = ensemble_alg_bytes(cache, res, args.precision) * steps
    achieved = bytes_alg / (kernel_ms * 1e-3) / 1e9 if kernel_ms > 0 else 0.0
    alg_per_launch = bytes_alg / max(launches, 1)
    check = None
    if remaining() > 150:
        check = fp64_check(rs, specs, cache, res, args.tol, args.max_iterations)
        print("bench: FP64 check done", file=sys.stderr, flush=True)
    else:
        check = {"skipped": "not enough of the time budget left"}
    cpu = None
    if not args.no_cpu and remaining() > 60:
        cpu = cpu_baseline(specs, [dict(outer=r.simple_iters, cells=m.n_cells(),
                                        converged=str(r.status).split(".")[-1] == "Converged")
                                   for r, m in zip(res, cache[0])])
        print("bench: cpu baseline done", file=sys.stderr, flush=True)
    traffic, traffic_info = None, None
    if not args.no_traffic:
        if remaining() > 150:
            try:
                traffic_info = measure_traffic(args)
                traffic = traffic_info["hbm_over_algorithmic"] * alg_per_launch
            except Exception as e:   # counters unavailable: report null with the reason
                traffic_info = {"error": str(e)[:300]}
            print("bench: traffic passes done", file=sys.stderr, flush=True)
        else:
            traffic_info = {"skipped": "not enough of the time budget left"}
    out = {
        "metric": METRIC,
        "value": conv / total,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "ms_per_step": 1e3 * total / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-inner/f64-defect" if args.precision == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": "C2 batched ensemble: %d periodic channels 4x(96-384), Re_tau logU[180,2000], "
                               "betaStar/a1/sigma_k1/sigma_w1/beta1 +-20%%, SIMPLE+SST to tol %g, %s"
                               % (args.members, args.tol,
                                  "FP32 inner solves with FP64 refinement" if args.precision == "mixed" else "FP64"),
                   "members": args.members, "tol": args.tol, "seed": args.seed, "precision": args.precision,
                   "max_iterations": args.max_iterations, "iters_per_launch": args.iters_per_launch,
                   "block": args.block, "slice_ms": args.slice_ms, "schedule": args.schedule},
        "steps_requested": args.steps,
        "warmup_requested": args.warmup,
        "step_definition": "one step = the whole ensemble solved cold (every member submitted at once, timed to "
                           "the last retirement); steps beyond the first run only while they fit %.0f s"
                           % args.step_budget_s,
        "p50_latency_s": float(np.percentile(lat, 50)),
        "p99_latency_s": float(np.percentile(lat, 99)),
        "converged": int(conv),
        "solves": len(lat),
        "outer_iters_median": float(np.median(outer)),
        "max_rel_l2_vs_fp64": (check or {}).get("max_rel_l2"),
        "fp64_check": check,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_simple_chunk", "launches": launches,
                     "avg_launch_ms": kernel_ms / max(launches, 1),
                     "alg_bytes_per_launch": alg_per_launch, "alg_bytes_total": bytes_alg,
                     "traffic_measurement": traffic_info},
        "cpu_baseline": cpu,
        "autotune": tuned,
        "bench_wall_s": time.perf_counter() - t_begin,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--members", type=int, default=512)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64",
                    help="member inner solves: FP64, or FP32 with FP64 defect and refinement")
    ap.add_argument("--iters-per-launch", type=int, default=500)
    ap.add_argument("--block", type=int, default=0, help="0: engine default (512)")
    ap.add_argument("--slice-ms", type=float, default=200.0,
                    help="time slice per launch in ms (0: fixed --iters-per-launch budget)")
    ap.add_argument("--schedule", choices=("round_robin", "largest_first"), default="largest_first",
                    help="which members a launch runs when more are active than the device holds")
    ap.add_argument("--budget-s", type=float, default=480.0,
                    help="headline: wall-clock budget of the whole run; optional legs are skipped when it runs out")
    ap.add_argument("--step-budget-s", type=float, default=300.0,
                    help="headline: timed steps after the first run only while the total stays within this")
    ap.add_argument("--autotune", type=int, default=1,
                    help="headline: pick the workgroup size and time slice by measuring candidates on the ensemble "
                         "(1, the default) or keep --block / --slice-ms (0)")
    ap.add_argument("--tune-cache", default="", help="JSON file caching autotuner picks (empty: none)")
    ap.add_argument("--graphs", type=int, default=1,
                    help="c1/c4p/bump/c4: replay AMG V-cycles from hipGraphs (1, the default) or launch directly (0)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 counter passes that measure the kernel's HBM traffic")
    ap.add_argument("--traffic-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--c4-members", type=int, default=65, help="c4 workload: ensemble members")
    ap.add_argument("--c4-iters", type=int, default=40, help="c4 workload: outer iterations per member and precision")
    ap.add_argument("--comp-members", type=int, default=256, help="comp workload: theta samples in the batch")
    ap.add_argument("--comp-cells", type=int, default=76800, help="comp workload: cells of the finer channel")
    ap.add_argument("--comp-fine-iters", type=int, default=100,
                    help="comp workload: outer iterations timed on the finer channel")
    ap.add_argument("--workload", choices=("c2", "c1", "c0", "walldist", "c4p", "bump", "c4", "comp"), default="c2",
                    help="c2: the 512-member ensemble (default); c1: the 261,600-cell pressure system; "
                         "c0: single-solve latency of the Re_tau 550 channel; walldist: the wall-distance "
                         "kernel on a 2M-cell mesh; c4p: the pressure system at 2M cells; bump: team solves "
                         "of one bump-channel member at 40k, 160k and 2M cells; c4: the 65-member 2M-cell bump "
                         "ensemble, mixed precision against FP64 per member")
    ap.add_argument("--bump-sizes", default="", help="comma-separated cell counts of the bump workload to run")
    ap.add_argument("--bump-members", type=int, default=1,
                    help="bump workload: also run this many members of the smallest size at once")
    ap.add_argument("--team-blocks", default="0",
                    help="bump workload: comma-separated team sizes in workgroups (0: the engine's choice)")
    ap.add_argument("--no-member", action="store_true", help="bump workload: skip the one-workgroup leg")
    ap.add_argument("--grid-levels", default="5",
                    help="bump workload: comma-separated p' AMG level caps of the kernel-per-phase path (5: the "
                         "default, coarsest inverted on the matrix cores; 0: the reference's 20-level cap)")
    ap.add_argument("--bump-paths", default="grid,team",
                    help="bump workload: comma-separated large-member paths to time (grid: kernel per phase; "
                         "team: cooperative workgroup team)")
    ap.add_argument("--max-iterations", type=int, default=200000,
                    help="outer-iteration cap per member (profiling runs shorten it)")
    args = ap.parse_args()
    if args.traffic_child:
        traffic_child(args)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU (--gpus 1): ensemble members would shard across GPUs with no "
              "exchange between them, and runs on several GPUs are not part of this benchmark", file=sys.stderr)
        sys.exit(2)
    {"c1": bench_c1, "c0": bench_c0, "walldist": bench_walldist, "bump": bench_bump,
     "c4p": lambda a: bench_c1(a, big=True), "c2": bench_c2, "c4": bench_c4, "comp": bench_comp}[args.workload](args)


if __name__ == "__main__":
    main()
