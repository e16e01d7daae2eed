"""Served mixed workload: Poisson request streams from 16 calibration chains
against two or three meshes, every request solved to tol 1e-8 from a cold start.

Meshes (a request picks one uniformly):
  * channel column: the periodic Re_tau 550 channel reduced to one column of
    128 cells (geometric 1.06 stretching, nu = 1/550, unit body force);
  * Couette 8 x 192: makeChannel2D(8, 192, 1.5, 2, Re = Uwall/nu, y+ 0.3) with
    couetteDefaults(Uwall = 2), nu = 1/20000, alpha_p 0.5, initial state
    (Ub, 0, 0), kIn = 1.5 (0.05 Ub)^2, omIn = kIn / (100 nu), as the
    reference's Couette forward model builds it
    (python/UQ/datasets/couette_forward.py:49-75);
  * (--meshes all) the periodic Gaussian-bump channel, 200 x 200 = 40,000
    cells (bump 0.3 x 1, length 8, height 2, nu 6e-4, unit body force),
    solved on the kernel-per-phase path (several bump members at once) and
    capped at --bump-max-iterations outer iterations: it does not reach 1e-8
    within 40,000 outer iterations, so its requests retire unconverged at the
    cap.

Usage: python serve_c3.py [--rates 50,100] [--requests N] [--max-iterations M] [--slice-ms T]
                          [--meshes both|column|all] [--bump-max-iterations B]
                          [--schedule round_robin|largest_first]
Prints one JSON line per rate: throughput (converged solves/s over the span
from first arrival to last result), p50/p99 latency (arrival -> result).
"""
import argparse
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)

import serving  # noqa: E402

THETA_INDICES = [8, 9, 0, 1, 2]      # betaStar, a1, sigma_k1, sigma_w1, beta1


def bump_mesh(rs, nx=200, ny=200, length=8.0, height=2.0, bump_h=0.3, bump_w=1.0, re=1667.0):
    """Streamwise-periodic channel with a Gaussian bump on the lower wall (the
    curved periodic factory, tanh wall clustering for y+ = 1 at Re)."""
    x = np.linspace(0.0, length, nx + 1)
    g = np.exp(-((x - 0.5 * length) / bump_w) ** 2)
    yb = bump_h * (g - g[0])
    yb[-1] = yb[0]
    return rs.Mesh.make_curved_channel_periodic_2d(x, yb, height, ny, re, 1.0)


def build_models(rs, tol, max_iterations, meshes="both", bump_max_iterations=3000):
    ps = rs.InferenceParameterSet.from_indices("c3_five", THETA_INDICES)
    keep, models, names = [], [], []
    # channel column, 128 cells
    y = rs.Mesh.geometric_channel_nodes(64, 1.06, 1.0)
    col = rs.Mesh.make_channel_column_1d(np.array(y), 0.25)
    col.compute_wall_distance()
    s = rs.SolverSettings()
    s.max_iterations = max_iterations; s.convergence_tol = tol; s.verbose = False
    s.body_force = rs.Vec3(1.0, 0.0, 0.0)
    bcs = rs.FlowBoundaryConditions.channel_defaults(col, 1.0, 1e-4, 1.0)
    models.append(rs.ForwardModel(col, ps, rs.ObservationOperator(), bcs, 1.0 / 550.0, s))
    keep.append(col); names.append("channel_column_128")
    if meshes == "column":
        return keep, models, names
    # Couette 8 x 192
    nu, uw = 1.0 / 20000.0, 2.0
    ub = 0.5 * uw
    cou = rs.Mesh.make_channel_2d(8, 192, 1.5, 2.0, uw / nu, 0.3)
    cou.compute_wall_distance()
    kin = 1.5 * (ub * 0.05) ** 2
    omin = kin / (nu * 100.0)
    s = rs.SolverSettings()
    s.max_iterations = max_iterations; s.convergence_tol = tol; s.verbose = False
    s.alpha_u = 0.7; s.alpha_p = 0.5
    bcs = rs.FlowBoundaryConditions.couette_defaults(cou, uw, kin, omin)
    models.append(rs.ForwardModel(cou, ps, rs.ObservationOperator(), bcs, nu, s, rs.Vec3(ub, 0.0, 0.0),
                                  0.0, kin, omin))
    keep.append(cou); names.append("couette_8x192")
    if meshes == "all":
        bump = bump_mesh(rs)
        bump.compute_wall_distance(True)
        s = rs.SolverSettings()
        s.max_iterations = min(max_iterations, bump_max_iterations); s.convergence_tol = tol; s.verbose = False
        s.body_force = rs.Vec3(1.0, 0.0, 0.0)
        bcs = rs.FlowBoundaryConditions.channel_defaults(bump, 1.0, 1e-4, 1.0)
        models.append(rs.ForwardModel(bump, ps, rs.ObservationOperator(), bcs, 6e-4, s))
        keep.append(bump); names.append("bump_40k")
    return keep, models, names


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--rates", default="50,100,200,400")
    ap.add_argument("--requests", type=int, default=64)
    ap.add_argument("--chains", type=int, default=16)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--max-iterations", type=int, default=200000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--slice-ms", type=float, default=100.0,
                    help="engine time slice per launch (0: fixed iteration budget)")
    ap.add_argument("--meshes", choices=("both", "column", "all"), default="both",
                    help="channel column + Couette, the channel column alone (short solves: a load test), or "
                         "all three C3 meshes (with the 40k-cell bump channel)")
    ap.add_argument("--bump-max-iterations", type=int, default=3000,
                    help="outer-iteration cap of bump requests (--meshes all)")
    ap.add_argument("--schedule", choices=("round_robin", "largest_first"), default="round_robin")
    args = ap.parse_args()
    import rans_sst_py as rs
    rs.engine_set_time_slice(int(args.slice_ms * 1000))
    rs.engine_set_schedule(args.schedule)
    if args.meshes == "all":
        rs.engine_set_grid(10000, 20)        # bump members: kernel-per-phase, concurrently
    keep, models, names = build_models(rs, args.tol, args.max_iterations, args.meshes, args.bump_max_iterations)
    for rate in [float(r) for r in args.rates.split(",")]:
        trace = serving.arrival_trace(rate, args.requests, len(models), n_chains=args.chains, seed=args.seed)
        done = [0]

        def progress(i, rec):
            done[0] += 1
            if done[0] % 16 == 0 or done[0] == len(trace):
                print("serve: rate %g: %d/%d retired at %.1f s" % (rate, done[0], len(trace), rec["done"]),
                      file=sys.stderr, flush=True)

        out = serving.serve(models, trace, rs.ticket_states, on_retire=progress)
        per_mesh = {}
        for mi, nm in enumerate(names):
            lat = [r["latency"] for r in out["records"] if r["mesh"] == mi]
            its = [r["iterations"] for r in out["records"] if r["mesh"] == mi]
            conv = sum(r["status"] == "Converged" for r in out["records"] if r["mesh"] == mi)
            per_mesh[nm] = dict(requests=len(lat), converged=int(conv), p50_s=serving.percentile(lat, 50),
                                p99_s=serving.percentile(lat, 99),
                                iterations_median=float(np.median(its)) if its else None)
        print(json.dumps(dict(workload="served mixed (C3): %d chains, Poisson %g req/s, meshes %s, tol %g"
                              % (args.chains, rate, "+".join(names), args.tol),
                              rate=rate, requests=out["n"], converged=out["converged"],
                              throughput_solves_per_s=out["throughput"], p50_latency_s=out["p50"],
                              p99_latency_s=out["p99"], per_mesh=per_mesh)), flush=True)


if __name__ == "__main__":
    main()
