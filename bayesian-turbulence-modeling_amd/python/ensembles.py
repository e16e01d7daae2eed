"""Synthetic ensembles of channel forward solves (the calibration layer's
member-parallel call pattern, reference python/UQ/datasets/channel_calibration.py:160-243,
expressed as one device batch).

channel_ensemble(n, seed): n streamwise-periodic plane channels (half-height
1, 4 x ny quads, unit body force).  Each member draws Re_tau log-uniform in
[180, 2000]; ny maps linearly in ln(Re_tau) onto [96, 384] (multiple of 8);
the symmetric geometric stretching ratio is solved so the first cell centre
sits at y+ = 0.5 (first cell height 1/Re_tau); nu = 1/Re_tau.  The SST
coefficients betaStar, a1, sigma_k1, sigma_w1, beta1 are drawn uniformly
within +-20 % of their defaults.
"""
import math

import numpy as np

THETA_INDICES = [8, 9, 0, 1, 2]                 # betaStar, a1, sigma_k1, sigma_w1, beta1
THETA_DEFAULTS = [0.09, 0.31, 0.85, 0.5, 0.075]


def stretch_ratio(n_half, y1):
    """Geometric ratio r with first cell y1 for n_half cells over unit height."""
    if y1 >= 1.0 / n_half:
        return 1.0
    lo, hi = 1.0 + 1e-12, 2.0
    for _ in range(200):
        r = 0.5 * (lo + hi)
        if (r - 1.0) / (r ** n_half - 1.0) > y1:
            lo = r
        else:
            hi = r
    return 0.5 * (lo + hi)


def member_spec(re_tau):
    lo, hi = math.log(180.0), math.log(2000.0)
    frac = (math.log(re_tau) - lo) / (hi - lo)
    ny = int(round((96 + frac * (384 - 96)) / 8.0)) * 8
    ny = min(384, max(96, ny))
    return ny, stretch_ratio(ny // 2, 1.0 / re_tau)


def channel_ensemble(n, seed=0):
    """List of dicts: re_tau, ny, ratio, nu, theta (5 values)."""
    rng = np.random.default_rng(seed)
    re = np.exp(rng.uniform(math.log(180.0), math.log(2000.0), n))
    th = rng.uniform(0.8, 1.2, (n, 5)) * np.array(THETA_DEFAULTS)
    out = []
    for i in range(n):
        ny, r = member_spec(float(re[i]))
        out.append(dict(re_tau=float(re[i]), ny=ny, ratio=r, nu=1.0 / float(re[i]), theta=[float(v) for v in th[i]]))
    return out


def build_models(rs, specs, tol=1e-8, max_iterations=200000):
    """One (mesh, ForwardModel) per member; meshes must be kept alive by the caller."""
    ps = rs.InferenceParameterSet.from_indices("c2_five", THETA_INDICES)
    obs = rs.ObservationOperator()
    meshes, fms = [], []
    for sp in specs:
        y = rs.Mesh.geometric_channel_nodes(sp["ny"] // 2, sp["ratio"], 1.0)
        mesh = rs.Mesh.make_channel_periodic_2d(np.linspace(0.0, 1.0, 5), np.array(y))
        mesh.compute_wall_distance()
        s = rs.SolverSettings()
        s.max_iterations = max_iterations
        s.convergence_tol = tol
        s.body_force = rs.Vec3(1.0, 0.0, 0.0)
        s.verbose = False
        bcs = rs.FlowBoundaryConditions.channel_defaults(mesh, 1.0, 1e-4, 1.0)
        meshes.append(mesh)
        fms.append(rs.ForwardModel(mesh, ps, obs, bcs, sp["nu"], s))
    return meshes, fms
