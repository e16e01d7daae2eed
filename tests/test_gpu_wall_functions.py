"""Compressible channels with and without wall functions on the device, with
the contract of the reference's tests/python/test_wall_functions.py:158-272:
a resolved 40 x 30 channel (first-cell y+ target 1) keeps its first-cell y+
below 5 on both walls; a coarse 24 x 14 channel (y+ target 30) with wall
functions does not diverge, has a finite, positive skin friction along the
bottom wall, a larger first-cell y+ than the resolved one, and a skin
friction at x = 5 within a factor 2 of the resolved one; the same coarse mesh
without wall functions does not diverge either.

The two diagnostics are those of the reference's python/wall_diagnostics.py
(:47-100), restated: y+ = beta*^(1/4) sqrt(k_P) delta / nu at the wall
face's owner cell, and Cf = nu |U_t| / delta / (U_ref^2 / 2) with U_t the
owner cell's velocity tangential to the face."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _channel(rs, nx, ny, y_plus_target, wall_functions, iters, Ma=0.1):
    eos = rs.IdealGasEOS()
    T_in, p_ref = 300.0, 101325.0
    rho_in, mu_in = eos.density(p_ref, T_in), eos.viscosity(T_in)
    Uin = Ma * eos.sound_speed(T_in)
    nu_in = mu_in / rho_in
    mesh = rs.Mesh.make_channel_2d(nx, ny, 10.0, 1.0, Re=rho_in * Uin / mu_in, yPlusTarget=y_plus_target)
    mesh.compute_wall_distance()
    k_in = 1.5 * (0.05 * Uin) ** 2
    om_in = k_in / (100.0 * nu_in)
    bcs = rs.CompressibleBoundaryConditions.channel_defaults(mesh, Uin, T_in, p_ref, k_in, om_in)
    obs = rs.ObservationOperator()
    obs.add_skin_friction(wall_patch="bottom", location=rs.Vec3(5.0, 0, 0), cf_obs=0.005, sigma=0.001, ref_vel=Uin)
    s = rs.SolverSettings()
    s.max_iterations = iters; s.convergence_tol = 1e-3; s.divergence_limit = 1e10
    s.alpha_u = 0.5; s.alpha_p = 0.2; s.alpha_t = 0.7; s.alpha_k = 0.4; s.alpha_omega = 0.4
    s.inner_iterations = 200; s.turb_start_iter = 50; s.turb_update_interval = 2; s.verbose = False
    s.use_wall_functions = wall_functions
    ps = rs.InferenceParameterSet.a1_betaStar()
    fm = rs.CompressibleForwardModel(mesh=mesh, param_set=ps, obs_op=obs, bcs=bcs, eos=eos, settings=s,
                                     u_init=rs.Vec3(Uin, 0, 0), p_init=p_ref, T_init=T_in, k_init=k_in,
                                     omega_init=om_in)
    print("compressible channel %d x %d, y+ target %g, wall functions %s" % (nx, ny, y_plus_target, wall_functions),
          flush=True)
    res = fm.evaluate(list(ps.pack(rs.SSTCoefficients())))
    return dict(mesh=mesh, fm=fm, res=res, fields=fm.last_fields() if fm.has_last_fields() else None, Uin=Uin,
                nu=nu_in)


def _y_plus(run, patch, beta_star=0.09):
    info = run["mesh"].wall_patch_data(patch)
    own = np.asarray(info["owner"], int)
    k_p = np.maximum(np.asarray(run["fields"]["k"], float)[own], 1e-30)
    delta = np.maximum(np.asarray(info["delta"], float), 1e-30)
    return beta_star ** 0.25 * np.sqrt(k_p) * delta / run["nu"]


def _cf(run, patch="bottom"):
    info = run["mesh"].wall_patch_data(patch)
    own = np.asarray(info["owner"], int)
    n = np.asarray(info["normal"], float).reshape(-1, 3)
    U = np.asarray(run["fields"]["U"], float).reshape(-1, 3)[own]
    Ut = U - np.sum(U * n, axis=1)[:, None] * n
    delta = np.maximum(np.asarray(info["delta"], float), 1e-30)
    x = np.asarray(info["center"], float).reshape(-1, 3)[:, 0]
    return x, run["nu"] * np.linalg.norm(Ut, axis=1) / delta / (0.5 * run["Uin"] ** 2)


@pytest.fixture(scope="module")
def fine(rs):
    run = _channel(rs, 40, 30, 1.0, False, 3000)
    assert run["fields"] is not None
    return run


@pytest.fixture(scope="module")
def coarse_wf(rs):
    run = _channel(rs, 24, 14, 30.0, True, 4000)
    assert run["fields"] is not None
    return run


def test_resolved_first_cell_y_plus(fine):
    for patch in ("bottom", "top"):
        yp = _y_plus(fine, patch)
        assert np.all(np.isfinite(yp)) and yp.max() < 5.0, (patch, yp.max())


def test_coarse_wall_functions(fine, coarse_wf):
    assert "Diverged" not in str(coarse_wf["res"].status)
    x_c, cf_c = _cf(coarse_wf)
    assert np.all(np.isfinite(cf_c)) and np.all(cf_c > 0)
    assert _y_plus(coarse_wf, "bottom").max() > _y_plus(fine, "bottom").max()
    x_f, cf_f = _cf(fine)
    ratio = cf_c[np.argmin(np.abs(x_c - 5.0))] / cf_f[np.argmin(np.abs(x_f - 5.0))]
    print("Cf(x = 5) coarse with wall functions / resolved: %.3f" % ratio, flush=True)
    assert 0.5 < ratio < 2.0


def test_coarse_without_wall_functions_does_not_diverge(rs):
    run = _channel(rs, 24, 14, 30.0, False, 4000)
    assert "Diverged" not in str(run["res"].status)
