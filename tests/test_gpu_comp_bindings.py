"""rans_sst_py.CompressibleForwardModel end to end on a tiny channel, with
the contract of the reference's binding tests (reference tests/python/
test_compressible_bindings.py:63-200): a 12 x 10 Ma 0.1 channel, 800 outer
iterations at most, one skin-friction observation.  The result carries a
documented status, a positive iteration count, a finite log-likelihood and
one finite float prediction; last_fields() holds the thermodynamic fields
beside the incompressible ones, with the shapes of the mesh, positive
density, temperature and pressure, a subsonic Mach number and a temperature
that stays near the inlet's."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tiny_case(rs):
    eos = rs.IdealGasEOS()
    T_in, p_ref = 300.0, 101325.0
    rho_in, mu_in = eos.density(p_ref, T_in), eos.viscosity(T_in)
    Uin = 0.1 * eos.sound_speed(T_in)
    nu_in = mu_in / rho_in
    k_in = 1.5 * (0.05 * Uin) ** 2
    om_in = k_in / (100.0 * nu_in)
    mesh = rs.Mesh.make_channel_2d(12, 10, 10.0, 1.0, Re=rho_in * Uin / mu_in, yPlusTarget=1.0)
    mesh.compute_wall_distance()
    bcs = rs.CompressibleBoundaryConditions.channel_defaults(mesh, Uin, T_in, p_ref, kIn=k_in, omIn=om_in)
    obs = rs.ObservationOperator()
    obs.add_skin_friction(wall_patch="bottom", location=rs.Vec3(5.0, 0.0, 0.0), cf_obs=0.005, sigma=0.001,
                          ref_vel=Uin)
    s = rs.SolverSettings()
    s.max_iterations = 800; s.convergence_tol = 1e-3; s.divergence_limit = 1e10
    s.alpha_u = 0.5; s.alpha_p = 0.2; s.alpha_t = 0.7; s.alpha_k = 0.4; s.alpha_omega = 0.4
    s.inner_iterations = 200; s.turb_start_iter = 50; s.turb_update_interval = 2; s.verbose = False
    ps = rs.InferenceParameterSet.a1_betaStar()
    fm = rs.CompressibleForwardModel(mesh=mesh, param_set=ps, obs_op=obs, bcs=bcs, eos=eos, settings=s,
                                     u_init=rs.Vec3(Uin, 0, 0), p_init=p_ref, T_init=T_in, k_init=k_in,
                                     omega_init=om_in)
    print("tiny compressible channel 12 x 10", flush=True)
    res = fm.evaluate(list(ps.pack(rs.SSTCoefficients())))
    return dict(fm=fm, res=res, mesh=mesh, eos=eos)


def test_result_contract(rs, tiny_case):
    r = tiny_case["res"]
    assert r.status in (rs.EvaluationStatus.Converged, rs.EvaluationStatus.Unconverged)
    assert r.simple_iters > 0
    assert np.isfinite(r.log_lik)
    assert len(r.predictions) == 1
    assert all(isinstance(v, float) and np.isfinite(v) for v in r.predictions)


def test_last_fields(tiny_case):
    fm, mesh, eos = tiny_case["fm"], tiny_case["mesh"], tiny_case["eos"]
    assert fm.has_last_fields()
    f = fm.last_fields()
    n = mesh.n_cells()
    for k in ("U", "p", "T", "rho", "k", "omega", "nuT", "F1", "F2", "Pk"):
        assert k in f, k
    assert np.asarray(f["U"]).shape == (n, 3)
    for k in ("p", "T", "rho", "k", "omega", "nuT"):
        assert np.asarray(f[k]).shape == (n,), k
    T, rho, p = np.asarray(f["T"]), np.asarray(f["rho"]), np.asarray(f["p"])
    assert np.all(rho > 0) and np.all(T > 0) and np.all(p > 0)
    Ma = np.linalg.norm(np.asarray(f["U"]), axis=1) / np.sqrt(eos.gamma * eos.R * T)
    assert Ma.max() < 0.8
    assert 200.0 < T.min() and T.max() < 400.0
