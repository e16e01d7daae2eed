"""The warm-started finite-difference true-model gradient
(ParameterSensitivity.eta_jacobian_warm_fd) on the engine: the 22 perturbed
solves of one gradient run as one batch on the device.

* the reference's own contract for this call (tests/python/
  test_warm_fd_gradient.py in the reference): warm-FD columns match a cold
  full FD on the strong coefficients, the log-likelihood gradient is
  consistent, kappa has no effect;
* parity with the oracle: the same warm-FD columns from the oracle's warm
  solves (om_simple_solve_warm from the engine's stored state, observations by
  om_observe), with the Jacobi-PCG pressure solver so both sides run the same
  algorithm and differ only by reduction-order round-off.
"""
import ctypes as C
import time

import numpy as np
import pytest

import rsg_abi as A

pytestmark = pytest.mark.gpu

NAMES = ["sigma_k1", "sigma_w1", "beta1", "alpha1", "sigma_k2", "sigma_w2",
         "beta2", "alpha2", "betaStar", "a1", "kappa"]
NONLINEAR = (8, 9)
STRONG = (2, 8, 9)
UB, RE_B = 1.0, 6800.0
CF_X = np.linspace(3.0, 7.0, 4)


def channel_args(rs, pressure_solver="AMG", max_iter=1500):
    """Inlet-velocity / outlet-pressure channel, 20 x 14 cells, four skin-friction
    observations on the bottom wall (the reference test's case)."""
    h, Lx = 1.0, 10.0
    nu = UB * h / RE_B
    cf = 0.073 * RE_B ** -0.25
    mesh = rs.Mesh.make_channel_2d(nx=20, ny=14, Lx=Lx, Ly=2.0 * h, Re=RE_B, yPlusTarget=1.0)
    mesh.compute_wall_distance()
    k_in = 1.5 * (UB * 0.05) ** 2
    om_in = k_in / (nu * 100.0)
    bcs = rs.FlowBoundaryConditions.channel_defaults(mesh, UB, k_in, om_in)
    obs = rs.ObservationOperator()
    for x in CF_X:
        obs.add_skin_friction("bottom", rs.Vec3(float(x), 0, 0), cf, 0.05 * cf, UB)
    s = rs.SolverSettings()
    s.max_iterations = max_iter; s.convergence_tol = 1e-6; s.verbose = False
    s.alpha_u = 0.7; s.alpha_p = 0.5
    s.pressure_solver = pressure_solver
    return (mesh, obs, bcs, nu, s, rs.Vec3(UB, 0, 0), 0.0, k_in, om_in)


def menter(rs):
    return list(rs.InferenceParameterSet.all11().pack(rs.SSTCoefficients()))


@pytest.fixture(scope="module")
def case(rs):
    args = channel_args(rs)
    sens = rs.ParameterSensitivity(*args)
    th = menter(rs)
    # as in the reference's test, the state is whatever 1500 outer iterations
    # reach (the reference solver itself does not meet tol 1e-6 on this case
    # within 5000 iterations)
    st = sens.solve_state(th)
    assert str(st).split(".")[-1] in ("Converged", "Unconverged") and sens.has_state()
    return rs, sens, args[0], th


def cold_eta(rs, theta, pressure_solver="AMG"):
    s = rs.ParameterSensitivity(*channel_args(rs, pressure_solver))
    s.solve_state(list(theta))
    return np.asarray(s.observe(list(theta)), float)


def test_warm_fd_matches_cold_fd(case):
    rs, sens, _, th = case
    t0 = time.perf_counter()
    res = sens.eta_jacobian_warm_fd(th, h_rel=5e-4, h_floor=1e-7, warm_max_iter=400, warm_tol=1e-5)
    t_warm = time.perf_counter() - t0
    J = np.asarray(res.d_obs_d_theta)
    assert J.shape == (4, 11)
    assert all(res.krylov_converged)
    assert res.n_residual_evals > 0
    th0 = np.asarray(th, float)
    for j in STRONG:
        hr = 1e-5 if j in NONLINEAR else 5e-4
        hj = max(hr * abs(th0[j]), 1e-7)
        e = np.zeros(11); e[j] = hj
        col = (cold_eta(rs, th0 + e) - cold_eta(rs, th0 - e)) / (2 * hj)
        rel = np.linalg.norm(J[:, j] - col) / max(np.linalg.norm(col), 1e-30)
        assert rel < 2e-2, (NAMES[j], rel)
    print("warm-FD gradient (22 solves in one batch): %.3f s, %d outer iterations in total"
          % (t_warm, res.n_residual_evals), flush=True)


def test_warm_fd_loglik_gradient_consistent(case):
    rs, sens, _, th = case
    g = np.asarray(sens.eta_jacobian_warm_fd(th, warm_max_iter=400, warm_tol=1e-5).log_lik_gradient)
    assert g.shape == (11,) and np.all(np.isfinite(g))
    assert abs(g[10]) < 1e-6                          # kappa does not enter the closure
    j = 8
    th0 = np.asarray(th, float); hj = max(1e-5 * abs(th0[j]), 1e-7)
    e = np.zeros(11); e[j] = hj
    sp = rs.ParameterSensitivity(*channel_args(rs)); sp.solve_state(list(th0 + e))
    sm = rs.ParameterSensitivity(*channel_args(rs)); sm.solve_state(list(th0 - e))
    g_fd = (sp.log_lik(list(th0 + e)) - sm.log_lik(list(th0 - e))) / (2 * hj)
    assert abs(g[j] - g_fd) <= 0.05 * abs(g_fd) + 1e-6


def test_warm_fd_requires_state_and_full_theta(rs):
    sens = rs.ParameterSensitivity(*channel_args(rs))
    r = sens.eta_jacobian_warm_fd(menter(rs))
    assert not sens.has_state() and not any(r.krylov_converged)
    assert np.all(np.asarray(r.d_obs_d_theta) == 0.0)
    with pytest.raises(ValueError):
        sens.solve_state([0.09, 0.31])


def _oracle_state(fields, nC, nB):
    U = np.asarray(fields["U"]); Ub = np.asarray(fields["U_bnd"])
    return {"U": U, "Ub": Ub, "p": fields["p"], "pb": fields["p_bnd"], "k": fields["k"], "kb": fields["k_bnd"],
            "om": fields["omega"], "omb": fields["omega_bnd"], "nuT": fields["nuT"], "F1": fields["F1"],
            "F2": fields["F2"], "Pk": fields["Pk"], "CDkw": fields["CDkw"]}


def test_warm_fd_matches_oracle(rs, oracle, tmp_path):
    from conftest import state_ptrs
    args = channel_args(rs, pressure_solver="PCG")
    mesh, nu = args[0], args[3]
    sens = rs.ParameterSensitivity(*args)
    th = menter(rs)
    sens.solve_state(th)
    res = sens.eta_jacobian_warm_fd(th, warm_max_iter=400, warm_tol=1e-5)
    J = np.asarray(res.d_obs_d_theta)

    path = str(tmp_path / "chan.msh")
    mesh.save_binary(path)
    msh = A.Msh(path)
    nC, nB = msh.nC, msh.nF - msh.nIF
    f = sens.state_fields()
    state = _oracle_state(f, nC, nB)
    turb = 1 if f["turb_established"] else 0
    bottom = msh.pnames.index("bottom")
    nob = len(CF_X)
    typ = np.full(nob, 3, np.int32); patch = np.full(nob, bottom, np.int32)
    loc = np.zeros(3 * nob); loc[0::3] = CF_X
    comp = np.zeros(nob, np.int32); ones = np.ones(nob); zeros = np.zeros(nob)
    s = A.default_settings(max_iterations=400, convergence_tol=1e-5, alpha_u=0.7, alpha_p=0.5, pressure_solver=1,
                           body_force=(0.0, 0.0, 0.0))
    D = lambda a: (C.c_double * len(a))(*a)
    k_in = 1.5 * (UB * 0.05) ** 2
    vals = [UB, k_in, k_in / (nu * 100.0)]
    m = oracle.om_mesh_load(path.encode())
    th0 = np.asarray(th, float)
    try:
        for j in STRONG:
            hr = 1e-5 if j in NONLINEAR else 5e-4
            hj = max(hr * abs(th0[j]), 1e-7)
            eta = []
            for sgn in (1.0, -1.0):
                t = th0.copy(); t[j] += sgn * hj
                out = {k: np.zeros(3 * nC if k == "U" else nC) for k in ("U", "p", "k", "om", "nuT", "F1", "F2",
                                                                       "Pk", "CDkw")}
                bnd = {"Ub": np.zeros(3 * nB), "pb": np.zeros(nB), "kb": np.zeros(nB), "omb": np.zeros(nB)}
                info = (C.c_int * 5)()
                keep, ptrs = state_ptrs(state)
                oracle.om_simple_solve_warm(C.c_void_p(m), 0, D(vals), C.c_double(nu), D(list(t)),
                                            D(A.packed_settings(s)), ptrs, turb,
                                            *[A.dptr(out[k]) for k in ("U", "p", "k", "om", "nuT", "F1", "F2",
                                                                       "Pk", "CDkw")],
                                            *[A.dptr(bnd[k]) for k in ("Ub", "pb", "kb", "omb")], None, 0, info)
                pred = np.zeros(nob)
                oracle.om_observe(C.c_void_p(m), nob, A.iptr(typ), A.iptr(patch), A.dptr(loc), A.iptr(comp),
                                  A.dptr(ones), A.dptr(ones * UB), A.dptr(zeros), A.dptr(ones),
                                  A.dptr(out["U"]), A.dptr(bnd["Ub"]), A.dptr(out["p"]), C.c_double(nu),
                                  A.dptr(pred))
                eta.append(pred)
            col = (eta[0] - eta[1]) / (2 * hj)
            rel = np.linalg.norm(J[:, j] - col) / max(np.linalg.norm(col), 1e-30)
            assert rel < 1e-3, (NAMES[j], rel)
    finally:
        oracle.om_mesh_free(C.c_void_p(m))
