"""Observables evaluated on the device (rsg_problem_set_observations; the
reference's ObservationOperator::evaluate and logLikelihood,
core/include/ObservationOperator.hpp:159-378, as ForwardModel::evaluate calls
them, ForwardModel.cpp:95-97).

Every observable type on four meshes: the predictions and the log-likelihood
the device returns with a finished solve must equal, bit for bit, what the
oracle (om_observe / om_loglik, themselves bit-exact against the compiled
reference in test_oracle_vs_ref.py) computes from the fields the same solve
returns.  The solves stop at a 300-iteration cap (status MAXITER), so the
fields are mid-trajectory, with recirculation on the step and hill walls.
"""
import ctypes as C

import numpy as np
import pytest

import rsg_abi as A
from conftest import settings_for
from test_gpu_parity import Problem, chk

pytestmark = pytest.mark.gpu


def run_observed(lib, ctx, oracle, path, case, seed):
    P = Problem(lib, ctx, path, case, settings_for(case, max_iterations=300))
    try:
        msh = P.msh
        o = A.observation_set(msh, np.random.default_rng(seed))
        recs, faces, fx = A.resolve_observations(msh, o)
        n = len(o["type"])
        chk(lib, ctx, lib.rsg_problem_set_observations(ctx, P.pid, n, recs, len(faces), A.iptr(faces),
                                                        A.dptr(fx)))
        print("observed solve", case, n, "observables", flush=True)
        t = P.submit()
        arr = (C.c_int64 * 1)(t)
        chk(lib, ctx, lib.rsg_wait(ctx, arr, 1, C.c_double(600.0)))
        r, a = A.new_result(msh.nC, msh.nF - msh.nIF)
        pred = np.full(n, np.nan)
        r.predictions = A.dptr(pred); r.predictions_cap = n
        chk(lib, ctx, lib.rsg_fetch(ctx, C.c_int64(t), C.byref(r)))
        lib.rsg_release(ctx, C.c_int64(t))
        assert r.status == 2 and r.n_predictions == n
        # the oracle on the same fields
        m = oracle.om_mesh_load(path.encode())
        try:
            po = np.zeros(n)
            oracle.om_observe(C.c_void_p(m), n, A.iptr(o["type"]), A.iptr(o["patch"]), A.dptr(o["loc"]),
                              A.iptr(o["comp"]), A.dptr(o["refArea"]), A.dptr(o["refVel"]), A.dptr(o["refP"]),
                              A.dptr(o["refRho"]), A.dptr(a["U"]), A.dptr(a["U_bnd"]), A.dptr(a["p"]),
                              C.c_double(case_nu(case)), A.dptr(po))
        finally:
            oracle.om_mesh_free(C.c_void_p(m))
        llo = oracle.om_loglik(n, A.dptr(po), A.dptr(o["observed"]), A.dptr(o["sigma"]), A.dptr(o["sigmaModel"]))
        return o, pred, r.log_lik, po, llo
    finally:
        P.close()


def case_nu(case):
    from conftest import CASES
    return CASES[case][2]


@pytest.mark.parametrize("case", ["bfs", "hill", "channel_bounded", "plate"])
def test_device_observables_match_oracle(rsg_lib, gpu_ctx, oracle, meshes, case):
    o, pred, ll, po, llo = run_observed(rsg_lib, gpu_ctx, oracle, meshes[case], case, 3)
    bad = np.nonzero(pred != po)[0]
    assert bad.size == 0, [(int(o["type"][i]), pred[i], po[i]) for i in bad]
    assert ll == llo
    assert np.all(np.isfinite(pred))


def test_observation_arguments(rsg_lib, gpu_ctx, meshes):
    P = Problem(rsg_lib, gpu_ctx, meshes["bfs"], "bfs", settings_for("bfs", max_iterations=10))
    try:
        rec = (A.Observation * 1)()
        rec[0].type = 2; rec[0].index = P.msh.nC; rec[0].component = 0        # cell out of range
        assert rsg_lib.rsg_problem_set_observations(gpu_ctx, P.pid, 1, rec, 0, None, None) == -1
        rec[0].type = 0; rec[0].index = -1; rec[0].face_begin = 0; rec[0].face_end = 1
        f = np.array([0], np.int32); x = np.zeros(1)                          # face 0 is internal
        assert rsg_lib.rsg_problem_set_observations(gpu_ctx, P.pid, 1, rec, 1, A.iptr(f), A.dptr(x)) == -1
        rec[0].type = 7
        assert rsg_lib.rsg_problem_set_observations(gpu_ctx, P.pid, 1, rec, 0, None, None) == -1
        rec[0].type = 2; rec[0].index = 0
        assert rsg_lib.rsg_problem_set_observations(gpu_ctx, P.pid, 1, rec, 0, None, None) == 0
        assert rsg_lib.rsg_problem_set_observations(gpu_ctx, P.pid, 1, rec, 0, None, None) != 0   # once only
    finally:
        P.close()


def test_empty_patch_observables(rsg_lib, gpu_ctx, meshes):
    """Observables on an empty patch take the reference's defaults
    (ObservationOperator.hpp: skin friction with no face 0, drag over no face 0,
    separation 0, reattachment 0 with no last face); the log-likelihood sums
    -0.5 (0 - observed)^2 / var."""
    P = Problem(rsg_lib, gpu_ctx, meshes["bfs"], "bfs", settings_for("bfs", max_iterations=5))
    try:
        recs = (A.Observation * 4)()
        observed = [0.3, -0.2, 0.5, 1.5]
        for i, typ in enumerate((3, 0, 4, 5)):
            recs[i].type = typ; recs[i].index = -1; recs[i].face_begin = 0; recs[i].face_end = 0
            recs[i].ref_velocity = 1.0; recs[i].ref_area = 1.0; recs[i].ref_density = 1.0
            recs[i].observed = observed[i]; recs[i].sigma = 0.1 * (i + 1); recs[i].sigma_model = 0.0
        chk(rsg_lib, gpu_ctx, rsg_lib.rsg_problem_set_observations(gpu_ctx, P.pid, 4, recs, 0, None, None))
        t = P.submit()
        chk(rsg_lib, gpu_ctx, rsg_lib.rsg_wait(gpu_ctx, (C.c_int64 * 1)(t), 1, C.c_double(120.0)))
        r, _ = A.new_result(P.msh.nC, P.msh.nF - P.msh.nIF)
        pred = np.full(4, np.nan)
        r.predictions = A.dptr(pred); r.predictions_cap = 4
        chk(rsg_lib, gpu_ctx, rsg_lib.rsg_fetch(gpu_ctx, C.c_int64(t), C.byref(r)))
        rsg_lib.rsg_release(gpu_ctx, C.c_int64(t))
        assert r.n_predictions == 4 and np.all(pred == 0.0)
        ll = 0.0
        for i in range(4):
            sg = 0.1 * (i + 1)
            var = sg * sg + 0.0 * 0.0
            ll += -0.5 * (0.0 - observed[i]) * (0.0 - observed[i]) / var
        assert r.log_lik == ll
    finally:
        P.close()
