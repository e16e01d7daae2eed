"""Team solves (rsg_set_team): the member SIMPLE code spread over many
workgroups of one cooperative launch, for meshes beyond one workgroup.

* trajectories against the oracle on the parity meshes, as the member tests
  do (Jacobi-PCG pressure; differences only from reduction order), with a
  team of 4 workgroups and with the whole device, and on the bump channel
  (3,200 cells, the bump workload's geometry) with 16 workgroups and the
  whole device;
* the team and the one-workgroup solve of the same member agree to round-off,
  with the AMG pressure solver.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import oracle_solve, settings_for
from test_gpu_parity import Problem, chk, rel

pytestmark = pytest.mark.gpu


def team_solve(lib, ctx, path, case, s, blocks, iters_per_launch=25):
    chk(lib, ctx, lib.rsg_set_team(ctx, 1, blocks, iters_per_launch))
    try:
        print("team solve", case, "blocks", blocks, flush=True)
        P = Problem(lib, ctx, path, case, s)
        try:
            return P.solve()
        finally:
            P.close()
    finally:
        lib.rsg_set_team(ctx, 0, 0, 20)


@pytest.mark.parametrize("case,iters,blocks", [("bfs", 80, 4), ("tri_quad", 80, 4), ("hill", 120, 4),
                                               ("c0", 200, 4), ("channel_bounded", 150, 0), ("bump", 120, 16),
                                               ("bump", 120, 0)])
def test_team_trajectory_matches_oracle(rsg_lib, gpu_ctx, oracle, meshes, case, iters, blocks):
    s = settings_for(case, max_iterations=iters, convergence_tol=1e-12, pressure_solver=1,
                     nut_floor_iters=min(500, iters // 2), record_history=1)
    r, a = team_solve(rsg_lib, gpu_ctx, meshes[case], case, s, blocks)
    o, ho, io = oracle_solve(oracle, meshes[case], case, s, hist_cap=iters)
    assert r.final_iter == io[0] and r.diverged == io[2]
    for k_gpu, k_or in (("U", "U"), ("k", "k"), ("omega", "om"), ("nuT", "nuT")):
        assert rel(a[k_gpu], o[k_or]) < 1e-4, (case, k_gpu, rel(a[k_gpu], o[k_or]))
    q = np.sqrt(np.mean(np.sum(o["U"].reshape(-1, 3) ** 2, axis=1) ** 2))
    dp = float(np.linalg.norm(a["p"] - o["p"]) / np.sqrt(len(o["p"])))
    assert dp < 1e-4 * q, (case, "p", dp / q)


def test_team_equals_member_with_amg(rsg_lib, gpu_ctx, meshes):
    s = settings_for("bfs", max_iterations=150, convergence_tol=1e-12, pressure_solver=0)
    r_t, a_t = team_solve(rsg_lib, gpu_ctx, meshes["bfs"], "bfs", s, 8)
    P = Problem(rsg_lib, gpu_ctx, meshes["bfs"], "bfs", s)
    try:
        r_m, a_m = P.solve()
    finally:
        P.close()
    assert r_t.final_iter == r_m.final_iter
    for k in ("U", "k", "omega", "nuT"):
        assert rel(a_t[k], a_m[k]) < 1e-6, (k, rel(a_t[k], a_m[k]))


def test_team_arguments(rsg_lib, gpu_ctx):
    assert rsg_lib.rsg_set_team(gpu_ctx, -1, 0, 20) == -1
    assert rsg_lib.rsg_set_team(gpu_ctx, 100, 257, 20) == -1
    assert rsg_lib.rsg_set_team(gpu_ctx, 100, 0, 0) == -1
    assert rsg_lib.rsg_set_team(gpu_ctx, 0, 0, 20) == 0


def _bump_mesh(rs, nx, ny):
    """Periodic channel (height 2, length 8) with a Gaussian bump 0.3 high and
    1 wide on the lower wall, nx x ny quads."""
    x = np.linspace(0.0, 8.0, nx + 1)
    g = np.exp(-((x - 4.0) / 1.0) ** 2)
    yb = 0.3 * (g - g[0])
    yb[-1] = yb[0]
    return rs.Mesh.make_curved_channel_periodic_2d(x, yb, 2.0, ny, 1667.0, 1.0)


def test_team_drop_in_matches_member(rs):
    """rans_sst_py.engine_set_team: a 16,000-cell bump channel solved by an
    automatically sized team (16 workgroups) through ForwardModel.evaluate
    matches the one-workgroup solve, AMG pressure solve included."""
    mesh = _bump_mesh(rs, 160, 100)
    mesh.compute_wall_distance(True)
    s = rs.SolverSettings()
    s.max_iterations = 60; s.convergence_tol = 1e-12; s.verbose = False
    s.body_force = rs.Vec3(1.0, 0.0, 0.0)
    bcs = rs.FlowBoundaryConditions.channel_defaults(mesh, 1.0, 1e-4, 1.0)
    ps = rs.InferenceParameterSet.a1_betaStar()
    out = {}
    for name, min_cells in (("member", 0), ("team", 10000)):
        rs.engine_set_team(min_cells, 0, 20)
        try:
            print("drop-in bump 160x100", name, flush=True)
            fm = rs.ForwardModel(mesh, ps, rs.ObservationOperator(), bcs, 6e-4, s)
            r = fm.evaluate([0.31, 0.09])
            out[name] = (r.simple_iters, str(r.status).split(".")[-1], fm.last_fields())
        finally:
            rs.engine_set_team(0, 0, 20)
    assert out["team"][0] == out["member"][0] == 60
    assert out["team"][1] == out["member"][1] == "Unconverged"
    for k in ("U", "k", "omega", "nuT"):
        a, b = out["team"][2][k].ravel(), out["member"][2][k].ravel()
        assert np.all(np.isfinite(a))
        assert rel(a, b) < 1e-6, (k, rel(a, b))
