"""Team solves (rsg_set_team): the member SIMPLE code spread over many
workgroups of one cooperative launch, for meshes beyond one workgroup.

* trajectories against the oracle on the parity meshes, as the member tests
  do (Jacobi-PCG pressure; differences only from reduction order), with a
  team of 4 workgroups and with the whole device;
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
                                               ("c0", 200, 4), ("channel_bounded", 150, 0)])
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
