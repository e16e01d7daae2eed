"""Kernel-per-phase solves (rsg_set_grid): one member's outer iteration as a
sequence of grid-wide launches (per-cell assembly / gradients / closure, CSR
BiCGSTAB / PCG / AMG-PCG with block-partial dot products, the p' hierarchy
aggregated on the host), for meshes too large for one workgroup.

* trajectories against the oracle on the parity meshes and the bump channel
  (Jacobi-PCG pressure: every algorithm is the reference's; only the
  summation order of the CSR products and dot products differs);
* with the AMG pressure solver the grid solve and the one-workgroup solve of
  the same member agree to round-off where both build the reference's
  hierarchy;
* through rans_sst_py (engine_set_grid) on a 16,000-cell bump channel the
  grid solve tracks the one-workgroup solve, and a finished member's
  observables are evaluated.
"""
import numpy as np
import pytest

import rsg_abi as A
from conftest import oracle_solve, settings_for
from test_gpu_parity import Problem, chk, rel

pytestmark = pytest.mark.gpu


def grid_solve(lib, ctx, path, case, s, iters_per_launch=25):
    chk(lib, ctx, lib.rsg_set_grid(ctx, 1, iters_per_launch))
    try:
        print("grid solve", case, flush=True)
        P = Problem(lib, ctx, path, case, s)
        try:
            return P.solve()
        finally:
            P.close()
    finally:
        lib.rsg_set_grid(ctx, 0, 20)


@pytest.mark.parametrize("case,iters", [("bfs", 80), ("tri_quad", 80), ("hill", 120), ("c0", 200),
                                        ("channel_bounded", 150), ("couette", 150), ("plate", 80), ("bump", 120)])
def test_grid_trajectory_matches_oracle(rsg_lib, gpu_ctx, oracle, meshes, case, iters):
    s = settings_for(case, max_iterations=iters, convergence_tol=1e-12, pressure_solver=1,
                     nut_floor_iters=min(500, iters // 2), record_history=1)
    r, a = grid_solve(rsg_lib, gpu_ctx, meshes[case], case, s)
    o, ho, io = oracle_solve(oracle, meshes[case], case, s, hist_cap=iters)
    assert r.final_iter == io[0] and r.diverged == io[2]
    he = a["history"][:6 * r.history_len].reshape(-1, 6)
    assert he.shape == ho.shape
    for k_gpu, k_or in (("U", "U"), ("k", "k"), ("omega", "om"), ("nuT", "nuT")):
        assert rel(a[k_gpu], o[k_or]) < 1e-4, (case, k_gpu, rel(a[k_gpu], o[k_or]))
    q = np.sqrt(np.mean(np.sum(o["U"].reshape(-1, 3) ** 2, axis=1) ** 2))
    dp = float(np.linalg.norm(a["p"] - o["p"]) / np.sqrt(len(o["p"])))
    assert dp < 1e-4 * q, (case, "p", dp / q)


# bfs: the reference's aggregation coarsens every level, so both paths build
# the same hierarchy and differ by round-off; bump: the reference's
# aggregation stalls on a coarse level of the p' matrix, where the grid path
# switches to pairwise matching (the member path keeps the stalled level), so
# the preconditioners differ and the transients agree to the inner tolerance
@pytest.mark.parametrize("case,bound", [("bfs", 1e-6), ("bump", 1e-3)])
def test_grid_equals_member_with_amg(rsg_lib, gpu_ctx, meshes, case, bound):
    s = settings_for(case, max_iterations=150, convergence_tol=1e-12, pressure_solver=0, amg_refresh=50,
                     coarse_refresh=10)
    r_g, a_g = grid_solve(rsg_lib, gpu_ctx, meshes[case], case, s)
    P = Problem(rsg_lib, gpu_ctx, meshes[case], case, s)
    try:
        r_m, a_m = P.solve()
    finally:
        P.close()
    assert r_g.final_iter == r_m.final_iter
    errs = {k: rel(a_g[k], a_m[k]) for k in ("U", "k", "omega", "nuT")}
    print("grid vs member (%s, AMG): %s" % (case, " ".join("%s %.1e" % kv for kv in errs.items())))
    for k, e in errs.items():
        assert e < bound, (k, e)


def test_grid_arguments(rsg_lib, gpu_ctx):
    assert rsg_lib.rsg_set_grid(gpu_ctx, -1, 20) == -1
    assert rsg_lib.rsg_set_grid(gpu_ctx, 100, 0) == -1
    assert rsg_lib.rsg_set_grid(gpu_ctx, 0, 20) == 0


def test_grid_drop_in_matches_member(rs):
    """rans_sst_py.engine_set_grid on a 16,000-cell bump channel through
    ForwardModel.evaluate, with a skin-friction observable: fields and
    prediction track the one-workgroup solve."""
    from test_gpu_team import _bump_mesh
    mesh = _bump_mesh(rs, 160, 100)
    mesh.compute_wall_distance(True)
    s = rs.SolverSettings()
    s.max_iterations = 60; s.convergence_tol = 1e-12; s.verbose = False
    s.body_force = rs.Vec3(1.0, 0.0, 0.0)
    bcs = rs.FlowBoundaryConditions.channel_defaults(mesh, 1.0, 1e-4, 1.0)
    ps = rs.InferenceParameterSet.a1_betaStar()
    obs = rs.ObservationOperator()
    obs.add_skin_friction("bottom_wall", rs.Vec3(4.0, 0.3, 0.0), 0.01, 0.001, 1.0)
    out = {}
    for name, min_cells in (("member", 0), ("grid", 10000)):
        rs.engine_set_grid(min_cells, 20)
        try:
            print("drop-in bump 160x100", name, flush=True)
            fm = rs.ForwardModel(mesh, ps, obs, bcs, 6e-4, s)
            r = fm.evaluate([0.31, 0.09])
            out[name] = (r.simple_iters, str(r.status).split(".")[-1], fm.last_fields(), list(r.predictions))
        finally:
            rs.engine_set_grid(0, 20)
    assert out["grid"][0] == out["member"][0] == 60
    assert out["grid"][1] == out["member"][1] == "Unconverged"
    # the two paths' p' hierarchies differ below the level where the
    # reference's aggregation stalls (see test_grid_equals_member_with_amg)
    for k in ("U", "k", "omega", "nuT"):
        a, b = out["grid"][2][k].ravel(), out["member"][2][k].ravel()
        assert np.all(np.isfinite(a))
        assert rel(a, b) < 1e-3, (k, rel(a, b))
    assert len(out["grid"][3]) == 1
    assert out["grid"][3][0] == pytest.approx(out["member"][3][0], rel=1e-3)


@pytest.mark.parametrize("levels", [5, 0])
def test_grid_mixed_tracks_fp64(rsg_lib, gpu_ctx, meshes, levels):
    """RSG_PREC_MIXED on the kernel-per-phase path (FP32 BiCGSTAB / AMG-PCG
    inner solves in row-sum form under FP64 defects and refinement, omega in
    FP64) against the FP64 grid solve of the same bump member, both with the
    capped (MFMA-inverted coarsest) and the reference's deep p' hierarchy:
    the transients agree to the inner solves' tolerance."""
    s = settings_for("bump", max_iterations=200, convergence_tol=1e-12, pressure_solver=0, amg_refresh=50,
                     coarse_refresh=10)
    chk(rsg_lib, gpu_ctx, rsg_lib.rsg_set_grid_levels(gpu_ctx, levels))
    chk(rsg_lib, gpu_ctx, rsg_lib.rsg_set_grid(gpu_ctx, 1, 50))
    P = Problem(rsg_lib, gpu_ctx, meshes["bump"], "bump", s)
    try:
        print("grid mixed vs fp64, bump, levels %d" % levels, flush=True)
        t64 = P.submit(prec=0)
        t32 = P.submit(prec=1)
        r64, a64 = P.fetch(t64)
        r32, a32 = P.fetch(t32)
    finally:
        P.close()
        rsg_lib.rsg_set_grid(gpu_ctx, 0, 20)
        rsg_lib.rsg_set_grid_levels(gpu_ctx, 5)
    assert r64.final_iter == r32.final_iter == 200
    errs = {k: rel(a32[k], a64[k]) for k in ("U", "k", "omega", "nuT")}
    print("grid mixed vs fp64 (levels %d): %s" % (levels, " ".join("%s %.1e" % kv for kv in errs.items())))
    for k, e in errs.items():
        assert np.all(np.isfinite(a32[k]))
        assert e < 1e-3, (k, e)


def test_grid_levels_arguments(rsg_lib, gpu_ctx):
    assert rsg_lib.rsg_set_grid_levels(gpu_ctx, 1) == -1
    assert rsg_lib.rsg_set_grid_levels(gpu_ctx, 21) == -1
    assert rsg_lib.rsg_set_grid_levels(gpu_ctx, 0) == 0
    assert rsg_lib.rsg_set_grid_levels(gpu_ctx, 5) == 0


def test_grid_vcycle_graphs_identical(rsg_lib, gpu_ctx, meshes):
    """Kernel-per-phase member with AMG pressure: V-cycles replayed from
    hipGraphs (captured once per hierarchy, invalidated when it is rebuilt
    every amg_refresh iterations) reproduce the directly launched solve bit for
    bit over 150 outer iterations (three hierarchy rebuilds)."""
    s = settings_for("bump", max_iterations=150, convergence_tol=1e-12, pressure_solver=0, amg_refresh=50,
                     coarse_refresh=10)
    out = {}
    try:
        for graphs in (0, 1):
            chk(rsg_lib, gpu_ctx, rsg_lib.rsg_set_graphs(gpu_ctx, graphs))
            out[graphs] = grid_solve(rsg_lib, gpu_ctx, meshes["bump"], "bump", s)
    finally:
        rsg_lib.rsg_set_graphs(gpu_ctx, 1)
    (r0, a0), (r1, a1) = out[0], out[1]
    assert r0.final_iter == r1.final_iter == 150
    for k in ("U", "p", "k", "omega", "nuT"):
        assert np.array_equal(a0[k], a1[k]), k


def test_grid_members_concurrent_equal_alone(rsg_lib, gpu_ctx, meshes):
    """Six kernel-per-phase members submitted together run concurrently (their
    own streams and host threads, four at a time); each one's fields equal its
    solve alone bit for bit."""
    s = settings_for("bump", max_iterations=60, convergence_tol=1e-12, pressure_solver=0, amg_refresh=50,
                     coarse_refresh=10)
    rng = np.random.default_rng(5)
    coeffs = []
    for _ in range(6):
        c = list(A.DEFAULT_COEFFS)
        for i in (8, 9, 0, 1, 2):
            c[i] *= rng.uniform(0.8, 1.2)
        coeffs.append(c)
    chk(rsg_lib, gpu_ctx, rsg_lib.rsg_set_grid(gpu_ctx, 1, 20))
    P = Problem(rsg_lib, gpu_ctx, meshes["bump"], "bump", s)
    try:
        print("grid members concurrently", flush=True)
        tickets = [P.submit(coeffs=c) for c in coeffs]
        together = [P.fetch(t) for t in tickets]
        print("grid members alone", flush=True)
        alone = [P.fetch(P.submit(coeffs=c)) for c in coeffs]
    finally:
        P.close()
        rsg_lib.rsg_set_grid(gpu_ctx, 0, 20)
    for (r1, a1), (r2, a2) in zip(together, alone):
        assert r1.final_iter == r2.final_iter == 60
        for k in ("U", "p", "k", "omega", "nuT"):
            assert np.array_equal(a1[k], a2[k]), k


def test_c4_shape_members_mixed_vs_fp64(rs):
    """The large-member ensemble (C4) at a tenth of its size: three members of
    a 160,000-cell bump channel (defaults and two +-20 % perturbations of the
    five SST coefficients), kernel-per-phase path with the 5-level p'
    hierarchy (coarsest inverted on the matrix cores), 30 outer iterations in
    mixed precision and in FP64; each member's fields track its FP64 solve."""
    from test_gpu_team import _bump_mesh
    import ensembles
    mesh = _bump_mesh(rs, 400, 400)
    mesh.compute_wall_distance(True)
    ps = rs.InferenceParameterSet.from_indices("c2_five", ensembles.THETA_INDICES)
    obs = rs.ObservationOperator()
    rng = np.random.default_rng(11)
    thetas = [list(ensembles.THETA_DEFAULTS)] + [[v * rng.uniform(0.8, 1.2) for v in ensembles.THETA_DEFAULTS]
                                                  for _ in range(2)]
    rs.engine_set_grid(100000, 10)
    rs.engine_set_grid_levels(5)
    out = {}
    try:
        for prec in ("mixed", "fp64"):
            s = rs.SolverSettings()
            s.max_iterations = 30; s.convergence_tol = 1e-12; s.verbose = False
            s.body_force = rs.Vec3(1.0, 0.0, 0.0)
            bcs = rs.FlowBoundaryConditions.channel_defaults(mesh, 1.0, 1e-4, 1.0)
            fms = [rs.ForwardModel(mesh, ps, obs, bcs, 6e-4, s) for _ in thetas]
            for fm in fms:
                fm.set_precision(prec)
            print("c4 shape: 3 members x 160k cells,", prec, flush=True)
            res = rs.evaluate_many(fms, thetas)
            assert all(r.simple_iters == 30 for r in res)
            out[prec] = [fm.last_fields() for fm in fms]
    finally:
        rs.engine_set_grid(0, 20)
    for i in range(len(thetas)):
        a, b = out["mixed"][i], out["fp64"][i]
        errs = {k: rel(np.asarray(a[k]).ravel(), np.asarray(b[k]).ravel()) for k in ("U", "k", "omega", "nuT")}
        print("member %d mixed vs fp64: %s" % (i, " ".join("%s %.1e" % kv for kv in errs.items())))
        for k, e in errs.items():
            assert e < 1e-4, (i, k, e)
