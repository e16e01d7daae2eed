"""GPU parity of the HIP engine (librsg.so, called through its C-ABI) against
the oracle and the reference's golden fields.

* linear solvers: device BiCGSTAB / Jacobi-PCG / AMG-PCG vs dense solves;
* short SIMPLE trajectories on six meshes (Jacobi-PCG pressure, the
  reference's own algorithms): only reduction order and libm ulps differ;
* C0 (periodic channel, Re_tau 550) driven to convergence with the device
  AMG: vs the reference's converged fields at tol 1e-8 and at 2e-10;
* members are bit-identical alone and in a mixed batch; admission under a
  small HBM budget; status classification forced from the host.
"""
import ctypes as C
import os
import time

import numpy as np
import pytest

import rsg_abi as A
from conftest import CASES, TESTS, oracle_solve, settings_for

pytestmark = pytest.mark.gpu
D = lambda a: (C.c_double * len(a))(*a)


def chk(lib, ctx, rc):
    if rc != 0:
        raise RuntimeError("rc=%d %s" % (rc, lib.rsg_last_error(ctx)))


class Problem:
    def __init__(self, lib, ctx, path, case, settings):
        self.lib, self.ctx = lib, ctx
        self.msh = A.Msh(path)
        kind, vals, nu, _ = CASES[case]
        self.mid = C.c_int32()
        chk(lib, ctx, lib.rsg_mesh_upload(ctx, C.byref(self.msh.desc()), C.byref(self.mid)))
        self.pid = C.c_int32()
        chk(lib, ctx, lib.rsg_problem_create(ctx, self.mid, C.byref(self.msh.bc(kind, vals)), C.c_double(nu),
                                             C.byref(settings), C.byref(self.pid)))
        self.settings = settings

    def submit(self, coeffs=None, warm=None, turb=0, prec=0):
        init = A.Init()
        init.kind = 0; init.U[0] = 1.0; init.k = 1e-4; init.omega = 1.0
        if warm is not None:   # oracle-layout state dict: U Ub p pb k kb om omb nuT F1 F2 Pk CDkw
            init.kind = 1
            self._keep = {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in warm.items()}
            w = self._keep
            init.U_cells = A.dptr(w["U"]); init.p_cells = A.dptr(w["p"]); init.k_cells = A.dptr(w["k"])
            init.om_cells = A.dptr(w["om"]); init.nuT = A.dptr(w["nuT"]); init.F1 = A.dptr(w["F1"])
            init.F2 = A.dptr(w["F2"]); init.Pk = A.dptr(w["Pk"]); init.CDkw = A.dptr(w["CDkw"])
            init.U_bnd = A.dptr(w["Ub"]); init.p_bnd = A.dptr(w["pb"]); init.k_bnd = A.dptr(w["kb"])
            init.om_bnd = A.dptr(w["omb"]); init.turb_established = turb
        t = C.c_int64()
        chk(self.lib, self.ctx, self.lib.rsg_submit(self.ctx, self.pid, D(coeffs or A.DEFAULT_COEFFS),
                                                    C.byref(init), prec, C.byref(t)))
        return t.value

    def fetch(self, ticket):
        arr = (C.c_int64 * 1)(ticket)
        chk(self.lib, self.ctx, self.lib.rsg_wait(self.ctx, arr, 1, C.c_double(600.0)))
        cap = self.settings.max_iterations if self.settings.record_history else 0
        r, a = A.new_result(self.msh.nC, self.msh.nF - self.msh.nIF, cap)
        chk(self.lib, self.ctx, self.lib.rsg_fetch(self.ctx, C.c_int64(ticket), C.byref(r)))
        self.lib.rsg_release(self.ctx, C.c_int64(ticket))
        assert r.status == 2, "device solve failed: %s" % self.lib.rsg_last_error(self.ctx)
        return r, a

    def solve(self, coeffs=None, warm=None, turb=0):
        return self.fetch(self.submit(coeffs, warm, turb))

    def close(self):
        self.lib.rsg_problem_release(self.ctx, self.pid)
        self.lib.rsg_mesh_release(self.ctx, self.mid)


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("solver", [2, 1, 0])
def test_device_linear_solvers(rsg_lib, gpu_ctx, oracle, meshes, solver):
    msh = A.Msh(meshes["c0"])
    n, nif = msh.nC, msh.nIF
    rng = np.random.default_rng(3)
    own = msh.owner[:nif].copy(); nbr = msh.neighbor.copy()
    cf = np.exp(rng.normal(size=nif))
    up = -cf.copy(); lo = -(cf + rng.uniform(0, 0.5, nif)) if solver == 2 else -cf.copy()
    dg = np.zeros(n); np.add.at(dg, own, -lo); np.add.at(dg, nbr, -up); dg[0] *= 2.0
    if solver == 2:
        dg *= 1.1   # the skewed periodic operator is otherwise numerically singular (cond ~1e15)
    b = rng.normal(size=n)
    Ad = np.diag(dg); Ad[own, nbr] += up; Ad[nbr, own] += lo
    xs = np.linalg.solve(Ad, b)
    x = np.zeros(n); st = np.zeros(4)
    chk(rsg_lib, gpu_ctx, rsg_lib.rsg_ldu_solve(gpu_ctx, n, nif, A.iptr(own), A.iptr(nbr), A.dptr(dg), A.dptr(up),
                                                A.dptr(lo), A.dptr(b), A.dptr(x), solver, 0, 2000, C.c_double(1e-11),
                                                A.dptr(st)))
    assert st[3] == 1
    assert st[2] == pytest.approx(np.linalg.norm(b), rel=1e-13)
    assert rel(x, xs) < 1e-8
    if solver != 0:   # same Krylov recurrence as the oracle: same iteration count
        xo = np.zeros(n); so = np.zeros(4)
        oracle.om_ldu_solve(solver, n, nif, A.iptr(own), A.iptr(nbr), A.dptr(dg), A.dptr(up), A.dptr(lo),
                            A.dptr(b), A.dptr(xo), 2000, C.c_double(1e-11), A.dptr(so))
        assert abs(st[0] - so[0]) <= 2


@pytest.mark.parametrize("case,iters", [("c0", 400), ("channel_bounded", 150), ("couette", 150),
                                        ("bfs", 80), ("plate", 80), ("hill", 120), ("tri_quad", 80)])
def test_trajectory_matches_oracle(rsg_lib, gpu_ctx, oracle, meshes, case, iters):
    s = settings_for(case, max_iterations=iters, convergence_tol=1e-12, pressure_solver=1,
                     nut_floor_iters=min(500, iters // 2), record_history=1)
    P = Problem(rsg_lib, gpu_ctx, meshes[case], case, s)
    try:
        r, a = P.solve()
    finally:
        P.close()
    o, ho, io = oracle_solve(oracle, meshes[case], case, s, hist_cap=iters)
    assert r.final_iter == io[0] and r.diverged == io[2]
    he = a["history"][:6 * r.history_len].reshape(-1, 6)
    assert he.shape == ho.shape
    # The device sums dot products in a block tree, the oracle left to right:
    # over a few hundred transient iterations those round-off differences
    # grow to ~1e-6 (U) .. ~1e-5 (k, nuT) relative.  The pressure of the
    # periodic and near-hydrostatic cases is tiny against its dynamic scale,
    # so it is compared against rms(|U|^2).
    for k_gpu, k_or in (("U", "U"), ("k", "k"), ("omega", "om"), ("nuT", "nuT")):
        assert rel(a[k_gpu], o[k_or]) < 1e-4, (case, k_gpu, rel(a[k_gpu], o[k_or]))
    q = np.sqrt(np.mean(np.sum(o["U"].reshape(-1, 3) ** 2, axis=1) ** 2))
    dp = float(np.linalg.norm(a["p"] - o["p"]) / np.sqrt(len(o["p"])))
    assert dp < 1e-4 * q, (case, "p", dp / q)


def _golden(tol):
    g = np.load(os.path.join(TESTS, "golden", "c0_ref_tol%.0e.npz" % tol))
    return {k: g[k] for k in g.files}


@pytest.mark.parametrize("block", [256, 512])
@pytest.mark.parametrize("tol,bound", [(1e-8, 1e-10), (2e-10, 1e-10)])
def test_c0_converged_fields_vs_reference_golden(rsg_lib, gpu_ctx, meshes, tol, bound, block):
    """Device FP64 path (device AMG) to convergence vs the reference's converged
    fields, with both member workgroup shapes; at tol 2e-10 both sit within
    round-off of the discrete fixed point, so the fields must agree to 1e-8
    relative L2."""
    s = settings_for("c0", max_iterations=80000, convergence_tol=tol, pressure_solver=0,
                     amg_refresh=1000, coarse_refresh=10)
    rsg_lib.rsg_set_launch(gpu_ctx, 2000, block)
    P = Problem(rsg_lib, gpu_ctx, meshes["c0"], "c0", s)
    try:
        r, a = P.solve()
    finally:
        P.close()
    g = _golden(tol)
    assert r.converged == 1
    assert abs(r.final_iter - int(g["final_iter"])) < 0.05 * int(g["final_iter"])
    errs = {k_gpu: rel(a[k_gpu], g[k_g]) for k_gpu, k_g in (("U", "U"), ("k", "k"), ("omega", "omega"),
                                                              ("nuT", "nuT"))}
    print("c0 tol %g block %d: %d iterations (reference %d), rel-L2 %s"
          % (tol, block, r.final_iter, int(g["final_iter"]), " ".join("%s %.2e" % kv for kv in errs.items())))
    for k_gpu, e in errs.items():
        assert e < bound, (k_gpu, e)
    # The body-force-driven periodic channel has no mean pressure gradient:
    # the converged p is round-off (|p| ~ 1e-12), so it is held against the
    # dynamic pressure scale rms(|U|^2) rather than against itself.
    q = np.sqrt(np.mean(np.sum(g["U"].reshape(-1, 3) ** 2, axis=1) ** 2))
    dp = float(np.linalg.norm(a["p"] - g["p"]) / np.sqrt(len(g["p"])))
    assert dp < bound * q, ("p", dp / q)


def test_member_identical_alone_and_in_mix(rsg_lib, gpu_ctx, meshes):
    s = settings_for("c0", max_iterations=1500, convergence_tol=1e-12, pressure_solver=0, amg_refresh=1000)
    rsg_lib.rsg_set_launch(gpu_ctx, 300, 256)
    P = Problem(rsg_lib, gpu_ctx, meshes["c0"], "c0", s)
    Q = Problem(rsg_lib, gpu_ctx, meshes["channel_bounded"], "channel_bounded",
                settings_for("channel_bounded", max_iterations=700, pressure_solver=0))
    try:
        alone, a0 = P.solve()
        rng = np.random.default_rng(0)
        tickets = []
        for i in range(12):
            c = list(A.DEFAULT_COEFFS)
            c[8] *= rng.uniform(0.8, 1.2); c[9] *= rng.uniform(0.8, 1.2)
            tickets.append((Q if i % 2 else P).submit(c))
            if i == 5:
                mine = P.submit()
        mix, a1 = P.fetch(mine)
        for i, t in enumerate(tickets):
            (Q if i % 2 else P).fetch(t)
    finally:
        P.close(); Q.close()
    for k in ("U", "p", "k", "omega", "nuT"):
        assert np.array_equal(a0[k], a1[k]), k
    assert alone.final_iter == mix.final_iter


def test_time_sliced_launches_do_not_change_results(rsg_lib, gpu_ctx, meshes):
    """rsg_set_time_slice: launches split a member at wall-clock boundaries
    (here ~2 ms, a few iterations each) instead of every 300 iterations; the
    member's fields and iteration count are bit-identical."""
    s = settings_for("c0", max_iterations=1200, convergence_tol=1e-12, pressure_solver=0, amg_refresh=1000)
    P = Problem(rsg_lib, gpu_ctx, meshes["c0"], "c0", s)
    try:
        rsg_lib.rsg_set_launch(gpu_ctx, 300, 0)
        st0 = np.zeros(5); rsg_lib.rsg_stats(gpu_ctx, A.dptr(st0))
        fixed, a0 = P.solve()
        st1 = np.zeros(5); rsg_lib.rsg_stats(gpu_ctx, A.dptr(st1))
        chk(rsg_lib, gpu_ctx, rsg_lib.rsg_set_time_slice(gpu_ctx, 2000))
        sliced, a1 = P.solve()
        st2 = np.zeros(5); rsg_lib.rsg_stats(gpu_ctx, A.dptr(st2))
    finally:
        rsg_lib.rsg_set_time_slice(gpu_ctx, 0)
        P.close()
    assert st1[0] - st0[0] == 4                    # 1200 iterations in launches of 300
    assert st2[0] - st1[0] > 4                     # more, shorter launches
    assert fixed.final_iter == sliced.final_iter == 1200
    for k in ("U", "p", "k", "omega", "nuT"):
        assert np.array_equal(a0[k], a1[k]), k
    assert rsg_lib.rsg_set_time_slice(gpu_ctx, -1) != 0


@pytest.mark.parametrize("policy", [0, 1])
def test_more_members_than_resident_slots(rsg_lib, gpu_ctx, meshes, policy):
    """300 members on two meshes, more than the device holds at once: each
    launch takes a rotating window (policy 0) or the largest meshes first
    (policy 1); every member's fields equal its solve alone."""
    s_small = settings_for("channel_bounded", max_iterations=60, pressure_solver=0)
    s_big = settings_for("c0", max_iterations=80, convergence_tol=1e-12, pressure_solver=0, amg_refresh=1000)
    Q = Problem(rsg_lib, gpu_ctx, meshes["channel_bounded"], "channel_bounded", s_small)
    P = Problem(rsg_lib, gpu_ctx, meshes["c0"], "c0", s_big)
    try:
        rsg_lib.rsg_set_launch(gpu_ctx, 20, 0)
        _, q_alone = Q.solve()
        _, p_alone = P.solve()
        chk(rsg_lib, gpu_ctx, rsg_lib.rsg_set_schedule(gpu_ctx, policy))
        tickets = [(Q if i % 10 else P).submit() for i in range(300)]
        res = [(Q if i % 10 else P).fetch(t) for i, t in enumerate(tickets)]
    finally:
        rsg_lib.rsg_set_schedule(gpu_ctx, 0)
        rsg_lib.rsg_set_launch(gpu_ctx, 200, 0)
        P.close(); Q.close()
    for i, (r, a) in enumerate(res):
        ref = q_alone if i % 10 else p_alone
        for k in ("U", "p", "k", "omega", "nuT"):
            assert np.array_equal(a[k], ref[k]), (i, k)
    assert rsg_lib.rsg_set_schedule(gpu_ctx, 7) != 0


@pytest.mark.parametrize("case", ["c0", "bfs"])
def test_warm_start_matches_oracle(rsg_lib, gpu_ctx, oracle, meshes, case):
    """rsg_init kind 1 (ForwardModel's warm start from a cached state, with its
    turbEstablished flag) against the oracle's warm start: same iteration
    count, fields within the block-reduction round-off of a trajectory."""
    from conftest import oracle_solve_warm
    s0 = settings_for(case, max_iterations=120, convergence_tol=1e-12, pressure_solver=1, nut_floor_iters=40)
    base, _, _ = oracle_solve(oracle, meshes[case], case, s0)
    coeffs = list(A.DEFAULT_COEFFS)
    coeffs[9] *= 1.05; coeffs[8] *= 0.97
    s = settings_for(case, max_iterations=90, convergence_tol=1e-12, pressure_solver=1, nut_floor_iters=40)
    for turb in (1, 0):
        o, _, io = oracle_solve_warm(oracle, meshes[case], case, s, base, turb, coeffs)
        P = Problem(rsg_lib, gpu_ctx, meshes[case], case, s)
        try:
            r, a = P.solve(coeffs, warm=base, turb=turb)
        finally:
            P.close()
        assert r.final_iter == io[0] and r.turb_established == io[4]
        for k_gpu, k_or in (("U", "U"), ("k", "k"), ("omega", "om"), ("nuT", "nuT")):
            assert rel(a[k_gpu], o[k_or]) < 1e-4, (case, turb, k_gpu, rel(a[k_gpu], o[k_or]))


def test_admission_by_hbm_budget(rsg_lib, meshes):
    ctx = C.c_void_p()
    # a C0 member's workspace is ~0.57 MB: 1.75 MiB holds the mesh and two members
    chk(rsg_lib, None, rsg_lib.rsg_context_create(0, C.c_uint64(1792 << 10), C.byref(ctx)))
    try:
        s = settings_for("c0", max_iterations=300, convergence_tol=1e-12)
        P = Problem(rsg_lib, ctx, meshes["c0"], "c0", s)
        tickets = [P.submit() for _ in range(10)]
        results = [P.fetch(t) for t in tickets]
        st = np.zeros(5)
        rsg_lib.rsg_stats(ctx, A.dptr(st))
        assert st[2] == 10 and 1 <= st[4] < 10        # retired all, never all at once
        for r, a in results[1:]:
            assert np.array_equal(a["U"], results[0][1]["U"])
        P.close()
    finally:
        rsg_lib.rsg_context_destroy(ctx)


def test_status_forced_from_host(rsg_lib, gpu_ctx, meshes):
    s = settings_for("channel_bounded", max_iterations=20, convergence_tol=1e-12)
    P = Problem(rsg_lib, gpu_ctx, meshes["channel_bounded"], "channel_bounded", s)
    r, _ = P.solve()
    P.close()
    assert r.converged == 0 and r.diverged == 0 and r.final_iter == 20
    s = settings_for("channel_bounded", max_iterations=200, divergence_limit=1e-9)
    P = Problem(rsg_lib, gpu_ctx, meshes["channel_bounded"], "channel_bounded", s)
    r, _ = P.solve()
    P.close()
    assert r.diverged == 1 and r.final_iter == 0


def test_admission_hold(rsg_lib, gpu_ctx, meshes, rs):
    """rsg_hold_admission: tickets submitted under a hold stay pending while
    any hold is left (holds nest, one release each); released, they start
    together in one launch (60 iterations each, 200 per launch) and give the
    results of solves submitted without a hold.  Releasing with none held and
    an out-of-range argument are errors."""
    lib, ctx = rsg_lib, gpu_ctx
    s = settings_for("channel_bounded", max_iterations=60, convergence_tol=1e-12)
    P = Problem(lib, ctx, meshes["channel_bounded"], "channel_bounded", s)
    try:
        r0, a0 = P.solve()
        held = 0
        try:
            for _ in range(2):
                chk(lib, ctx, lib.rsg_hold_admission(ctx, 1))
                held += 1
            t = (C.c_int64 * 2)(P.submit(), P.submit())
            st = (C.c_int32 * 2)()
            for _ in range(2):
                time.sleep(0.3)
                chk(lib, ctx, lib.rsg_poll(ctx, t, 2, st))
                assert list(st) == [0, 0], "admitted while held"
                if held == 2:
                    chk(lib, ctx, lib.rsg_hold_admission(ctx, 0))
                    held -= 1
            st0 = (C.c_double * 5)()
            chk(lib, ctx, lib.rsg_stats(ctx, st0))
        finally:
            for _ in range(held):
                lib.rsg_hold_admission(ctx, 0)
        chk(lib, ctx, lib.rsg_wait(ctx, t, 2, C.c_double(600.0)))
        st1 = (C.c_double * 5)()
        chk(lib, ctx, lib.rsg_stats(ctx, st1))
        assert st1[0] - st0[0] == 1 and st1[1] - st0[1] == 2, (list(st0), list(st1))
        for x in t:
            r, a = P.fetch(x)
            assert (r.final_iter, r.converged) == (r0.final_iter, r0.converged)
            for k in ("U", "p", "k", "omega", "nuT"):
                assert np.array_equal(a[k], a0[k]), k
        assert lib.rsg_hold_admission(ctx, 0) == -4
        assert lib.rsg_hold_admission(ctx, 2) == -1
    finally:
        P.close()
    with pytest.raises(RuntimeError):
        rs.engine_hold_admission(False)


def test_forward_model_end_to_end(rs, oracle):
    """rans_sst_py drop-in: evaluate -> status, predictions, last_fields; the
    warm-start cache restarts a nearby theta from the stored state."""
    y = rs.Mesh.geometric_channel_nodes(64, 1.06, 1.0)
    mesh = rs.Mesh.make_channel_periodic_2d(np.linspace(0.0, 1.0, 5), np.array(y))
    mesh.compute_wall_distance()
    s = rs.SolverSettings()
    s.max_iterations = 32000; s.convergence_tol = 1e-8; s.verbose = False
    s.body_force = rs.Vec3(1.0, 0.0, 0.0)
    obs = rs.ObservationOperator()
    obs.add_skin_friction("bottom_wall", rs.Vec3(0.5, 0.0, 0.5), 0.005, 0.001, 1.0)
    obs.add_velocity_profile(rs.Vec3(0.5, 1.0, 0.5), 0, 20.0, 1.0)
    bcs = rs.FlowBoundaryConditions.channel_defaults(mesh, 1.0, 1e-4, 1.0)
    fm = rs.ForwardModel(mesh, rs.InferenceParameterSet.a1_betaStar(), obs, bcs, 1.0 / 550.0, s)
    r = fm.evaluate([0.31, 0.09])
    assert str(r.status).split(".")[-1] == "Converged"
    f = fm.last_fields()
    assert f["U"].shape == (512, 3) and np.all(np.isfinite(f["U"]))
    g = _golden(1e-8)
    assert rel(f["U"].ravel(), g["U"]) < 1e-5
    # Cf = nu (U_c - 0) / delta / (0.5 U_ref^2) at the bottom face nearest x = 0.5
    wd = mesh.wall_patch_data("bottom_wall")
    fi = int(np.argmin(np.sum((wd["center"] - [0.5, 0.0, 0.5]) ** 2, axis=1)))
    cf = (1.0 / 550.0) * f["U"][wd["owner"][fi], 0] / wd["delta"][fi] / 0.5
    assert r.predictions[0] == pytest.approx(cf, rel=1e-12)
    assert fm.cache_size() == 1
    # theta within the 0.5 warm-start radius restarts from the cached state.
    # The reference normalises residuals by the restart's own iteration-0
    # residuals (~1e-14 for Uy here), so this restart never meets tol 1e-8:
    # the reference solver compiled from source returns Unconverged at the
    # iteration cap on exactly this case, with the fields still at the fixed point.
    r2 = fm.evaluate([0.32, 0.09])
    assert str(r2.status).split(".")[-1] == "Unconverged"
    assert r2.simple_iters == 32000
    f2 = fm.last_fields()
    assert 0.0 < rel(f2["U"].ravel(), f["U"].ravel()) < 0.05


def test_fields_left_on_device(rs, rsg_lib, gpu_ctx, meshes):
    """RSG_INIT_NO_FIELDS: a solve whose fields stay on the device returns the
    status, iterations, history and device-evaluated predictions of the same
    solve with fields; rsg_fetch with field pointers is refused.  Through
    rans_sst_py (set_collect_fields(False)): the same predictions and
    log-likelihood, no last_fields and no warm-start cache entry."""
    lib, ctx = rsg_lib, gpu_ctx
    s = settings_for("channel_bounded", max_iterations=60, convergence_tol=1e-12, record_history=1)
    P = Problem(lib, ctx, meshes["channel_bounded"], "channel_bounded", s)
    try:
        r0, a0 = P.solve()
        init = A.Init()
        init.kind = 0x100; init.U[0] = 1.0; init.k = 1e-4; init.omega = 1.0
        t = C.c_int64()
        chk(lib, ctx, lib.rsg_submit(ctx, P.pid, D(A.DEFAULT_COEFFS), C.byref(init), 0, C.byref(t)))
        chk(lib, ctx, lib.rsg_wait(ctx, (C.c_int64 * 1)(t.value), 1, C.c_double(600.0)))
        r, a = A.new_result(P.msh.nC, P.msh.nF - P.msh.nIF, s.max_iterations)
        assert lib.rsg_fetch(ctx, t, C.byref(r)) == -1
        bare = A.Result()
        hist = np.zeros(6 * s.max_iterations)
        bare.history = A.dptr(hist); bare.history_cap = s.max_iterations
        chk(lib, ctx, lib.rsg_fetch(ctx, t, C.byref(bare)))
        lib.rsg_release(ctx, t)
        assert (bare.status, bare.final_iter, bare.converged) == (2, r0.final_iter, r0.converged)
        assert np.array_equal(hist[:6 * bare.history_len], a0["history"][:6 * r0.history_len])
        bad = A.Init(); bad.kind = 2
        assert lib.rsg_submit(ctx, P.pid, D(A.DEFAULT_COEFFS), C.byref(bad), 0, C.byref(t)) == -1
    finally:
        P.close()

    y = rs.Mesh.geometric_channel_nodes(32, 1.06, 1.0)
    mesh = rs.Mesh.make_channel_periodic_2d(np.linspace(0.0, 1.0, 5), np.array(y))
    mesh.compute_wall_distance()
    st = rs.SolverSettings()
    st.max_iterations = 30000; st.convergence_tol = 1e-6; st.verbose = False
    st.body_force = rs.Vec3(1.0, 0.0, 0.0)
    obs = rs.ObservationOperator()
    obs.add_skin_friction("bottom_wall", rs.Vec3(0.5, 0.0, 0.5), 0.005, 0.001, 1.0)
    obs.add_velocity_profile(rs.Vec3(0.5, 1.0, 0.5), 0, 20.0, 1.0)
    bcs = rs.FlowBoundaryConditions.channel_defaults(mesh, 1.0, 1e-4, 1.0)
    ps = rs.InferenceParameterSet.a1_betaStar()
    with_f = rs.ForwardModel(mesh, ps, obs, bcs, 1.0 / 180.0, st)
    without = rs.ForwardModel(mesh, ps, obs, bcs, 1.0 / 180.0, st)
    without.set_collect_fields(False)
    assert with_f.collect_fields() and not without.collect_fields()
    ra, rb = with_f.evaluate([0.31, 0.09]), without.evaluate([0.31, 0.09])
    print("with fields", ra.status, ra.simple_iters, list(ra.predictions), flush=True)
    assert (str(ra.status), ra.simple_iters) == (str(rb.status), rb.simple_iters)
    assert list(ra.predictions) == list(rb.predictions) and ra.log_lik == rb.log_lik
    assert str(ra.status).split(".")[-1] == "Converged"
    assert with_f.has_last_fields() and not without.has_last_fields()
    assert (with_f.cache_size(), without.cache_size()) == (1, 0)


@pytest.mark.parametrize("path", ["member", "grid"])
def test_gauss_seidel_trajectory_vs_oracle(rsg_lib, gpu_ctx, oracle, meshes, path):
    """Gauss-Seidel (LinearSolver.hpp:258-299) as the momentum and turbulence
    solver, on the one-workgroup member path and on the kernel-per-phase
    path: the trajectory tracks the oracle's (the sweeps are the reference's
    sequential recurrence; only the residual norms are reduced in another
    order)."""
    case, iters = "channel_bounded", 120
    s = settings_for(case, max_iterations=iters, convergence_tol=1e-12, pressure_solver=1, momentum_solver=3,
                     turbulence_solver=3, nut_floor_iters=60)
    if path == "grid":
        chk(rsg_lib, gpu_ctx, rsg_lib.rsg_set_grid(gpu_ctx, 1, 40))
    P = Problem(rsg_lib, gpu_ctx, meshes[case], case, s)
    try:
        print("gauss-seidel trajectory", path, flush=True)
        r, a = P.solve()
    finally:
        P.close()
        rsg_lib.rsg_set_grid(gpu_ctx, 0, 20)
    o, _, io = oracle_solve(oracle, meshes[case], case, s)
    assert r.final_iter == io[0]
    for k_gpu, k_or in (("U", "U"), ("k", "k"), ("omega", "om"), ("nuT", "nuT")):
        assert rel(a[k_gpu], o[k_or]) < 1e-4, (path, k_gpu, rel(a[k_gpu], o[k_or]))


def test_gauss_seidel_standalone_solve(rsg_lib, gpu_ctx, oracle, meshes):
    """rsg_ldu_solve with RSG_SOLVER_GAUSS_SEIDEL against a dense solve and
    the oracle's sweep count."""
    msh = A.Msh(meshes["hill"])
    n, nif = msh.nC, msh.nIF
    rng = np.random.default_rng(3)
    own = np.ascontiguousarray(msh.owner[:nif], np.int32); nbr = np.ascontiguousarray(msh.neighbor, np.int32)
    cf = np.exp(rng.normal(size=nif))
    up, lo = -cf.copy(), -cf.copy()
    dg = np.zeros(n); np.add.at(dg, own, cf); np.add.at(dg, nbr, cf); dg *= 1.5
    b = rng.normal(size=n)
    x = np.zeros(n); st = np.zeros(4)
    chk(rsg_lib, gpu_ctx, rsg_lib.rsg_ldu_solve(gpu_ctx, n, nif, A.iptr(own), A.iptr(nbr), A.dptr(dg), A.dptr(up),
                                                A.dptr(lo), A.dptr(b), A.dptr(x), 3, 0, 1000, C.c_double(1e-10),
                                                A.dptr(st)))
    Ad = np.diag(dg); Ad[own, nbr] += up; Ad[nbr, own] += lo
    assert st[3] == 1 and np.linalg.norm(b - Ad @ x) / np.linalg.norm(b) < 1e-9
    xo = np.zeros(n); so = np.zeros(4)
    oracle.om_ldu_solve(3, n, nif, A.iptr(own), A.iptr(nbr), A.dptr(dg), A.dptr(up), A.dptr(lo), A.dptr(b),
                        A.dptr(xo), 1000, C.c_double(1e-10), A.dptr(so))
    assert st[0] == so[0]
    assert rel(x, xo) < 1e-12
