"""The product member solver (csrc/rsg_member.h) compiled sequentially
(tests/emu) against the oracle.  With a Jacobi-PCG pressure solve every
algorithm on the path is the reference's, so the whole SIMPLE trajectory must
agree bit for bit; with the device AMG (damped-Jacobi V-cycle, exact dense
coarsest solve) the trajectory differs only through the pressure
preconditioner."""
import ctypes as C

import numpy as np
import pytest

import rsg_abi as A
from conftest import CASES, oracle_solve, settings_for

D = lambda a: (C.c_double * len(a))(*a)


def emu_solve(emu, path, case, s):
    kind, vals, nu, _ = CASES[case]
    msh = A.Msh(path)
    r, arr = A.new_result(msh.nC, msh.nF - msh.nIF, s.max_iterations if s.record_history else 0)
    rc = emu.emu_solve(C.byref(msh.desc()), C.byref(msh.bc(kind, vals)), C.c_double(nu), C.byref(s),
                       D(A.DEFAULT_COEFFS), D([1, 0, 0, 0, 1e-4, 1.0]), 0, C.byref(r))
    assert rc == 0
    return r, arr


@pytest.mark.parametrize("resident", [1, 0])
@pytest.mark.parametrize("case,iters", [("c0", 700), ("channel_bounded", 150), ("couette", 150),
                                        ("bfs", 80), ("plate", 80), ("hill", 120), ("tri_quad", 80)])
def test_emulated_device_trajectory_bit_exact(oracle, emu, meshes, case, iters, resident):
    """Both linear-solver paths of the device code: the register/LDS-resident
    solvers and the global-memory ones."""
    s = settings_for(case, max_iterations=iters, convergence_tol=1e-12, pressure_solver=1,
                     nut_floor_iters=min(500, iters // 2), record_history=1)
    emu.emu_set_resident(resident)
    try:
        r, a = emu_solve(emu, meshes[case], case, s)
    finally:
        emu.emu_set_resident(1)
    o, ho, io = oracle_solve(oracle, meshes[case], case, s, hist_cap=iters)
    assert r.final_iter == io[0] and r.converged == io[1] and r.diverged == io[2]
    he = a["history"][:6 * r.history_len].reshape(-1, 6)
    assert np.array_equal(he, ho)
    for k_emu, k_or in (("U", "U"), ("p", "p"), ("k", "k"), ("omega", "om"), ("nuT", "nuT"),
                        ("F1", "F1"), ("F2", "F2"), ("Pk", "Pk"), ("CDkw", "CDkw"),
                        ("U_bnd", "Ub"), ("p_bnd", "pb"), ("k_bnd", "kb"), ("om_bnd", "omb")):
        assert np.array_equal(a[k_emu], o[k_or]), k_emu


@pytest.mark.parametrize("case,iters", [("channel_bounded", 120), ("c0", 300), ("tri_quad", 60)])
def test_emulated_gauss_seidel_trajectory_bit_exact(oracle, emu, meshes, case, iters):
    """Gauss-Seidel (LinearSolver.hpp:258-299) as the momentum and turbulence
    solver, Jacobi-PCG pressure: the whole trajectory equals the oracle's."""
    s = settings_for(case, max_iterations=iters, convergence_tol=1e-12, pressure_solver=1, momentum_solver=3,
                     turbulence_solver=3, nut_floor_iters=min(500, iters // 2), record_history=1)
    r, a = emu_solve(emu, meshes[case], case, s)
    o, ho, io = oracle_solve(oracle, meshes[case], case, s, hist_cap=iters)
    assert r.final_iter == io[0] and r.converged == io[1] and r.diverged == io[2]
    he = a["history"][:6 * r.history_len].reshape(-1, 6)
    assert np.array_equal(he, ho)
    for k_emu, k_or in (("U", "U"), ("p", "p"), ("k", "k"), ("omega", "om"), ("nuT", "nuT")):
        assert np.array_equal(a[k_emu], o[k_or]), k_emu


@pytest.mark.parametrize("resident", [1, 0])
def test_emulated_amg_tracks_oracle(oracle, emu, meshes, resident):
    # 700 iterations of C0 with the device AMG vs the oracle's reference AMG:
    # same discretisation, different p' preconditioner
    s = settings_for("c0", max_iterations=700, convergence_tol=1e-12, pressure_solver=0, amg_refresh=1)
    emu.emu_set_resident(resident)
    try:
        r, a = emu_solve(emu, meshes["c0"], "c0", s)
    finally:
        emu.emu_set_resident(1)
    o, _, _ = oracle_solve(oracle, meshes["c0"], "c0", s)
    for k_emu, k_or in (("U", "U"), ("k", "k"), ("omega", "om"), ("nuT", "nuT")):
        rel = np.linalg.norm(a[k_emu] - o[k_or]) / np.linalg.norm(o[k_or])
        assert rel < 1e-4, (k_emu, rel)
    assert r.inner_iters[3] < 0.5 * 700 * 30   # the V-cycle keeps p' solves short


def test_emulated_warm_restart_is_a_continuation(oracle, emu, meshes):
    # a solve split at an arbitrary outer iteration equals the unsplit solve
    # (state, normalisation and carried norms all persist across launches)
    msh = A.Msh(meshes["c0"])
    s = settings_for("c0", max_iterations=300, convergence_tol=1e-12, pressure_solver=1)
    kind, vals, nu, _ = CASES["c0"]
    r1, a1 = A.new_result(msh.nC, msh.nF - msh.nIF)
    emu.emu_solve(C.byref(msh.desc()), C.byref(msh.bc(kind, vals)), C.c_double(nu), C.byref(s),
                  D(A.DEFAULT_COEFFS), D([1, 0, 0, 0, 1e-4, 1.0]), 37, C.byref(r1))
    r2, a2 = A.new_result(msh.nC, msh.nF - msh.nIF)
    emu.emu_solve(C.byref(msh.desc()), C.byref(msh.bc(kind, vals)), C.c_double(nu), C.byref(s),
                  D(A.DEFAULT_COEFFS), D([1, 0, 0, 0, 1e-4, 1.0]), 300, C.byref(r2))
    assert np.array_equal(a1["U"], a2["U"]) and np.array_equal(a1["k"], a2["k"])


@pytest.mark.parametrize("solver", [0, 1, 2, 3])
def test_emulated_standalone_linear_solve(oracle, emu, meshes, solver):
    """The stand-alone solve path (topology-only mesh with no patches, as
    rsg_ldu_solve builds it): Jacobi-PCG, BiCGSTAB and Gauss-Seidel equal the
    oracle bit for bit; AMG-PCG solves to the requested tolerance."""
    msh = A.Msh(meshes["hill"])
    n, nif = msh.nC, msh.nIF
    rng = np.random.default_rng(solver)
    own = msh.owner[:nif].copy(); nbr = msh.neighbor.copy()
    cf = np.exp(rng.normal(size=nif))
    up = -cf.copy(); lo = -(cf + rng.uniform(0, 0.5, nif)) if solver == 2 else -cf.copy()
    dg = np.zeros(n); np.add.at(dg, own, -lo); np.add.at(dg, nbr, -up); dg[0] *= 2.0
    if solver == 3:       # Gauss-Seidel converges slowly on the pinned Laplacian: a dominant diagonal
        dg = dg * 1.5
    b = rng.normal(size=n)
    x = np.zeros(n); st = np.zeros(5)
    assert emu.emu_ldu_solve(n, nif, A.iptr(own), A.iptr(nbr), A.dptr(dg), A.dptr(up), A.dptr(lo), A.dptr(b),
                             A.dptr(x), solver, 1000, C.c_double(1e-10), A.dptr(st)) == 0
    assert st[3] == 1
    Ad = np.diag(dg); Ad[own, nbr] += up; Ad[nbr, own] += lo
    assert np.linalg.norm(b - Ad @ x) / np.linalg.norm(b) < 1e-9
    if solver != 0:
        xo = np.zeros(n); so = np.zeros(4)
        oracle.om_ldu_solve(solver, n, nif, A.iptr(own), A.iptr(nbr), A.dptr(dg), A.dptr(up), A.dptr(lo),
                            A.dptr(b), A.dptr(xo), 1000, C.c_double(1e-10), A.dptr(so))
        assert np.array_equal(x, xo) and np.array_equal(st[:4], so)
    else:
        assert st[4] >= 2


@pytest.mark.parametrize("case", ["c0", "bfs"])
def test_emulated_warm_start_bit_exact(oracle, emu, meshes, case):
    """A warm start (cached state + its turbEstablished flag, other
    coefficients) on the device algorithm equals the oracle's bit for bit."""
    from conftest import oracle_solve_warm, state_ptrs
    s0 = settings_for(case, max_iterations=120, convergence_tol=1e-12, pressure_solver=1, nut_floor_iters=40)
    base, _, _ = oracle_solve(oracle, meshes[case], case, s0)
    coeffs = list(A.DEFAULT_COEFFS)
    coeffs[9] *= 1.05; coeffs[8] *= 0.97
    s = settings_for(case, max_iterations=90, convergence_tol=1e-12, pressure_solver=1, nut_floor_iters=40,
                     record_history=1)
    kind, vals, nu, _ = CASES[case]
    msh = A.Msh(meshes[case])
    for turb in (1, 0):
        o, ho, io = oracle_solve_warm(oracle, meshes[case], case, s, base, turb, coeffs, hist_cap=90)
        r, a = A.new_result(msh.nC, msh.nF - msh.nIF, 90)
        keep, ptrs = state_ptrs(base)
        assert emu.emu_solve_warm(C.byref(msh.desc()), C.byref(msh.bc(kind, vals)), C.c_double(nu), C.byref(s),
                                  D(coeffs), ptrs, turb, C.byref(r)) == 0
        assert r.final_iter == io[0] and r.turb_established == io[4]
        assert np.array_equal(a["history"][:6 * r.history_len].reshape(-1, 6), ho)
        for k_emu, k_or in (("U", "U"), ("p", "p"), ("k", "k"), ("omega", "om"), ("nuT", "nuT"),
                            ("U_bnd", "Ub"), ("p_bnd", "pb"), ("k_bnd", "kb"), ("om_bnd", "omb")):
            assert np.array_equal(a[k_emu], o[k_or]), (turb, k_emu)


def test_emulated_vcycle_level_images_are_exact(emu, meshes):
    """The resident V-cycle's coarse levels through their ELL images round
    exactly like the CSR levels (same entries, same order): whole AMG
    trajectories agree bit for bit."""
    s = settings_for("c0", max_iterations=120, convergence_tol=1e-12, pressure_solver=0, amg_refresh=7,
                     coarse_refresh=3)
    outs = []
    for images, stage in ((1, 1), (1, 0), (0, 0)):
        emu.emu_set_coarse_images(images)
        emu.emu_set_stage(stage)
        try:
            outs.append(emu_solve(emu, meshes["c0"], "c0", s)[1])
        finally:
            emu.emu_set_coarse_images(1)
            emu.emu_set_stage(1)
    for o in outs[1:]:
        for k in ("U", "p", "k", "omega", "nuT"):
            assert np.array_equal(outs[0][k], o[k]), k


@pytest.mark.parametrize("ny,staging", [(384, "full"), (512, "partial")])
def test_emulated_staging_is_exact(emu, rs, tmp_path, ny, staging):
    """LDS copies of the coarse images (16-bit indices): a 4 x 384 periodic
    channel (the largest ensemble member) has every coarse level staged; a
    4 x 512 one exceeds the staging capacity, so only the deepest levels are
    copied.  Either way the trajectory equals the unstaged one bit for bit."""
    import ensembles
    r = ensembles.stretch_ratio(ny // 2, 1.0 / 2000.0)
    y = rs.Mesh.geometric_channel_nodes(ny // 2, r, 1.0)
    m = rs.Mesh.make_channel_periodic_2d(np.linspace(0.0, 1.0, 5), np.array(y))
    path = str(tmp_path / "c2max.msh")
    m.save_binary(path)
    s = settings_for("c0", max_iterations=25, convergence_tol=1e-12, pressure_solver=0, amg_refresh=1000,
                     coarse_refresh=3)
    outs = []
    for stage in (1, 0):
        emu.emu_set_stage(stage)
        try:
            outs.append(emu_solve(emu, path, "c0", s)[1])
        finally:
            emu.emu_set_stage(1)
    for k in ("U", "p", "k", "omega", "nuT"):
        assert np.array_equal(outs[0][k], outs[1][k]), k
    # run order (stage, then no stage): the observed value is the staged run's
    first = emu.emu_last_staged_first()
    if staging == "full":
        assert first == 1
    else:
        assert 1 < first <= 20
