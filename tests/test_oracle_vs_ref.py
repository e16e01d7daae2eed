"""The oracle against the reference solver compiled from its own sources
(oracle/_ref/libref.so): bit-for-bit equality of whole SIMPLE trajectories
on every parity case, of each linear solver, of the AMG hierarchy, of the
Green-Gauss gradient, of the SST closure and of the operator assembly."""
import ctypes as C

import numpy as np
import pytest

import rsg_abi as A
from conftest import CASES, FIELD_KEYS, oracle_solve, settings_for

D = lambda a: (C.c_double * len(a))(*a)


def ref_solve(ref, path, case, settings, hist_cap):
    kind, vals, nu, _ = CASES[case]
    msh = A.Msh(path)
    n, nB = msh.nC, msh.nF - msh.nIF
    out = {k: np.zeros(3 * n if k == "U" else n) for k in FIELD_KEYS}
    bnd = {"Ub": np.zeros(3 * nB), "pb": np.zeros(nB), "kb": np.zeros(nB), "omb": np.zeros(nB)}
    hist = np.zeros(6 * max(hist_cap, 1))
    info = (C.c_int * 5)()
    rc = ref.ref_simple_solve(path.encode(), kind, D(list(vals)), C.c_double(nu), D(A.DEFAULT_COEFFS),
                              D(A.packed_settings(settings)), D([1, 0, 0, 0, 1e-4, 1.0]),
                              *[A.dptr(out[k]) for k in FIELD_KEYS],
                              *[A.dptr(bnd[k]) for k in ("Ub", "pb", "kb", "omb")], A.dptr(hist), hist_cap, info)
    assert rc == 0
    out.update(bnd)
    return out, hist[:6 * info[3]].reshape(-1, 6), list(info)


@pytest.mark.parametrize("case,iters", [("c0", 600), ("channel_bounded", 150), ("couette", 150),
                                        ("bfs", 80), ("plate", 80), ("hill", 120), ("tri_quad", 80)])
@pytest.mark.parametrize("psolver", [0, 1])
def test_trajectory_bit_exact(oracle, ref, meshes, case, iters, psolver):
    s = settings_for(case, max_iterations=iters, convergence_tol=1e-12, pressure_solver=psolver,
                     nut_floor_iters=min(500, iters // 2))
    o, ho, io = oracle_solve(oracle, meshes[case], case, s, hist_cap=iters)
    r, hr, ir = ref_solve(ref, meshes[case], case, s, iters)
    assert io == ir
    assert np.array_equal(ho, hr)
    for k in list(FIELD_KEYS) + ["Ub", "pb", "kb", "omb"]:
        assert np.array_equal(o[k], r[k]), k


def test_variants_and_wall_functions_bit_exact(oracle, ref, meshes):
    for variant, wallfn in ((1, 0), (2, 0), (0, 1)):
        s = settings_for("channel_bounded", max_iterations=60, sst_variant=variant, use_wall_functions=wallfn,
                         pressure_solver=1, nut_floor_iters=20)
        o, ho, io = oracle_solve(oracle, meshes["channel_bounded"], "channel_bounded", s, hist_cap=60)
        r, hr, ir = ref_solve(ref, meshes["channel_bounded"], "channel_bounded", s, 60)
        assert np.array_equal(ho, hr) and np.array_equal(o["U"], r["U"]) and np.array_equal(o["om"], r["om"])


def _system(path, seed, nonsym):
    msh = A.Msh(path)
    rng = np.random.default_rng(seed)
    nif, n = msh.nIF, msh.nC
    own = msh.owner[:nif].copy(); nbr = msh.neighbor.copy()
    cf = np.exp(rng.normal(size=nif))
    up = -cf.copy(); lo = -(cf + rng.uniform(0, 0.5, nif)) if nonsym else -cf.copy()
    dg = np.zeros(n); np.add.at(dg, own, -lo); np.add.at(dg, nbr, -up); dg[0] *= 2.0
    b = rng.normal(size=n)
    return n, nif, own, nbr, dg, up, lo, b


@pytest.mark.parametrize("solver", [0, 1, 2, 3])
def test_linear_solvers_bit_exact(oracle, ref, meshes, solver):
    n, nif, own, nbr, dg, up, lo, b = _system(meshes["c0"], 7, solver in (2, 3))
    xo = np.zeros(n); xr = np.zeros(n); so = np.zeros(4); sr = np.zeros(4)
    args = (n, nif, A.iptr(own), A.iptr(nbr), A.dptr(dg), A.dptr(up), A.dptr(lo), A.dptr(b))
    oracle.om_ldu_solve(solver, *args, A.dptr(xo), 300, C.c_double(1e-9), A.dptr(so))
    ref.ref_ldu_solve(solver, *args, A.dptr(xr), 300, C.c_double(1e-9), A.dptr(sr))
    assert np.array_equal(so, sr) and np.array_equal(xo, xr)


def test_amg_hierarchy_bit_exact(oracle, ref, meshes):
    for case in ("c0", "bfs", "hill"):
        n, nif, own, nbr, dg, up, lo, _ = _system(meshes[case], 3, False)
        so, sr = np.zeros(21, np.int32), np.zeros(21, np.int32)
        zo, zr = np.zeros(21, np.int32), np.zeros(21, np.int32)
        mo, mr = np.zeros(4 * n, np.int32), np.zeros(4 * n, np.int32)
        args = (n, nif, A.iptr(own), A.iptr(nbr), A.dptr(dg), A.dptr(up), A.dptr(lo))
        lo_ = oracle.om_amg_hierarchy(*args, A.iptr(so), A.iptr(zo), 21, A.iptr(mo), 4 * n)
        lr_ = ref.ref_amg_hierarchy(*args, A.iptr(sr), A.iptr(zr), 21, A.iptr(mr), 4 * n)
        assert lo_ == lr_ and lo_ > 1
        assert np.array_equal(so, sr) and np.array_equal(zo, zr) and np.array_equal(mo, mr)


def test_green_gauss_wall_distance_sst_bit_exact(oracle, ref, meshes):
    rng = np.random.default_rng(1)
    for case in ("c0", "bfs", "hill"):
        path = meshes[case]
        msh = A.Msh(path)
        m = oracle.om_mesh_load(path.encode())
        phi = rng.normal(size=msh.nC); phib = rng.normal(size=msh.nF - msh.nIF)
        go, gr = np.zeros(3 * msh.nC), np.zeros(3 * msh.nC)
        oracle.om_green_gauss(C.c_void_p(m), A.dptr(phi), A.dptr(phib), A.dptr(go))
        ref.ref_green_gauss(path.encode(), A.dptr(phi), A.dptr(phib), A.dptr(gr))
        assert np.array_equal(go, gr)
        wo, wr = np.zeros(msh.nC), np.zeros(msh.nC)
        oracle.om_wall_distance(C.c_void_p(m), A.dptr(wo))
        ref.ref_wall_distance(path.encode(), A.dptr(wr))
        assert np.array_equal(wo, wr) and np.array_equal(wo, msh.wd)
        oracle.om_mesh_free(C.c_void_p(m))
    for which, args in ((0, (0.01, 50.0, 0.02, 1e-5, 1e-3)), (1, (0.3, 2.0, 0.1, 1e-4)),
                        (2, (0.1, 3.0, 20.0, 0.8)), (3, (0.2, 5.0, 2.0, 30.0)), (4, (1.0, 100.0, 1.0, 1.0))):
        for variant in (0, 1, 2):
            assert oracle.om_sst_point(which, D(A.DEFAULT_COEFFS), variant, D(list(args))) == \
                ref.ref_sst_point(which, D(A.DEFAULT_COEFFS), variant, D(list(args)))


def _state(path, seed):
    msh = A.Msh(path)
    rng = np.random.default_rng(seed)
    n, nB = msh.nC, msh.nF - msh.nIF
    U = np.column_stack([1 + 0.1 * rng.normal(size=n), 0.05 * rng.normal(size=n), np.zeros(n)]).ravel()
    Ub = np.column_stack([0.1 * rng.normal(size=nB), 0.01 * rng.normal(size=nB), np.zeros(nB)]).ravel()
    p = rng.normal(size=n); pb = rng.normal(size=nB)
    k = np.exp(rng.normal(size=n) - 5); kb = np.exp(rng.normal(size=nB) - 5)
    om = np.exp(rng.normal(size=n) + 3); omb = np.exp(rng.normal(size=nB) + 3)
    return msh, [U, Ub, p, pb, k, kb, om, omb]


@pytest.mark.parametrize("case", ["c0", "channel_bounded", "couette", "bfs", "hill"])
def test_assembly_bit_exact(oracle, ref, meshes, case):
    path = meshes[case]
    msh, st = _state(path, 5)
    kind, vals, nu, _ = CASES[case]
    s = settings_for(case)
    n, nif = msh.nC, msh.nIF
    sz = 2 * n + 2 * nif
    outs_o = [np.zeros(sz), np.zeros(sz), np.zeros(sz), np.zeros(n), np.zeros(5 * n)]
    outs_r = [np.zeros(sz), np.zeros(sz), np.zeros(sz), np.zeros(n), np.zeros(5 * n)]
    m = oracle.om_mesh_load(path.encode())
    common = (kind, D(list(vals)), C.c_double(nu), D(A.DEFAULT_COEFFS), D(A.packed_settings(s)),
              *[A.dptr(a) for a in st])
    oracle.om_precon_blocks(C.c_void_p(m), *common, *[A.dptr(a) for a in outs_o])
    ref.ref_precon_blocks(path.encode(), *common, *[A.dptr(a) for a in outs_r])
    for a, b in zip(outs_o, outs_r):
        assert np.array_equal(a, b)
    Ro, Rr = np.zeros(4 * n), np.zeros(4 * n)
    oracle.om_residual(C.c_void_p(m), *common, A.dptr(Ro))
    ref.ref_residual(path.encode(), *common, A.dptr(Rr))
    assert np.array_equal(Ro, Rr)
    # transpose stress, both boundary treatments
    for comp in (0, 1):
        for wz in (0, 1):
            so, sr = np.zeros(n), np.zeros(n)
            coeff = st[4] * 10
            oracle.om_transpose_stress(C.c_void_p(m), A.dptr(coeff), A.dptr(st[0]), A.dptr(st[1]), comp, wz, A.dptr(so))
            ref.ref_transpose_stress(path.encode(), A.dptr(coeff), A.dptr(st[0]), A.dptr(st[1]), comp, wz, A.dptr(sr))
            assert np.array_equal(so, sr)
    oracle.om_mesh_free(C.c_void_p(m))


@pytest.mark.parametrize("case", ["c0", "bfs"])
def test_warm_start_bit_exact(oracle, ref, meshes, case):
    """ForwardModel's warm start (ForwardModel.cpp:56-60): the solve restarted
    from a cached state with other coefficients, the state's turbEstablished
    flag carried along, equals the reference's bit for bit."""
    from conftest import oracle_solve_warm, state_ptrs
    s0 = settings_for(case, max_iterations=120, convergence_tol=1e-12, pressure_solver=1, nut_floor_iters=40)
    base, _, _ = oracle_solve(oracle, meshes[case], case, s0)
    coeffs = list(A.DEFAULT_COEFFS)
    coeffs[9] *= 1.05; coeffs[8] *= 0.97
    s = settings_for(case, max_iterations=90, convergence_tol=1e-12, pressure_solver=1, nut_floor_iters=40)
    for turb in (1, 0):
        o, ho, io = oracle_solve_warm(oracle, meshes[case], case, s, base, turb, coeffs, hist_cap=90)
        kind, vals, nu, _ = CASES[case]
        msh = A.Msh(meshes[case])
        n, nB = msh.nC, msh.nF - msh.nIF
        r = {k: np.zeros(3 * n if k == "U" else n) for k in FIELD_KEYS}
        rb = {"Ub": np.zeros(3 * nB), "pb": np.zeros(nB), "kb": np.zeros(nB), "omb": np.zeros(nB)}
        hist = np.zeros(6 * 90)
        info = (C.c_int * 5)()
        keep, ptrs = state_ptrs(base)
        assert ref.ref_simple_solve_warm(meshes[case].encode(), kind, D(list(vals)), C.c_double(nu), D(coeffs),
                                         D(A.packed_settings(s)), ptrs, turb,
                                         *[A.dptr(r[k]) for k in FIELD_KEYS],
                                         *[A.dptr(rb[k]) for k in ("Ub", "pb", "kb", "omb")],
                                         A.dptr(hist), 90, info) == 0
        assert io == list(info)
        assert np.array_equal(ho, hist[:6 * info[3]].reshape(-1, 6))
        for k in FIELD_KEYS:
            assert np.array_equal(o[k], r[k]), (turb, k)
        for k in ("Ub", "pb", "kb", "omb"):
            assert np.array_equal(o[k], rb[k]), (turb, k)


@pytest.mark.parametrize("case", ["bfs", "hill", "channel_bounded", "plate"])
def test_observation_operator_bit_exact(oracle, ref, meshes, case):
    """om_observe / om_loglik against ObservationOperator::evaluate and
    logLikelihood of the reference, every observable type."""
    path = meshes[case]
    msh = A.Msh(path)
    rng = np.random.default_rng(5)
    o = A.observation_set(msh, rng)
    U, Ub, p = A.wavy_state(msh, rng)
    n = len(o["type"])
    nu = 1.7e-4
    po, pr, llr = np.zeros(n), np.zeros(n), np.zeros(1)
    m = oracle.om_mesh_load(path.encode())
    try:
        oracle.om_observe(C.c_void_p(m), n, A.iptr(o["type"]), A.iptr(o["patch"]), A.dptr(o["loc"]),
                          A.iptr(o["comp"]), A.dptr(o["refArea"]), A.dptr(o["refVel"]), A.dptr(o["refP"]),
                          A.dptr(o["refRho"]), A.dptr(U), A.dptr(Ub), A.dptr(p), C.c_double(nu), A.dptr(po))
    finally:
        oracle.om_mesh_free(C.c_void_p(m))
    llo = oracle.om_loglik(n, A.dptr(po), A.dptr(o["observed"]), A.dptr(o["sigma"]), A.dptr(o["sigmaModel"]))
    assert ref.ref_observe(path.encode(), n, A.iptr(o["type"]), A.iptr(o["patch"]), A.dptr(o["loc"]),
                           A.iptr(o["comp"]), A.dptr(o["refArea"]), A.dptr(o["refVel"]), A.dptr(o["refP"]),
                           A.dptr(o["refRho"]), A.dptr(o["observed"]), A.dptr(o["sigma"]),
                           A.dptr(o["sigmaModel"]), A.dptr(U), A.dptr(Ub), A.dptr(p), C.c_double(nu),
                           A.dptr(pr), A.dptr(llr)) == 0
    assert np.array_equal(po, pr)
    assert llo == llr[0]
    # the crossings were found (interpolated x inside the patch), not the defaults
    sep = po[o["type"] == 4]; rea = po[o["type"] == 5]
    assert np.any(sep != 0.0) and np.all(np.isfinite(rea))
