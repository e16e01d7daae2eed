"""Pins the oracle (oracle/rsg_oracle.c) with what does not need the
reference binary: the known answers and closed forms of the reference's own
C++ tests, brute-force dense solves, exactness of Green-Gauss on linear fields
and hand-computable aggregations.  Each check is chosen so that a dropped
term, a wrong sign or a transposed operand in the restatement fails it."""
import ctypes as C

import numpy as np
import pytest

import rsg_abi as A

COEFFS = A.DEFAULT_COEFFS


def sst(oracle, which, *args, variant=0, coeffs=COEFFS):
    D = lambda a: (C.c_double * len(a))(*a)
    return oracle.om_sst_point(which, D(coeffs), variant, D(list(args)))


def close(a, b, rtol):
    return abs(a - b) <= rtol * max(abs(a), abs(b), 1e-300)


def test_gauss_seidel_known_answer(oracle):
    # tests/cpp/test_infra_fixes.cpp:45-63: one in-place sweep on a 3-cell chain
    own = np.array([0, 1], np.int32); nbr = np.array([1, 2], np.int32)
    dg = np.array([2.0, 2.0, 2.0]); up = np.array([-1.0, -1.0]); lo = np.array([-1.0, -1.0])
    b = np.array([1.0, 0.0, 1.0]); x = np.zeros(3); out = np.zeros(4)
    oracle.om_ldu_solve(3, 3, 2, A.iptr(own), A.iptr(nbr), A.dptr(dg), A.dptr(up), A.dptr(lo), A.dptr(b),
                        A.dptr(x), 1, C.c_double(1e-30), A.dptr(out))
    assert np.array_equal(x, [0.5, 0.25, 0.625])


def test_production_omega_closed_forms(oracle):
    # tests/cpp/test_sst_production_2003.cpp:51-145
    bs = COEFFS[8]
    nuT, S, k, w = 0.01, 1.0, 1.0, 100.0
    assert nuT * S * S < 10 * bs * k * w
    assert close(sst(oracle, 4, nuT, S, k, w), S * S, 1e-14)              # limiter inactive
    nuT, S, k, w = 1.0, 100.0, 1.0, 1.0
    assert close(sst(oracle, 4, nuT, S, k, w), 0.9, 1e-14)                # 10 b* k w / nuT
    for nuT in (1e-12, 1e-6, 1e-2, 0.5, 3.0):                              # bounded by S^2
        for S in (0.0, 0.3, 3.0, 300.0):
            for k in (1e-8, 1e-2, 1.0, 50.0):
                for w in (1e-3, 1.0, 1e4):
                    assert sst(oracle, 4, nuT, S, k, w) <= S * S + 1e-14
    assert close(sst(oracle, 4, 0.0, 2.0, 0.5, 200.0), 4.0, 1e-14)        # nuT -> 0 selects S^2
    for st in ((0.01, 1.0, 1.0, 100.0), (1.0, 100.0, 1.0, 1.0), (0.2, 5.0, 2.0, 30.0)):
        assert close(sst(oracle, 4, *st), sst(oracle, 3, *st) / st[0], 1e-12)
    rho, muT, S, k, w = 1.7, 0.4, 40.0, 3.0, 8.0
    assert close(min(S * S, 10 * bs * rho * k * w / muT), sst(oracle, 4, muT / rho, S, k, w), 1e-12)


def test_blending_and_eddy_viscosity_closed_forms(oracle):
    # SSTModel.cpp:9-65: F1 = tanh(arg1^4), F2 = tanh(arg2^2), Bradshaw limiter
    k, w, y, nu = 0.01, 50.0, 0.02, 1e-5
    bs, sw2, a1 = COEFFS[8], COEFFS[5], COEFFS[9]
    cd = 1e-3
    t1 = np.sqrt(k) / (bs * w * y); t2 = 500 * nu / (y * y * w); t3 = 4 * sw2 * k / (cd * y * y)
    assert close(sst(oracle, 0, k, w, y, nu, cd), np.tanh(min(max(t1, t2), t3) ** 4), 1e-13)
    t1b = 2 * np.sqrt(k) / (bs * w * y)
    assert close(sst(oracle, 1, k, w, y, nu), np.tanh(max(t1b, t2) ** 2), 1e-13)
    S, F2 = 500.0, 0.7
    assert close(sst(oracle, 2, k, w, S, F2), a1 * k / max(a1 * w, S * F2), 1e-14)      # limiter active
    assert close(sst(oracle, 2, k, w, 1e-3, F2), k / w, 1e-14)                           # a1 w branch
    assert close(sst(oracle, 2, k, w, S, F2, variant=1), k / w, 1e-14)                   # NoLimiter


def _mesh(oracle, rs, tmp_path, mesh):
    path = str(tmp_path / "m.msh")
    mesh.save_binary(path)
    return oracle.om_mesh_load(path.encode()), path


def test_green_gauss_exact_on_linear_fields(oracle, rs, tmp_path):
    mesh = rs.Mesh.make_channel_2d(12, 14, 2.0, 1.0)
    m, _ = _mesh(oracle, rs, tmp_path, mesh)
    cc = mesh.cell_centers(); fc = mesh.face_centers()
    nif = mesh.n_internal_faces()
    a, b = 0.7, -1.3
    phi = a * cc[:, 0] + b * cc[:, 1] + 0.4
    phib = a * fc[nif:, 0] + b * fc[nif:, 1] + 0.4
    g = np.zeros(3 * mesh.n_cells())
    oracle.om_green_gauss(C.c_void_p(m), A.dptr(phi), A.dptr(phib), A.dptr(g))
    g = g.reshape(-1, 3)
    # face interpolation is exact for linear fields only where the face sits
    # midway between the two centres (uniform x spacing): check x everywhere
    # and y on cells whose y-neighbours are equally spaced
    assert np.allclose(g[:, 0], a, atol=1e-11)
    i = np.arange(mesh.n_cells()) // 12
    mid = (i == 7)
    assert np.allclose(g[mid, 1], b, atol=0.05)


def test_transpose_stress_manufactured(oracle, rs, tmp_path):
    # tests/cpp/test_transpose_stress.cpp:70-136
    nx, ny = 12, 14
    mesh = rs.Mesh.make_channel_2d(nx, ny, 2.0, 1.0)
    m, _ = _mesh(oracle, rs, tmp_path, mesh)
    cc = mesh.cell_centers(); fc = mesh.face_centers(); nif = mesh.n_internal_faces()
    vol = mesh.cell_volumes(); n = mesh.n_cells()
    ci = np.arange(n); ii = ci % nx; jj = ci // nx
    interior = (ii >= 2) & (ii < nx - 2) & (jj >= 2) & (jj < ny - 2)

    def run(ufn, cfn, comp, wall_zero):
        U = np.array([ufn(x, y) for x, y in cc[:, :2]]).ravel()
        Ub = np.array([ufn(x, y) for x, y in fc[nif:, :2]]).ravel()
        coeff = np.array([cfn(x, y) for x, y in cc[:, :2]])
        src = np.zeros(n)
        oracle.om_transpose_stress(C.c_void_p(m), A.dptr(coeff), A.dptr(U), A.dptr(Ub), comp, wall_zero, A.dptr(src))
        return src / vol

    a, c0 = 0.8, 0.35
    s = run(lambda x, y: (a * x * x, 0, 0), lambda x, y: c0, 0, 0)
    assert np.all(np.abs(s[interior] - 2 * a * c0) < 1e-11)
    b_, g_, d_ = 0.6, 1.3, 0.4
    sx = run(lambda x, y: (g_ * y, d_ * x, 0), lambda x, y: b_ * y, 0, 1)
    sy = run(lambda x, y: (g_ * y, d_ * x, 0), lambda x, y: b_ * y, 1, 1)
    wall_row = (jj == 0) & (ii >= 2) & (ii < nx - 2)
    assert np.all(np.abs(sx[wall_row] - b_ * d_) < 1e-11)
    assert np.all(np.abs(sx[interior] - b_ * d_) < 1e-11)
    assert np.all(np.abs(sx[interior] - b_ * g_) > 0.1 * abs(b_ * (g_ - d_)))
    assert np.all(np.abs(sy[interior]) < 1e-11)
    for comp in (0, 1):
        s = run(lambda x, y: (np.tanh(4 * (y - 0.5)), 0, 0), lambda x, y: 0.02 + 0.05 * y * (1 - y), comp, 0)
        assert np.all(np.abs(s * vol) < 1e-14)


@pytest.mark.parametrize("solver", [0, 1, 2, 3])
def test_linear_solvers_against_dense(oracle, rs, tmp_path, solver):
    mesh = rs.Mesh.make_channel_2d(9, 7, 1.0, 1.0)
    path = str(tmp_path / "l.msh"); mesh.save_binary(path)
    msh = A.Msh(path)
    rng = np.random.default_rng(solver)
    nif = msh.nIF; n = msh.nC
    own = msh.owner[:nif].copy(); nbr = msh.neighbor.copy()
    cf = np.exp(rng.normal(size=nif))
    up = -cf.copy(); lo = -cf.copy()
    if solver == 2:   # BiCGSTAB on a non-symmetric convection-diffusion-like matrix
        lo = -(cf + rng.uniform(0, 0.5, nif))
    dg = np.zeros(n); np.add.at(dg, own, -lo); np.add.at(dg, nbr, -up); dg += 0.05
    b = rng.normal(size=n)
    Ad = np.diag(dg); Ad[own, nbr] += up; Ad[nbr, own] += lo
    xs = np.linalg.solve(Ad, b)
    x = np.zeros(n); out = np.zeros(4)
    oracle.om_ldu_solve(solver, n, nif, A.iptr(own), A.iptr(nbr), A.dptr(dg), A.dptr(up), A.dptr(lo),
                        A.dptr(b), A.dptr(x), 5000, C.c_double(1e-12), A.dptr(out))
    assert out[3] == 1
    assert out[2] == pytest.approx(np.linalg.norm(b), rel=1e-14)     # initialRes = |b - A 0|
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-8
    assert np.linalg.norm(b - Ad @ x) / np.linalg.norm(b) < 1e-10


def test_aggregation_hand_cases(oracle):
    # 1D chain, diag 2, off-diagonal -1: strength 0.5 > 0.25 pairs neighbours
    n = 7
    rp, ci, v, dg = [0], [], [], np.full(n, 2.0)
    for i in range(n):
        for j in (i - 1, i, i + 1):
            if 0 <= j < n:
                ci.append(j); v.append(2.0 if j == i else -1.0)
        rp.append(len(ci))
    m = np.zeros(n, np.int32)
    nc = oracle.om_build_aggregates(n, A.iptr(np.array(rp, np.int32)), A.iptr(np.array(ci, np.int32)),
                                    A.dptr(np.array(v)), A.dptr(dg), A.iptr(m))
    assert nc == 4 and list(m) == [0, 0, 1, 1, 2, 2, 3]
    # weak couplings (strength 0.1) are never merged in pass 1; every cell seeds
    v2 = [(-0.2 if c != i else 2.0) for i in range(n) for c in ci[rp[i]:rp[i + 1]]]
    nc = oracle.om_build_aggregates(n, A.iptr(np.array(rp, np.int32)), A.iptr(np.array(ci, np.int32)),
                                    A.dptr(np.array(v2)), A.dptr(dg), A.iptr(m))
    assert nc == n and list(m) == list(range(n))
