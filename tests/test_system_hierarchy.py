"""Host side of the large-system solve (rsg_system_solve): the AMG hierarchy
it builds is the oracle's (AMGSolver::setup, core/include/LinearSolver.hpp:
516-550) level for level -- sizes, nonzeros and every aggregate map -- and
the oracle's is the compiled reference's where that library is built.
Runs on the CPU: rsg_system_hierarchy makes no device call.
"""
import ctypes as C

import numpy as np
import pytest

import c1_system
import rsg_abi as A

SMALL = dict(nx_up=15, nx_dn=45, ny_up=20, ny_low=4, wall_layers=3)


def _args(s):
    return (s["n"], s["nif"], A.iptr(s["own"]), A.iptr(s["nbr"]), A.dptr(s["diag"]), A.dptr(s["upper"]),
            A.dptr(s["lower"]))


def product_hierarchy(lib, s):
    n = s["n"]
    sizes = np.zeros(32, np.int32); nnz = np.zeros(32, np.int32); maps = np.zeros(4 * n, np.int32)
    nl = lib.rsg_system_hierarchy(*_args(s), A.iptr(sizes), A.iptr(nnz), 32, A.iptr(maps), C.c_int64(4 * n))
    assert nl > 0
    return nl, sizes[:nl], nnz[:nl], maps[:int(sizes[:nl - 1].sum())]


def checker_hierarchy(fn, s):
    n = s["n"]
    sizes = np.zeros(32, np.int32); nnz = np.zeros(32, np.int32); maps = np.zeros(4 * n, np.int32)
    nl = fn(*_args(s), A.iptr(sizes), A.iptr(nnz), 32, A.iptr(maps), 4 * n)
    return nl, sizes[:nl], nnz[:nl], maps[:int(sizes[:nl - 1].sum())]


@pytest.fixture(scope="module")
def host_lib():
    import os
    from conftest import PKG
    path = os.path.join(PKG, "librsg.so")
    if not os.path.exists(path):
        raise RuntimeError("librsg.so is not built: run __graft_entry__.build()")
    return C.CDLL(path)


def test_generator_shapes():
    s = c1_system.make_system()
    assert s["n"] == 261600 and s["n_tri"] == 247200 and s["n_quad"] == 14400
    assert s["nif"] == 398760 and s["n_outlet"] == 240
    assert np.all(s["own"] < s["nbr"])
    # SPD M-matrix: positive diagonal, non-positive off-diagonals, row sums >= 0
    rs = s["diag"].copy()
    np.add.at(rs, s["own"], s["upper"]); np.add.at(rs, s["nbr"], s["lower"])
    assert np.all(s["upper"] < 0) and np.all(rs >= -1e-9 * s["diag"]) and np.sum(rs > 1e-12) == 240
    assert abs(s["b"].sum()) < 1e-9


@pytest.mark.parametrize("size", ["small", "full"])
def test_hierarchy_matches_oracle(host_lib, oracle, size):
    s = c1_system.make_system(**(SMALL if size == "small" else {}))
    nl, sz, nz, mp = product_hierarchy(host_lib, s)
    nlo, szo, nzo, mpo = checker_hierarchy(oracle.om_amg_hierarchy, s)
    assert nl == nlo
    np.testing.assert_array_equal(sz, szo)
    np.testing.assert_array_equal(nz, nzo)
    np.testing.assert_array_equal(mp, mpo)
    # coarsening stops at <= 50 rows, after 20 coarsenings, or where the
    # greedy aggregation stalls (the full system's weakly coupled rows leave a
    # coarsest level of ~8k rows, in the reference as here)
    assert np.all(np.diff(sz) < 0) and nl <= 21
    if size == "full":
        assert nl == 17 and sz[-1] == 7779


def test_hierarchy_matches_reference(host_lib, ref):
    s = c1_system.make_system(**SMALL)
    nl, sz, nz, mp = product_hierarchy(host_lib, s)
    nlr, szr, nzr, mpr = checker_hierarchy(ref.ref_amg_hierarchy, s)
    assert nl == nlr
    np.testing.assert_array_equal(sz, szr)
    np.testing.assert_array_equal(nz, nzr)
    np.testing.assert_array_equal(mp, mpr)


def test_hierarchy_rejects_bad_faces(host_lib):
    s = c1_system.make_system(**SMALL)
    bad = dict(s); bad["nbr"] = s["nbr"].copy(); bad["nbr"][3] = s["n"]
    sizes = np.zeros(32, np.int32)
    assert host_lib.rsg_system_hierarchy(*_args(bad), A.iptr(sizes), None, 32, None, C.c_int64(0)) < 0
