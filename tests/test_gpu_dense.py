"""GPU parity of the coarsest-level dense inverse (rsg_dense_inverse through
the C-ABI): blocked Gauss-Jordan whose block products run on the FP64 matrix
cores.  It stands where the reference solves its coarsest AMG level with
Gauss-Seidel sweeps (core/include/LinearSolver.hpp:508-512, 557-562).

Bars: against numpy's LAPACK inverse of the same matrix, relative Frobenius
error below 1e-12 times the condition number (FP64 round-off of an
elimination); A X = I to the same bound; sizes that are and are not multiples
of the 64-row block; a matrix with a non-positive pivot is rejected (the
flag is raised by the kernel's pivot check, no fault)."""
import ctypes as C

import numpy as np
import pytest

import rsg_abi as A
from test_gpu_parity import chk

pytestmark = pytest.mark.gpu


def spd(n, seed):
    """A Galerkin-like SPD M-matrix: a random sparse graph Laplacian plus a
    small positive diagonal (as a coarsest AMG level of a pressure system)."""
    rng = np.random.default_rng(seed)
    W = np.zeros((n, n))
    for i in range(n):
        nb = rng.choice(n, size=min(n - 1, 6), replace=False)
        for j in nb:
            if j != i:
                W[i, j] = W[j, i] = rng.uniform(0.1, 2.0)
    L = np.diag(W.sum(axis=1)) - W
    return L + np.diag(rng.uniform(1e-3, 1e-1, n))


def inverse(lib, ctx, M):
    n = M.shape[0]
    a = np.ascontiguousarray(M, dtype=np.float64)
    inv = np.zeros_like(a)
    sec = C.c_double()
    print("dense inverse n=%d" % n, flush=True)
    chk(lib, ctx, lib.rsg_dense_inverse(ctx, n, A.dptr(a), A.dptr(inv), C.byref(sec)))
    return inv, sec.value


@pytest.mark.parametrize("n", [1, 5, 64, 65, 130, 700])
def test_matches_lapack_inverse(rsg_lib, gpu_ctx, n):
    M = spd(n, seed=n)
    X, sec = inverse(rsg_lib, gpu_ctx, M)
    ref = np.linalg.inv(M)
    cond = np.linalg.cond(M)
    err = np.linalg.norm(X - ref) / np.linalg.norm(ref)
    res = np.linalg.norm(M @ X - np.eye(n)) / np.sqrt(n)
    print("n=%d cond %.2e rel err %.2e residual %.2e kernel %.3f ms" % (n, cond, err, res, 1e3 * sec))
    assert err < 1e-12 * cond and res < 1e-12 * cond
    assert sec > 0.0


def test_symmetric_result(rsg_lib, gpu_ctx):
    M = spd(300, seed=3)
    X, _ = inverse(rsg_lib, gpu_ctx, M)
    assert np.linalg.norm(X - X.T) / np.linalg.norm(X) < 1e-12 * np.linalg.cond(M)


def test_rejects_non_positive_pivot(rsg_lib, gpu_ctx):
    M = spd(100, seed=1)
    M[70, 70] = -5.0          # the pivot of row 70 is negative: not SPD
    a = np.ascontiguousarray(M)
    inv = np.zeros_like(a)
    ERR_ARG = -1
    assert rsg_lib.rsg_dense_inverse(gpu_ctx, 100, A.dptr(a), A.dptr(inv), None) == ERR_ARG
    assert rsg_lib.rsg_dense_inverse(gpu_ctx, 0, A.dptr(a), A.dptr(inv), None) == ERR_ARG
