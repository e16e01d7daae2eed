"""rsg_ldu_solve_noalloc: the stand-alone LDU solve with its device workspace
in the library's static scratch array (no hipMalloc / hipFree), against
rsg_ldu_solve on the same inputs -- two calls in a row, the second on a
smaller system, every solver and both precisions -- and its size limit."""
import ctypes as C

import numpy as np
import pytest

import rsg_abi as A

pytestmark = pytest.mark.gpu


def channel_system(nx, ny, solver, seed):
    """A periodic-in-x quad channel's face graph with log-normal face
    coefficients: a pinned Laplacian (AMG, PCG) or a relaxed non-symmetric
    operator (BiCGSTAB)."""
    rng = np.random.default_rng(seed)
    own, nbr = [], []
    for j in range(ny):
        for i in range(nx):
            c = j * nx + i
            own.append(c); nbr.append(j * nx + (i + 1) % nx)
            if j + 1 < ny:
                own.append(c); nbr.append(c + nx)
    own = np.array(own, np.int32); nbr = np.array(nbr, np.int32)
    n, nif = nx * ny, len(own)
    cf = np.exp(rng.normal(size=nif))
    if solver == 2:
        up, lo = -1.05 * cf, -0.95 * cf
        dg = np.zeros(n); np.add.at(dg, own, cf); np.add.at(dg, nbr, cf); dg = dg / 0.7
    else:
        up, lo = -cf, -cf.copy()
        dg = np.zeros(n); np.add.at(dg, own, cf); np.add.at(dg, nbr, cf); dg[0] *= 2.0
    b = rng.normal(size=n)
    return n, nif, own, nbr, dg, np.ascontiguousarray(up), np.ascontiguousarray(lo), b


def solve(lib, ctx, fn, sysm, solver, prec):
    n, nif, own, nbr, dg, up, lo, b = sysm
    x = np.zeros(n); st = np.zeros(4)
    rc = getattr(lib, fn)(ctx, n, nif, A.iptr(own), A.iptr(nbr), A.dptr(dg), A.dptr(up), A.dptr(lo), A.dptr(b),
                          A.dptr(x), solver, prec, 5000, C.c_double(1e-10), A.dptr(st))
    assert rc == 0, (fn, rc, lib.rsg_last_error(ctx))
    return x, st


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("solver", [0, 1, 2])
def test_noalloc_matches_ldu_solve(rsg_lib, gpu_ctx, solver, prec):
    for nx, ny, seed in ((4, 384, 1), (4, 96, 2)):          # second call smaller than the first
        sysm = channel_system(nx, ny, solver, seed)
        print("noalloc solver %d prec %d: %d cells" % (solver, prec, sysm[0]), flush=True)
        x0, s0 = solve(rsg_lib, gpu_ctx, "rsg_ldu_solve", sysm, solver, prec)
        x1, s1 = solve(rsg_lib, gpu_ctx, "rsg_ldu_solve_noalloc", sysm, solver, prec)
        assert s0[3] == 1 and s1[3] == 1
        # same kernels on the same inputs: bit-identical
        assert np.array_equal(x0, x1), float(np.abs(x0 - x1).max())
        assert np.array_equal(s0, s1)


def test_noalloc_rejects_oversize_system(rsg_lib, gpu_ctx):
    sysm = channel_system(4, 2048, 1, 3)                     # 8192 cells: ~12.5 MB of workspace
    n, nif, own, nbr, dg, up, lo, b = sysm
    x = np.zeros(n); st = np.zeros(4)
    rc = rsg_lib.rsg_ldu_solve_noalloc(gpu_ctx, n, nif, A.iptr(own), A.iptr(nbr), A.dptr(dg), A.dptr(up),
                                       A.dptr(lo), A.dptr(b), A.dptr(x), 1, 0, 100, C.c_double(1e-8), A.dptr(st))
    assert rc == -1
    assert b"scratch array" in rsg_lib.rsg_last_error(gpu_ctx)
    # and the library still solves afterwards
    x0, s0 = solve(rsg_lib, gpu_ctx, "rsg_ldu_solve_noalloc", channel_system(4, 128, 1, 4), 1, 0)
    assert s0[3] == 1
