"""Mixed-precision member solves on the GPU (RSG_PREC_MIXED through the
C-ABI): FP32 inner solves (BiCGSTAB, Jacobi-PCG, AMG-PCG with an FP32 V-cycle
around the FP64 coarsest inverse) with FP64 defect and iterative refinement.

* stand-alone mixed linear solves meet the tolerance in the true FP64
  residual (dense check);
* C0 (Re_tau 550 channel) solved to tol 1e-8 in FP64 and mixed in the same
  engine batch: every field within 1e-6 relative L2 of the FP64 GPU solve and
  of the reference's converged fields;
* a sample of the C2 ensemble (the six lowest-Re_tau members, through the
  rans_sst_py drop-in): per-member errors of U, p, k, omega, nu_t within 1e-6
  of the FP64 GPU solve of the same member.
"""
import ctypes as C
import os

import numpy as np
import pytest

import rsg_abi as A
from conftest import TESTS, settings_for
from test_gpu_parity import Problem, chk, rel

pytestmark = pytest.mark.gpu


def p_err(a, b):
    """p against the dynamic-pressure scale rms(|U|^2) of b (the periodic
    channel's p is round-off, so it cannot be held against itself)."""
    q = np.sqrt(np.mean(np.sum(np.asarray(b["U"]).reshape(-1, 3) ** 2, axis=1) ** 2))
    return float(np.linalg.norm(np.asarray(a["p"]) - np.asarray(b["p"])) / np.sqrt(len(b["p"])) / q)


@pytest.mark.parametrize("solver", [2, 1, 0])
def test_mixed_linear_solves_meet_tolerance_in_fp64(rsg_lib, gpu_ctx, meshes, solver):
    msh = A.Msh(meshes["c0"])
    n, nif = msh.nC, msh.nIF
    rng = np.random.default_rng(13)
    own = np.ascontiguousarray(msh.owner[:nif], np.int32); nbr = np.ascontiguousarray(msh.neighbor, np.int32)
    cf = msh.area[:nif] / msh.delta[:nif]
    if solver == 2:   # relaxed, non-symmetric (momentum-like)
        up, lo = -1.05 * cf, -0.95 * cf
        dg = np.zeros(n); np.add.at(dg, own, cf); np.add.at(dg, nbr, cf); dg = dg / 0.7
    else:             # pinned Laplacian (p'-like)
        up, lo = -cf, -cf.copy()
        dg = np.zeros(n); np.add.at(dg, own, cf); np.add.at(dg, nbr, cf); dg[0] *= 2.0
    up, lo = np.ascontiguousarray(up), np.ascontiguousarray(lo)
    Ad = np.diag(dg); Ad[own, nbr] += up; Ad[nbr, own] += lo
    b = rng.normal(size=n)
    for tol in (1e-3, 1e-9):
        x = np.zeros(n); st = np.zeros(4)
        chk(rsg_lib, gpu_ctx, rsg_lib.rsg_ldu_solve(gpu_ctx, n, nif, A.iptr(own), A.iptr(nbr), A.dptr(dg),
                                                    A.dptr(up), A.dptr(lo), A.dptr(b), A.dptr(x), solver, 1,
                                                    20000, C.c_double(tol), A.dptr(st)))
        true = np.linalg.norm(b - Ad @ x) / np.linalg.norm(b)
        assert st[3] == 1, (tol, st)
        assert st[2] == pytest.approx(np.linalg.norm(b), rel=1e-12)   # the FP64 first defect
        assert true < tol, (tol, true)


def test_c0_mixed_matches_fp64_gpu_and_reference(rsg_lib, gpu_ctx, meshes):
    s = settings_for("c0", max_iterations=80000, convergence_tol=1e-8, pressure_solver=0,
                     amg_refresh=1000, coarse_refresh=10)
    rsg_lib.rsg_set_launch(gpu_ctx, 2000, 0)
    P = Problem(rsg_lib, gpu_ctx, meshes["c0"], "c0", s)
    try:
        t64 = P.submit(prec=0)
        t32 = P.submit(prec=1)            # both in the engine at once: two kernels per round
        r64, a64 = P.fetch(t64)
        r32, a32 = P.fetch(t32)
    finally:
        P.close()
    g = np.load(os.path.join(TESTS, "golden", "c0_ref_tol1e-08.npz"))
    assert r64.converged == 1 and r32.converged == 1
    assert abs(r32.final_iter - r64.final_iter) < 0.01 * r64.final_iter
    errs = {}
    for k_gpu, k_g in (("U", "U"), ("k", "k"), ("omega", "omega"), ("nuT", "nuT")):
        errs[k_gpu] = (rel(a32[k_gpu], a64[k_gpu]), rel(a32[k_gpu], g[k_g]))
    errs["p"] = (p_err(a32, a64), p_err(a32, g))
    print("c0 mixed: %d iterations (fp64 %d); rel-L2 vs fp64 / vs reference: %s"
          % (r32.final_iter, r64.final_iter, " ".join("%s %.1e/%.1e" % (k, *v) for k, v in errs.items())))
    for k, (e64, eref) in errs.items():
        assert e64 < 1e-6 and eref < 1e-6, (k, e64, eref)


def test_c2_sample_mixed_matches_fp64(rs):
    """Six members of the 512-member ensemble (the lowest Re_tau, ny 96-112),
    each solved to tol 1e-8 in FP64 and in mixed precision, all twelve in one
    batch through rans_sst_py (ForwardModel.set_precision)."""
    import ensembles
    specs = sorted(ensembles.channel_ensemble(512, 0), key=lambda sp: sp["re_tau"])[:6]
    _, f64 = ensembles.build_models(rs, specs, tol=1e-8, max_iterations=200000)
    meshes, f32 = ensembles.build_models(rs, specs, tol=1e-8, max_iterations=200000)
    for fm in f32:
        fm.set_precision("mixed")
    rs.engine_set_time_slice(200000)
    res = rs.evaluate_many(f64 + f32, [sp["theta"] for sp in specs] * 2)
    worst = 0.0
    for i, sp in enumerate(specs):
        r64, r32 = res[i], res[i + len(specs)]
        assert str(r64.status).endswith("Converged") and str(r32.status).endswith("Converged"), (r64.status, r32.status)
        a, b = f32[i].last_fields(), f64[i].last_fields()
        e = {k: rel(np.asarray(a[k]), np.asarray(b[k])) for k in ("U", "k", "omega", "nuT")}
        e["p"] = p_err(a, b)
        print("member Re_tau %.0f ny %d: %d / %d iterations, %s" % (sp["re_tau"], sp["ny"], r64.simple_iters,
                                                                     r32.simple_iters,
                                                                     " ".join("%s %.1e" % kv for kv in e.items())))
        worst = max(worst, max(e.values()))
    assert worst < 1e-6, worst
