"""The compressible member code (rsg_comp_dev.h: comp_iteration /
run_comp_member, the code the compressible member kernel runs) in the
sequential emulation build, against the reference compressible solver
compiled from its own sources (oracle/_ref/libref.so, ref_comp_solve):
whole trajectories bit for bit -- every residual-history entry and every
field -- on the 16 x 12 Ma 0.1 channel of the reference's compressible tests
with Jacobi-PCG, BiCGSTAB and Gauss-Seidel pressure solves, with wall
functions, and the reference's own 40 x 30 regression case run to
convergence (Jacobi-PCG pressure: the member path's AMG is a different
preconditioner)."""
import ctypes as C
import importlib.util
import os

import numpy as np
import pytest

import comp_cases as CC
import rsg_abi as A

HERE = os.path.dirname(os.path.abspath(__file__))
D = lambda a: (C.c_double * len(a))(*a)


def _golden_module():
    spec = importlib.util.spec_from_file_location("make_comp_golden", os.path.join(HERE, "golden", "make_comp_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def emu_comp(emu, path, s, Ma=0.1):
    msh = A.Msh(path)
    f = CC.inflow(Ma)
    n, nB = msh.nC, msh.nF - msh.nIF
    r, a = A.new_result(n, nB, s.max_iterations if s.record_history else 0)
    for k in ("T", "rho", "p_mech"):
        a[k] = np.zeros(n)
    bc, cd = msh.comp_bc(f["Uin"], CC.T_IN, CC.P_REF, f["kIn"], f["omIn"])
    rc = emu.emu_comp_solve(C.byref(msh.desc()), C.byref(bc), C.byref(cd), C.byref(s), D(A.DEFAULT_COEFFS),
                            D(CC.init_state(Ma)), C.byref(r), A.dptr(a["T"]), A.dptr(a["rho"]), A.dptr(a["p_mech"]))
    assert rc == 0
    a["hist"] = a["history"][:6 * r.history_len].reshape(-1, 6)
    return r, a


def _check_bit_exact(r, a, o, h, info):
    assert (r.final_iter, r.converged, r.diverged, r.turb_established) == (info[0], info[1], info[2], info[4])
    assert np.array_equal(a["hist"], h)
    for k_e, k_r in (("U", "U"), ("p_mech", "p"), ("T", "T"), ("rho", "rho"), ("k", "k"), ("omega", "om"),
                     ("nuT", "nuT"), ("F1", "F1"), ("F2", "F2"), ("Pk", "Pk")):
        assert np.array_equal(a[k_e], o[k_r]), k_e


@pytest.fixture(scope="module")
def chan16(rs, tmp_path_factory):
    path = str(tmp_path_factory.mktemp("emucomp") / "c16.msh")
    CC.mesh(rs, 16, 12, 0.1).save_binary(path)
    return path


@pytest.mark.parametrize("label,kw", [
    ("pcg", dict(pressure_solver=1)),
    ("bicgstab", dict(pressure_solver=2)),
    ("gauss_seidel", dict(pressure_solver=3, momentum_solver=3, turbulence_solver=3)),
    ("wall_functions", dict(pressure_solver=1, use_wall_functions=1)),
    ("interval1", dict(pressure_solver=1, turb_update_interval=1, turb_start_iter=20)),
])
def test_emulated_comp_trajectory_bit_exact(emu, ref, chan16, label, kw):
    s = CC.settings(300, record_history=1, convergence_tol=1e-12, **kw)
    r, a = emu_comp(emu, chan16, s)
    o, h, info = _golden_module().ref_comp_solve(ref, chan16, s)
    _check_bit_exact(r, a, o, h, info)


def test_emulated_comp_regression_case_bit_exact(emu, ref, rs, tmp_path):
    """The reference's regression channel (40 x 30, its settings) to
    convergence, with Jacobi-PCG pressure solves."""
    path = str(tmp_path / "c40.msh")
    CC.mesh(rs, 40, 30, 0.1).save_binary(path)
    s = CC.settings(3000, record_history=1, pressure_solver=1)
    r, a = emu_comp(emu, path, s)
    o, h, info = _golden_module().ref_comp_solve(ref, path, s)
    assert info[1] == 1
    _check_bit_exact(r, a, o, h, info)
    # the exported pressure is p_mech - (2/3) rho max(k, 0)
    assert np.array_equal(a["p"], a["p_mech"] - (2.0 / 3.0) * a["rho"] * np.maximum(a["k"], 0.0))
