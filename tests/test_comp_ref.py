"""The compressible path's oracle: the reference compressible SIMPLE solver
compiled from its own sources (oracle/_ref/libref.so, ref_comp_solve in
oracle/ref_driver.cpp), pinned against the reference's own regression data
(tests/regression/compressible_channel_ma01.json, copied to tests/golden/)
on the channel built by this package's mesh factory; the golden trajectories
of tests/golden/comp_channel_16x12.npz reproduced; and the host layer's
closed forms (IdealGasEOS, channel_defaults, argument checks) without a GPU."""
import json
import math
import os

import numpy as np
import pytest

import comp_cases as CC
import rsg_abi as A

HERE = os.path.dirname(os.path.abspath(__file__))


def _golden_module():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_comp_golden", os.path.join(HERE, "golden", "make_comp_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _ref_solve(ref, path, s):
    return _golden_module().ref_comp_solve(ref, path, s)


def test_reference_build_matches_reference_regression(ref, rs, tmp_path):
    base = json.load(open(os.path.join(HERE, "golden", "compressible_channel_ma01.json")))
    exp, tol = base["expected"], base["tolerances"]
    m = CC.mesh(rs, 40, 30, 0.1)
    path = str(tmp_path / "c40.msh"); m.save_binary(path)
    out, hist, info = _ref_solve(ref, path, CC.settings(3000))
    assert info[1] == 1 and info[2] == 0
    assert abs(info[0] - exp["simple_iters"]) <= tol["simple_iters_abs"]
    rho = out["rho"]; k = out["k"]; T = out["T"]
    U = out["U"].reshape(-1, 3); Um = np.linalg.norm(U, axis=1)
    Ma = Um / np.sqrt(1.4 * 287.0 * T)
    assert abs(Ma.max() - exp["Ma_max"]) / exp["Ma_max"] < tol["Ma_max_rel"]
    assert abs(T.min() - exp["T_min"]) <= tol["T_extrema_abs_K"] and abs(T.max() - exp["T_max"]) <= tol["T_extrema_abs_K"]
    assert abs(rho.min() - exp["rho_min"]) / exp["rho_min"] < tol["rho_extrema_rel"]
    assert abs(rho.max() - exp["rho_max"]) / exp["rho_max"] < tol["rho_extrema_rel"]
    assert abs(Um.max() - exp["U_max"]) / exp["U_max"] < tol["U_max_rel"]
    imb = CC.column_mass_imbalance(np.asarray(m.cell_centers()), np.asarray(m.cell_volumes()), rho, U[:, 0])
    assert imb < tol["mass_flux_rel_imbalance_max"]
    # the state obeys the two-pressure equation of state exactly
    p_th = out["p"] - (2.0 / 3.0) * rho * np.maximum(k, 0.0)
    assert np.allclose(p_th, rho * 287.0 * np.maximum(T, 1.0), rtol=1e-12)
    g = dict(np.load(os.path.join(HERE, "golden", "comp_channel_40x30.npz")))
    assert np.array_equal(info, g["info"])
    for key in ("U", "p", "T", "rho", "k", "om", "nuT"):
        assert np.array_equal(out[key], g[key]), key


def test_reference_build_reproduces_golden_trajectories(ref, rs, tmp_path):
    g = dict(np.load(os.path.join(HERE, "golden", "comp_channel_16x12.npz")))
    path = str(tmp_path / "c16.msh"); CC.mesh(rs, 16, 12, 0.1).save_binary(path)
    mod = _golden_module()
    for case in ("pcg52t",):
        out, hist, info = mod.ref_comp_solve(ref, path, mod.case_settings(case))
        assert np.array_equal(hist, g[case + "_hist"]) and np.array_equal(info, g[case + "_info"])
        for k in ("U", "p", "T", "rho", "k", "om"):
            assert np.array_equal(out[k], g[case + "_" + k]), (case, k)


def test_ideal_gas_closed_forms(rs):
    e = rs.IdealGasEOS()
    assert e.Cp() == pytest.approx(1004.5, rel=1e-15) and e.Cv() == pytest.approx(717.5, rel=1e-15)
    assert e.viscosity(e.T_ref) == pytest.approx(e.mu_ref, rel=1e-15)   # Sutherland at T_ref
    assert e.viscosity(400.0) == pytest.approx(1.716e-5 * (400 / 273.15) ** 1.5 * 383.55 / 510.4, rel=1e-14)
    assert e.density(101325.0, 300.0) * 287.0 * 300.0 == pytest.approx(101325.0, rel=1e-15)
    assert e.temperature(101325.0, e.density(101325.0, 300.0)) == pytest.approx(300.0, rel=1e-15)
    assert e.sound_speed(300.0) == pytest.approx(math.sqrt(1.4 * 287.0 * 300.0), rel=1e-15)
    assert e.conductivity(2e-5, 0.0) == pytest.approx(2e-5 * 1004.5 / 0.72, rel=1e-14)


def test_channel_defaults_and_argument_checks(rs):
    f = CC.inflow(0.1)
    m = CC.mesh(rs, 8, 6, 0.1)
    bcs = rs.CompressibleBoundaryConditions.channel_defaults(m, f["Uin"], 300.0, 101325.0, f["kIn"], f["omIn"])
    types = {t.patch_name: str(t.type).split(".")[-1] for t in bcs.temperature_bc}
    assert types == {"inlet": "Dirichlet", "outlet": "Neumann", "bottom": "AdiabaticWall", "top": "AdiabaticWall"}
    ps = rs.InferenceParameterSet.a1_betaStar()
    obs = rs.ObservationOperator()
    obs.add_drag("bottom", 0.01, 0.001, 1.0, f["Uin"])
    with pytest.raises(ValueError):
        rs.CompressibleForwardModel(m, ps, obs, bcs, rs.IdealGasEOS())
    fm = rs.CompressibleForwardModel(m, ps, rs.ObservationOperator(), bcs, rs.IdealGasEOS())
    r = fm.evaluate([10.0, 10.0])          # out of bounds: no engine needed
    assert "InvalidParameters" in str(r.status) and r.log_lik == -1e30
    assert not fm.has_last_fields()
