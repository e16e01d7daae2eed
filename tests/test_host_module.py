"""rans_sst_py host behaviour that needs no GPU: parameter sets, boundary
condition profiles, settings defaults, mesh queries, the odd-even
diagnostic and the ForwardModel's argument-error conventions."""
import numpy as np
import pytest


def test_parameter_sets(rs):
    ps = rs.InferenceParameterSet.a1_betaStar()
    assert ps.in_bounds([0.31, 0.09])
    assert not ps.in_bounds([0.31]) and not ps.in_bounds([0.31, 0.09, 0.5]) and not ps.in_bounds([])
    assert not ps.in_bounds([float("nan"), 0.09]) and not ps.in_bounds([0.9, 0.09])
    assert rs.InferenceParameterSet.live10().n_active() == 10
    assert rs.InferenceParameterSet.all11().active_names()[10] == "kappa"
    c = ps.unpack([0.4, 0.1])
    assert c.a1 == 0.4 and c.betaStar == 0.1 and c.beta1 == 0.075
    assert ps.pack(c) == [0.4, 0.1]
    with pytest.raises(RuntimeError):
        ps.unpack([0.4])
    with pytest.raises(RuntimeError):
        rs.InferenceParameterSet.from_indices("x", [1, 1])
    nw = rs.InferenceParameterSet.near_wall4()
    assert nw.lower_bounds() == [0.1, 0.01, 0.01, 0.1] and nw.upper_bounds() == [0.8, 0.2, 0.2, 2.0]


def test_settings_defaults_match_reference(rs):
    s = rs.SolverSettings()
    assert (s.max_iterations, s.convergence_tol, s.divergence_limit) == (500, 1e-5, 1e6)
    assert (s.alpha_u, s.alpha_p, s.alpha_k, s.alpha_omega) == (0.7, 0.3, 0.5, 0.5)
    assert (s.inner_iterations, s.inner_tolerance, s.turb_start_iter, s.turb_update_interval) == (200, 1e-3, 5, 1)
    assert (s.nut_floor_iters, s.k_min, s.omega_min, s.von_karman) == (500, 1e-10, 1e-6, 0.41)
    assert s.pressure_solver == "AMG" and s.momentum_solver == "BiCGSTAB"


def test_bc_profiles(rs):
    m = rs.Mesh.make_channel_2d(4, 6, 1.0, 1.0)
    bc = rs.FlowBoundaryConditions.channel_defaults(m, 1.0, 1e-4, 1.0)
    bc.set_velocity_profile(m, "inlet", np.ones((6, 3)))
    bc.set_k_profile(m, "inlet", np.ones(6))
    with pytest.raises(RuntimeError):
        bc.set_k_profile(m, "inlet", np.ones(5))
    with pytest.raises(RuntimeError):
        bc.set_omega_profile(m, "nope", np.ones(6))
    with pytest.raises(RuntimeError):
        bc.set_velocity_profile(m, "inlet", np.ones((6, 2)))


def test_mesh_queries(rs):
    m = rs.Mesh.make_channel_2d(4, 6, 1.0, 2.0)
    assert m.patch_names() == ["bottom", "top", "inlet", "outlet"]
    assert m.patch_types() == ["wall", "wall", "inlet", "outlet"]
    d = m.wall_patch_data("bottom")
    assert d["n_faces"] == 4 and np.allclose(d["normal"][:, 1], -1.0)
    assert m.cell_centers().shape == (24, 3) and m.node_coords().shape == (70, 3)
    m.compute_wall_distance()
    wd_walls = m.wall_distance().copy()
    m.set_patch_type("top", "symmetry")
    m.compute_wall_distance()
    assert np.all(m.wall_distance() >= wd_walls - 1e-15)
    with pytest.raises(RuntimeError):
        m.wall_patch_data("nope")


def test_odd_even_discrimination(rs):
    # tests/cpp/test_infra_fixes.cpp:95-111
    m = rs.Mesh.make_channel_2d(12, 10, 2.0, 1.0)
    cc = m.cell_centers()
    smooth = list(0.3 * cc[:, 0] + 0.1 * cc[:, 1])
    ci = np.arange(m.n_cells())
    checker = list(np.where(((ci % 12) + (ci // 12)) % 2 == 0, 1.0, -1.0))
    r_s = rs.odd_even_energy_ratio(m, smooth)
    r_c = rs.odd_even_energy_ratio(m, checker)
    assert r_s < 0.05 and r_c > 0.99 and r_c > 20 * r_s
    with pytest.raises(ValueError):
        rs.odd_even_energy_ratio(m, [1.0])


def test_forward_model_rejects_invalid_theta_without_solving(rs):
    m = rs.Mesh.make_channel_2d(4, 6, 1.0, 1.0)
    m.compute_wall_distance()
    fm = rs.ForwardModel(m, rs.InferenceParameterSet.a1_betaStar(), rs.ObservationOperator(),
                         rs.FlowBoundaryConditions.channel_defaults(m, 1.0, 1e-4, 1.0), 1e-3)
    r = fm.evaluate([5.0, 0.09])
    assert str(r.status).split(".")[-1] == "InvalidParameters" and r.log_lik == -1e30
    assert fm.penalized_log_likelihood([0.31]) == -1e30
    assert not fm.has_last_fields()
    with pytest.raises(RuntimeError):
        fm.last_fields()
