"""A streamwise-periodic curved channel through rans_sst_py, with the
contract of the reference's tests/python/test_periodic_mesh_binding.py:15-55:
the mesh factory from node arrays (a cosine hill of height 1 over a 4.5-long
period, 30 x 20 cells), its two wall patches, the body-force setting, a
reattachment-length observation and a body-force-driven solve that ends
Converged or Unconverged with finite fields and a bulk flow established."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_periodic_curved_channel_with_body_force(rs):
    nx, ny, Lx, y_top = 30, 20, 4.5, 3.0
    x_nodes = np.linspace(0.0, Lx, nx + 1)
    y_bottom = 0.5 * (1.0 + np.cos(2.0 * np.pi * x_nodes / Lx))
    mesh = rs.Mesh.make_curved_channel_periodic_2d(x_nodes, y_bottom, y_top, ny, Re=2800.0, yPlusTarget=1.0)
    mesh.compute_wall_distance()
    assert set(mesh.patch_names()) == {"bottom_wall", "top_wall"}
    assert set(mesh.patch_types()) == {"wall"}
    assert mesh.n_cells() == nx * ny
    k_in, om_in, nu = 1e-4, 10.0, 3.0e-4
    bcs = rs.FlowBoundaryConditions.channel_defaults(mesh, 1.0, k_in, om_in)
    s = rs.SolverSettings()
    s.max_iterations = 15000; s.convergence_tol = 1e-4
    s.alpha_u, s.alpha_p = 0.4, 0.25
    s.verbose = False
    s.body_force = rs.Vec3(4.0e-3, 0.0, 0.0)
    obs = rs.ObservationOperator()
    obs.add_reattachment_length("bottom_wall", xr_obs=2.0, sigma=1.0)
    ps = rs.InferenceParameterSet.a1_betaStar()
    fm = rs.ForwardModel(mesh, ps, obs, bcs, nu, s, rs.Vec3(0.3, 0.0, 0.0), 0.0, k_in, om_in)
    print("periodic curved channel 30 x 20", flush=True)
    r = fm.evaluate(list(ps.pack(rs.SSTCoefficients())))
    print("status", r.status, "iterations", r.simple_iters, "predictions", list(r.predictions), flush=True)
    assert str(r.status).split(".")[-1] in ("Converged", "Unconverged")
    assert fm.has_last_fields()
    U = np.asarray(fm.last_fields()["U"]).reshape(-1, 3)
    vol = np.asarray(mesh.cell_volumes())
    assert np.all(np.isfinite(U))
    assert float(np.sum(U[:, 0] * vol) / np.sum(vol)) > 0.05
