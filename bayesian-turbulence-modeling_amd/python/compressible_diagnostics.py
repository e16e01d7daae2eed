"""Diagnostics of the compressible forward model (drop-in for the reference's
python/compressible_diagnostics.py): the Ma-scaled channel validation case,
its run summary (status, iterations, skin friction, Mach / temperature /
density extrema, positivity, column mass-flux imbalance), on the MI355X
build of rans_sst_py.

  make_validation_case(name, Ma, nx, ny, ...)   python/compressible_diagnostics.py:134-199
  run_validation_case(case)                     :202-266
  summarise_fields(fields, eos)                 :36-59
  positivity_ok(fields)                         :62-70
  column_mass_imbalance(mesh, fields)           :73-131
"""
from __future__ import annotations

import time
from typing import Any

import numpy as np

from solver_bindings import _rs


def summarise_fields(fields: dict, eos) -> dict:
    U = np.asarray(fields["U"]); p = np.asarray(fields["p"]); T = np.asarray(fields["T"])
    rho = np.asarray(fields["rho"]); nuT = np.asarray(fields["nuT"])
    Umag = np.linalg.norm(U, axis=1)
    Ma = Umag / np.sqrt(eos.gamma * eos.R * T)
    return {"U_max": float(Umag.max()), "U_mean": float(Umag.mean()), "p_min": float(p.min()),
            "p_max": float(p.max()), "T_min": float(T.min()), "T_max": float(T.max()),
            "rho_min": float(rho.min()), "rho_max": float(rho.max()), "Ma_max": float(Ma.max()),
            "Ma_mean": float(Ma.mean()), "nuT_max": float(nuT.max())}


def positivity_ok(fields: dict) -> tuple[bool, dict]:
    info = {"rho_positive": bool(np.all(np.asarray(fields["rho"]) > 0.0)),
            "p_positive": bool(np.all(np.asarray(fields["p"]) > 0.0)),
            "T_positive": bool(np.all(np.asarray(fields["T"]) > 0.0))}
    return all(info.values()), info


def column_mass_imbalance(mesh, fields, x_tol: float = 1e-9) -> dict:
    """Spread of the per-column mass flux sum(rho Ux V) / dx over its mean
    (uniform column spacing detected from the cell centres)."""
    centers = np.asarray(mesh.cell_centers()); volumes = np.asarray(mesh.cell_volumes())
    rho = np.asarray(fields["rho"]); Ux = np.asarray(fields["U"])[:, 0]
    xs = np.sort(centers[:, 0])
    ux = [xs[0]]
    for x in xs[1:]:
        if x - ux[-1] > x_tol * max(abs(ux[-1]), 1.0):
            ux.append(float(x))
    ux = np.asarray(ux)
    if ux.size < 2:
        return {"max_flux": 0.0, "min_flux": 0.0, "mean_flux": 0.0, "rel_imbalance": float("nan"),
                "n_columns": int(ux.size)}
    dx = float(np.median(np.diff(ux)))
    flux = np.zeros(ux.size)
    for i, x in enumerate(ux):
        m = np.abs(centers[:, 0] - x) < 0.5 * dx
        if np.any(m):
            flux[i] = float(np.sum(rho[m] * Ux[m] * volumes[m]) / dx)
    nz = flux[flux != 0.0]
    if nz.size == 0:
        return {"max_flux": 0.0, "min_flux": 0.0, "mean_flux": 0.0, "rel_imbalance": float("nan")}
    mean = float(np.mean(nz))
    spread = float(nz.max() - nz.min())
    return {"max_flux": float(nz.max()), "min_flux": float(nz.min()), "mean_flux": mean,
            "rel_imbalance": spread / abs(mean) if mean != 0 else float("nan"), "n_columns": int(nz.size)}


def make_validation_case(name: str, Ma: float, nx: int = 32, ny: int = 24, Lx: float = 10.0, H: float = 1.0,
                         max_iterations: int = 4000, convergence_tol: float = 1e-3, turb_intensity: float = 0.05,
                         nut_floor_iters: int | None = None) -> dict[str, Any]:
    """Mesh, gas, BCs, settings and forward model of an Ma channel: inlet at
    Ma a(T_in), T_in = 300 K, p_out = 101325 Pa, wall-graded to y+ 1 at the
    inlet Reynolds number, skin friction observed at x = Lx/4, Lx/2, 3Lx/4."""
    rs = _rs()
    eos = rs.IdealGasEOS()
    T_in, p_ref = 300.0, 101325.0
    rho_in = eos.density(p_ref, T_in)
    mu_in = eos.viscosity(T_in)
    Uin = Ma * eos.sound_speed(T_in)
    nu_in = mu_in / rho_in
    Re = rho_in * Uin * H / mu_in
    mesh = rs.Mesh.make_channel_2d(nx, ny, Lx, H, Re=Re, yPlusTarget=1.0)
    mesh.compute_wall_distance()
    kIn = 1.5 * (Uin * turb_intensity) ** 2
    omIn = kIn / (nu_in * 100.0)
    bcs = rs.CompressibleBoundaryConditions.channel_defaults(mesh, Uin, T_in, p_ref, kIn, omIn)
    obs = rs.ObservationOperator()
    for x in (0.25 * Lx, 0.5 * Lx, 0.75 * Lx):
        obs.add_skin_friction(wall_patch="bottom", location=rs.Vec3(x, 0.0, 0.0), cf_obs=0.005, sigma=0.001,
                              ref_vel=Uin)
    s = rs.SolverSettings()
    s.max_iterations = max_iterations; s.convergence_tol = convergence_tol; s.divergence_limit = 1e10
    s.alpha_u = 0.5 if Ma <= 0.3 else 0.4; s.alpha_p = 0.2; s.alpha_t = 0.7
    s.alpha_k = 0.4; s.alpha_omega = 0.4
    s.inner_iterations = 200; s.inner_tolerance = 1e-4
    s.turb_start_iter = 50; s.turb_update_interval = 2
    s.verbose = False; s.report_interval = 500
    if nut_floor_iters is not None:
        s.nut_floor_iters = nut_floor_iters
    param_set = rs.InferenceParameterSet.a1_betaStar()
    fm = rs.CompressibleForwardModel(mesh=mesh, param_set=param_set, obs_op=obs, bcs=bcs, eos=eos, settings=s,
                                     u_init=rs.Vec3(Uin, 0, 0), p_init=p_ref, T_init=T_in, k_init=kIn,
                                     omega_init=omIn)
    return {"name": name, "Ma": Ma, "Re": Re, "Uin": Uin, "T_in": T_in, "p_ref": p_ref, "rho_in": rho_in,
            "mu_in": mu_in, "mesh": mesh, "eos": eos, "obs": obs, "bcs": bcs, "settings": s,
            "param_set": param_set, "forward": fm, "nx": nx, "ny": ny, "Lx": Lx, "H": H,
            "max_iterations": max_iterations}


def run_validation_case(case: dict[str, Any]) -> dict[str, Any]:
    """Solve at the default SST coefficients; the run's summary (fields that
    are missing after a failed solve are reported as NaN)."""
    rs = _rs()
    fm, eos = case["forward"], case["eos"]
    theta = list(case["param_set"].pack(rs.SSTCoefficients()))
    t0 = time.time()
    result = fm.evaluate(theta)
    elapsed = time.time() - t0
    status = str(result.status)
    summary: dict[str, Any] = {
        "name": case["name"], "Ma": case["Ma"], "Re": case["Re"], "mesh_cells": int(case["mesh"].n_cells()),
        "max_iterations": case["max_iterations"], "simple_iters": int(result.simple_iters),
        "elapsed_s": float(elapsed), "status": status, "converged": "Converged" in status,
        "diverged": "Diverged" in status or "Divergence" in status, "log_lik": float(result.log_lik),
        "Cf_at_stations": list(result.predictions)}
    if not fm.has_last_fields():
        nan = float("nan")
        summary.update({"fields_available": False, "positivity_ok": False,
                        "positivity": {"rho_positive": False, "p_positive": False, "T_positive": False},
                        **{k: nan for k in ("U_max", "U_mean", "p_min", "p_max", "T_min", "T_max", "rho_min",
                                            "rho_max", "Ma_max", "Ma_mean", "nuT_max")},
                        "mass_flux": {"max_flux": nan, "min_flux": nan, "mean_flux": nan, "rel_imbalance": nan,
                                      "n_columns": 0}})
        return summary
    fields = fm.last_fields()
    ok, info = positivity_ok(fields)
    summary.update({"fields_available": True, "positivity_ok": ok, "positivity": info,
                    **summarise_fields(fields, eos), "mass_flux": column_mass_imbalance(case["mesh"], fields)})
    return summary
