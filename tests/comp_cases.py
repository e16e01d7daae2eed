"""The compressible channel cases of the reference's compressible tests
(python/compressible_diagnostics.py:134-199, make_validation_case): an
Ma-scaled inlet on a wall-graded channel at T_in = 300 K, p_out = 101325 Pa,
5 % turbulence intensity, and the settings that case uses.  Shared by the
golden-vector script and the tests (test infrastructure)."""
import math

import numpy as np

import rsg_abi as A

T_IN, P_REF = 300.0, 101325.0


def inflow(Ma, H=1.0, eos=A.EOS_AIR):
    """Uin, Re, kIn, omIn of make_validation_case (IdealGasEOS defaults)."""
    g, R, _, _, mu_ref, T_ref, S = eos
    rho_in = P_REF / (R * T_IN)
    r = T_IN / T_ref
    mu_in = mu_ref * r * math.sqrt(r) * (T_ref + S) / (T_IN + S)
    Uin = Ma * math.sqrt(g * R * T_IN)
    nu_in = mu_in / rho_in
    kIn = 1.5 * (Uin * 0.05) ** 2
    return dict(Uin=Uin, Re=rho_in * Uin * H / mu_in, kIn=kIn, omIn=kIn / (nu_in * 100.0), rho_in=rho_in)


def mesh(rs, nx, ny, Ma, Lx=10.0, H=1.0):
    m = rs.Mesh.make_channel_2d(nx, ny, Lx, H, inflow(Ma, H)["Re"], 1.0)
    m.compute_wall_distance()
    return m


def settings(max_iterations, Ma=0.1, **kw):
    """make_validation_case's SolverSettings (convergence_tol 1e-3, alpha_u 0.5,
    alpha_p 0.2, alpha_t 0.7, alpha_k = alpha_omega = 0.4, 200 inner iterations
    to 1e-4, turbulence from iteration 50 every second iteration); the engine's
    AMG refreshes every outer iteration like the reference's AMG solver."""
    base = dict(max_iterations=max_iterations, convergence_tol=1e-3, divergence_limit=1e10,
                alpha_u=0.5 if Ma <= 0.3 else 0.4, alpha_p=0.2, alpha_k=0.4, alpha_omega=0.4,
                inner_iterations=200, inner_tolerance=1e-4, turb_start_iter=50, turb_update_interval=2)
    base.update(kw)
    return A.default_settings(**base)


def init_state(Ma):
    f = inflow(Ma)
    return (f["Uin"], 0.0, 0.0, P_REF, T_IN, f["kIn"], f["omIn"])


def column_mass_imbalance(centers, volumes, rho, Ux, x_tol=1e-9):
    """compressible_diagnostics.column_mass_imbalance (:73-131): spread of the
    per-column mass flux sum(rho Ux V) / dx over its mean."""
    xs = np.sort(centers[:, 0])
    ux = [xs[0]]
    for x in xs[1:]:
        if x - ux[-1] > x_tol * max(abs(ux[-1]), 1.0):
            ux.append(float(x))
    ux = np.asarray(ux)
    dx = float(np.median(np.diff(ux)))
    fl = np.array([np.sum((rho * Ux * volumes)[np.abs(centers[:, 0] - x) < 0.5 * dx]) / dx for x in ux])
    nz = fl[fl != 0.0]
    return float((nz.max() - nz.min()) / abs(nz.mean()))
