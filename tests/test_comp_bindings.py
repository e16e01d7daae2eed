"""The compressible part of rans_sst_py's surface on the host (no GPU): the
symbols the reference's binding tests import and the ideal-gas relations
(reference tests/python/test_compressible_bindings.py:25-58;
IdealGasEOS.hpp:6-38).  Each relation is checked against its closed form
with the gas of the reference's defaults (air: gamma 1.4, R 287, Sutherland
mu_ref 1.716e-5 at 273.15 K, S 110.4 K)."""
import math

import pytest


def test_compressible_symbols_present(rs):
    for name in ("ForwardModel", "CompressibleForwardModel", "IdealGasEOS", "CompressibleBoundaryConditions",
                 "TBCType", "PatchTBC"):
        assert hasattr(rs, name), name


def test_eos_constants(rs):
    eos = rs.IdealGasEOS()
    assert (eos.gamma, eos.R) == (1.4, 287.0)
    assert eos.Cp() == pytest.approx(1.4 * 287.0 / 0.4, rel=1e-14)
    assert eos.Cv() == pytest.approx(287.0 / 0.4, rel=1e-14)
    assert eos.Cp() - eos.Cv() == pytest.approx(eos.R, rel=1e-12)


def test_eos_state_round_trip(rs):
    eos = rs.IdealGasEOS()
    p, T = 101325.0, 300.0
    rho = eos.density(p, T)
    assert rho == pytest.approx(p / (287.0 * T), rel=1e-15)
    assert eos.temperature(p, rho) == pytest.approx(T, rel=1e-12)
    assert eos.pressure(rho, T) == pytest.approx(p, rel=1e-12)


def test_eos_sound_speed_and_mach(rs):
    eos = rs.IdealGasEOS()
    a = eos.sound_speed(300.0)
    assert a == pytest.approx(math.sqrt(1.4 * 287.0 * 300.0), rel=1e-14)   # 347.2 m/s
    assert eos.mach_number(0.1 * a, 300.0) == pytest.approx(0.1, rel=1e-12)


def test_eos_sutherland_viscosity(rs):
    eos = rs.IdealGasEOS()
    mus = [eos.viscosity(T) for T in (250.0, 300.0, 400.0, 600.0)]
    assert all(m1 < m2 for m1, m2 in zip(mus, mus[1:]))
    # the reference temperature returns the reference viscosity
    assert eos.viscosity(eos.T_ref) == pytest.approx(eos.mu_ref, rel=1e-14)
    # Sutherland's law at 300 K
    T, Tr, S = 300.0, eos.T_ref, eos.S_suth
    assert eos.viscosity(T) == pytest.approx(eos.mu_ref * (T / Tr) ** 1.5 * (Tr + S) / (T + S), rel=1e-13)
