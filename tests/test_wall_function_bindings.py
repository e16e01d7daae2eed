"""Wall-function settings and wall-patch introspection through rans_sst_py on
the host (no GPU), with the contract of the reference's tests/python/
test_wall_functions.py:93-156 (SolverSettings wall-function fields,
FlowFields.hpp; Mesh::wallPatchData, python/Bindings.cpp)."""
import numpy as np
import pytest


def test_wall_function_settings_defaults_and_setters(rs):
    s = rs.SolverSettings()
    assert s.use_wall_functions is False
    assert s.von_karman == pytest.approx(0.41) and s.wall_fn_E == pytest.approx(9.0)
    s.use_wall_functions = True
    s.von_karman = 0.42
    s.wall_fn_E = 8.5
    assert (s.use_wall_functions, s.von_karman, s.wall_fn_E) == (True, 0.42, 8.5)


@pytest.fixture(scope="module")
def channel(rs):
    eos = rs.IdealGasEOS()
    T, p = 300.0, 101325.0
    Uin = 0.1 * eos.sound_speed(T)
    Re = eos.density(p, T) * Uin / eos.viscosity(T)
    return rs.Mesh.make_channel_2d(20, 14, 10.0, 1.0, Re=Re, yPlusTarget=1.0)


def test_wall_patches(channel):
    names, types = channel.patch_names(), channel.patch_types()
    assert "bottom" in names and "top" in names
    assert all(t == "wall" for n, t in zip(names, types) if n in ("bottom", "top"))


def test_wall_patch_data(channel):
    info = channel.wall_patch_data("bottom")
    for k in ("name", "type", "n_faces", "owner", "delta", "center", "normal", "area"):
        assert k in info, k
    assert (info["name"], info["type"], info["n_faces"]) == ("bottom", "wall", 20)
    owners = np.asarray(info["owner"], int)
    assert np.all((owners >= 0) & (owners < channel.n_cells()))
    # the bottom wall's faces face -y, and their owners are the first cell row
    assert np.allclose(np.asarray(info["normal"]).reshape(-1, 3)[:, 1], -1.0)
    assert np.all(np.asarray(info["delta"]) > 0)


def test_unknown_patch_raises(channel):
    with pytest.raises(RuntimeError, match="unknown patch"):
        channel.wall_patch_data("nope")
