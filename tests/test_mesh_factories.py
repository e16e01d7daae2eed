"""Product mesh factories (rans_sst_py.Mesh) against the reference's own
factories: the MSH1 files must be byte-identical (same face order,
owner/neighbour, patch lists and every geometric double).  Plus the
geometric invariants the reference's tests check."""
import ctypes as C
import filecmp

import numpy as np
import pytest

import rsg_abi as A

D = lambda a: (C.c_double * len(a))(*a)
I = lambda a: (C.c_int * len(a))(*a)


@pytest.mark.parametrize("kind,ip,dp,make", [
    (0, [12, 10], [2.0, 1.0], lambda rs: rs.Mesh.make_channel_2d(12, 10, 2.0, 1.0)),
    (1, [16, 12], [4.0, 1.0, 5000.0, 1.0], lambda rs: rs.Mesh.make_channel_2d(16, 12, 4.0, 1.0, 5000.0, 1.0)),
    (1, [8, 48], [1.5, 2.0, 40000.0, 0.3], lambda rs: rs.Mesh.make_channel_2d(8, 48, 1.5, 2.0, 40000.0, 0.3)),
    (1, [4, 8], [1.0, 2.0, 50.0, 1.0], lambda rs: rs.Mesh.make_channel_2d(4, 8, 1.0, 2.0, 50.0, 1.0)),
    (2, [12, 16], [2.0, 1.0, 1e5, 1.0], lambda rs: rs.Mesh.make_plate_2d(12, 16, 2.0, 1.0, 1e5, 1.0)),
    (3, [6, 12, 8, 6], [2.0, 6.0, 1.0, 2.0], lambda rs: rs.Mesh.make_backward_facing_step_2d(6, 12, 8, 6, 2.0, 6.0, 1.0, 2.0)),
    (4, [6, 12, 8, 6], [2.0, 6.0, 1.0, 2.0, 5000.0, 1.0],
     lambda rs: rs.Mesh.make_backward_facing_step_2d(6, 12, 8, 6, 2.0, 6.0, 1.0, 2.0, 5000.0, 1.0)),
])
def test_factory_byte_identical(ref, rs, tmp_path, kind, ip, dp, make):
    pr, pp = str(tmp_path / "ref.msh"), str(tmp_path / "prod.msh")
    assert ref.ref_make_mesh(kind, I(ip), D(dp), pr.encode()) == 0
    make(rs).save_binary(pp)
    assert filecmp.cmp(pr, pp, shallow=False)


def test_curved_periodic_byte_identical(ref, rs, tmp_path):
    x = np.linspace(0.0, 2.0, 9)
    yb = 0.1 * (1.0 - np.cos(2 * np.pi * x / 2.0))
    pr, pp = str(tmp_path / "ref.msh"), str(tmp_path / "prod.msh")
    dp = [1.0, 2000.0, 1.0] + list(x) + list(yb)
    assert ref.ref_make_mesh(5, I([8, 10]), D(dp), pr.encode()) == 0
    rs.Mesh.make_curved_channel_periodic_2d(x, yb, 1.0, 10, 2000.0, 1.0).save_binary(pp)
    assert filecmp.cmp(pr, pp, shallow=False)


def test_periodic_channel_invariants(rs, tmp_path):
    # tests/cpp/test_periodic_channel.cpp:49-87
    nx, Lx = 36, 3.0
    x = np.linspace(0.0, Lx, nx + 1)
    yb = 0.2 * 0.5 * (1.0 - np.cos(2 * np.pi * x / Lx))
    m = rs.Mesh.make_curved_channel_periodic_2d(x, yb, 1.0, 24, 1000.0, 1.0)
    vol = m.cell_volumes().sum()
    trap = np.sum(0.5 * ((1.0 - yb[:-1]) + (1.0 - yb[1:])) * np.diff(x))
    assert abs(vol - trap) < 1e-10
    path = str(tmp_path / "p.msh"); m.save_binary(path)
    msh = A.Msh(path)
    assert np.allclose(np.linalg.norm(msh.fn, axis=1), 1.0, atol=1e-14)
    wrap = slice((nx - 1) * 24, (nx - 1) * 24 + 24)
    assert np.all(msh.delta[wrap] < 2 * Lx / nx)


def test_geometric_channel_c0(rs):
    y = np.array(rs.Mesh.geometric_channel_nodes(64, 1.06, 1.0))
    h = np.diff(y)
    assert len(y) == 129 and y[0] == 0.0 and y[64] == 1.0 and y[-1] == 2.0
    assert np.allclose(h[1:64] / h[:63], 1.06, rtol=1e-12)
    assert np.allclose(h, h[::-1], rtol=0, atol=1e-15)
    assert abs(h[0] - 1.0 / sum(1.06 ** j for j in range(64))) < 1e-16
    m = rs.Mesh.make_channel_periodic_2d(np.linspace(0, 1, 5), y)
    assert (m.n_cells(), m.n_faces(), m.n_internal_faces()) == (512, 1028, 1020)


def test_bfs_counts(rs):
    # tests/cpp/test_mesh_bfs.cpp:31-111
    nxu, nxd, nyu, nyd = 6, 12, 8, 6
    m = rs.Mesh.make_backward_facing_step_2d(nxu, nxd, nyu, nyd, 2.0, 6.0, 1.0, 2.0)
    assert m.n_cells() == (nxu + nxd) * nyu + nxd * nyd
    assert m.n_patches() == 6
    assert m.n_internal_faces() == (nxu + nxd) * (nyu - 1) + nxd * (nyd - 1) + nxd + (nxu + nxd - 1) * nyu + (nxd - 1) * nyd
    assert np.all(m.cell_volumes() > 0)
    m.compute_wall_distance()
    assert np.all(np.isfinite(m.wall_distance()))


def test_column_mesh_drops_streamwise_faces(rs):
    y = np.array(rs.Mesh.geometric_channel_nodes(64, 1.06, 1.0))
    m = rs.Mesh.make_channel_column_1d(y, 0.25)
    assert m.n_cells() == 128 and m.n_internal_faces() == 127 and m.n_faces() == 129
