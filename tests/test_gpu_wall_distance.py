"""Wall distance on the device (rsg_wall_distance; Mesh.compute_wall_distance(
on_device=True)) against Mesh::computeWallDistance (core/src/Mesh.cpp:345-373).

* meshes of every factory: the device distances equal, bit for bit, the host
  mirror's and the oracle's (om_wall_distance, itself checked against the
  compiled reference in test_oracle_vs_ref.py);
* point clouds through the C-ABI at sizes around the kernel's tiles (512 cells
  per workgroup, 1024 wall points per LDS tile): bit-equal to a brute-force
  search in the reference's operation order (norm of the difference, no
  fused multiply-add, minimum of the norms);
* the no-wall case (every distance 1e10) and an empty mesh.
"""
import ctypes as C

import numpy as np
import pytest

import rsg_abi as A

pytestmark = pytest.mark.gpu


def brute(cc, wp, chunk=256):
    out = np.empty(len(cc))
    for s in range(0, len(cc), chunk):
        d = cc[s:s + chunk, None, :] - wp[None, :, :]
        out[s:s + chunk] = np.min(np.sqrt(d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1] + d[..., 2] * d[..., 2]),
                                  axis=1)
    return out


@pytest.mark.parametrize("case", ["c0", "channel_bounded", "couette", "bfs", "plate", "hill"])
def test_mesh_wall_distance_matches_host_and_oracle(rs, oracle, meshes, case):
    path = meshes[case]
    m = rs.Mesh.load_from_file(path)
    m.compute_wall_distance()
    host = np.array(m.wall_distance(), copy=True)
    m.compute_wall_distance(on_device=True)
    dev = np.array(m.wall_distance(), copy=True)
    om = oracle.om_mesh_load(path.encode())
    try:
        wo = np.zeros(m.n_cells())
        oracle.om_wall_distance(C.c_void_p(om), A.dptr(wo))
    finally:
        oracle.om_mesh_free(C.c_void_p(om))
    assert np.array_equal(dev, host), np.max(np.abs(dev - host))
    assert np.array_equal(dev, wo)


@pytest.mark.parametrize("nc,nw", [(1, 1), (255, 3), (513, 1023), (1000, 1024), (1025, 1025), (3000, 2500),
                                   (70001, 3001)])
def test_point_cloud_matches_brute_force(rsg_lib, gpu_ctx, nc, nw):
    rng = np.random.default_rng(nc + 7 * nw)
    cc = rng.uniform(-1.0, 1.0, (nc, 3)) * np.array([3.0, 0.5, 0.01])
    wp = rng.uniform(-1.0, 1.0, (nw, 3)) * np.array([3.0, 0.5, 0.01])
    wp[: nw // 2, 1] = -0.5                                        # points on the two walls
    wp[nw // 2:, 1] = 0.5
    out = np.full(nc, -1.0)
    secs = np.zeros(1)
    print("point cloud", nc, nw, flush=True)
    rc = rsg_lib.rsg_wall_distance(gpu_ctx, nc, A.dptr(np.ascontiguousarray(cc)), nw,
                                   A.dptr(np.ascontiguousarray(wp)), A.dptr(out), A.dptr(secs))
    assert rc == 0, rsg_lib.rsg_last_error(gpu_ctx)
    ref = brute(cc, wp)
    assert np.array_equal(out, ref), (np.argmax(np.abs(out - ref)), np.max(np.abs(out - ref)))
    assert secs[0] > 0.0


def test_no_wall_and_empty(rsg_lib, gpu_ctx, rs):
    cc = np.zeros((5, 3)); out = np.zeros(5)
    assert rsg_lib.rsg_wall_distance(gpu_ctx, 5, A.dptr(cc), 0, None, A.dptr(out), None) == 0
    assert np.all(out == 1e10)
    assert rsg_lib.rsg_wall_distance(gpu_ctx, 0, None, 0, None, None, None) == 0
    assert rsg_lib.rsg_wall_distance(gpu_ctx, 5, None, 1, A.dptr(cc), A.dptr(out), None) == -1
    # a mesh with its walls retyped: no wall face left
    m = rs.Mesh.make_channel_2d(4, 6, 1.0, 1.0)
    for name in m.patch_names():
        m.set_patch_type(name, "patch")
    m.compute_wall_distance(on_device=True)
    assert np.all(np.asarray(m.wall_distance()) == 1e10)
