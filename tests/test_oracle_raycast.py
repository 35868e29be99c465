"""Pins of the oracle's raycast (oracle/oracle.cpp orc_raycast_f32,
restating kernels/tsdf_kernels.cu:236-361 and sample_tsdf_trilinear :14-64)
by closed forms:

* a plane z = z0 whose TSDF is linear in the sample position origin + k s
  (the reference's trilinear reads voxel k there): every ray's vertex lies on
  the plane and its normal is (0, 0, -1), the gradient direction;
* the same plane stored at voxel centres (origin + (k + 1/2) s, where the
  integration evaluates the TSDF): the reference's sampling puts the surface
  half a voxel nearer the camera, at z0 - s/2;
* an unobserved slab in front of the plane: the rays through it still find
  the plane behind it (a failed sample only resets the previous one), and
  rays whose crossing falls into an unobserved region find nothing;
* a single node moving every voxel by a rigid motion: the warped vertex is
  the moved unwarped vertex and the normal is rotated with it.
"""
import numpy as np

from cudafusion_amd import _lib as L

DIMS = (40, 40, 48)
S = 0.01
TAU = 0.03
ORIGIN = (-0.2, -0.2, 0.5)
Z0 = 0.83
CAM = L.camera(100.0, 100.0, 16.0, 12.0, 32, 24)
POSE = np.eye(4, dtype=np.float32).reshape(16)


def _plane(center_offset=0.0, weight=1.0):
    vp = L.volume_params(DIMS, S, TAU, ORIGIN, 1.0, 50.0)
    z = np.float32(ORIGIN[2]) + (np.arange(DIMS[2], dtype=np.float32) + np.float32(center_offset)) * np.float32(S)
    t = np.clip((np.float32(Z0) - z) / np.float32(TAU), -1, 1).astype(np.float32)
    vox = np.zeros(int(np.prod(DIMS)), L.VOXEL)
    vox["tsdf"] = np.broadcast_to(t[:, None, None], DIMS[::-1]).reshape(-1)
    vox["weight"] = weight
    return vp, vox


def test_raycast_linear_plane(orc):
    vp, vox = _plane()
    v, n, idx = orc.raycast_f32(vp, vox, CAM, POSE)
    v, n = v.reshape(-1, 3), n.reshape(-1, 3)
    hit = idx >= 0
    assert hit.all()
    np.testing.assert_allclose(v[:, 2], Z0, atol=2e-6)
    np.testing.assert_allclose(n, np.broadcast_to([0, 0, -1], n.shape), atol=1e-6)
    # the vertex lies on the pixel's ray: x / z = (u - cx) / fx
    u = np.arange(CAM.width)[None, :].repeat(CAM.height, 0).reshape(-1)
    np.testing.assert_allclose(v[:, 0] / v[:, 2], (u - CAM.cx) / CAM.fx, atol=1e-6)
    # the base voxel of the crossing sample contains the point just behind the plane
    zi = idx // (DIMS[0] * DIMS[1])
    assert (np.abs(ORIGIN[2] + zi * S - Z0) <= S + 1e-6).all()


def test_raycast_half_voxel_offset_of_centre_values(orc):
    vp, vox = _plane(center_offset=0.5)
    v, _, idx = orc.raycast_f32(vp, vox, CAM, POSE)
    assert (idx >= 0).all()
    np.testing.assert_allclose(v.reshape(-1, 3)[:, 2], Z0 - S / 2, atol=2e-6)


def test_raycast_unobserved_regions(orc):
    vp, vox = _plane()
    w = vox["weight"].reshape(DIMS[::-1])
    w[10:14] = 0.0                      # a slab far in front of the plane (z 0.6-0.64 m)
    w[31:35, :, : DIMS[0] // 2] = 0.0   # the crossing region (z = 0.81-0.85 m) of the x < 0 half
    vox["weight"] = w.reshape(-1)
    v, _, idx = orc.raycast_f32(vp, vox, CAM, POSE)
    v = v.reshape(-1, 3)
    x_vox = np.floor((v[:, 0] - ORIGIN[0]) / S)
    hit = idx >= 0
    # rays whose crossing lies in the x >= 0 half hit the plane; the others find nothing
    u = np.arange(CAM.width)[None, :].repeat(CAM.height, 0).reshape(-1)
    left = (u - CAM.cx) / CAM.fx * Z0 < -0.02
    right = (u - CAM.cx) / CAM.fx * Z0 > 0.02
    assert hit[right].all() and not hit[left].any()
    np.testing.assert_allclose(v[hit, 2], Z0, atol=2e-6)
    assert (x_vox[hit] >= DIMS[0] // 2 - 1).all()
    assert (v[~hit] == 0).all()


def test_raycast_warp_single_node(orc):
    vp, vox = _plane()
    V = len(vox)
    node = np.zeros(1, L.NODE)
    node["pos"] = (0.0, 0.0, Z0)
    node["radius"] = 0.05
    knn = np.tile(np.array([0, -1, -1, -1], np.int32), V)
    knn_w = np.tile(np.array([1, 0, 0, 0], np.float32), V)
    T = np.zeros(1, L.DUALQUAT)
    T["real"][0, 3] = 1
    tw = np.array([0.0, 0.0, 0.1, 0.01, -0.02, 0.005], np.float32)
    out = np.zeros(1, L.DUALQUAT)
    orc.load().orc_twist_update_f32(tw.ctypes.data, T.ctypes.data, 0.0, 0.0, 1.0, out.ctypes.data)
    v0, n0, _ = orc.raycast_f32(vp, vox, CAM, POSE)
    v1, n1, idx = orc.raycast_f32(vp, vox, CAM, POSE, node, out, knn, knn_w)
    from cudafusion_amd.harness import dq_to_rt
    R, t = dq_to_rt(out.view(np.float32).reshape(8))
    c = np.array([0.0, 0.0, Z0])
    ref_v = (v0.reshape(-1, 3) - c) @ R.T + c + t   # rotation about the node, then its translation
    np.testing.assert_allclose(v1.reshape(-1, 3), ref_v, atol=2e-6)
    np.testing.assert_allclose(n1.reshape(-1, 3), n0.reshape(-1, 3) @ R.T, atol=2e-6)
