"""GPU raycast (nrf_raycast, TSDFVolume::raycast) against the FP32 oracle
(oracle/oracle.cpp orc_raycast_f32) on the 128^3 sphere after the rigid
frame 0, unwarped and warped through the voxel kNN with perturbed node
transforms.  Same FP32 operation order without contraction on both sides:
hit voxels bit-exact, vertices and normals within 1e-6."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_raycast_matches_oracle(gpu, orc):
    import torch

    from cudafusion_amd import _lib as L
    from cudafusion_amd import scene as S
    from cudafusion_amd.fusion import TSDFVolume, WarpField
    sc = S.make_scene("sphere", frames=1)
    w = sc.w
    fr = sc.render(0)
    vol = TSDFVolume(TSDFVolume.Params(w.dims, w.voxel_size, w.truncation, w.origin, w.decay_alpha, w.max_weight))
    cam = L.camera(w.fx, w.fy, w.cx, w.cy, w.width, w.height)
    pose = np.eye(4, dtype=np.float32).reshape(16)
    vol.integrate(torch.from_numpy(fr.depth.reshape(-1)).cuda(), None, cam, pose)
    nodes = S.scene_nodes(sc)
    wf = WarpField(w.node_spacing)
    wf.upload_nodes(nodes)
    knn, knn_w = wf.compute_voxel_knn(vol)
    vox = vol.device_data().cpu().numpy().view(L.VOXEL)
    hk, hw = knn.cpu().numpy(), knn_w.cpu().numpy()
    N = len(nodes)
    rng = np.random.default_rng(7)
    T = np.zeros(N, L.DUALQUAT)
    T["real"][:, 3] = 1
    for i in range(N):
        tw = np.concatenate([rng.normal(0, 0.02, 3), rng.normal(0, 0.004, 3)]).astype(np.float32)
        out = np.zeros(1, L.DUALQUAT)
        orc.load().orc_twist_update_f32(tw.ctypes.data, T[i:i + 1].copy().ctypes.data, 0.0, 0.0, 1.0,
                                        out.ctypes.data)
        T[i] = out[0]
    d_nodes = torch.from_numpy(nodes.view(np.uint8).copy()).cuda()
    d_T = torch.from_numpy(T.view(np.uint8).copy()).cuda()
    for warped in (False, True):
        if warped:
            gv, gn, gi = vol.raycast(cam, pose, d_nodes.data_ptr(), d_T.data_ptr(), N, knn.data_ptr(),
                                     knn_w.data_ptr(), out_canonical_vidx=True)
            rv, rn, ri = orc.raycast_f32(vol.vp, vox, cam, pose, nodes, T, hk, hw)
        else:
            gv, gn, gi = vol.raycast(cam, pose, out_canonical_vidx=True)
            rv, rn, ri = orc.raycast_f32(vol.vp, vox, cam, pose)
        torch.cuda.synchronize()
        gi = gi.cpu().numpy()
        hits = int((gi >= 0).sum())
        print(f"[raycast warped={warped}] hits {hits} (oracle {int((ri >= 0).sum())}), valid px "
              f"{int(fr.valid.sum())}", flush=True)
        assert hits > 0.6 * int(fr.valid.sum())  # grazing and silhouette rays meet unobserved voxels
        np.testing.assert_array_equal(gi, ri)
        np.testing.assert_allclose(gv.cpu().numpy().reshape(-1), rv, atol=1e-6)
        np.testing.assert_allclose(gn.cpu().numpy().reshape(-1), rn, atol=1e-6)
        if not warped:
            # the surface is the scene's sphere (r = 0.25 m at z = 1, bent by at most 1 cm)
            p = gv.cpu().numpy()[gi >= 0] - np.array([0.0, 0.0, 1.0])
            assert np.abs(np.linalg.norm(p, axis=1) - 0.25).max() < 0.01 + 2 * w.voxel_size
