"""GPU parity of the projective data association (nrf_depth_normals,
nrf_pixel_knn, nrf_find_correspondences, nrf_mark_optimized_voxels) against
the FP32 oracle (oracle/association.cpp) on the 128^3 / 320x240 sphere
workload: the live surface is the canonical surface of frame f (the live
surface prediction of a frame whose warp is still the identity), the depth
is frame f's noisy depth.  Bars: ids, validity flags, counts and stats
bit-exact; vectors and weights to 1e-6 (same FP32 operation order; the only
differences are the FP32 division / square-root implementations, both
correctly rounded)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scene(gpu, orc):
    import torch

    from cudafusion_amd import _lib as L
    from cudafusion_amd import scene as S
    from cudafusion_amd.fusion import TSDFVolume, WarpField
    sc = S.make_scene("sphere", frames=3)
    w = sc.w
    frames = [sc.render(f) for f in range(3)]
    nodes = S.scene_nodes(sc)
    vol = TSDFVolume(TSDFVolume.Params(w.dims, w.voxel_size, w.truncation, w.origin, w.decay_alpha, w.max_weight))
    cam = L.camera(w.fx, w.fy, w.cx, w.cy, w.width, w.height)
    pose = np.eye(4, dtype=np.float32).reshape(16)
    vol.integrate(torch.from_numpy(frames[0].depth.reshape(-1)).cuda(), None, cam, pose)
    wf = WarpField(w.node_spacing)
    wf.upload_nodes(nodes)
    knn = wf.compute_voxel_knn(vol)
    return dict(w=w, frames=frames, cam=cam, pose=pose, vol=vol, knn=knn, nodes=nodes, torch=torch)


def _live_maps(sc, f):
    w = sc["w"]
    fr = sc["frames"][f]
    valid = fr.valid.reshape(-1).astype(bool)
    live = np.where(valid[:, None], fr.canonical.reshape(-1, 3), 0).astype(np.float32)
    nrm = np.where(valid[:, None], fr.normal.reshape(-1, 3), 0).astype(np.float32)
    o = np.asarray(w.origin, np.float32)
    ijk = np.floor((fr.canonical.reshape(-1, 3) - o) / np.float32(w.voxel_size)).astype(np.int64)
    ok = valid & np.all((ijk >= 0) & (ijk < np.asarray(w.dims)), axis=1)
    hit = np.where(ok, ijk[:, 0] + w.dims[0] * (ijk[:, 1] + w.dims[1] * ijk[:, 2]), -1).astype(np.int32)
    return live, nrm, hit


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("radius", [3, 8])
def test_association_matches_oracle(scene, orc, radius):
    from cudafusion_amd import _lib as L
    torch = scene["torch"]
    lib = L.load()
    w, cam, pose = scene["w"], scene["cam"], scene["pose"]
    npx = w.width * w.height
    vknn, vknn_w = (t.cpu().numpy() for t in scene["knn"])
    V = int(np.prod(w.dims))
    counts = torch.zeros(V, dtype=torch.int32, device="cuda")
    last = torch.full((V,), -1, dtype=torch.int32, device="cuda")
    h_counts = np.zeros(V, np.int32)
    h_last = np.full(V, -1, np.int32)
    ap = L.assoc_params(0.05, 0.7, 0.2, radius)
    for f in (1, 2):
        depth = scene["frames"][f].depth.reshape(-1).astype(np.float32)
        live, nrm, hit = _live_maps(scene, f)
        d_depth, d_live, d_nrm, d_hit = (_dev(torch, a) for a in (depth, live, nrm, hit))
        d_dn = torch.empty(npx * 3, dtype=torch.float32, device="cuda")
        L.check(lib.nrf_depth_normals(d_depth.data_ptr(), C.byref(cam), d_dn.data_ptr(), None), "depth_normals")
        d_pk = torch.empty(npx * 4, dtype=torch.int32, device="cuda")
        d_pw = torch.empty(npx * 4, dtype=torch.float32, device="cuda")
        L.check(lib.nrf_pixel_knn(d_hit.data_ptr(), npx, scene["knn"][0].data_ptr(), scene["knn"][1].data_ptr(),
                                  d_pk.data_ptr(), d_pw.data_ptr(), None), "pixel_knn")
        d_corr = torch.empty(npx * 76, dtype=torch.uint8, device="cuda")
        d_nv = torch.zeros(1, dtype=torch.int32, device="cuda")
        d_st = torch.zeros(6, dtype=torch.int32, device="cuda")
        L.check(lib.nrf_find_correspondences(d_live.data_ptr(), d_nrm.data_ptr(), d_depth.data_ptr(), d_dn.data_ptr(),
                                             C.byref(cam), pose.ctypes.data, d_pk.data_ptr(), d_pw.data_ptr(),
                                             C.byref(ap), d_corr.data_ptr(), d_nv.data_ptr(), d_st.data_ptr(), None),
                "find_correspondences")
        L.check(lib.nrf_mark_optimized_voxels(d_corr.data_ptr(), d_hit.data_ptr(), npx, C.byref(scene["vol"].vp), f,
                                              counts.data_ptr(), last.data_ptr(), None), "mark_optimized_voxels")
        torch.cuda.synchronize()
        # oracle
        dn = orc.depth_normals_f32(depth, cam)
        np.testing.assert_allclose(d_dn.cpu().numpy(), dn, atol=1e-6)
        pk, pw = orc.pixel_knn(hit, vknn, vknn_w)
        np.testing.assert_array_equal(d_pk.cpu().numpy(), pk)
        np.testing.assert_array_equal(d_pw.cpu().numpy(), pw)
        rc, rn, rs = orc.find_correspondences_f32(live, nrm, depth, dn, cam, pose, pk, pw, ap)
        gc = d_corr.cpu().numpy().view(L.CORR)
        print(f"[assoc r={radius} f{f}] valid {int(d_nv.item())} (oracle {rn}), stats {d_st.cpu().numpy().tolist()}",
              flush=True)
        assert int(d_nv.item()) == rn and rn > 0.5 * int(scene["frames"][f].valid.sum())
        np.testing.assert_array_equal(d_st.cpu().numpy(), rs)
        np.testing.assert_array_equal(gc["valid"], rc["valid"])
        np.testing.assert_array_equal(gc["node_ids"], rc["node_ids"])
        np.testing.assert_array_equal(gc["node_ws"], rc["node_ws"])
        for field in ("src", "dst", "normal", "weight"):
            np.testing.assert_allclose(gc[field], rc[field], atol=1e-6, err_msg=field)
        orc.mark_optimized_voxels(rc, hit, w.dims, f, h_counts, h_last)
        np.testing.assert_array_equal(counts.cpu().numpy(), h_counts)
        np.testing.assert_array_equal(last.cpu().numpy(), h_last)
    assert h_counts.max() == 2


def test_skin_and_raster_match_oracle(scene, orc):
    """The frame-0 mesh skinned through the voxel kNN with perturbed node
    transforms (nrf_skin_vertices) and rasterised into the camera
    (nrf_rasterize_mesh), against the oracle: skinned vertices to 1e-6; with
    the device's skinned vertices as input, the winning triangle of every
    pixel is the oracle's (hit voxels bit-exact, covered pixels equal) and
    the live vertices, normals and canonical points agree to 1e-6."""
    from cudafusion_amd import _lib as L
    torch = scene["torch"]
    lib = L.load()
    cam = scene["cam"]
    vol = scene["vol"]
    verts, tris = vol.extract_surface_device()
    nv, nt = verts.shape[0], tris.shape[0]
    assert nt > 1000
    nodes = scene["nodes"]
    N = len(nodes)
    rng = np.random.default_rng(4)
    T = np.zeros(N, L.DUALQUAT)
    T["real"][:, 3] = 1
    for i in range(N):  # small random node motions
        tw = np.concatenate([rng.normal(0, 0.01, 3), rng.normal(0, 0.003, 3)]).astype(np.float32)
        out = np.zeros(1, L.DUALQUAT)
        orc.load().orc_twist_update_f32(tw.ctypes.data, T[i:i + 1].copy().ctypes.data, 0.0, 0.0, 1.0,
                                        out.ctypes.data)
        T[i] = out[0]
    d_nodes = _dev(torch, nodes.view(np.uint8))
    d_T = _dev(torch, T.view(np.uint8))
    d_cen = torch.empty(N * 32, dtype=torch.uint8, device="cuda")
    L.check(lib.nrf_warp_centered_table(d_nodes.data_ptr(), d_T.data_ptr(), N, d_cen.data_ptr(), None), "centered")
    knn, knn_w = scene["knn"]
    d_warp = torch.empty(nv * 3, dtype=torch.float32, device="cuda")
    L.check(lib.nrf_skin_vertices(verts.data_ptr(), nv, C.byref(vol.vp), d_cen.data_ptr(), N, knn.data_ptr(),
                                  knn_w.data_ptr(), d_warp.data_ptr(), None), "skin")
    pose = scene["pose"]
    npx = cam.width * cam.height
    ws = torch.empty(lib.nrf_raster_workspace_size(C.byref(cam)), dtype=torch.uint8, device="cuda")
    lv, ln, cp = (torch.empty(npx * 3, dtype=torch.float32, device="cuda") for _ in range(3))
    hit = torch.empty(npx, dtype=torch.int32, device="cuda")
    L.check(lib.nrf_rasterize_mesh(verts.data_ptr(), d_warp.data_ptr(), tris.data_ptr(), nv, nt, C.byref(cam),
                                   pose.ctypes.data, C.byref(vol.vp), ws.data_ptr(), ws.numel(), lv.data_ptr(),
                                   ln.data_ptr(), hit.data_ptr(), cp.data_ptr(), None), "raster")
    torch.cuda.synchronize()
    h_verts = verts.cpu().numpy().reshape(-1)
    ref_w = orc.skin_vertices_f32(vol.vp, h_verts, nodes, T, knn.cpu().numpy(), knn_w.cpu().numpy())
    g_w = d_warp.cpu().numpy()
    moved = np.abs(ref_w - h_verts).max()
    np.testing.assert_allclose(g_w, ref_w, atol=1e-6)
    rlv, rln, rhit, rcp = orc.rasterize_f32(h_verts, g_w, tris.cpu().numpy().reshape(-1), cam, pose, vol.vp)
    ghit = hit.cpu().numpy()
    print(f"[raster] {nt} triangles, skinning moved vertices by up to {moved:.2e} m, covered pixels "
          f"{int((ghit >= 0).sum())} (oracle {int((rhit >= 0).sum())})", flush=True)
    assert moved > 1e-4 and (ghit >= 0).sum() > 5000
    np.testing.assert_array_equal(ghit, rhit)
    np.testing.assert_allclose(lv.cpu().numpy(), rlv, atol=1e-6)
    np.testing.assert_allclose(ln.cpu().numpy(), rln, atol=1e-6)
    np.testing.assert_allclose(cp.cpu().numpy(), rcp, atol=1e-6)
