"""The pipeline's marching cubes skips the bricks its integrations never
updated (brick occupancy).  Every cube of such a brick has a zero-weight
corner, so the culled extraction must equal the full extraction of the same
volume bit for bit: vertices, triangle count and order."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _full_mc(h):
    """nrf_mc_extract / nrf_mc16_extract over the pipeline's volume (every
    brick classified)."""
    import torch

    from cudafusion_amd import _lib as L
    lib = h.lib
    ws = torch.empty(lib.nrf_mc_workspace_size(C.byref(h.vp)), dtype=torch.uint8, device="cuda")
    T = h.cfg.max_triangles
    v = torch.empty(T * 9, dtype=torch.float32, device="cuda")
    t = torch.empty(T * 3, dtype=torch.int32, device="cuda")
    cnt = torch.zeros(2, dtype=torch.int32, device="cuda")
    if h.storage16:
        t16, w8 = C.c_void_p(), C.c_void_p()
        L.check(lib.nrf_pipeline_volume16(h.pipe, C.byref(t16), C.byref(w8), None, None), "volume16")
        L.check(lib.nrf_mc16_extract(C.byref(h.vp), t16, w8, ws.data_ptr(), ws.numel(), v.data_ptr(), t.data_ptr(),
                                     T, cnt.data_ptr(), None), "mc16")
    else:
        L.check(lib.nrf_mc_extract(C.byref(h.vp), lib.nrf_pipeline_voxels(h.pipe), ws.data_ptr(), ws.numel(),
                                   v.data_ptr(), None, t.data_ptr(), T, cnt.data_ptr(), None), "mc")
    torch.cuda.synchronize()
    nt = int(cnt[1].item())
    assert 0 < nt <= T
    return v[: nt * 9].cpu().numpy().reshape(-1, 3)


@pytest.mark.parametrize("kind,storage16", [("sphere", False), ("body", False), ("body", True)])
def test_culled_mc_equals_full_extraction(gpu, kind, storage16):
    import torch

    from cudafusion_amd.harness import Harness
    kw = dict(jtj_precision=3, precond_fp64=True, refine_jacobi=True) if storage16 else {}
    h = Harness(kind, frames=3, arap_consistent_sign=True, validate=False, pcg_pipelined=3, storage16=storage16,
                **kw)
    h.setup()
    for f in (1, 2, 3):
        h.step(f)
        torch.cuda.synchronize()
        culled = h.mesh()
        full = _full_mc(h)
        print(f"[occupancy] {kind} storage16={storage16} frame {f}: {len(full) // 3} triangles", flush=True)
        assert culled.shape == full.shape, (culled.shape, full.shape)
        assert np.array_equal(culled.view(np.uint32), full.view(np.uint32))
    h.close()


@pytest.mark.parametrize("kind", ["sphere", "body"])
def test_culled_integration_equals_full(gpu, kind):
    """The pipeline's integration skips the waves whose voxels integrate
    unwarped and cannot reach a usable depth within the truncation (depth
    block maxima).  Its volume must equal nrf_tsdf_integrate (every voxel
    visited) applied to the volume before the step with the step's
    transforms, bit for bit, for the rigid frame 0 and warped frames."""
    import torch

    from cudafusion_amd import _lib as L
    from cudafusion_amd.harness import Harness, identity_pose
    h = Harness(kind, frames=3, arap_consistent_sign=True, validate=False, pcg_pipelined=3)
    h.setup()
    lib = h.lib
    dev = torch.device("cuda")
    V = int(np.prod(h.w.dims))
    st = torch.cuda.current_stream().cuda_stream
    # frame 0 (rigid): the pipeline's integration against a fresh full one
    ref = torch.empty(V * 12, dtype=torch.uint8, device=dev)
    L.check(lib.nrf_tsdf_reset(C.byref(h.vp), ref.data_ptr(), st), "reset")
    pose = identity_pose()
    L.check(lib.nrf_tsdf_integrate(C.byref(h.vp), ref.data_ptr(), h.depth[0].data_ptr(), C.byref(h.cam),
                                   pose.ctypes.data, None, None, 0, None, None, None, None, 0, None, st), "rigid")
    torch.cuda.synchronize()
    got = h.download_voxels()
    exp = ref.cpu().numpy().view(L.VOXEL)
    assert np.array_equal(got.view(np.uint8), exp.view(np.uint8)), "rigid frame 0"
    knn = lib.nrf_pipeline_knn(h.pipe)
    knn_w = lib.nrf_pipeline_knn_w(h.pipe)
    cen = torch.empty(h.N * 32, dtype=torch.uint8, device=dev)
    for f in (1, 2, 3):
        before = torch.from_numpy(h.download_voxels().view(np.uint8).copy()).to(dev)
        h.step(f)
        torch.cuda.synchronize()
        nodes = lib.nrf_pipeline_nodes(h.pipe)
        T = lib.nrf_pipeline_transforms(h.pipe)
        L.check(lib.nrf_warp_centered_table(nodes, T, h.N, cen.data_ptr(), st), "centered")
        L.check(lib.nrf_tsdf_integrate(C.byref(h.vp), before.data_ptr(), h.depth[f].data_ptr(), C.byref(h.cam),
                                       pose.ctypes.data, nodes, T, h.N, cen.data_ptr(), knn, knn_w, None,
                                       h.cfg.min_opt_count, None, st), "warped")
        torch.cuda.synchronize()
        got = h.download_voxels()
        exp = before.cpu().numpy().view(L.VOXEL)
        upd = int((got["weight"] != 0).sum())
        print(f"[cull] {kind} frame {f}: {upd} voxels with weight", flush=True)
        assert np.array_equal(got.view(np.uint8), exp.view(np.uint8)), f"frame {f}"
    h.close()
