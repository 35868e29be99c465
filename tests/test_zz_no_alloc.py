"""nrf_mc_extract_noalloc (the workspace-free marching cubes) against
nrf_mc_extract with a caller workspace: identical counts, vertices, normals
and triangles, bit for bit, on two calls in a row, the second on a smaller,
odd-sized volume (so the scratch holds the first call's rows), and the
capacity clamp.  All calls go through the C ABI."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _volume(torch, L, dims, voxel, centre, radii, seed):
    """Analytic ellipsoid TSDF (normalised, truncated at 3 voxels) with
    weight 0 outside the band and in a few random holes."""
    X, Y, Z = dims
    vp = L.volume_params(dims, voxel, 3 * voxel, (0.0, 0.0, 0.0))
    z, y, x = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    p = np.stack([(x + 0.5) * voxel, (y + 0.5) * voxel, (z + 0.5) * voxel], -1) - np.asarray(centre)
    q = np.sqrt(((p / np.asarray(radii)) ** 2).sum(-1))
    sdf = (1.0 - q) * min(radii)
    t = np.clip(sdf / (3 * voxel), -1, 1).astype(np.float32)
    w = (np.abs(sdf) < 4 * voxel).astype(np.float32)
    rng = np.random.default_rng(seed)
    w.reshape(-1)[rng.integers(0, w.size, w.size // 200)] = 0.0
    vox = np.zeros(X * Y * Z, L.VOXEL)
    vox["tsdf"] = t.reshape(-1)
    vox["weight"] = w.reshape(-1)
    dev = torch.from_numpy(vox.view(np.uint8).copy()).cuda()
    return vp, dev


def _extract(torch, lib, L, vp, vox, noalloc, max_tris):
    counts = torch.zeros(4, dtype=torch.int32, device="cuda")
    verts = torch.full((max(max_tris, 1) * 9,), -7.0, dtype=torch.float32, device="cuda")
    nrm = torch.full((max(max_tris, 1) * 9,), -7.0, dtype=torch.float32, device="cuda")
    tris = torch.full((max(max_tris, 1) * 3,), -7, dtype=torch.int32, device="cuda")
    if noalloc:
        rc = lib.nrf_mc_extract_noalloc(C.byref(vp), vox.data_ptr(), verts.data_ptr(), nrm.data_ptr(),
                                        tris.data_ptr(), max_tris, counts.data_ptr(), None)
        L.check(rc, "nrf_mc_extract_noalloc")
    else:
        size = lib.nrf_mc_workspace_size(C.byref(vp))
        ws = torch.empty(size, dtype=torch.uint8, device="cuda")
        rc = lib.nrf_mc_extract(C.byref(vp), vox.data_ptr(), ws.data_ptr(), size, verts.data_ptr(), nrm.data_ptr(),
                                tris.data_ptr(), max_tris, counts.data_ptr(), None)
        L.check(rc, "nrf_mc_extract")
    torch.cuda.synchronize()
    return counts.cpu().numpy(), verts.cpu().numpy(), nrm.cpu().numpy(), tris.cpu().numpy()


def _count(torch, lib, L, vp, vox):
    return int(_extract(torch, lib, L, vp, vox, False, 0)[0][1])


def test_noalloc_matches_workspace_form(gpu):
    import torch

    from cudafusion_amd import _lib as L
    lib = L.load()
    cases = [((128, 128, 128), 0.006, (0.39, 0.38, 0.40), (0.25, 0.22, 0.30), 1),
             ((70, 61, 53), 0.008, (0.27, 0.25, 0.21), (0.18, 0.15, 0.12), 2)]
    for dims, voxel, centre, radii, seed in cases:
        vp, vox = _volume(torch, L, dims, voxel, centre, radii, seed)
        T = _count(torch, lib, L, vp, vox)
        print(f"[noalloc] dims {dims}: {T} triangles", flush=True)
        assert T > 1000
        ref = _extract(torch, lib, L, vp, vox, False, T)
        got = _extract(torch, lib, L, vp, vox, True, T)
        assert got[0][0] == ref[0][0] == 3 * T and got[0][1] == ref[0][1] == T
        for a, b in zip(got[1:], ref[1:]):
            np.testing.assert_array_equal(a, b)
        # capacity clamp: the full count is reported, only max_tris written
        cap = T // 3
        ref_c = _extract(torch, lib, L, vp, vox, False, cap)
        got_c = _extract(torch, lib, L, vp, vox, True, cap)
        assert got_c[0][1] == ref_c[0][1] == T
        for a, b in zip(got_c[1:], ref_c[1:]):
            np.testing.assert_array_equal(a, b)


def test_noalloc_rejects_oversized_volume(gpu):
    from cudafusion_amd import _lib as L
    lib = L.load()
    vp = L.volume_params((8, 600, 600), 0.01, 0.03, (0, 0, 0))
    counts = C.c_int32 * 2
    rc = lib.nrf_mc_extract_noalloc(C.byref(vp), C.c_void_p(256), None, None, None, 0, C.cast(counts(), C.c_void_p),
                                    None)
    assert rc == 2  # NRF_ERR_TOO_LARGE, before any device work
    assert b"262144" in lib.nrf_last_error()
