"""Pins the oracle's TSDF integration (kernels/tsdf_kernels.cu:85-174) and
marching cubes (src/marching_cubes.cu) with closed forms, invariants and
the reference's own tables and scenes."""
import os
import re

import numpy as np
import pytest

from cudafusion_amd import _lib as L

REF_MC = "/root/reference/src/marching_cubes.cu"


def _vp(dims=(8, 8, 8), s=0.01, tr=0.03, org=(-0.04, -0.04, 0.5), decay=1.0, wmax=50.0):
    return L.volume_params(dims, s, tr, org, decay, wmax)


def _fresh(n):
    v = np.zeros(n, L.VOXEL)
    v["tsdf"] = 1
    return v


def test_plane_closed_form(orc):
    """Fronto-parallel plane at depth d, identity pose: sdf = d - z_voxel."""
    vp = _vp((16, 16, 16), 0.01, 0.03, (-0.08, -0.08, 0.45))
    cam = L.camera(100.0, 100.0, 40.0, 40.0, 80, 80)
    d = 0.53
    depth = np.full(80 * 80, d, np.float32)
    pose = np.eye(4, dtype=np.float32).reshape(16)
    vox = _fresh(16 ** 3)
    n = orc.integrate_f32(vp, vox, depth, cam, pose)
    t64, w64 = np.ones(16 ** 3), np.zeros(16 ** 3)
    orc.integrate_f64(vp, t64, w64, depth, cam, pose)
    # the FP64 oracle evaluates in double from the float32 parameters
    z = np.float64(np.float32(0.45)) + (np.arange(16) + 0.5) * np.float64(np.float32(0.01))
    sdf = np.float64(np.float32(d)) - z
    exp_w = (sdf >= -0.03).astype(np.float64)
    exp_t = np.where(exp_w > 0, np.minimum(1.0, sdf / 0.03), 1.0)
    # every voxel of a z-slice projects inside the 80x80 image here
    tz = t64.reshape(16, 16, 16)
    wz = w64.reshape(16, 16, 16)
    for k in range(16):
        assert np.all(wz[k] == exp_w[k])
        np.testing.assert_allclose(tz[k], exp_t[k], atol=1e-12)
    assert n == int(exp_w.sum()) * 256
    np.testing.assert_allclose(vox["tsdf"], t64, atol=1e-6)
    np.testing.assert_array_equal(vox["weight"], w64.astype(np.float32))


def test_running_average_decay_and_cap(orc):
    vp = _vp((4, 4, 4), 0.01, 0.03, (-0.02, -0.02, 0.5), decay=0.5, wmax=1.6)
    cam = L.camera(100.0, 100.0, 20.0, 20.0, 40, 40)
    pose = np.eye(4, dtype=np.float32).reshape(16)
    t = np.ones(64)
    w = np.zeros(64)
    depths = [0.53, 0.54, 0.52]
    exp_t = np.ones(64)
    exp_w = np.zeros(64)
    zc = 0.5 + (np.arange(4) + 0.5) * 0.01
    for d in depths:
        orc.integrate_f64(vp, t, w, np.full(1600, d, np.float32), cam, pose)
        for k in range(4):
            sdf = np.float32(d) - zc[k]
            if sdf < -0.03:
                continue
            obs = min(1.0, sdf / 0.03)
            sl = slice(16 * k, 16 * k + 16)
            dw = exp_w[sl] * 0.5
            nw = dw + 1
            exp_t[sl] = (exp_t[sl] * dw + obs) / nw
            exp_w[sl] = np.minimum(nw, 1.6)
    np.testing.assert_allclose(t, exp_t, atol=1e-7)
    np.testing.assert_allclose(w, exp_w, atol=1e-12)
    # weights 1 -> 1.5 -> 1.75 are capped at 1.6 on the third frame
    assert np.all(exp_w == pytest.approx(1.6))


def test_projection_rounds_half_to_even(orc):
    """__float2int_rn: a voxel projecting to u = 2.5 samples pixel 2, 3.5 samples 4."""
    vp = _vp((1, 1, 1), 0.01, 0.05, (0.0, 0.0, 1.0))
    pose = np.eye(4, dtype=np.float32).reshape(16)
    # voxel centre (0.005, 0.005, 1.005); choose cx so fx * x / z + cx = 2.5 exactly
    x, z = np.float32(0.005), np.float32(1.005)
    fx = np.float32(201.0)
    base = np.float32(fx * x / z)
    for target, pix in ((2.5, 2), (3.5, 4)):
        cxv = np.float32(target) - base
        assert np.float32(fx * x / z + cxv) == np.float32(target)
        cam = L.camera(float(fx), float(fx), float(cxv), float(cxv), 8, 8)
        depth = np.zeros(64, np.float32)
        depth[pix * 8 + pix] = 1.02  # only (u, v) = (pix, pix) has depth
        vox = _fresh(1)
        orc.integrate_f32(vp, vox, depth, cam, pose)
        assert vox["weight"][0] == 1.0


def test_warped_gate_and_unobserved(orc):
    """opt-count gate applies only with a warp; voxels without kNN integrate unwarped."""
    vp = _vp((4, 4, 4), 0.01, 0.03, (-0.02, -0.02, 0.5))
    cam = L.camera(100.0, 100.0, 20.0, 20.0, 40, 40)
    pose = np.eye(4, dtype=np.float32).reshape(16)
    depth = np.full(1600, 0.53, np.float32)
    nodes = np.zeros(1, L.NODE)
    nodes["pos"] = [0, 0, 0.52]
    nodes["radius"] = 0.02
    T = np.zeros(1, L.DUALQUAT)
    T["real"][0, 3] = 1
    T["dual"][0, 2] = 0.5 * 0.01  # translation 1 cm along z (dual = t r / 2)
    knn = np.full(64 * 4, -1, np.int32)
    knn_w = np.zeros(64 * 4, np.float32)
    knn[0::4] = 0
    knn_w[0::4] = 1
    knn[: 4 * 16] = -1  # first slice: no nodes -> unwarped
    opt = np.full(64, 5, np.int32)
    opt[:8] = 0
    a = _fresh(64)
    orc.integrate_f32(vp, a, depth, cam, pose, nodes, T, knn, knn_w, opt, 1)
    b = _fresh(64)
    orc.integrate_f32(vp, b, depth, cam, pose)
    assert np.all(a["weight"][:8] == 0)               # gated out
    np.testing.assert_array_equal(a["tsdf"][8:16], b["tsdf"][8:16])  # unwarped slice
    c = _fresh(64)
    orc.integrate_f32(vp, c, depth, cam, pose, nodes, T, knn, knn_w, None, 0)
    # warped slices: voxel moved 1 cm closer to the depth -> sdf smaller by 0.01
    k = 2
    sl = slice(16 * k, 16 * k + 16)
    zc = np.float32(0.5) + np.float32(k + 0.5) * np.float32(0.01)
    np.testing.assert_allclose(c["tsdf"][sl], min(1.0, (0.53 - (zc + 0.01)) / 0.03), atol=1e-5)


# ------------------------------------------------------------------ MC

def _table(orc):
    rows = []
    with open(orc.TABLE) as f:
        for line in f:
            line = line.strip()
            if not line or line.startswith("#"):
                continue
            rows.append([] if line == "." else [int(ch, 16) for ch in line])
    return rows


EDGES = [(0, 1), (1, 2), (3, 2), (0, 3), (4, 5), (5, 6), (7, 6), (4, 7), (0, 4), (1, 5), (2, 6), (3, 7)]


def test_table_triangulates_exactly_the_cut_edges(orc):
    rows = _table(orc)
    assert len(rows) == 256
    for c, r in enumerate(rows):
        assert len(r) % 3 == 0
        cut = {e for e, (a, b) in enumerate(EDGES) if ((c >> a) & 1) != ((c >> b) & 1)}
        assert set(r) == cut, c
        for t in range(0, len(r), 3):
            assert len({r[t], r[t + 1], r[t + 2]}) == 3
    assert rows[0] == [] and rows[255] == []


def test_table_equals_reference_copy(orc):
    if not os.path.exists(REF_MC):
        pytest.skip("reference tree not present")
    src = open(REF_MC).read()
    body = src[src.index("triangle_table[256][16]"):]
    body = body[: body.index("};")]
    ref = [[int(x) for x in r.split(",")] for r in re.findall(r"\{([-0-9,\s]+)\}", body)]
    ref = [r[: r.index(-1)] if -1 in r else r for r in ref]
    assert _table(orc) == ref


def test_case_index_closed_form(orc):
    rng = np.random.default_rng(5)
    dims = (9, 7, 6)
    vp = _vp(dims, 0.01, 0.03, (0, 0, 0))
    V = int(np.prod(dims))
    vox = np.zeros(V, L.VOXEL)
    vox["tsdf"] = rng.uniform(-1, 1, V)
    vox["weight"] = (rng.uniform(0, 1, V) > 0.05).astype(np.float32)
    cases, verts, nrm = orc.mc_f32(vp, vox)
    t = vox["tsdf"].reshape(dims[2], dims[1], dims[0])
    w = vox["weight"].reshape(dims[2], dims[1], dims[0])
    corners = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    exp = np.zeros((dims[2], dims[1], dims[0]), np.uint8)
    ok = np.ones((dims[2] - 1, dims[1] - 1, dims[0] - 1), bool)
    idx = np.zeros_like(ok, dtype=np.int32)
    for b, (dx, dy, dz) in enumerate(corners):
        sl = (slice(dz, dz + dims[2] - 1), slice(dy, dy + dims[1] - 1), slice(dx, dx + dims[0] - 1))
        ok &= w[sl] > 0
        idx |= (t[sl] > 0).astype(np.int32) << b
    idx = np.where(ok & (idx > 0) & (idx < 255), idx, 0)
    exp[:-1, :-1, :-1] = idx
    np.testing.assert_array_equal(cases, exp.reshape(-1))
    rows = _table(orc)
    assert len(verts) == 3 * sum(len(rows[c]) // 3 for c in exp.reshape(-1))


def test_reference_sphere_scene_area_and_orientation(orc):
    """tests/test_mesh_extraction.cu:180-219: r = 0.3 at (0, 0, 1), 128^3 at 1 cm."""
    vp = _vp((128, 128, 128), 0.01, 0.03, (-0.64, -0.64, 0.36))
    X = np.arange(128)
    z, y, x = np.meshgrid(X, X, X, indexing="ij")
    wx = np.float32(-0.64) + x * np.float32(0.01)
    wy = np.float32(-0.64) + y * np.float32(0.01)
    wz = np.float32(0.36) + z * np.float32(0.01)
    d = np.sqrt(wx ** 2 + wy ** 2 + (wz - 1.0) ** 2) - 0.3
    vox = np.zeros(128 ** 3, L.VOXEL)
    vox["tsdf"] = np.clip(d, -0.03, 0.03).reshape(-1)
    vox["weight"] = 1
    cases, verts, nrm = orc.mc_f32(vp, vox)
    tri = verts.reshape(-1, 3, 3)
    assert len(tri) > 50 and len(verts) > 100  # the reference test's pass bounds
    area = 0.5 * np.linalg.norm(np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0]), axis=1).sum()
    assert abs(area - 4 * np.pi * 0.09) < 2e-3 * 4 * np.pi * 0.09
    # surface points sit on the sphere (centre shifted by the +0.5-voxel corner offset)
    c = np.array([0.005, 0.005, 1.005])
    r = np.linalg.norm(verts - c, axis=1)
    assert np.abs(r - 0.3).max() < 0.01
    # tsdf > 0 sets the bit (inverted convention): face normals still point out
    cen = tri.mean(1)
    dots = (nrm.reshape(-1, 3, 3)[:, 0] * (cen - c)).sum(1)
    assert (dots > 0).mean() > 0.98 and (dots < 0).sum() == 0
    # f32 and f64 oracles agree on cases and closely on vertices
    c64, v64, _ = orc.mc_f64(vp, vox["tsdf"].astype(np.float64), vox["weight"].astype(np.float64))
    np.testing.assert_array_equal(c64, cases)
    np.testing.assert_allclose(verts, v64, atol=1e-6)


def test_weight_zero_corner_skips_cube(orc):
    vp = _vp((2, 2, 2), 0.01, 0.03, (0, 0, 0))
    vox = np.zeros(8, L.VOXEL)
    vox["tsdf"] = [-1, 1, 1, 1, 1, 1, 1, 1]
    vox["weight"] = 1
    cases, verts, _ = orc.mc_f32(vp, vox)
    assert cases[0] == 254 and len(verts) == 3
    # the single triangle cuts edges 0, 3, 8 at the midpoints
    np.testing.assert_allclose(sorted(map(tuple, np.round(verts / 0.01 - 0.5, 6))),
                               sorted([(0.5, 0, 0), (0, 0.5, 0), (0, 0, 0.5)]), atol=1e-6)
    vox["weight"][5] = 0
    cases, verts, _ = orc.mc_f32(vp, vox)
    assert cases[0] == 0 and len(verts) == 0
