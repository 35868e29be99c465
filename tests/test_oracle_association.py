"""Pins of the oracle's projective data association (oracle/association.cpp,
restating kernels/tsdf_kernels.cu:180-230, 386-651) by closed forms:

* depth normals of fronto-parallel and tilted planes (the plane's unit
  normal, pointing to the camera) and the zero normal where a neighbour's
  depth is missing or in the last row and column;
* the pixel kNN gather;
* the correspondence search on a plane seen by an identity camera, whose
  live surface is the depth map's own back projection: every candidate of
  the window lies on the plane (point-to-plane distance 0), so the strict <
  rule keeps the first one in row-major window order, the top-left pixel,
  with the reference's confidence weight; and each rejection reason once;
* the optimised-voxel counts (a 3x3x3 neighbourhood once per frame).
"""
import numpy as np

from cudafusion_amd import _lib as L

W, H = 40, 30
CAM = L.camera(50.0, 55.0, 19.5, 14.5, W, H)


def _unproject(u, v, z, cam=CAM):
    return np.stack([z * (u - cam.cx) / cam.fx, z * (v - cam.cy) / cam.fy, z], -1)


def _grid():
    v, u = np.meshgrid(np.arange(H, dtype=np.float64), np.arange(W, dtype=np.float64), indexing="ij")
    return u, v


def test_depth_normals_planes(orc):
    u, v = _grid()
    # fronto-parallel plane at 1.25 m
    d = np.full((H, W), 1.25, np.float32)
    n = orc.depth_normals_f32(d, CAM).reshape(H, W, 3)
    np.testing.assert_allclose(n[:-1, :-1], np.broadcast_to([0, 0, -1], (H - 1, W - 1, 3)), atol=1e-6)
    assert (n[-1] == 0).all() and (n[:, -1] == 0).all()
    # tilted plane z = 1 + 0.4 x: z = 1 / (1 - 0.4 (u - cx) / fx); normal -(-0.4, 0, 1) / |.|
    a = 0.4
    zt = (1.0 / (1.0 - a * (u - CAM.cx) / CAM.fx)).astype(np.float32)
    n = orc.depth_normals_f32(zt, CAM).reshape(H, W, 3)
    ref = np.array([a, 0.0, -1.0]) / np.sqrt(1 + a * a)
    np.testing.assert_allclose(n[:-1, :-1], np.broadcast_to(ref, (H - 1, W - 1, 3)), atol=2e-4)
    # a missing depth voids the pixel and the pixels left of / above it
    d2 = d.copy()
    d2[10, 12] = 0
    n = orc.depth_normals_f32(d2, CAM).reshape(H, W, 3)
    for (y, x) in ((10, 12), (10, 11), (9, 12)):
        assert (n[y, x] == 0).all()
    assert (n[9, 11] != 0).any() and (n[11, 12] != 0).any()


def test_pixel_knn_gather(orc):
    rng = np.random.default_rng(1)
    knn = rng.integers(0, 50, 40 * 4).astype(np.int32)
    w = rng.random(40 * 4).astype(np.float32)
    hit = np.array([3, -1, 0, 39, -5], np.int32)
    ids, ws = orc.pixel_knn(hit, knn, w)
    ids, ws = ids.reshape(-1, 4), ws.reshape(-1, 4)
    np.testing.assert_array_equal(ids[0], knn[12:16])
    np.testing.assert_array_equal(ids[3], knn[156:160])
    assert (ids[1] == -1).all() and (ws[1] == 0).all() and (ids[4] == -1).all()
    np.testing.assert_array_equal(ws[2], w[0:4])


def _plane_case(R=2, depth=1.0):
    u, v = _grid()
    d = np.full((H, W), depth, np.float32)
    live = _unproject(u, v, d.astype(np.float64)).astype(np.float32).reshape(-1, 3)
    nrm = np.zeros((H * W, 3), np.float32)
    nrm[:, 2] = -1
    return d, live, nrm


def test_correspondences_plane_closed_form(orc):
    R = 2
    d, live, nrm = _plane_case(R)
    dn = orc.depth_normals_f32(d, CAM)
    pose = np.eye(4, dtype=np.float32).reshape(16)
    ap = L.assoc_params(0.05, 0.7, 0.2, R)
    pk = np.arange(H * W * 4, dtype=np.int32)
    pw = np.full(H * W * 4, 0.25, np.float32)
    corrs, n, stats = orc.find_correspondences_f32(live, nrm, d, dn, CAM, pose, pk, pw, ap)
    c = corrs.reshape(H, W)
    # pixel (u, v): the window's first candidate (u - R, v - R) with a depth
    # normal (not the last row / column) and on the image
    cap = max(3 * 0.05, 0.05 + 0.03)
    view_eps = max(1 - 0.2, 1e-4)
    for (y, x) in ((10, 10), (14, 20), (2, 2), (0, 5), (29, 39)):
        cand = [(yy, xx) for yy in range(y - R, y + R + 1) for xx in range(x - R, x + R + 1)
                if 0 <= yy < H - 1 and 0 <= xx < W - 1]
        yy, xx = cand[0]
        p = _unproject(float(x), float(y), 1.0)
        q = _unproject(float(xx), float(yy), 1.0)
        ed = np.linalg.norm(p - q)
        vc = abs(p[2] / np.linalg.norm(p))
        conf = ((max(0, 1 - ed / cap) + 1.0 + max(0, 1 - (1 - vc) / view_eps)) / 3) ** 2
        cc = c[y, x]
        assert cc["valid"] == 1
        np.testing.assert_allclose(cc["dst"], q, atol=1e-6)
        np.testing.assert_allclose(cc["src"], p, atol=1e-6)
        np.testing.assert_allclose(cc["normal"], [0, 0, -1], atol=1e-6)
        assert abs(cc["weight"] - conf) < 1e-5
        np.testing.assert_array_equal(cc["node_ids"], pk.reshape(H, W, 4)[y, x])
    assert n == H * W and stats[5] == n and stats[:5].sum() == 0


def test_correspondences_rejections(orc):
    R = 1
    d, live, nrm = _plane_case(R)
    dn = orc.depth_normals_f32(d, CAM)
    pose = np.eye(4, dtype=np.float32).reshape(16)
    ap = L.assoc_params(0.05, 0.7, 0.2, R)
    lv, ln, dd = live.copy(), nrm.copy(), d.copy()
    px = lambda y, x: y * W + x  # noqa: E731
    ln[px(3, 3)] = 0                                  # 0 invalid live normal
    lv[px(4, 4)] = lv[px(4, 4)] * np.float32(-1)      # 1 behind the camera
    dd[20:27, 20:27] = 0                              # 2 no depth around (23, 23)
    lv[px(8, 8), 2] += np.float32(0.2)                # 3 farther than the caps (with (12, 12) below)
    c60 = np.float32(0.5)
    # angle 60 deg > acos(0.7) (view cos ~0.5 passes): no candidate passes the
    # angle test, so none is close either, and the reference tests "close"
    # first (kernels/tsdf_kernels.cu:531-548): counted under 3
    ln[px(12, 12)] = [np.sqrt(1 - c60 * c60), 0, -c60]
    ln[px(15, 19)] = [1, 0, 0]                        # 4 normal across the view ray (view cos ~0.01 < 0.2)
    corrs, n, stats = orc.find_correspondences_f32(lv, ln, dd, orc.depth_normals_f32(dd, CAM), CAM, pose, None, None,
                                                   ap)
    c = corrs.reshape(H, W)
    for (y, x) in ((3, 3), (4, 4), (23, 23), (8, 8), (12, 12), (15, 19)):
        assert c[y, x]["valid"] == 0 and c[y, x]["weight"] == 0 and (c[y, x]["node_ids"] == -1).all()
    # the hole is 7 x 7: the 5 x 5 pixels inside it see no depth in their 3 x 3 windows
    assert stats[0] == 1 and stats[1] == 1 and stats[2] == 25 and stats[3] >= 2 and stats[4] == 1
    assert stats.sum() == H * W and stats[5] == n
    assert c[5, 5]["valid"] == 1 and (c[5, 5]["node_ids"] == -1).all()  # no pixel kNN given


def test_mark_optimized_voxels_counts(orc):
    dims = (6, 5, 4)
    V = 6 * 5 * 4
    counts = np.zeros(V, np.int32)
    last = np.full(V, -1, np.int32)
    corrs = np.zeros(4, L.CORR)
    corrs["valid"] = [1, 1, 1, 0]
    corrs["weight"] = [1.0, 0.5, 0.0, 1.0]
    idx = lambda x, y, z: (z * 5 + y) * 6 + x  # noqa: E731
    hit = np.array([idx(2, 2, 1), idx(2, 2, 1), idx(5, 0, 0), idx(1, 1, 1)], np.int32)
    orc.mark_optimized_voxels(corrs, hit, dims, 7, counts, last)
    c = counts.reshape(4, 5, 6)
    # the two valid corrs share a voxel: its neighbourhood counts once; weight 0 and invalid ones count nothing
    assert c[0:3, 1:4, 1:4].min() == 1 and c.sum() == 27
    orc.mark_optimized_voxels(corrs, hit, dims, 8, counts, last)
    assert counts.reshape(4, 5, 6)[0:3, 1:4, 1:4].min() == 2 and counts.sum() == 54
    # border: a corner voxel's neighbourhood is clipped to 2 x 2 x 2
    corrs2 = np.zeros(1, L.CORR)
    corrs2["valid"], corrs2["weight"] = 1, 1.0
    c2 = np.zeros(V, np.int32)
    orc.mark_optimized_voxels(corrs2, np.array([idx(5, 4, 3)], np.int32), dims, 0, c2, np.full(V, -1, np.int32))
    assert c2.sum() == 8


# ------------------------------------------------- live surface prediction

def test_raster_fronto_parallel_triangles(orc):
    """Two fronto-parallel triangles at 1.0 m and 0.8 m overlapping in the
    image: every pixel centre inside a triangle's projection gets that
    triangle (the nearer where both cover it), its live vertex is the
    pixel-centre ray's point on the plane (screen-space and object-space
    barycentrics agree on a fronto-parallel triangle), the face normal, the
    canonical point interpolated the same way and that point's voxel."""
    vp = L.volume_params((64, 64, 64), 0.0213, 0.06, (-0.6413, -0.6387, 0.2071))
    pose = np.eye(4, dtype=np.float32).reshape(16)

    def tri_at(z, u0, v0, u1, v1, u2, v2):
        return [_unproject(float(u), float(v), z) for u, v in ((u0, v0), (u1, v1), (u2, v2))]

    A = tri_at(1.0, 4, 4, 30, 6, 8, 26)
    B = tri_at(0.8, 15, 10, 36, 12, 20, 28)
    verts = np.array(A + B, np.float32).reshape(-1)
    canon = verts.copy()
    tris = np.array([0, 1, 2, 3, 4, 5], np.int32)
    lv, ln, hit, cp = orc.rasterize_f32(canon, verts, tris, CAM, pose, vp)
    lv, ln, cp = lv.reshape(H, W, 3), ln.reshape(H, W, 3), cp.reshape(H, W, 3)
    hit = hit.reshape(H, W)

    def inside(tri, x, y):
        P = np.array([[CAM.fx * p[0] / p[2] + CAM.cx, CAM.fy * p[1] / p[2] + CAM.cy] for p in tri])
        def e(a, b, q):
            return (q[0] - a[0]) * (b[1] - a[1]) - (q[1] - a[1]) * (b[0] - a[0])
        q = (x + 0.5, y + 0.5)
        ar = e(P[0], P[1], P[2])
        w = [e(P[1], P[2], q) / ar, e(P[2], P[0], q) / ar, e(P[0], P[1], q) / ar]
        return min(w) > 1e-4, min(w) < -1e-4  # clearly inside / clearly outside
    nA = nB = 0
    for y in range(H):
        for x in range(W):
            inA, outA = inside(A, x, y)
            inB, outB = inside(B, x, y)
            if not (inA or outA) or not (inB or outB):
                continue  # on an edge: left to the -1e-5 test
            z = 0.8 if inB else (1.0 if inA else None)
            if z is None:
                assert hit[y, x] == -1 and (lv[y, x] == 0).all() and (ln[y, x] == 0).all()
                continue
            nA += z == 1.0
            nB += z == 0.8
            ref = _unproject(x + 0.5, y + 0.5, z)
            np.testing.assert_allclose(lv[y, x], ref, atol=2e-6)
            np.testing.assert_allclose(cp[y, x], ref, atol=2e-6)
            np.testing.assert_allclose(np.abs(ln[y, x]), [0, 0, 1], atol=1e-6)
            # the voxel of the canonical point, floor((p - origin) / s) in FP32
            q = cp[y, x].astype(np.float32)
            vox = np.floor((q - np.array(vp.origin[:], np.float32)) / np.float32(vp.voxel_size)).astype(int)
            assert hit[y, x] == (vox[2] * 64 + vox[1]) * 64 + vox[0]
    assert nA > 50 and nB > 50


def test_raster_equal_depth_lowest_triangle_wins(orc):
    vp = L.volume_params((8, 8, 8), 0.1, 0.3, (-0.4, -0.4, 0.6))
    pose = np.eye(4, dtype=np.float32).reshape(16)
    t = [_unproject(float(u), float(v), 1.0) for u, v in ((5, 5), (25, 5), (5, 25))]
    verts = np.array(t + t, np.float32).reshape(-1)
    canon = verts.copy()
    canon[9:] += np.float32(0.05)  # the second copy's canonical points differ
    lv, ln, hit, cp = orc.rasterize_f32(canon, verts, np.arange(6, dtype=np.int32), CAM, pose, vp)
    cp = cp.reshape(H, W, 3)
    np.testing.assert_allclose(cp[10, 10], _unproject(10.5, 10.5, 1.0), atol=2e-6)  # triangle 0, not 1


def test_skinning_identity_and_rigid_node(orc):
    """Identity transforms leave every vertex in place; one node moved by a
    rigid transform moves every vertex whose voxel has that node as its only
    neighbour by exactly that transform; vertices outside the volume stay."""
    vp = L.volume_params((10, 10, 10), 0.05, 0.15, (-0.25, -0.25, 0.75))
    rng = np.random.default_rng(5)
    canon = rng.uniform(-0.3, 0.3, (200, 3)).astype(np.float32) + np.float32([0, 0, 1.0])
    nodes = np.zeros(1, L.NODE)
    nodes["pos"] = [0.0, 0.0, 1.0]
    nodes["radius"] = 0.1
    V = 1000
    knn = np.full(V * 4, -1, np.int32)
    knn[0::4] = 0
    knn_w = np.zeros(V * 4, np.float32)
    knn_w[0::4] = 1.0
    T = np.zeros(1, L.DUALQUAT)
    T["real"][0] = [0, 0, 0, 1]
    np.testing.assert_array_equal(orc.skin_vertices_f32(vp, canon, nodes, T, knn, knn_w), canon)
    # a rotation of 0.1 rad about z through the node, plus a shift
    th = 0.1
    Rz = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1.0]])
    t = np.array([0.01, -0.02, 0.005])
    T64 = np.zeros(8)
    m = np.hstack([Rz, t.reshape(3, 1)])
    orc.load().orc_mat_to_dq_f64(np.ascontiguousarray(m).ctypes.data, T64.ctypes.data)
    T["real"][0], T["dual"][0] = T64[:4], T64[4:]
    out = orc.skin_vertices_f32(vp, canon, nodes, T, knn, knn_w)
    c = np.array([0, 0, 1.0])
    inside = np.all((canon >= -0.25 + np.float32([0, 0, 1.0])) & (canon < 0.25 + np.float32([0, 0, 1.0])), axis=1)
    # T applied about the node position (dq_centered): R (p - c) + c + t
    ref = (canon.astype(np.float64) - c) @ Rz.T + c + t
    np.testing.assert_allclose(out[inside], ref[inside], atol=2e-6)
    np.testing.assert_array_equal(out[~inside], canon[~inside])
    assert inside.sum() > 20 and (~inside).sum() > 20
