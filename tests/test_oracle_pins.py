"""Independent pins of two oracle functions that the other tests only use:

* the voxel 4-NN snapshot (src/warp_field.cu:299-405), against a numpy brute
  force written from the reference's rules: voxels with weight <= 0 (or no
  nodes) get ids -1 and weights 0; otherwise the four nearest nodes by
  squared distance from the voxel centre, ties to the lower node index
  (strict < insertion), fewer than four nodes leaving -1 slots, and weights
  exp(-d^2 / 2 max(r, 1e-6)^2) normalised by their sum (0 when the sum is
  <= 1e-8).  Exact distance ties are built from dyadic coordinates, so the
  tie rule decides the ids;
* the caller's step validation (include/fusion_app.h:409-491), on crafted
  steps with the accept / reject outcome of each rule written out by hand:
  non-finite components, the 5x-clamp spike (boundary included), the mean
  bound on the clamped step, and the disabled clamps and mean test.
"""
import ctypes as C

import numpy as np
import pytest


def _nodes(pos, radius):
    from cudafusion_amd import _lib as L
    n = np.zeros(len(pos), L.NODE)
    n["pos"] = np.asarray(pos, np.float32)
    n["radius"] = np.asarray(radius, np.float32)
    return n


def _vox(weights):
    from cudafusion_amd import _lib as L
    v = np.zeros(len(weights), L.VOXEL)
    v["tsdf"] = 1.0
    v["weight"] = np.asarray(weights, np.float32)
    return v


def _brute_knn(vp, weight, nodes, dtype):
    """numpy restatement of the reference rules in arithmetic `dtype`."""
    X, Y, Z = vp.dims[:]
    s = dtype(vp.voxel_size)
    o = [dtype(v) for v in vp.origin[:]]
    V = X * Y * Z
    ids = np.full((V, 4), -1, np.int32)
    ws = np.zeros((V, 4), np.float64)
    N = len(nodes)
    npos = nodes["pos"].astype(dtype)
    for idx in range(V):
        if weight[idx] <= 0 or N == 0:
            continue
        x, y, z = idx % X, (idx // X) % Y, idx // (X * Y)
        p = [o[0] + (dtype(x) + dtype(0.5)) * s, o[1] + (dtype(y) + dtype(0.5)) * s,
             o[2] + (dtype(z) + dtype(0.5)) * s]
        dx, dy, dz = p[0] - npos[:, 0], p[1] - npos[:, 1], p[2] - npos[:, 2]
        d2 = dx * dx + dy * dy + dz * dz
        order = np.lexsort((np.arange(N), d2))[:4]  # by distance, then lower index
        k = len(order)
        ids[idx, :k] = order
        r = np.maximum(nodes["radius"][order].astype(np.float64), np.float64(np.float32(1e-6)))
        w = np.exp(-d2[order].astype(np.float64) / (2 * r * r))
        tot = w.sum()
        ws[idx, :k] = w / tot if tot > 1e-8 else 0.0
    return ids.reshape(-1), ws.reshape(-1)


def _oracle_knn(orc, vp, vox, nodes, fp64):
    V = int(np.prod(vp.dims[:]))
    knn = np.zeros(V * 4, np.int32)
    w = np.zeros(V * 4, np.float64 if fp64 else np.float32)
    lib = orc.load()
    fn = lib.orc_knn_f64 if fp64 else lib.orc_knn_f32
    nd = nodes if len(nodes) else np.zeros(1, nodes.dtype)
    fn(C.byref(vp), orc.ptr(vox), orc.ptr(nd), len(nodes), orc.ptr(knn), orc.ptr(w))
    return knn, w


def _check(orc, vp, vox, nodes):
    """Oracle (FP32 and FP64 forms) against the brute force in the same
    arithmetic: ids exactly, weights to rounding."""
    for fp64, dt in ((False, np.float32), (True, np.float64)):
        ids_ref, w_ref = _brute_knn(vp, vox["weight"], nodes, dt)
        ids, w = _oracle_knn(orc, vp, vox, nodes, fp64)
        np.testing.assert_array_equal(ids, ids_ref)
        np.testing.assert_allclose(w, w_ref, rtol=1e-12 if fp64 else 2e-6, atol=1e-12 if fp64 else 1e-7)
    return ids_ref


def test_knn_exact_ties_lower_index_wins(orc):
    """Nodes on the voxel grid's corners: every voxel centre is equidistant
    (d^2 = 3 (s/2)^2, exact in FP32) from its eight corner nodes, so the tie
    rule picks the four lowest indices; the node order is a random
    permutation, and two nodes are duplicated (same position, different
    index)."""
    from cudafusion_amd import _lib as L
    s = 0.25
    vp = L.volume_params((4, 3, 3), s, 3 * s, (0.0, 0.0, 0.0))
    corners = np.array([(i * s, j * s, k * s) for k in range(4) for j in range(4) for i in range(5)], np.float32)
    rng = np.random.default_rng(3)
    perm = rng.permutation(len(corners))
    pos = np.concatenate([corners[perm], corners[perm[:2]]])
    radius = rng.uniform(0.1, 0.4, len(pos))
    nodes = _nodes(pos, radius)
    weights = np.ones(36, np.float32)
    weights[[1, 7, 20]] = 0.0  # unobserved voxels
    weights[30] = -1.0
    vox = _vox(weights)
    ids = _check(orc, vp, vox, nodes)
    ids = ids.reshape(-1, 4)
    assert (ids[[1, 7, 20, 30]] == -1).all()
    # voxel 0 at (s/2, s/2, s/2): its eight corners are the nearest, tied; the
    # chosen four are the four lowest node indices among them
    c0 = [n for n in range(len(pos)) if np.all(pos[n] <= s + 1e-6)]
    assert len(c0) >= 8
    assert list(ids[0]) == sorted(c0)[:4]


def test_knn_radius_clamp_and_zero_sum(orc):
    """Radius <= 0 is clamped to 1e-6: a node at the voxel centre keeps weight
    1, nodes farther away get exp(-huge) = 0, and a voxel whose four weights
    all vanish gets all-zero weights (sum <= 1e-8) with its ids kept."""
    from cudafusion_amd import _lib as L
    s = 0.125
    vp = L.volume_params((3, 2, 2), s, 3 * s, (0.0, 0.0, 0.0))
    pos = [(s / 2, s / 2, s / 2), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), (1.0, 1.0, 1.0)]
    nodes = _nodes(pos, [0.0, -1.0, 0.0, 0.0, 0.0])
    vox = _vox(np.ones(12, np.float32))
    ids = _check(orc, vp, vox, nodes).reshape(-1, 4)
    _, w = _oracle_knn(orc, vp, vox, nodes, True)
    w = w.reshape(-1, 4)
    assert ids[0, 0] == 0 and w[0, 0] == 1.0 and (w[0, 1:] == 0).all()
    assert (ids[1] >= 0).all() and (w[1] == 0).all()


def test_knn_fewer_nodes_than_neighbours_and_none(orc):
    from cudafusion_amd import _lib as L
    vp = L.volume_params((3, 3, 2), 0.1, 0.3, (-0.1, 0.05, 0.3))
    nodes = _nodes([(0.0, 0.1, 0.35), (0.1, 0.2, 0.4)], [0.05, 0.08])
    vox = _vox(np.ones(18, np.float32))
    ids = _check(orc, vp, vox, nodes).reshape(-1, 4)
    assert (ids[:, 2:] == -1).all() and (ids[:, :2] >= 0).all()
    _, w = _oracle_knn(orc, vp, vox, nodes, True)
    np.testing.assert_allclose(w.reshape(-1, 4).sum(1), 1.0, rtol=1e-12)
    empty = _nodes(np.zeros((0, 3)), np.zeros(0))
    ids0, w0 = _oracle_knn(orc, vp, vox, empty, False)
    assert (ids0 == -1).all() and (w0 == 0).all()


def test_knn_random_nodes(orc):
    from cudafusion_amd import _lib as L
    rng = np.random.default_rng(11)
    vp = L.volume_params((9, 7, 6), 0.02, 0.06, (-0.09, -0.07, 0.5))
    pos = rng.uniform([-0.1, -0.08, 0.48], [0.1, 0.08, 0.64], (23, 3))
    nodes = _nodes(pos, rng.uniform(0.01, 0.05, 23))
    weights = (rng.random(9 * 7 * 6) < 0.6).astype(np.float32) * 3.0
    _check(orc, vp, _vox(weights), nodes)


# ------------------------------------------------------------ validation

def _gn(scale=0.5, rot=0.0625, trans=0.03125, mean=0.05):
    from cudafusion_amd import _lib as L
    return L.gn_params(5, scale, rot, trans, 1, mean)


def _dx(rows):
    d = np.zeros((4, 6))
    for i, r in enumerate(rows):
        d[i] = r
    return d.reshape(-1)


CASES = [
    # (name, gn kwargs, per-node dx rows, accept)
    ("zero step", {}, [], True),
    # scale * |omega| = 0.5 * 0.625 = 0.3125 = 5 * max_rot exactly: not a spike
    ("rotation spike at the boundary", {}, [(0.625, 0, 0, 0, 0, 0)], True),
    ("rotation spike above the boundary", {}, [(0.6251, 0, 0, 0, 0, 0)], False),
    # scale * |v| = 0.5 * 0.3125 = 0.15625 = 5 * max_trans
    ("translation spike at the boundary", {}, [(0, 0, 0, 0, 0.3125, 0)], True),
    ("translation spike above the boundary", {}, [(0, 0, 0, 0.2, 0.25, 0.1)], False),
    ("NaN component", {}, [(0, 0, 0, 0, 0, 0), (0, np.nan, 0, 0, 0, 0)], False),
    ("infinite component", {}, [(0, 0, 0, np.inf, 0, 0)], False),
    # every node clamped to max_rot: mean of the clamped norms 0.0625 > 0.05
    ("mean bound exceeded", {}, [(0.2, 0, 0, 0, 0, 0)] * 4, False),
    ("mean bound met with equality", {"mean": 0.0625}, [(0.2, 0, 0, 0, 0, 0)] * 4, True),
    ("mean test disabled", {"mean": 0.0}, [(0.2, 0, 0, 0, 0, 0)] * 4, True),
    # one node far above the others (ratio spike) but below 5x the clamp: a
    # diagnostic only in the reference, not a rejection
    ("ratio spike is not a rejection", {}, [(0.5, 0, 0, 0, 0, 0), (1e-4, 0, 0, 0, 0, 0),
                                            (0, 1e-4, 0, 0, 0, 0), (0, 0, 1e-4, 0, 0, 0)], True),
    # clamps disabled: no spike test, the mean uses the scaled raw norms
    ("no clamps, mean exceeded", {"rot": 0.0, "trans": 0.0}, [(0.2, 0, 0, 0, 0, 0)] * 4, False),
    ("no clamps, mean met", {"rot": 0.0, "trans": 0.0}, [(0.08, 0, 0, 0, 0, 0)] * 4, True),
    ("no clamps, no mean test", {"rot": 0.0, "trans": 0.0, "mean": 0.0}, [(1e3, 0, 0, 0, 1e3, 0)] * 4, True),
    ("no clamps, overflow to inf", {"rot": 0.0, "trans": 0.0, "mean": 0.0}, [(3e38, 3e38, 0, 0, 0, 0)], False),
]


@pytest.mark.parametrize("name,kw,rows,accept", CASES, ids=[c[0] for c in CASES])
def test_validate_step_rules(orc, name, kw, rows, accept):
    gn = _gn(**kw)
    dx = _dx(rows)
    assert orc.validate(dx, gn, fp64=False) is accept
    if name != "no clamps, overflow to inf":  # FP64 does not overflow there
        assert orc.validate(dx, gn, fp64=True) is accept
