"""Pins the oracle's SE(3) / dual-quaternion algebra (include/se3_math.cuh)
against the matrix exponential (scipy golden vectors) and algebraic
invariants, and its DQ blend (warp_point, se3_math.cuh:450-510) against
special cases that reduce to a single rigid transform."""
import ctypes as C
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def golden():
    return np.load(os.path.join(HERE, "golden", "se3_expm.npz"))


def exp64(orc, tw):
    out = np.zeros(12)
    t = np.ascontiguousarray(tw, np.float64)
    orc.load().orc_exp_se3_f64(orc.ptr(t), orc.ptr(out))
    return out.reshape(3, 4)


def mat2dq(orc, m34):
    out = np.zeros(8)
    m = np.ascontiguousarray(m34, np.float64)
    orc.load().orc_mat_to_dq_f64(orc.ptr(m), orc.ptr(out))
    return out


def dq2mat(orc, q):
    out = np.zeros(12)
    qq = np.ascontiguousarray(q, np.float64)
    orc.load().orc_dq_to_mat_f64(orc.ptr(qq), orc.ptr(out))
    return out.reshape(3, 4)


def test_exp_se3_matches_expm(orc, golden):
    for tw, ref in zip(golden["twists"], golden["expm"]):
        np.testing.assert_allclose(exp64(orc, tw), ref[:3], atol=1e-12)


def test_exp_se3_fp32_close_to_expm(orc, golden):
    for tw, ref in zip(golden["twists"], golden["expm"]):
        out = np.zeros(12, np.float32)
        t32 = tw.astype(np.float32)
        orc.load().orc_exp_se3_f32(orc.ptr(t32), orc.ptr(out))
        np.testing.assert_allclose(out.reshape(3, 4), ref[:3], atol=2e-6)


def test_dq_roundtrip_all_branches(orc, golden):
    # mat4_to_dq then dq_to_mat4 is the identity on rigid transforms; the
    # golden set contains traces > 0 and near-pi rotations about x, y, z.
    branches = set()
    for ref in golden["expm"]:
        m = ref[:3]
        tr = np.trace(m[:, :3])
        if tr > 0:
            branches.add(0)
        elif m[0, 0] > m[1, 1] and m[0, 0] > m[2, 2]:
            branches.add(1)
        elif m[1, 1] > m[2, 2]:
            branches.add(2)
        else:
            branches.add(3)
        q = mat2dq(orc, m)
        assert abs(np.linalg.norm(q[:4]) - 1) < 1e-12
        np.testing.assert_allclose(dq2mat(orc, q), m, atol=1e-12)
    assert branches == {0, 1, 2, 3}


def test_dq_mul_is_matrix_product(orc, golden):
    mats = golden["expm"]
    for a, b in zip(mats[:-1], mats[1:]):
        qa, qb = mat2dq(orc, a[:3]), mat2dq(orc, b[:3])
        out = np.zeros(8)
        orc.load().orc_dq_mul_f64(orc.ptr(qa), orc.ptr(qb), orc.ptr(out))
        np.testing.assert_allclose(dq2mat(orc, out), (a @ b)[:3], atol=1e-12)


def test_dq_centered_rotates_about_center(orc, golden):
    rng = np.random.default_rng(1)
    for ref in golden["expm"][:10]:
        q = mat2dq(orc, ref[:3])
        c = rng.normal(0, 0.5, 3)
        qc = np.zeros(8)
        orc.load().orc_dq_centered_f64(orc.ptr(q), orc.ptr(c), orc.ptr(qc))
        p = rng.normal(0, 0.5, 3)
        out = np.zeros(3)
        orc.load().orc_dq_transform_f64(orc.ptr(qc), orc.ptr(p), orc.ptr(out))
        R, t = ref[:3, :3], ref[:3, 3]
        np.testing.assert_allclose(out, R @ (p - c) + c + t, atol=1e-12)


def test_twist_update_left_composes_exponential(orc, golden):
    mats = golden["expm"]
    tws = golden["twists"]
    for k in range(10):
        T0 = mat2dq(orc, mats[k][:3])
        out = np.zeros(8)
        orc.load().orc_twist_update_f64(orc.ptr(tws[k + 1]), orc.ptr(T0), 0.0, 0.0, 1.0, orc.ptr(out))
        np.testing.assert_allclose(dq2mat(orc, out), (mats[k + 1] @ mats[k])[:3], atol=1e-11)


def test_twist_update_clamps_rotation_and_translation_separately(orc):
    T0 = np.array([0, 0, 0, 1, 0, 0, 0, 0.0])
    tw = np.array([0.3, 0.0, 0.0, 0.0, 0.2, 0.0])
    out = np.zeros(8)
    orc.load().orc_twist_update_f64(orc.ptr(tw), orc.ptr(T0), 0.1, 0.05, 1.0, orc.ptr(out))
    m = dq2mat(orc, out)
    ang = np.arccos(np.clip((np.trace(m[:, :3]) - 1) / 2, -1, 1))
    assert abs(ang - 0.1) < 1e-12
    # translation of exp(0.1 e_x, 0.05 e_y): V v with V from the clamped omega
    w = np.array([0.1, 0, 0])
    hat = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    th = 0.1
    V = np.eye(3) + (1 - np.cos(th)) / th ** 2 * hat + (th - np.sin(th)) / th ** 3 * hat @ hat
    np.testing.assert_allclose(m[:, 3], V @ np.array([0, 0.05, 0]), atol=1e-12)


def _nodes(P):
    from cudafusion_amd._lib import NODE
    n = np.zeros(len(P), NODE)
    n["pos"] = P
    n["radius"] = 0.05
    n["neighbors"] = -1
    return n


def warp64(orc, pts, ids, ws, nodes, T8):
    out = np.zeros(pts.shape, np.float64)
    # keep every converted array referenced until the call returns
    p = np.ascontiguousarray(pts, np.float32)
    i = np.ascontiguousarray(ids, np.int32)
    w = np.ascontiguousarray(ws, np.float32)
    t = np.ascontiguousarray(T8, np.float64)
    orc.load().orc_warp_points_f64(orc.ptr(p), orc.ptr(i), orc.ptr(w), len(pts), orc.ptr(nodes), orc.ptr(t),
                                   len(nodes), orc.ptr(out))
    return out


def test_warp_identity_and_empty(orc):
    rng = np.random.default_rng(2)
    P = rng.normal(0, 0.1, (6, 3)).astype(np.float32)
    nodes = _nodes(P)
    pts = rng.normal(0, 0.1, (20, 3)).astype(np.float32)
    ids = rng.integers(0, 6, (20, 4))
    ws = rng.uniform(0.1, 1, (20, 4))
    T = orc.identity_dq64(6)
    np.testing.assert_allclose(warp64(orc, pts, ids, ws, nodes, T), pts, atol=1e-7)
    # no valid neighbour (ids -1 or weights below 1e-8): the point is returned unchanged
    T2 = np.tile(mat2dq(orc, np.hstack([np.eye(3), [[0.1], [0], [0]]])), (6, 1))
    ids2 = np.full((20, 4), -1)
    np.testing.assert_allclose(warp64(orc, pts, ids2, ws, nodes, T2), pts, atol=0)
    ws3 = np.full((20, 4), 1e-9)
    np.testing.assert_allclose(warp64(orc, pts, ids, ws3, nodes, T2), pts, atol=0)


def test_warp_single_node_and_uniform_transform(orc, golden):
    rng = np.random.default_rng(3)
    P = rng.normal(0, 0.1, (5, 3)).astype(np.float32)
    nodes = _nodes(P)
    m = golden["expm"][4][:3]
    T = np.tile(mat2dq(orc, m), (5, 1))
    pts = rng.normal(0, 0.1, (10, 3)).astype(np.float32)
    # one neighbour: rotation about that node's centre
    ids = np.tile([2, -1, -1, -1], (10, 1))
    ws = np.tile([0.7, 0, 0, 0], (10, 1))
    exp = (pts - P[2]) @ m[:, :3].T + P[2] + m[:, 3]
    np.testing.assert_allclose(warp64(orc, pts, ids, ws, nodes, T), exp, atol=1e-6)


def test_warp_antipodal_invariance(orc, golden):
    rng = np.random.default_rng(4)
    P = rng.normal(0, 0.1, (4, 3)).astype(np.float32)
    nodes = _nodes(P)
    T = np.array([mat2dq(orc, golden["expm"][k][:3] @ np.eye(4)) for k in range(4)])
    small = [mat2dq(orc, np.hstack([np.eye(3), rng.normal(0, 0.01, (3, 1))])) for _ in range(4)]
    T = np.array(small)
    pts = rng.normal(0, 0.1, (8, 3)).astype(np.float32)
    ids = np.tile([0, 1, 2, 3], (8, 1))
    ws = rng.uniform(0.1, 1, (8, 4))
    a = warp64(orc, pts, ids, ws, nodes, T)
    Tn = T.copy()
    Tn[2] *= -1  # q and -q are the same rigid transform
    b = warp64(orc, pts, ids, ws, nodes, Tn)
    np.testing.assert_allclose(a, b, atol=1e-12)


def warpn64(orc, nrm, ids, ws, T8):
    out = np.zeros(nrm.shape, np.float64)
    n = np.ascontiguousarray(nrm, np.float32)
    i = np.ascontiguousarray(ids, np.int32)
    w = np.ascontiguousarray(ws, np.float32)
    t = np.ascontiguousarray(T8, np.float64)
    orc.load().orc_warp_normals_f64(orc.ptr(n), orc.ptr(i), orc.ptr(w), len(nrm), orc.ptr(t), len(T8), orc.ptr(out))
    return out


def test_warp_normal_rotates_by_uncentred_blend(orc, golden):
    """warp_normal (se3_math.cuh:512-558): with every neighbour carrying the
    same rotation R the normal becomes R n (translations and centres play no
    part); with no valid neighbour it is only normalised."""
    rng = np.random.default_rng(6)
    m = golden["expm"][7][:3]
    T = np.tile(mat2dq(orc, m), (4, 1))
    nrm = rng.normal(0, 1, (12, 3))
    ids = np.tile([0, 1, 2, 3], (12, 1))
    ws = rng.uniform(0.1, 1, (12, 4))
    out = warpn64(orc, nrm, ids, ws, T)
    exp = nrm.astype(np.float32).astype(np.float64) @ m[:, :3].T
    exp /= np.linalg.norm(exp, axis=1, keepdims=True)
    np.testing.assert_allclose(out, exp, atol=1e-6)
    T2 = T.copy()
    T2[1] *= -1  # antipodal copy of the same rotation
    np.testing.assert_allclose(warpn64(orc, nrm, ids, ws, T2), exp, atol=1e-6)
    none = warpn64(orc, nrm, np.full((12, 4), -1), ws, T)
    n32 = nrm.astype(np.float32).astype(np.float64)
    np.testing.assert_allclose(none, n32 / np.linalg.norm(n32, axis=1, keepdims=True), atol=1e-7)
