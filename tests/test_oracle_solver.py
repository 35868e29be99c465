"""Pins the oracle's GN solver (kernels/solver_kernels.cu) without retyping it:

* the assembled BSR system equals the dense Gauss-Newton normal equations
  built from finite-difference Jacobians of the true residuals (data term:
  n.(warp(p) - dst); ARAP: T_i(x_j) - T_j(x_j)), with the reference's three
  structural quirks applied explicitly - Q1 the ARAP rotation Jacobian has
  the opposite sign, Q2 data-term blocks outside the graph pattern are
  dropped, Q3 an edge i->j is used only from the node with the smaller index;
* the block-Jacobi blocks are inverses of (A_ii + 1e-4 I);
* PCG reproduces numpy's dense solve and a textbook PCG trace;
* the GN loop with re-warped sources recovers a known rigid motion.
"""
import ctypes as C

import numpy as np
import pytest

from cudafusion_amd import _lib as L


def _hat(w):
    return np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])


def _dq64(orc, R, t):
    m = np.ascontiguousarray(np.hstack([R, t.reshape(3, 1)]), np.float64)
    out = np.zeros(8)
    orc.load().orc_mat_to_dq_f64(m.ctypes.data, out.ctypes.data)
    return out


def _twist(orc, tw, T8):
    tw = np.ascontiguousarray(tw, np.float64)
    T8 = np.ascontiguousarray(T8, np.float64)
    out = np.zeros(8)
    orc.load().orc_twist_update_f64(tw.ctypes.data, T8.ctypes.data, 0.0, 0.0, 1.0, out.ctypes.data)
    return out


def _xf_c(orc, T8, p, c):
    """dq_transform_point_centered(T, p, c) via dq_centered + transform."""
    T8 = np.ascontiguousarray(T8, np.float64)
    c = np.ascontiguousarray(c, np.float64)
    p = np.ascontiguousarray(p, np.float64)
    qc = np.zeros(8)
    orc.load().orc_dq_centered_f64(T8.ctypes.data, c.ctypes.data, qc.ctypes.data)
    out = np.zeros(3)
    orc.load().orc_dq_transform_f64(qc.ctypes.data, p.ctypes.data, out.ctypes.data)
    return out


def _graph():
    rng = np.random.default_rng(3)
    N = 7
    nodes = np.zeros(N, L.NODE)
    nodes["pos"] = rng.uniform(-0.05, 0.05, (N, 3)).astype(np.float32) + np.float32([0, 0, 1])
    nodes["radius"] = 0.04
    nodes["neighbors"] = -1
    lists = {0: [1, 2], 1: [0, 3], 2: [4], 3: [1, 5, 6], 4: [2], 5: [6], 6: [0]}  # 6->0 only: Q3
    ws = rng.uniform(0.01, 1.0, (N, 8)).astype(np.float32)
    ws[2, 0] = 0.01  # below the 0.05 floor
    for i, l in lists.items():
        nodes["neighbors"][i, : len(l)] = l
        nodes["neighbor_w"][i] = ws[i]
        nodes["num_neighbors"][i] = len(l)
    return nodes


def _corrs(nodes, n=40, seed=4):
    rng = np.random.default_rng(seed)
    N = len(nodes)
    c = np.zeros(n, L.CORR)
    P = nodes["pos"].astype(np.float64)
    for k in range(n):
        ids = rng.choice(N, 4, replace=False)
        if k % 7 == 3:
            ids[3] = -1
        ws = rng.uniform(0.1, 1.0, 4)
        ws[ids < 0] = 0
        if k % 11 == 5:
            ws[2] = 1e-9  # below the 1e-8 validity threshold
        ws = ws / ws[ws > 1e-8].sum()
        ws[ws <= 1e-8] = 0 if k % 11 != 5 else ws[ws <= 1e-8]
        src = P[ids[0]] + rng.normal(0, 0.01, 3)
        nrm = rng.normal(0, 1, 3)
        nrm /= np.linalg.norm(nrm)
        c[k]["src"] = src
        c[k]["dst"] = src - nrm * rng.uniform(-0.15, 0.15) + rng.normal(0, 0.002, 3)
        c[k]["normal"] = nrm
        c[k]["node_ids"] = ids
        c[k]["node_ws"] = ws
        c[k]["weight"] = [1.0, 0.5, -1.0, 2.0][k % 4] if k % 13 else 1.0
        c[k]["valid"] = 0 if k % 17 == 16 else 1
    return c


def _huber(r, d):
    return 1.0 if abs(r) <= d else d / abs(r)


def _brute_force(orc, nodes, corrs, T, lam_s, lam_d, consistent_sign=0):
    """Dense normal equations from finite-difference Jacobians.  Unless
    consistent_sign, the ARAP rotation columns are negated (Q1)."""
    N = len(nodes)
    P = nodes["pos"].astype(np.float64)
    rp, ci = orc.pattern(nodes)
    inpat = np.zeros((N, N), bool)
    for i in range(N):
        inpat[i, ci[rp[i]:rp[i + 1]]] = True
    A = np.zeros((6 * N, 6 * N))
    b = np.zeros(6 * N)
    I8 = np.array([0, 0, 0, 1, 0, 0, 0, 0.0])
    eps = 1e-6
    # data term: J_k by central differences of n.(warp(src) - dst) about identity
    for c in corrs:
        if not c["valid"]:
            continue
        src, dst, n = c["src"].astype(np.float64), c["dst"].astype(np.float64), c["normal"].astype(np.float64)
        r = float(n @ (src - dst))
        fw = max(float(c["weight"]), 0.0) * _huber(np.float32(r), 0.2)
        if fw <= 1e-8:
            continue
        J = {}
        for k in range(4):
            nid = int(c["node_ids"][k])
            w = float(c["node_ws"][k])
            if nid < 0 or nid >= N or c["node_ws"][k] <= 1e-8:
                continue
            g = np.zeros(6)
            for a in range(6):
                d = np.zeros(6)
                d[a] = eps
                vals = []
                for sgn in (1, -1):
                    # warp of src through node nid alone moved by exp(+-eps e_a), blended with weight w
                    # at identity the blend derivative is w times the single-node derivative
                    Tk = _twist(orc, sgn * d, I8)
                    q = _xf_c(orc, Tk, src, P[nid])
                    vals.append(float(n @ (q - dst)))
                g[a] = w * (vals[0] - vals[1]) / (2 * eps)
            J[k] = (nid, g)
        for k, (ni, gk) in J.items():
            b[6 * ni:6 * ni + 6] += -fw * gk * r
        for k, (ni, gk) in J.items():
            for l, (nj, gl) in J.items():
                if not inpat[ni, nj]:
                    continue  # Q2
                A[6 * ni:6 * ni + 6, 6 * nj:6 * nj + 6] += fw * np.outer(gk, gl)
    # ARAP: r = T_i(x_j) - T_j(x_j), edges i->j with j > i only (Q3)
    for i in range(N):
        for k in range(int(nodes["num_neighbors"][i])):
            j = int(nodes["neighbors"][i][k])
            if j < 0 or j >= N or j <= i:
                continue
            xi, xj = P[i], P[j]
            Tij = _xf_c(orc, T[i], xj, xi)
            Tii = _xf_c(orc, T[i], xi, xi)
            Tjj = _xf_c(orc, T[j], xj, xj)
            r = Tij - Tjj
            dr, dc = np.linalg.norm(xi - xj), np.linalg.norm(Tii - Tjj)
            strain = abs(dc - dr) / max(dr, 1e-6)
            lam = lam_s * max(float(nodes["neighbor_w"][i][k]), float(np.float32(0.05))) * \
                _huber(np.linalg.norm(r), 0.05) * np.exp(-(strain / 0.25) ** 2)
            if lam <= 1e-8:
                continue
            Jfull = np.zeros((3, 6 * N))
            for node, sign in ((i, 1.0), (j, -1.0)):
                for a in range(6):
                    d = np.zeros(6)
                    d[a] = eps
                    rr = []
                    for sg in (1, -1):
                        Tn = _twist(orc, sg * d, T[node])
                        if node == i:
                            rr.append(_xf_c(orc, Tn, xj, xi) - Tjj)
                        else:
                            rr.append(Tij - _xf_c(orc, Tn, xj, xj))
                    col = (rr[0] - rr[1]) / (2 * eps)
                    if a < 3 and not consistent_sign:
                        col = -col  # Q1: the reference's rotation Jacobian has the opposite sign
                    Jfull[:, 6 * node + a] = col
            A += lam * Jfull.T @ Jfull
            b += -lam * Jfull.T @ r
    for i in range(N):
        for k in range(6):
            d = A[6 * i + k, 6 * i + k]
            fl = 1e-3 if k < 3 else 1e-2
            A[6 * i + k, 6 * i + k] = d + lam_d * max(abs(d), float(np.float32(fl)))
    return A, b


def _bsr_to_dense(rp, ci, blocks, N):
    A = np.zeros((6 * N, 6 * N))
    for i in range(N):
        for q in range(rp[i], rp[i + 1]):
            A[6 * i:6 * i + 6, 6 * ci[q]:6 * ci[q] + 6] = blocks[q]
    return A


def test_pattern_symmetrised_sorted(orc):
    nodes = _graph()
    rp, ci = orc.pattern(nodes)
    N = len(nodes)
    exp = [set([i]) for i in range(N)]
    for i in range(N):
        for j in nodes["neighbors"][i][: nodes["num_neighbors"][i]]:
            exp[i].add(int(j))
            exp[int(j)].add(i)
    for i in range(N):
        assert list(ci[rp[i]:rp[i + 1]]) == sorted(exp[i])


@pytest.mark.parametrize("consistent", [0, 1])
@pytest.mark.parametrize("moved", [False, True])
def test_assembly_equals_finite_difference_normal_equations(orc, moved, consistent):
    nodes = _graph()
    corrs = _corrs(nodes)
    N = len(nodes)
    rng = np.random.default_rng(9)
    T = np.zeros((N, 8))
    T[:, 3] = 1
    if moved:
        for i in range(N):
            T[i] = _twist(orc, np.concatenate([rng.normal(0, 0.03, 3), rng.normal(0, 0.004, 3)]), T[i])
    rp, ci, A, b, M = orc.assemble_f64(corrs, nodes, T, 10.0, 1e-2, consistent_sign=consistent)
    Ad = _bsr_to_dense(rp, ci, A, N)
    Ab, bb = _brute_force(orc, nodes, corrs, T, 10.0, 1e-2, consistent)
    scale = np.abs(Ab).max()
    np.testing.assert_allclose(Ad, Ab, atol=1e-7 * scale)
    np.testing.assert_allclose(b, bb, atol=1e-7 * max(np.abs(bb).max(), 1e-12))
    if moved:
        assert np.abs(bb).max() > 1e-6  # the ARAP residuals contribute to b
    # FP32 restatement agrees with FP64 to single precision
    T32 = T.astype(np.float32).view(L.DUALQUAT).reshape(-1)
    _, _, A32, b32, _ = orc.assemble_f32(corrs, nodes, T32, 10.0, 1e-2, consistent_sign=consistent)
    np.testing.assert_allclose(A32, A, atol=1e-5 * scale)
    if moved:
        # the two modes differ exactly in the sign of the ARAP rotation rows:
        # rot-rot and trans-trans blocks agree, b differs in the rotation part
        _, _, A_o, b_o, _ = orc.assemble_f64(corrs, nodes, T, 10.0, 1e-2, consistent_sign=1 - consistent)
        np.testing.assert_allclose(A_o[:, 3:, 3:], A[:, 3:, 3:], atol=1e-12 * scale)
        assert np.abs(b_o.reshape(-1, 6)[:, :3] - b.reshape(-1, 6)[:, :3]).max() > 1e-9


def test_block_jacobi_inverse_and_fallback(orc):
    nodes = _graph()
    corrs = _corrs(nodes)
    N = len(nodes)
    T = np.zeros((N, 8))
    T[:, 3] = 1
    rp, ci, A, b, M = orc.assemble_f64(corrs, nodes, T, 10.0, 1e-2)
    for i in range(N):
        q = rp[i] + list(ci[rp[i]:rp[i + 1]]).index(i)
        D = A[q] + float(np.float32(1e-4)) * np.eye(6)  # compute_preconditioner_kernel adds 1e-4f I
        np.testing.assert_allclose(M[i] @ D, np.eye(6), atol=1e-9)
    # a non-SPD diagonal block falls back to diag(1 / (a_kk + 1e-4 + 1e-6));
    # a negative damping factor makes the diagonal negative
    _, _, A2, _, M2 = orc.assemble_f64(corrs, nodes, T, 10.0, -2.0)
    bad = 0
    for i in range(N):
        q = rp[i] + list(ci[rp[i]:rp[i + 1]]).index(i)
        D = A2[q] + float(np.float32(1e-4)) * np.eye(6)
        try:
            np.linalg.cholesky(D)
        except np.linalg.LinAlgError:
            bad += 1
            np.testing.assert_allclose(M2[i], np.diag(1.0 / (np.diag(D) + float(np.float32(1e-6)))), rtol=1e-12)
    assert bad > 0


def _textbook_pcg(A, b, Minv_blocks, maxit, tol):
    """Saad, Iterative Methods, Alg. 9.1, block-diagonal preconditioner, x0 = 0."""
    n = len(b)
    Mi = np.zeros((n, n))
    for i, blk in enumerate(Minv_blocks):
        Mi[6 * i:6 * i + 6, 6 * i:6 * i + 6] = blk
    x = np.zeros(n)
    r = b.copy()
    z = Mi @ r
    p = z.copy()
    rz = r @ z
    norms = []
    for _ in range(maxit):
        Ap = A @ p
        pAp = p @ Ap
        if abs(pAp) < 1e-12:
            break
        al = rz / pAp
        x += al * p
        r -= al * Ap
        norms.append(np.linalg.norm(r))
        if norms[-1] < tol:
            break
        z = Mi @ r
        rzn = r @ z
        p = z + rzn / rz * p
        rz = rzn
    return x, norms


def test_pcg_solves_and_follows_textbook_trace(orc):
    nodes = _graph()
    corrs = _corrs(nodes)
    N = len(nodes)
    T = np.zeros((N, 8))
    T[:, 3] = 1
    rp, ci, A, b, M = orc.assemble_f64(corrs, nodes, T, 10.0, 1e-2)
    b = b * 1e4  # keep p.Ap above the 1e-12 breakdown threshold down to a tight residual
    Ad = _bsr_to_dense(rp, ci, A, N)
    np.testing.assert_allclose(Ad, Ad.T, atol=1e-12 * np.abs(Ad).max())
    assert np.all(np.linalg.eigvalsh(Ad) > 0)
    x, it, fl, r, trace = orc.pcg_f64(rp, ci, A, b, M, 200, 1e-14, trace=True)
    np.testing.assert_allclose(Ad @ x, b, atol=1e-10 * np.linalg.norm(b))
    xt, norms = _textbook_pcg(Ad, b, M, 10, 0.0)
    x10, it10, _, _, tr10 = orc.pcg_f64(rp, ci, A, b, M, 10, 0.0, trace=True)
    assert it10 == 10
    np.testing.assert_allclose(tr10, norms, rtol=1e-9)
    np.testing.assert_allclose(x10, xt, rtol=1e-9, atol=1e-14)


def test_pcg_stopping_rules(orc):
    nodes = _graph()
    corrs = _corrs(nodes)
    N = len(nodes)
    T = np.zeros((N, 8))
    T[:, 3] = 1
    rp, ci, A, b, M = orc.assemble_f64(corrs, nodes, T, 10.0, 1e-2)
    # b = 0: p.Ap = 0 at the first iteration -> breakdown exit, x stays 0
    x, it, fl, r, _ = orc.pcg_f64(rp, ci, A, np.zeros_like(b), M, 10, 1e-4)
    assert it == 0 and fl & 1 and not x.any()
    # absolute tolerance (Q5): stop at the first ||r|| < tol
    _, norms = _textbook_pcg(_bsr_to_dense(rp, ci, A, N), b, M, 50, 0.0)
    tol = 0.5 * (norms[2] + norms[3])
    if norms[3] < norms[2]:
        x, it, fl, r, _ = orc.pcg_f64(rp, ci, A, b, M, 50, tol)
        assert fl & 2 and it == 1 + next(k for k, v in enumerate(norms) if v < tol)
    # relative tolerance option: tol * ||b||
    b = b * 1e4
    x, it, fl, r, _ = orc.pcg_f64(rp, ci, A, b, M, 200, 1e-6, rel=1)
    assert fl & 2 and r < 1e-6 * np.linalg.norm(b)


def test_gn_drives_point_to_plane_residuals_to_zero(orc):
    """Exact correspondences of a rigidly moved sphere cap: the GN loop with
    re-warped sources and the twist update drives every point-to-plane
    residual to zero (the data term alone, Levenberg damping 1; tangential
    sliding on a sphere is unobservable for point-to-plane, so only the
    residual along the normal is pinned)."""
    rng = np.random.default_rng(12)
    # canonical: a spherical cap, nodes on it
    u = rng.normal(0, 1, (4000, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    u = u[u[:, 2] < -0.2]
    c0 = np.array([0, 0, 1.0])
    pts = c0 + 0.2 * u
    nrm = u
    from cudafusion_amd.scene import sample_nodes
    nodes = sample_nodes(pts, nrm, 0.05, 0.05, (-0.5, -0.5, 0.5), 0.01)
    N = len(nodes)
    w = np.array([0.02, -0.015, 0.03])
    th = np.linalg.norm(w)
    R = np.eye(3) + np.sin(th) / th * _hat(w) + (1 - np.cos(th)) / th ** 2 * _hat(w) @ _hat(w)
    t = np.array([0.004, -0.003, 0.006])
    tgt = (pts - c0) @ R.T + c0 + t
    n = len(pts)
    corrs = np.zeros(n, L.CORR)
    P = nodes["pos"].astype(np.float64)
    for k in range(n):
        d2 = ((P - pts[k]) ** 2).sum(1)
        ids = np.argsort(d2, kind="stable")[:4]
        ws = np.exp(-d2[ids] / (2 * 0.05 ** 2))
        corrs[k]["node_ids"] = ids
        corrs[k]["node_ws"] = ws / ws.sum()
        corrs[k]["dst"] = tgt[k]
        corrs[k]["normal"] = R @ nrm[k]
        corrs[k]["weight"] = 1
        corrs[k]["valid"] = 1
    can = pts.astype(np.float32)
    res = []
    for it in (5, 20):
        gn = L.gn_params(it, 1.0, 0.12, 0.04, 0, 0.03)
        sp = L.solver_params(it, 100, 0.0, 0.0, 1.0, 1.0)
        T64, dx, tot, done, last = orc.solve_gn_f64(corrs, can, nodes, orc.identity_dq64(N), gn, sp)
        assert done == it
        out = np.zeros(n * 3)
        ids = np.ascontiguousarray(corrs["node_ids"])
        ws = np.ascontiguousarray(corrs["node_ws"])
        orc.load().orc_warp_points_f64(can.ctypes.data, ids.ctypes.data, ws.ctypes.data, n, nodes.ctypes.data,
                                       T64.ctypes.data, N, out.ctypes.data)
        res.append(np.abs(((out.reshape(-1, 3) - tgt) * (nrm @ R.T)).sum(1)))
    r0 = np.abs(((pts - tgt) * (nrm @ R.T)).sum(1))
    assert r0.max() > 5e-3
    assert res[0].max() < 1e-3 and res[1].max() < 2e-4 and res[1].mean() < 2e-5


def test_refinement_step_completes_the_solve(orc):
    """The optional refinement step of the oracle's GN loop: after a 2-iteration
    PCG, r = b - A x and A d = r by as many PCG iterations as unknowns from
    d = 0 (exact in exact arithmetic), x += d.  The returned step is then numpy's dense
    solve of the assembled system (a refinement that solved A d = b, or that
    replaced x by d, would not be), while the 2-iteration step alone is not."""
    nodes = _graph()
    corrs = _corrs(nodes)
    N = len(nodes)
    T64 = orc.identity_dq64(N)
    rp, ci, A, b, _ = orc.assemble_f64(corrs, nodes, T64, 10.0, 1e-2)
    dense = _bsr_to_dense(rp, ci, A, N)
    x_star = np.linalg.solve(dense, b)
    gn = L.gn_params(1, 1.0, 0.0, 0.0, 0, 0.0)
    steps = {}
    for refine in (0, 6 * N):
        sp = L.solver_params(1, 2, 0.0, 10.0, 1.0, 1e-2, refine_iterations=refine)
        _, dx, _, done, _ = orc.solve_gn_f64(corrs, None, nodes, T64, gn, sp)
        assert done == 1
        steps[refine] = dx
    rel = lambda v: np.linalg.norm(v - x_star) / np.linalg.norm(x_star)  # noqa: E731
    assert rel(steps[6 * N]) < 1e-7  # CG in n steps, FP64 rounding of an ill-conditioned system
    assert rel(steps[0]) > 1e-3


def test_refinement_step_block_jacobi_form(orc):
    """The refinement step with the block-Jacobi inverse as its correction
    solve: the returned step is x + M^-1 (b - A x) for the PCG's x and the
    oracle's own assembled A, b and M, and it lowers the residual of a
    2-iteration PCG on this system."""
    nodes = _graph()
    corrs = _corrs(nodes)
    N = len(nodes)
    T64 = orc.identity_dq64(N)
    rp, ci, A, b, M = orc.assemble_f64(corrs, nodes, T64, 10.0, 1e-2)
    dense = _bsr_to_dense(rp, ci, A, N)
    gn = L.gn_params(1, 1.0, 0.0, 0.0, 0, 0.0)
    out = {}
    for jac in (0, 1):
        sp = L.solver_params(1, 2, 0.0, 10.0, 1.0, 1e-2, refine_jacobi=jac)
        out[jac] = orc.solve_gn_f64(corrs, None, nodes, T64, gn, sp)[1]
    x = out[0]
    r = b - dense @ x
    z = np.concatenate([M[i] @ r[6 * i:6 * i + 6] for i in range(N)])
    assert np.linalg.norm(z) > 1e-3 * np.linalg.norm(x)  # the correction is not negligible
    # r = b - A x cancels: its rounding depends on the order of the sums
    np.testing.assert_allclose(out[1] - x, z, rtol=1e-5, atol=1e-6 * np.abs(z).max())
    assert np.linalg.norm(b - dense @ out[1]) < np.linalg.norm(r)


def _data_jacobians(orc, nodes, corrs):
    """Per valid correspondence: (fw, [(node, J_k)]) with J_k by central
    differences of n.(warp(src) - dst) about the identity."""
    N = len(nodes)
    P = nodes["pos"].astype(np.float64)
    I8 = np.array([0, 0, 0, 1, 0, 0, 0, 0.0])
    eps = 1e-6
    out = []
    for c in corrs:
        if not c["valid"]:
            continue
        src, dst, n = c["src"].astype(np.float64), c["dst"].astype(np.float64), c["normal"].astype(np.float64)
        r = float(n @ (src - dst))
        fw = max(float(c["weight"]), 0.0) * _huber(np.float32(r), 0.2)
        if fw <= 1e-8:
            continue
        Js = []
        for k in range(4):
            nid = int(c["node_ids"][k])
            if nid < 0 or nid >= N or c["node_ws"][k] <= 1e-8:
                continue
            g = np.zeros(6)
            for a in range(6):
                d = np.zeros(6)
                d[a] = eps
                vals = [float(n @ (_xf_c(orc, _twist(orc, sg * d, I8), src, P[nid]) - dst)) for sg in (1, -1)]
                g[a] = float(c["node_ws"][k]) * (vals[0] - vals[1]) / (2 * eps)
            Js.append((nid, g))
        out.append((fw, Js))
    return out


def test_psd_dropped_pairs_keeps_the_data_term_psd(orc):
    """psd_dropped_pairs (opt-in): a skinning pair (k, l) outside the graph
    pattern adds f J_k^T J_k to block (k, k) instead of being dropped (Q2).
    The data term is then the full J^T J plus, for every dropped unordered
    pair, v v^T with v = e_k J_k - e_l J_l (finite-difference Jacobians),
    hence positive semidefinite, while the reference's assembly of the same
    correspondences differs from it by exactly those terms."""
    nodes = _graph()
    corrs = _corrs(nodes)
    N = len(nodes)
    T64 = orc.identity_dq64(N)
    rp, ci, Aq, _, _ = orc.assemble_f64(corrs, nodes, T64, 0.0, 0.0)
    _, _, Ap, _, _ = orc.assemble_f64(corrs, nodes, T64, 0.0, 0.0, psd_pairs=1)
    Dq, Dp = _bsr_to_dense(rp, ci, Aq, N), _bsr_to_dense(rp, ci, Ap, N)
    inpat = np.zeros((N, N), bool)
    for i in range(N):
        inpat[i, ci[rp[i]:rp[i + 1]]] = True
    full = np.zeros((6 * N, 6 * N))
    corr = np.zeros((6 * N, 6 * N))
    ndrop = 0
    for fw, Js in _data_jacobians(orc, nodes, corrs):
        v = np.zeros(6 * N)
        for ni, g in Js:
            v[6 * ni:6 * ni + 6] += g
        full += fw * np.outer(v, v)
        for a in range(len(Js)):
            for b2 in range(a + 1, len(Js)):
                (ni, gi), (nj, gj) = Js[a], Js[b2]
                if inpat[ni, nj]:
                    continue
                ndrop += 1
                u = np.zeros(6 * N)
                u[6 * ni:6 * ni + 6] += gi
                u[6 * nj:6 * nj + 6] -= gj
                corr += fw * np.outer(u, u)
    assert ndrop > 10
    scale = np.abs(full).max()
    np.testing.assert_allclose(Dp, full + corr, atol=1e-6 * scale)
    assert np.linalg.eigvalsh(Dp).min() > -1e-9 * scale
    # the reference's system (Q2) is the full J^T J with the out-of-pattern blocks removed
    np.testing.assert_allclose(Dq, full * np.kron(inpat, np.ones((6, 6))), atol=1e-6 * scale)
