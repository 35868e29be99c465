"""The FP64 definition of the learned preconditioner (oracle/learned.cpp):
its analytic gradient against central finite differences (both variants),
the preconditioner application against a dense BJ^-1 + L L^T built here
from the blocks, and SPD-ness of the result."""
import numpy as np
import pytest

from oracle import pyoracle as po


def _graph(N=36, seed=0):
    """A small deformation graph (points on a noisy sheet, 8 nearest
    neighbours within reach) with a random SPD block system on its pattern."""
    from cudafusion_amd import _lib as L
    rng = np.random.default_rng(seed)
    P = np.stack([rng.uniform(0, 0.2, N), rng.uniform(0, 0.2, N), 1.0 + 0.01 * rng.normal(size=N)], 1)
    nodes = np.zeros(N, L.NODE)
    nodes["pos"] = P
    nodes["normal"] = [0, 0, -1]
    nodes["radius"] = 0.03
    nodes["neighbors"] = -1
    for i in range(N):
        d = ((P - P[i]) ** 2).sum(1)
        cand = [j for j in np.argsort(d, kind="stable") if j != i and d[j] < 0.08 ** 2][:8]
        nodes["neighbors"][i, :len(cand)] = cand
        nodes["neighbor_w"][i, :len(cand)] = np.exp(-d[cand] / (2 * 0.03 ** 2))
        nodes["num_neighbors"][i] = len(cand)
    rp, ci = po.pattern(nodes)
    nb = len(ci)
    A = np.zeros((nb, 6, 6))
    blocks = {}
    for i in range(N):
        for q in range(rp[i], rp[i + 1]):
            j = ci[q]
            if j < i:
                blocks[(i, j)] = 0.3 * rng.normal(size=(6, 6))
    for i in range(N):
        for q in range(rp[i], rp[i + 1]):
            j = ci[q]
            if j == i:
                B = rng.normal(size=(6, 6))
                A[q] = B @ B.T + 8 * np.eye(6) * 10 ** rng.uniform(-1, 1, 6)
            elif j < i:
                A[q] = blocks[(i, j)]
            else:
                A[q] = blocks[(j, i)].T
    Ad = po.lp_diag_blocks(rp, ci, A)
    Minv = np.linalg.inv(Ad.reshape(N, 6, 6) + 1e-4 * np.eye(6))
    return nodes, rp, ci, A, Ad, Minv


def _params(rng, scale=1.0):
    offs = po.lp_offsets()
    p = np.zeros(po.lp_param_count())
    fan = {"W0": 28, "Wm0": 68, "Wm1": 68, "Wm2": 68, "Wu0": 128, "Wu1": 128, "Wu2": 128, "Wo": 64, "We": 132}
    for name, (o, n) in offs.items():
        if name.startswith("b"):
            p[o:o + n] = 0.1 * rng.normal(size=n)
        else:
            p[o:o + n] = rng.normal(size=n) / np.sqrt(fan[name]) * (0.3 if name in ("Wo", "We") else 1.0) * scale
    return p


def _dense(nodes, rp, ci, A):
    N = len(nodes)
    D = np.zeros((6 * N, 6 * N))
    for i in range(N):
        for q in range(rp[i], rp[i + 1]):
            D[6 * i:6 * i + 6, 6 * ci[q]:6 * ci[q] + 6] = A[q]
    return D


@pytest.mark.parametrize("variant", [0, 1], ids=["block_diagonal", "graph_pattern"])
def test_loss_gradient_matches_finite_differences(variant):
    nodes, rp, ci, A, Ad, Minv = _graph()
    N = len(nodes)
    rng = np.random.default_rng(1)
    x, Cm = po.lp_features_f64(nodes, Ad, Minv)
    assert np.isfinite(x).all() and np.abs(Cm).sum() > 0
    p = _params(rng)
    V = rng.normal(size=(3, 6 * N))
    loss, g = po.lp_loss_grad_f64(p, variant, nodes, x, Cm, Minv, rp, ci, A, V)
    offs = po.lp_offsets()
    # every parameter group, including the edge head for the graph variant
    picks = []
    for name, (o, n) in offs.items():
        if variant == 0 and name in ("We", "be"):
            assert np.all(g[o:o + n] == 0)
            continue
        picks += list(o + rng.choice(n, size=min(n, 6), replace=False))
    # central differences at three step sizes: ReLU kinks near a sample
    # spoil the larger steps, round-off the smaller ones; the best must agree
    gmax = np.abs(g).max()
    bad = []
    for k in picks:
        errs = []
        for eps in (1e-7, 1e-6, 1e-5):
            pp, pm = p.copy(), p.copy()
            pp[k] += eps
            pm[k] -= eps
            lp_, _ = po.lp_loss_grad_f64(pp, variant, nodes, x, Cm, Minv, rp, ci, A, V, grad=False)
            lm_, _ = po.lp_loss_grad_f64(pm, variant, nodes, x, Cm, Minv, rp, ci, A, V, grad=False)
            errs.append(abs((lp_ - lm_) / (2 * eps) - g[k]))
        if min(errs) > 1e-5 * abs(g[k]) + 1e-7 * gmax:
            bad.append((k, g[k], min(errs)))
    print(f"[lp grad] variant {variant}: loss {loss:.4e}, {len(picks)} parameters checked, {len(bad)} off")
    assert not bad, bad[:5]
    assert gmax > 0


def test_apply_equals_dense_preconditioner_and_is_spd():
    nodes, rp, ci, A, Ad, Minv = _graph(N=24, seed=3)
    N = len(nodes)
    rng = np.random.default_rng(2)
    x, Cm = po.lp_features_f64(nodes, Ad, Minv)
    p = _params(rng, scale=2.0)
    Td, Te = po.lp_forward_f64(p, 1, nodes, x)
    # dense L: L_ii = C_i lower(Td_i), L_ij = C_i Te_is for neighbour slots j < i
    Ld = np.zeros((6 * N, 6 * N))
    tril = np.tril_indices(6)
    for i in range(N):
        C = Cm[i].reshape(6, 6)
        T = np.zeros((6, 6))
        T[tril] = Td[i]
        Ld[6 * i:6 * i + 6, 6 * i:6 * i + 6] = C @ T
        for s in range(nodes["num_neighbors"][i]):
            j = nodes["neighbors"][i, s]
            if 0 <= j < i:
                Ld[6 * i:6 * i + 6, 6 * j:6 * j + 6] = C @ Te[i, s].reshape(6, 6)
        assert np.allclose(C @ C.T, Minv[i], atol=1e-12)
    M = np.zeros((6 * N, 6 * N))
    for i in range(N):
        M[6 * i:6 * i + 6, 6 * i:6 * i + 6] = Minv[i]
    M += Ld @ Ld.T
    R = rng.normal(size=(2, 6 * N))
    Z = po.lp_apply_f64(1, nodes, Cm, Minv, Td, Te, R)
    np.testing.assert_allclose(Z, R @ M.T, rtol=1e-10, atol=1e-12)
    assert np.linalg.eigvalsh(0.5 * (M + M.T)).min() > 0
    # the block-diagonal variant keeps only the diagonal blocks
    Z0 = po.lp_apply_f64(0, nodes, Cm, Minv, Td, np.zeros_like(Te), R)
    Ld0 = np.zeros_like(Ld)
    for i in range(N):
        Ld0[6 * i:6 * i + 6, 6 * i:6 * i + 6] = Ld[6 * i:6 * i + 6, 6 * i:6 * i + 6]
    M0 = M - Ld @ Ld.T + Ld0 @ Ld0.T
    np.testing.assert_allclose(Z0, R @ M0.T, rtol=1e-10, atol=1e-12)


def test_learned_pcg_reduces_to_block_jacobi_and_solves():
    """With zero heads the preconditioner is block-Jacobi; with any heads the
    FP64 PCG still solves the system (numpy dense solve)."""
    nodes, rp, ci, A, Ad, Minv = _graph(N=30, seed=4)
    N = len(nodes)
    rng = np.random.default_rng(5)
    x, Cm = po.lp_features_f64(nodes, Ad, Minv)
    b = rng.normal(size=6 * N)
    D = _dense(nodes, rp, ci, A)
    xs = np.linalg.solve(D, b)
    zero = np.zeros((N, 21)), np.zeros((N, 8, 36))
    it0, tr0, x0 = po.lp_pcg_f64(0, nodes, rp, ci, A, b, Cm, Minv, *zero, 500, 1e-10)
    _, it_bj, _, _, _ = po.pcg_f64(rp, ci, A, b, Minv.reshape(N, 36), 500, 1e-10, rel=1)
    assert it0 == it_bj and tr0 < 1e-9
    np.testing.assert_allclose(x0, xs, rtol=1e-7, atol=1e-9)
    p = _params(rng)
    Td, Te = po.lp_forward_f64(p, 1, nodes, x)
    it1, tr1, x1 = po.lp_pcg_f64(1, nodes, rp, ci, A, b, Cm, Minv, Td, Te, 500, 1e-10)
    assert tr1 < 1e-9
    np.testing.assert_allclose(x1, xs, rtol=1e-7, atol=1e-9)
