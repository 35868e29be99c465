"""Learned preconditioner on the device (csrc/learned.hip) against its FP64
definition (oracle/learned.cpp) on a system the solver assembled: features,
Cholesky factors, network outputs, probe loss and every parameter's
gradient; then training reduces the loss, and the folded block-diagonal
preconditioner runs in the solver's PCG to the tolerance."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _system():
    import torch

    from cudafusion_amd.harness import Harness
    h = Harness("sphere", frames=2, arap_consistent_sign=True, validate=False, pcg_pipelined=True)
    h.setup()
    h.step(1)
    torch.cuda.synchronize()
    return h


@pytest.mark.parametrize("variant", [0, 1], ids=["block_diagonal", "graph_pattern"])
def test_forward_and_gradient_match_fp64_oracle(gpu, orc, variant):
    import torch

    from cudafusion_amd import learned as LP
    from oracle import pyoracle as po
    h = _system()
    lib, N = h.lib, h.N
    solver = lib.nrf_pipeline_solver(h.pipe)
    d_nodes = lib.nrf_pipeline_nodes(h.pipe)
    rp, ci, A, b, Minv = LP.solver_system(lib, solver, N)
    lp = LP.LearnedPreconditioner(variant, max_nodes=N, max_probes=4, seed=3)
    # heads at full scale so that L (and every gradient) is far from zero
    lp.set_params(LP.init_params(seed=3, head_scale=1.0))
    lp.set_graph(h.nodes)
    lp.forward(solver, d_nodes)
    Cg, Tdg, Teg, xg = lp.blocks()
    Ad = po.lp_diag_blocks(rp, ci, A.astype(np.float64))
    x64, C64 = po.lp_features_f64(h.nodes, Ad, Minv.astype(np.float64))
    ex = np.abs(xg - x64).max()
    ec = np.abs(Cg - C64).max() / np.abs(C64).max()
    pb = LP.bf16_round(lp.params())
    Td64, Te64 = po.lp_forward_f64(pb, variant, h.nodes, xg.astype(np.float64))
    etd = np.abs(Tdg - Td64).max() / np.abs(Td64).max()
    ete = np.abs(Teg - Te64).max() / max(np.abs(Te64).max(), 1e-30)
    rng = np.random.default_rng(7)
    V = rng.normal(size=(4, 6 * N)).astype(np.float32)
    loss = lp.loss_grad(solver, d_nodes, torch.from_numpy(V).cuda())
    g = lp.grads()
    loss64, g64 = po.lp_loss_grad_f64(pb, variant, h.nodes, xg.astype(np.float64), Cg.astype(np.float64),
                                      Minv.astype(np.float64), rp, ci, A.astype(np.float64), V.astype(np.float64))
    print(f"[lp gpu] variant {variant} N={N}: features {ex:.1e}, C {ec:.1e}, Td {etd:.1e}, Te {ete:.1e}, "
          f"loss {loss:.6e} vs {loss64:.6e}", flush=True)
    assert ex < 1e-5 and ec < 1e-5
    assert etd < 1e-4 and (variant == 0 and np.abs(Teg).max() == 0 or ete < 1e-4)
    assert abs(loss - loss64) < 1e-4 * loss64
    worst = 0.0
    for name, (o, n) in po.lp_offsets().items():
        ref = g64[o:o + n]
        scale = np.abs(ref).max()
        if variant == 0 and name in ("We", "be"):
            assert np.all(g[o:o + n] == 0)
            continue
        err = np.abs(g[o:o + n] - ref).max() / scale
        worst = max(worst, err)
        assert err < 2e-3, (name, err)
    print(f"[lp gpu] variant {variant}: worst gradient group error {worst:.1e} (relative to the group's largest)",
          flush=True)
    lp.close()
    h.close()


def test_training_reduces_loss_and_folded_pcg_converges(gpu):
    import ctypes as C

    import torch

    from cudafusion_amd import _lib as L
    from cudafusion_amd import learned as LP
    h = _system()
    lib, N = h.lib, h.N
    solver = lib.nrf_pipeline_solver(h.pipe)
    d_nodes = lib.nrf_pipeline_nodes(h.pipe)
    lp = LP.LearnedPreconditioner(0, max_nodes=N, max_probes=3, seed=0)
    lp.set_graph(h.nodes)
    # the training objective of train_lp: the solution x* of the system as the
    # main probe (|M^-1 b - x*|^2 up to scale) and two small Gaussian probes
    from cudafusion_amd.train_lp import pcg_solve
    xs = torch.zeros(6 * N, dtype=torch.float32, device="cuda")
    pcg_solve(lib, solver, h.sp, xs, 1e-7, 2000)
    v0 = xs / xs.square().mean().sqrt()
    gen = torch.Generator(device="cuda").manual_seed(0)
    losses = []
    for it in range(80):
        V = torch.cat([v0[None], 0.1 * torch.randn(2, 6 * N, device="cuda", generator=gen)], 0).contiguous()
        lp.forward(solver, d_nodes)
        losses.append(lp.loss_grad(solver, d_nodes, V))
        lp.adam(1e-3)
    print(f"[lp train] loss {losses[0]:.4f} -> {losses[-1]:.4f}", flush=True)
    assert np.mean(losses[-10:]) < 0.98 * np.mean(losses[:5])
    # fold: Minv_i + L_ii L_ii^T
    rp, ci, A, b, M0 = LP.solver_system(lib, solver, N)
    lp.forward(solver, d_nodes)
    Cg, Tdg, _, _ = lp.blocks()
    lp.fold(solver)
    torch.cuda.synchronize()
    M1 = LP.solver_system(lib, solver, N)[4]
    tril = np.tril_indices(6)
    for i in range(0, N, max(1, N // 50)):
        T = np.zeros((6, 6))
        T[tril] = Tdg[i]
        Li = Cg[i].reshape(6, 6).astype(np.float64) @ T
        np.testing.assert_allclose(M1[i], M0[i] + Li @ Li.T, rtol=1e-5, atol=1e-6 * np.abs(M0[i]).max())
    # the solver's PCG with the folded preconditioner, to 1e-6 relative
    sp = L.SolverParams.from_buffer_copy(bytes(h.sp))
    sp.pcg_iterations = 300
    sp.pcg_tolerance = 1e-6
    sp.relative_tolerance = 1
    L.check(lib.nrf_solver_set_params(solver, C.byref(sp)), "set_params")
    x = torch.zeros(6 * N, dtype=torch.float32, device="cuda")
    res = L.PcgResult()
    L.check(lib.nrf_solver_resolve(solver, x.data_ptr(), C.byref(res), torch.cuda.current_stream().cuda_stream),
            "resolve")
    out = (C.c_double * 2)()
    L.check(lib.nrf_solver_true_residual(solver, x.data_ptr(), out, torch.cuda.current_stream().cuda_stream),
            "true_residual")
    rel = out[0] / out[1]
    print(f"[lp pcg] learned block-diagonal: {res.iterations} iterations, true relative residual {rel:.2e}",
          flush=True)
    assert res.flags & 2 and rel < 1e-5
    lp.close()
    h.close()


@pytest.fixture(scope="module")
def held_out(gpu, orc):
    """The trained weights (cudafusion_amd/lp_weights_v0.npz, _v1.npz, made by
    cudafusion_amd.train_lp) on a held-out capsule-body sequence (seed 1000),
    frames 1 and 2: PCG iterations to 1e-6 relative residual with block-
    Jacobi and with each learned variant.  Block diagonal: the solver's PCG
    with the folded preconditioner (device), judged by the FP64 true
    residual; graph pattern: the FP64 oracle PCG with M^-1 = BJ^-1 + L L^T,
    L from the device forward."""
    import os

    import torch

    from cudafusion_amd import learned as LP
    from cudafusion_amd import scene as S
    from cudafusion_amd.harness import Harness
    from cudafusion_amd.train_lp import pcg_solve
    from oracle import pyoracle as po
    here = os.path.dirname(LP.__file__)
    paths = {v: os.path.join(here, f"lp_weights_v{v}.npz") for v in (0, 1)}
    if not all(os.path.exists(p) for p in paths.values()):
        pytest.skip("trained weights not present (python -m cudafusion_amd.train_lp)")
    w = S.body_instance_workload(2400, 256, 0.004, 640, frames=2, seed=1000)
    h = Harness(workload=w, frames=2, arap_consistent_sign=True, validate=False, pcg_pipelined=True)
    h.setup()
    lib, N = h.lib, h.N
    solver = lib.nrf_pipeline_solver(h.pipe)
    d_nodes = lib.nrf_pipeline_nodes(h.pipe)
    rows = []
    for f in (1, 2):
        h.step(f)
        torch.cuda.synchronize()
        rp, ci, A, b, Minv = LP.solver_system(lib, solver, N)
        x = torch.zeros(6 * N, dtype=torch.float32, device="cuda")
        it_bj, rel_bj, _ = pcg_solve(lib, solver, h.sp, x, 1e-6, 1000)
        out = {"N": N, "frame": f, "block_jacobi": it_bj, "rel_block_jacobi": rel_bj}
        for v in (1, 0):  # the block-diagonal fold changes the solver's inverses: last
            lp = LP.LearnedPreconditioner(v, max_nodes=N, max_probes=1)
            lp.load(paths[v])
            lp.set_graph(h.nodes)
            lp.forward(solver, d_nodes)
            if v == 1:
                Cm, Td, Te, _ = lp.blocks()
                z = np.zeros((N, 21)), np.zeros((N, 8, 36))
                it0, _, _ = po.lp_pcg_f64(1, h.nodes, rp, ci, A, b, Cm, Minv, *z, 1000, 1e-6)
                it1, r1, _ = po.lp_pcg_f64(1, h.nodes, rp, ci, A, b, Cm, Minv, Td, Te, 1000, 1e-6)
                out.update(fp64_block_jacobi=it0, fp64_graph_pattern=it1, rel_graph_pattern=r1)
            else:
                lp.fold(solver)
                out["fallbacks"] = lp.fallbacks()
                x.zero_()
                it_l, rel_l, _ = pcg_solve(lib, solver, h.sp, x, 1e-6, 1000)
                out.update(block_diagonal=it_l, rel_block_diagonal=rel_l)
            lp.close()
        print(f"[lp held-out] {out}", flush=True)
        rows.append(out)
    h.close()
    return rows


def test_held_out_iterations_both_variants(held_out):
    """Both learned variants converge to 1e-6 on the held-out systems, with
    no node falling back to block-Jacobi; the iteration counts are printed
    (the gain test below records that they do not halve)."""
    for out in held_out:
        assert out["rel_graph_pattern"] < 1e-6 and out["rel_block_diagonal"] < 2e-6 and out["rel_block_jacobi"] < 2e-6
        assert out["fallbacks"] == 0


@pytest.mark.xfail(strict=True, reason="target not met: the trained one-hop corrections BJ^-1 + L L^T raise the "
                                       "PCG iterations against block-Jacobi on held-out systems instead of halving "
                                       "them (the additive PSD form cannot lower the spectrum's large end)")
def test_learned_preconditioner_halves_iterations(held_out):
    """The learned preconditioner's target: at least 2x fewer PCG iterations to 1e-6 than
    block-Jacobi, for the better of the two variants, on every held-out
    system.  Kept as a strict expected failure: it documents the shortfall
    and fails if the target is reached without this record being updated."""
    for out in held_out:
        best = min(out["block_diagonal"], out["fp64_graph_pattern"])
        assert 2 * best <= out["block_jacobi"]


def test_fold_falls_back_to_block_jacobi_on_non_finite(gpu):
    """Network parameters with a NaN in the diagonal head's bias give a
    non-finite L_ii at every node: the fold keeps every block-Jacobi inverse
    unchanged and counts N fallbacks (a host-forced condition: no kernel
    faults or waits)."""
    import torch

    from cudafusion_amd import learned as LP
    from oracle import pyoracle as po
    h = _system()
    lib, N = h.lib, h.N
    solver = lib.nrf_pipeline_solver(h.pipe)
    d_nodes = lib.nrf_pipeline_nodes(h.pipe)
    lp = LP.LearnedPreconditioner(0, max_nodes=N, max_probes=1, seed=0)
    lp.set_graph(h.nodes)
    params = lp.params()
    o, n = po.lp_offsets()["bo"]
    params[o] = np.nan
    lp.set_params(params)
    M0 = LP.solver_system(lib, solver, N)[4].copy()
    lp.forward(solver, d_nodes)
    lp.fold(solver)
    torch.cuda.synchronize()
    M1 = LP.solver_system(lib, solver, N)[4]
    assert lp.fallbacks() == N
    np.testing.assert_array_equal(M1, M0)
    lp.close()
    h.close()
