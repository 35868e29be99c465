"""Training and held-out evaluation of the learned preconditioner.

    python -m cudafusion_amd.train_lp --variants 0,1 --out-dir cudafusion_amd

Data: capsule-body sequences (the full-scale configuration: 256^3 at 4 mm,
640x480, 5 GN iterations) with node counts log-uniform in [1000, 6000]
(node spacing scaled, seeds 0.. for training, 1000.. held out).  After
every frame the solver holds the JtJ system of the frame's last GN
iteration; its solution x* is computed once with the solver's PCG (block-
Jacobi, relative residual 1e-7).  Each training step takes one stored system
and minimises sum_p |M^-1 A v_p - v_p|^2 / (P 6N) over the probes
v = x* / rms(x*) (so the first probe's term is |M^-1 b - x*|^2 up to scale:
the preconditioner is trained to approximate A^-1 on the right-hand sides
the solver meets) and two Gaussian probes weighted 0.1 (which keep M^-1
reasonable in every direction).

Held-out evaluation (block-diagonal variant): per system, PCG iterations
and device ms to a 1e-6 relative residual with block-Jacobi and with the
learned preconditioner folded into the solver (the network's forward time
is reported separately).  The graph-pattern variant is trained here; its
PCG iterations are measured by tests/test_gpu_learned.py.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np
import torch

from . import _lib as L
from . import learned as LP
from . import scene as S
from .harness import Harness


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


class SystemStore:
    """Systems of one sequence, kept on the device: BSR values, RHS, block-
    Jacobi inverses and the solution, restorable into the sequence's solver."""

    def __init__(self, h: Harness):
        self.h = h
        self.lib = h.lib
        self.solver = h.lib.nrf_pipeline_solver(h.pipe)
        self.d_nodes = h.lib.nrf_pipeline_nodes(h.pipe)
        self.N = h.N
        ptrs = [C.c_void_p() for _ in range(5)]
        nb = C.c_int()
        L.check(self.lib.nrf_solver_system(self.solver, C.byref(ptrs[0]), C.byref(ptrs[1]), C.byref(ptrs[2]),
                                           C.byref(nb), C.byref(ptrs[3]), C.byref(ptrs[4])), "solver_system")
        self.nb = nb.value
        self.p_values, self.p_b, self.p_minv = ptrs[2].value, ptrs[3].value, ptrs[4].value
        self.items = []  # (values, b, Minv, x*) tensors

    def capture(self, solve_tol: float = 1e-7) -> int:
        """Stores the solver's current system and its solution."""
        N, nb = self.N, self.nb
        vals = torch.empty(nb * 36, dtype=torch.float32, device="cuda")
        b = torch.empty(6 * N, dtype=torch.float32, device="cuda")
        minv = torch.empty(N * 36, dtype=torch.float32, device="cuda")
        _d2d(vals, self.p_values)
        _d2d(b, self.p_b)
        _d2d(minv, self.p_minv)
        x = torch.zeros(6 * N, dtype=torch.float32, device="cuda")
        it, rel, _ = pcg_solve(self.lib, self.solver, self.h.sp, x, solve_tol, 2000)
        self.items.append((vals, b, minv, x))
        return it

    def restore(self, k: int) -> None:
        vals, b, minv, _ = self.items[k]
        _h2d_dev(self.p_values, vals)
        _h2d_dev(self.p_b, b)
        _h2d_dev(self.p_minv, minv)


_hip = None


def _hiplib():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so.7")
        _hip.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
        _hip.hipMemcpyAsync.restype = C.c_int
    return _hip


def _d2d(dst: torch.Tensor, src_ptr: int) -> None:
    rc = _hiplib().hipMemcpyAsync(C.c_void_p(dst.data_ptr()), C.c_void_p(src_ptr), dst.numel() * dst.element_size(),
                                  3, C.c_void_p(torch.cuda.current_stream().cuda_stream))
    if rc != 0:
        raise L.NrfError(f"hipMemcpyAsync D2D failed ({rc})")


def _h2d_dev(dst_ptr: int, src: torch.Tensor) -> None:
    rc = _hiplib().hipMemcpyAsync(C.c_void_p(dst_ptr), C.c_void_p(src.data_ptr()), src.numel() * src.element_size(),
                                  3, C.c_void_p(torch.cuda.current_stream().cuda_stream))
    if rc != 0:
        raise L.NrfError(f"hipMemcpyAsync D2D failed ({rc})")


def pcg_solve(lib, solver, sp0, x: torch.Tensor, tol: float, maxit: int):
    """The solver's PCG (its current preconditioner) on its current system to
    a relative residual; returns (iterations, true relative residual, ms)."""
    sp = L.SolverParams.from_buffer_copy(bytes(sp0))
    sp.pcg_iterations = maxit
    sp.pcg_tolerance = tol
    sp.relative_tolerance = 1
    L.check(lib.nrf_solver_set_params(solver, C.byref(sp)), "set_params")
    st = torch.cuda.current_stream()
    res = L.PcgResult()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    L.check(lib.nrf_solver_resolve(solver, x.data_ptr(), C.byref(res), st.cuda_stream), "resolve")
    e1.record()
    out = (C.c_double * 2)()
    L.check(lib.nrf_solver_true_residual(solver, x.data_ptr(), out, st.cuda_stream), "true_residual")
    L.check(lib.nrf_solver_set_params(solver, C.byref(sp0)), "set_params")
    torch.cuda.synchronize()
    return int(res.iterations), float(out[0] / max(out[1], 1e-300)), e0.elapsed_time(e1)


def make_sequences(seeds, frames: int, lo: int = 1000, hi: int = 6000):
    out = []
    rng = np.random.default_rng(12345 + seeds[0])
    for s in seeds:
        n = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        w = S.body_instance_workload(n, 256, 0.004, 640, frames=frames, seed=s)
        h = Harness(workload=w, frames=frames, arap_consistent_sign=True, validate=False, pcg_pipelined=True)
        h.setup()
        out.append(h)
        log(f"[data] sequence seed {s}: {h.N} nodes")
    return out


def collect(hs, frames: int, tol: float = 1e-7):
    stores = []
    for h in hs:
        st = SystemStore(h)
        for f in range(1, frames + 1):
            h.step(f)
            torch.cuda.synchronize()
            st.capture(tol)
        stores.append(st)
    return stores


def train(lp: LP.LearnedPreconditioner, stores, epochs: int, lr: float, seed: int = 0, rand_w: float = 0.1):
    gen = torch.Generator(device="cuda").manual_seed(seed)
    order = [(si, k) for si, st in enumerate(stores) for k in range(len(st.items))]
    rng = np.random.default_rng(seed)
    hist = []
    cur_graph = -1
    for ep in range(epochs):
        rng.shuffle(order)
        tot = 0.0
        for si, k in order:
            st = stores[si]
            if si != cur_graph:
                lp.set_graph(st.h.nodes)
                cur_graph = si
            st.restore(k)
            x = st.items[k][3]
            N = st.N
            v0 = x / (x.square().mean().sqrt() + 1e-30)
            V = torch.cat([v0[None], rand_w * torch.randn(2, 6 * N, device="cuda", generator=gen)], 0)
            lp.forward(st.solver, st.d_nodes)
            tot += lp.loss_grad(st.solver, st.d_nodes, V.contiguous())
            lp.adam(lr)
        hist.append(tot / len(order))
        log(f"[train] epoch {ep}: mean loss {hist[-1]:.5f}")
    return hist


def evaluate(lp: LP.LearnedPreconditioner, stores, tol: float = 1e-6, maxit: int = 1000):
    """Per held-out system: block-Jacobi vs learned (folded) PCG on the device
    (block-diagonal variant)."""
    rows = []
    for st in stores:
        lp.set_graph(st.h.nodes)
        for k in range(len(st.items)):
            st.restore(k)
            x = torch.zeros(6 * st.N, dtype=torch.float32, device="cuda")
            it_bj, rel_bj, ms_bj = pcg_solve(st.lib, st.solver, st.h.sp, x, tol, maxit)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            lp.forward(st.solver, st.d_nodes)
            lp.fold(st.solver)
            e1.record()
            torch.cuda.synchronize()
            row = {"N": st.N, "bj_iterations": it_bj, "bj_ms": round(ms_bj, 4), "bj_true_rel": rel_bj,
                   "network_ms": round(e0.elapsed_time(e1), 4)}
            x.zero_()
            it_l, rel_l, ms_l = pcg_solve(st.lib, st.solver, st.h.sp, x, tol, maxit)
            row.update(learned_iterations=it_l, learned_ms=round(ms_l, 4), learned_true_rel=rel_l)
            rows.append(row)
            log(f"[eval] {row}")
    return rows


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--variants", default="0,1", help="comma-separated: 0 block diagonal, 1 graph pattern")
    ap.add_argument("--train-seqs", type=int, default=48)
    ap.add_argument("--eval-seqs", type=int, default=20)
    ap.add_argument("--frames", type=int, default=8)
    ap.add_argument("--eval-frames", type=int, default=4)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out-dir", default="", help="writes lp_weights_v<variant>.npz there")
    ap.add_argument("--json", default="")
    args = ap.parse_args(argv)
    variants = [int(v) for v in args.variants.split(",") if v != ""]
    t0 = time.perf_counter()
    tr = make_sequences(list(range(args.train_seqs)), args.frames)
    st_tr = collect(tr, args.frames)
    log(f"[data] {len(tr)} training sequences, nodes {min(h.N for h in tr)}-{max(h.N for h in tr)}, "
        f"{sum(len(s.items) for s in st_tr)} systems ({time.perf_counter() - t0:.0f} s)")
    ev = make_sequences(list(range(1000, 1000 + args.eval_seqs)), args.eval_frames)
    st_ev = collect(ev, args.eval_frames)
    log(f"[data] {len(ev)} held-out sequences, {sum(len(s.items) for s in st_ev)} systems "
        f"({time.perf_counter() - t0:.0f} s)")
    summary = {"train_systems": sum(len(s.items) for s in st_tr), "eval_systems": sum(len(s.items) for s in st_ev),
               "train_nodes": [h.N for h in tr], "eval_nodes": [h.N for h in ev], "variants": {}}
    for v in variants:
        lp = LP.LearnedPreconditioner(v, max_nodes=8192, max_probes=3, seed=0)
        tt = time.perf_counter()
        hist = train(lp, st_tr, args.epochs, args.lr)
        train_s = time.perf_counter() - tt
        out = {"loss_per_epoch": [round(x, 6) for x in hist], "train_s": round(train_s, 1)}
        if v == 0:
            rows = evaluate(lp, st_ev)
            out["rows"] = rows
            out["mean_iterations"] = {"block_jacobi": float(np.mean([r["bj_iterations"] for r in rows])),
                                      "learned": float(np.mean([r["learned_iterations"] for r in rows]))}
            out["mean_ms"] = {"block_jacobi": float(np.mean([r["bj_ms"] for r in rows])),
                              "learned_pcg": float(np.mean([r["learned_ms"] for r in rows])),
                              "network": float(np.mean([r["network_ms"] for r in rows]))}
        if args.out_dir:
            lp.save(os.path.join(args.out_dir, f"lp_weights_v{v}.npz"))
        log(f"[summary] variant {v}: " + json.dumps({k: x for k, x in out.items() if k != "rows"}))
        summary["variants"][v] = out
        lp.close()
    if args.json:
        with open(args.json, "w") as f:
            json.dump(summary, f, indent=1)
    for h in tr + ev:
        h.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
