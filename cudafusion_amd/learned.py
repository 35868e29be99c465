"""Learned preconditioner (build-only component): the nrf_lp_* entry points
wrapped for training and evaluation.

The network and the preconditioner M^-1 = BJ^-1 + L L^T are defined in
oracle/learned.cpp and run on the device by csrc/learned.hip (BF16 weights,
FP32 accumulation).  Training data are the JtJ systems the GN solver
assembles on capsule-body sequences; each training step draws random probe
vectors v, and minimises sum_p |M^-1 A v_p - v_p|^2 / (P 6N) with Adam.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L
from .fusion import copy_device_to_numpy, copy_numpy_to_device

# flat parameter layout: name -> (offset, size); fan-in of each weight matrix
_LAYOUT = [("W0", 28 * 64, 28), ("b0", 64, 0)]
for _l in range(3):
    _LAYOUT += [(f"Wm{_l}", 68 * 64, 68), (f"bm{_l}", 64, 0), (f"Wu{_l}", 128 * 64, 128), (f"bu{_l}", 64, 0)]
_LAYOUT += [("Wo", 64 * 21, 64), ("bo", 21, 0), ("We", 132 * 36, 132), ("be", 36, 0)]


def layout() -> dict:
    out, o = {}, 0
    for name, n, fan in _LAYOUT:
        out[name] = (o, n, fan)
        o += n
    return out


def init_params(seed: int = 0, head_scale: float = 1e-2) -> np.ndarray:
    """He-style initialisation; the heads start small, so L is near zero and
    the preconditioner starts as block-Jacobi."""
    rng = np.random.default_rng(seed)
    lay = layout()
    p = np.zeros(sum(n for _, n, _ in _LAYOUT), np.float32)
    for name, (o, n, fan) in lay.items():
        if fan == 0:
            continue
        s = np.sqrt(2.0 / fan) * (head_scale if name in ("Wo", "We") else 1.0)
        p[o:o + n] = rng.normal(0.0, s, n)
    return p


class LearnedPreconditioner:
    def __init__(self, variant: int = 0, max_nodes: int = 8192, max_probes: int = 8, seed: int = 0):
        self.lib = L.load()
        self.variant = variant
        self.max_probes = max_probes
        self.lp = self.lib.nrf_lp_create(variant, max_nodes, max_probes)
        if not self.lp:
            raise L.NrfError("nrf_lp_create failed: " + self.lib.nrf_last_error().decode())
        self.n = int(self.lib.nrf_lp_param_count())
        assert self.n == len(init_params(0))
        self.set_params(init_params(seed))
        self.N = 0
        self.step_count = 0

    # -------------------------------------------------------- parameters
    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        assert p.shape == (self.n,)
        copy_numpy_to_device(self.lib.nrf_lp_params(self.lp), p)
        L.check(self.lib.nrf_lp_reset_optimizer(self.lp), "lp_reset_optimizer")
        self.step_count = 0

    def params(self) -> np.ndarray:
        return copy_device_to_numpy(self.lib.nrf_lp_params(self.lp), self.n, np.float32)

    def grads(self) -> np.ndarray:
        return copy_device_to_numpy(self.lib.nrf_lp_grads(self.lp), self.n, np.float32)

    def save(self, path: str) -> None:
        np.savez(path, params=self.params(), variant=np.int32(self.variant))

    def load(self, path: str) -> None:
        with np.load(path, allow_pickle=False) as z:
            if int(z["variant"]) != self.variant:
                raise ValueError(f"{path}: weights of variant {int(z['variant'])}, not {self.variant}")
            self.set_params(z["params"])

    # ------------------------------------------------------------ graph
    def set_graph(self, host_nodes: np.ndarray) -> None:
        L.check(self.lib.nrf_lp_set_graph(self.lp, host_nodes.ctypes.data, len(host_nodes)), "lp_set_graph")
        self.N = len(host_nodes)

    # ------------------------------------------------------ device work
    def forward(self, solver, d_nodes: int) -> None:
        L.check(self.lib.nrf_lp_forward(self.lp, solver, d_nodes, self.N, torch.cuda.current_stream().cuda_stream),
                "lp_forward")

    def loss_grad(self, solver, d_nodes: int, probes: torch.Tensor, want_loss: bool = True) -> float | None:
        assert probes.dtype == torch.float32 and probes.is_cuda and probes.shape[1] == 6 * self.N
        out = C.c_double()
        L.check(self.lib.nrf_lp_loss_grad(self.lp, solver, d_nodes, probes.data_ptr(), probes.shape[0],
                                          C.byref(out) if want_loss else None,
                                          torch.cuda.current_stream().cuda_stream), "lp_loss_grad")
        return out.value if want_loss else None

    def adam(self, lr: float) -> None:
        self.step_count += 1
        L.check(self.lib.nrf_lp_adam(self.lp, lr, self.step_count, torch.cuda.current_stream().cuda_stream),
                "lp_adam")

    def fold(self, solver) -> None:
        L.check(self.lib.nrf_lp_fold(self.lp, solver, torch.cuda.current_stream().cuda_stream), "lp_fold")

    def fallbacks(self) -> int:
        """Nodes that kept their block-Jacobi inverse in the last fold (the
        correction or the folded block was not finite)."""
        n = C.c_int()
        L.check(self.lib.nrf_lp_fallbacks(self.lp, C.byref(n), torch.cuda.current_stream().cuda_stream),
                "lp_fallbacks")
        return n.value

    def blocks(self):
        """(C, Td, Te, features) of the last forward, as numpy arrays."""
        ptrs = [C.c_void_p() for _ in range(4)]
        L.check(self.lib.nrf_lp_blocks(self.lp, *[C.byref(p) for p in ptrs]), "lp_blocks")
        N = self.N
        Cm = copy_device_to_numpy(ptrs[0].value, N * 36, np.float32).reshape(N, 36)
        Td = copy_device_to_numpy(ptrs[1].value, N * 21, np.float32).reshape(N, 21)
        Te = copy_device_to_numpy(ptrs[2].value, N * 8 * 36, np.float32).reshape(N, 8, 36)
        x = copy_device_to_numpy(ptrs[3].value, N * 28, np.float32).reshape(N, 28)
        return Cm, Td, Te, x

    def close(self) -> None:
        if self.lp:
            self.lib.nrf_lp_destroy(self.lp)
            self.lp = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def bf16_round(p: np.ndarray) -> np.ndarray:
    """The BF16 values the device kernels use for FP32 parameters (round to
    nearest even), as float64."""
    u = np.ascontiguousarray(p, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def solver_system(lib, solver, N: int):
    """(row_ptr, col_idx, values[nb, 6, 6], rhs, Minv[N, 6, 6]) of a solver, on the host."""
    rp, ci, vals, b, M = (C.c_void_p() for _ in range(5))
    nb = C.c_int()
    L.check(lib.nrf_solver_system(solver, C.byref(rp), C.byref(ci), C.byref(vals), C.byref(nb), C.byref(b),
                                  C.byref(M)), "solver_system")
    nbv = nb.value
    return (copy_device_to_numpy(rp.value, N + 1, np.int32), copy_device_to_numpy(ci.value, nbv, np.int32),
            copy_device_to_numpy(vals.value, nbv * 36, np.float32).reshape(nbv, 6, 6),
            copy_device_to_numpy(b.value, 6 * N, np.float32),
            copy_device_to_numpy(M.value, N * 36, np.float32).reshape(N, 6, 6))
