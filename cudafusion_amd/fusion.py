"""Python mirror of the reference's component API over the C ABI.

``TSDFVolume``, ``WarpField`` and ``GaussNewtonSolver`` keep the names,
nested ``Params``, defaults and method meanings of include/tsdf_volume.h,
include/warp_field.h and include/solver.h.  Device memory is held in torch
tensors on the current HIP device; every computation is a call into
``libnrf.so`` (there is no Python or CPU fallback).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib as L


def _dev() -> torch.device:
    if not torch.cuda.is_available():
        raise L.NrfError("no HIP device: the nrf kernels need an MI355X (gfx950)")
    return torch.device("cuda", torch.cuda.current_device())


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _ptr(t) -> int | None:
    if t is None:
        return None
    if isinstance(t, int):
        return t
    return t.data_ptr()


def to_device_records(arr: np.ndarray) -> torch.Tensor:
    """Copy a numpy record array to a device byte tensor (same bytes)."""
    raw = np.ascontiguousarray(arr).view(np.uint8).reshape(-1)
    return torch.from_numpy(raw.copy()).to(_dev())


def from_device_records(t: torch.Tensor, dtype: np.dtype) -> np.ndarray:
    return t.detach().cpu().numpy().view(np.uint8).reshape(-1).view(dtype)


def pose_array(pose) -> np.ndarray:
    p = np.ascontiguousarray(np.asarray(pose, dtype=np.float32).reshape(16))
    return p


class TSDFVolume:
    """include/tsdf_volume.h:10-81."""

    @dataclass
    class Params:
        dims: tuple = (256, 256, 256)
        voxel_size: float = 0.006
        truncation: float = 0.03
        origin: tuple = (-0.768, -0.768, 0.5)
        decay_alpha: float = 0.90
        max_weight: float = 50.0

    def __init__(self, p: "TSDFVolume.Params"):
        self._p = p
        self.vp = L.volume_params(p.dims, p.voxel_size, p.truncation, p.origin, p.decay_alpha, p.max_weight)
        self.lib = L.load()
        n = self.total_voxels()
        self._vox = torch.empty(n * 12, dtype=torch.uint8, device=_dev())
        L.check(self.lib.nrf_tsdf_reset(C.byref(self.vp), _ptr(self._vox), _stream()), "nrf_tsdf_reset")
        self._mc_ws = None
        self._centered = None

    def params(self):
        return self._p

    def total_voxels(self) -> int:
        d = self._p.dims
        return int(d[0]) * int(d[1]) * int(d[2])

    def device_data(self) -> torch.Tensor:
        return self._vox

    def idx_to_xyz(self, idx: int):
        X, Y, _ = self._p.dims
        return (idx % X, (idx // X) % Y, idx // (X * Y))

    def voxel_to_world(self, xyz):
        # no +0.5 here, as in tsdf_volume.h:162-166
        o, s = self._p.origin, self._p.voxel_size
        return tuple(np.float32(o[k]) + np.float32(xyz[k]) * np.float32(s) for k in range(3))

    def integrate(self, depth, normals, cam, camera_pose, nodes=None, transforms=None, num_nodes=0,
                  voxel_knn=None, voxel_knn_w=None, voxel_opt_counts=None, min_opt_count=0,
                  centered=None, updated_count=None):
        pose = pose_array(camera_pose)
        L.check(self.lib.nrf_tsdf_integrate(
            C.byref(self.vp), _ptr(self._vox), _ptr(depth), C.byref(cam), pose.ctypes.data,
            _ptr(nodes), _ptr(transforms), int(num_nodes), _ptr(centered), _ptr(voxel_knn), _ptr(voxel_knn_w),
            _ptr(voxel_opt_counts), int(min_opt_count), _ptr(updated_count), _stream()), "nrf_tsdf_integrate")

    def raycast(self, cam, camera_pose, nodes=None, transforms=None, num_nodes=0, voxel_knn=None, voxel_knn_w=None,
                out_canonical_vidx: bool = False):
        """TSDFVolume::raycast (include/tsdf_volume.h:38-44): device tensors of
        live vertices and normals (W*H x 3) and, when asked, the canonical
        voxel index per pixel."""
        npx = cam.width * cam.height
        v = torch.empty(npx * 3, dtype=torch.float32, device=_dev())
        n = torch.empty(npx * 3, dtype=torch.float32, device=_dev())
        idx = torch.empty(npx, dtype=torch.int32, device=_dev()) if out_canonical_vidx else None
        pose = pose_array(camera_pose)
        L.check(self.lib.nrf_raycast(C.byref(self.vp), _ptr(self._vox), C.byref(cam), pose.ctypes.data, _ptr(nodes),
                                     _ptr(transforms), int(num_nodes), _ptr(voxel_knn), _ptr(voxel_knn_w), _ptr(v),
                                     _ptr(n), _ptr(idx), _stream()), "nrf_raycast")
        return (v.view(-1, 3), n.view(-1, 3)) + ((idx,) if out_canonical_vidx else ())

    def _mc(self, normals: bool, max_triangles: int | None):
        if self._mc_ws is None:
            size = self.lib.nrf_mc_workspace_size(C.byref(self.vp))
            self._mc_ws = torch.empty(size, dtype=torch.uint8, device=_dev())
        counts = torch.zeros(4, dtype=torch.int32, device=_dev())
        # count pass first (no output) to size the buffers, like the reference
        L.check(self.lib.nrf_mc_extract(C.byref(self.vp), _ptr(self._vox), _ptr(self._mc_ws), self._mc_ws.numel(),
                                        None, None, None, 0, _ptr(counts), _stream()), "nrf_mc_extract")
        T = int(counts[1].item()) if max_triangles is None else int(max_triangles)
        verts = torch.empty(max(T, 1) * 9, dtype=torch.float32, device=_dev())
        nrm = torch.empty(max(T, 1) * 9, dtype=torch.float32, device=_dev()) if normals else None
        tris = torch.empty(max(T, 1) * 3, dtype=torch.int32, device=_dev())
        L.check(self.lib.nrf_mc_extract(C.byref(self.vp), _ptr(self._vox), _ptr(self._mc_ws), self._mc_ws.numel(),
                                        _ptr(verts), _ptr(nrm), _ptr(tris), T, _ptr(counts), _stream()),
                "nrf_mc_extract")
        n = min(T, int(counts[1].item()))
        return verts[: n * 9].view(-1, 3), (nrm[: n * 9].view(-1, 3) if nrm is not None else None), \
            tris[: n * 3].view(-1, 3)

    def extract_surface(self):
        """Host copies of (vertices[3T,3], normals[3T,3], triangles[T,3])."""
        v, n, t = self._mc(True, None)
        return v.cpu().numpy(), n.cpu().numpy(), t.cpu().numpy()

    def extract_surface_device(self):
        v, _, t = self._mc(False, None)
        return v, t

    def mc_cases(self) -> np.ndarray:
        if self._mc_ws is None:
            raise L.NrfError("run extract_surface first")
        return self._mc_ws[: self.total_voxels()].cpu().numpy()


class WarpField:
    """include/warp_field.h:12-126 (node sampling / pruning are not part of this library)."""

    MAX_NODES = 8192

    def __init__(self, node_radius: float, max_nodes: int = MAX_NODES):
        self.lib = L.load()
        self.node_radius = node_radius
        self.max_nodes = max_nodes
        self._n = 0
        self._nodes = torch.zeros(max_nodes * 96, dtype=torch.uint8, device=_dev())
        self._T = torch.zeros(max_nodes * 8, dtype=torch.float32, device=_dev())
        self._Tprev = torch.zeros(max_nodes * 8, dtype=torch.float32, device=_dev())
        self._h_nodes = np.zeros(0, L.NODE)

    def num_nodes(self) -> int:
        return self._n

    def device_nodes(self) -> torch.Tensor:
        return self._nodes

    def device_transforms(self) -> torch.Tensor:
        return self._T

    def device_transforms_prev(self) -> torch.Tensor:
        return self._Tprev

    def upload_nodes(self, nodes: np.ndarray) -> None:
        n = min(len(nodes), self.max_nodes)
        self._h_nodes = np.array(nodes[:n], dtype=L.NODE)
        self._nodes[: n * 96].copy_(to_device_records(self._h_nodes))
        if n != self._n:
            self._n = n
            self.reset_transforms()
            self._Tprev[: n * 8].copy_(self._T[: n * 8])

    def upload_transforms(self, transforms: np.ndarray) -> None:
        n = min(len(transforms), self.max_nodes)
        if self._n == 0:
            self._n = n
        n = min(n, self._n)
        t = torch.from_numpy(np.ascontiguousarray(transforms[:n]).view(np.float32).reshape(-1).copy()).to(_dev())
        self._T[: n * 8].copy_(t)
        self._Tprev[: n * 8].copy_(t)

    def download_nodes(self) -> np.ndarray:
        return self._h_nodes.copy()

    def download_transforms(self) -> np.ndarray:
        return self._T[: self._n * 8].cpu().numpy().view(L.DUALQUAT).copy()

    def reset_transforms(self) -> None:
        t = torch.zeros(self._n, 8, dtype=torch.float32)
        t[:, 3] = 1.0
        self._T[: self._n * 8].copy_(t.reshape(-1).to(_dev()))

    def save_transforms(self) -> None:
        self._Tprev[: self._n * 8].copy_(self._T[: self._n * 8])

    def restore_transforms(self) -> None:
        self._T[: self._n * 8].copy_(self._Tprev[: self._n * 8])

    def compute_voxel_knn(self, volume: TSDFVolume):
        if self._n == 0:
            return None, None
        V = volume.total_voxels()
        knn = torch.empty(V * 4, dtype=torch.int32, device=_dev())
        knn_w = torch.empty(V * 4, dtype=torch.float32, device=_dev())
        L.check(self.lib.nrf_voxel_knn(C.byref(volume.vp), _ptr(volume.device_data()), _ptr(self._nodes), self._n,
                                       _ptr(knn), _ptr(knn_w), _stream()), "nrf_voxel_knn")
        return knn, knn_w

    def apply_twist_increment(self, delta_x: torch.Tensor, max_rot: float, max_trans: float,
                              update_scale: float = 1.0) -> None:
        if self._n == 0:
            return
        L.check(self.lib.nrf_warp_apply_twist(_ptr(delta_x), _ptr(self._T), self._n, max_rot, max_trans,
                                              update_scale, None, None, _stream()), "nrf_warp_apply_twist")

    def centered_table(self) -> torch.Tensor:
        out = torch.empty(max(1, self._n) * 8, dtype=torch.float32, device=_dev())
        L.check(self.lib.nrf_warp_centered_table(_ptr(self._nodes), _ptr(self._T), self._n, _ptr(out), _stream()),
                "nrf_warp_centered_table")
        return out


class GaussNewtonSolver:
    """include/solver.h:10-89."""

    @dataclass
    class Params:
        gn_iterations: int = 3
        pcg_iterations: int = 10
        pcg_tolerance: float = 1e-4
        lambda_smooth: float = 1.0
        lambda_rot: float = 1.0
        lambda_damping: float = 1e-2
        debug: bool = False
        relative_tolerance: bool = False
        precond_fp64: bool = False
        arap_consistent_sign: bool = False  # opt-in: ARAP rotation rows with the data term's sign
        pcg_pipelined: int = 0  # opt-in: 1 A p by recurrence, product overlapping an all-reduce; 2 one all-reduce;
        # 3 pipelined: the one all-reduce overlapping the halo exchange and the product
        jtj_precision: int = 0  # 0 FP32, 1 BF16, 2 FP16, 3 / 4 split BF16 / FP16 JtJ blocks (RHS stays FP32)
        refine_iterations: int = 0  # > 0: one FP64-residual iterative-refinement step per solve
        refine_jacobi: bool = False  # one refinement step with the block-Jacobi inverse as its correction solve
        true_residual_rounds: int = 0  # > 0: convergence by the FP64 true residual, correction PCG rounds
        psd_dropped_pairs: bool = False  # opt-in: data-term pairs outside the pattern kept PSD on the diagonal

        def to_c(self) -> L.SolverParams:
            return L.solver_params(self.gn_iterations, self.pcg_iterations, self.pcg_tolerance, self.lambda_smooth,
                                   self.lambda_rot, self.lambda_damping, int(self.debug),
                                   int(self.relative_tolerance), int(self.precond_fp64),
                                   int(self.arap_consistent_sign), int(self.pcg_pipelined),
                                   int(self.jtj_precision), int(self.refine_iterations), int(self.refine_jacobi),
                                   int(self.true_residual_rounds), int(self.psd_dropped_pairs))

    def __init__(self, p: "GaussNewtonSolver.Params", max_nodes: int):
        self.lib = L.load()
        _dev()
        self.params = p
        self._cp = p.to_c()
        self.handle = self.lib.nrf_solver_create(C.byref(self._cp), int(max_nodes))
        if not self.handle:
            raise L.NrfError("nrf_solver_create failed: " + self.lib.nrf_last_error().decode())
        self.last_result = L.PcgResult()

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            self.lib.nrf_solver_destroy(h)
            self.handle = None

    def true_residual(self, x) -> tuple:
        """(||b - A x||, ||b||) of the last assembled system, in FP64."""
        out = (C.c_double * 2)()
        L.check(self.lib.nrf_solver_true_residual(self.handle, _ptr(x), out, None), "nrf_solver_true_residual")
        return out[0], out[1]

    def set_params(self, p: "GaussNewtonSolver.Params") -> None:
        self.params = p
        self._cp = p.to_c()
        L.check(self.lib.nrf_solver_set_params(self.handle, C.byref(self._cp)), "nrf_solver_set_params")

    def set_pattern(self, host_nodes: np.ndarray) -> None:
        arr = np.ascontiguousarray(host_nodes, dtype=L.NODE)
        L.check(self.lib.nrf_solver_set_pattern(self.handle, arr.ctypes.data, len(arr)), "nrf_solver_set_pattern")

    def solve(self, corrs, num_corrs, d_nodes, d_transforms, num_nodes, delta_x) -> L.PcgResult:
        res = L.PcgResult()
        L.check(self.lib.nrf_solver_solve(self.handle, _ptr(corrs), int(num_corrs), _ptr(d_nodes),
                                          _ptr(d_transforms), int(num_nodes), _ptr(delta_x), C.byref(res),
                                          _stream()), "nrf_solver_solve")
        self.last_result = res
        return res

    def solve_gn(self, corrs, canonical, num_corrs, d_nodes, d_transforms, centered, num_nodes, gn: L.GNParams,
                 delta_x, sync: bool = True) -> L.PcgResult:
        res = L.PcgResult()
        L.check(self.lib.nrf_solver_solve_gn(self.handle, _ptr(corrs), _ptr(canonical), int(num_corrs),
                                             _ptr(d_nodes), _ptr(d_transforms), _ptr(centered), int(num_nodes),
                                             C.byref(gn), _ptr(delta_x), C.byref(res) if sync else None,
                                             _stream()), "nrf_solver_solve_gn")
        self.last_result = res
        return res

    def system_matrix(self):
        """(row_ptr, col_idx, values[nb,6,6]) as host numpy arrays (syncs)."""
        rp, ci, va, rhs, pc = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
        nb = C.c_int()
        L.check(self.lib.nrf_solver_system(self.handle, C.byref(rp), C.byref(ci), C.byref(va), C.byref(nb),
                                           C.byref(rhs), C.byref(pc)), "nrf_solver_system")
        torch.cuda.synchronize()
        return rp.value, ci.value, va.value, nb.value, rhs.value, pc.value


def copy_device_to_numpy(ptr: int, count: int, dtype) -> np.ndarray:
    """Synchronous D2H copy of `count` elements of `dtype` from a raw device address."""
    dt = np.dtype(dtype)
    out = np.empty(count, dtype=dt)
    if count == 0:
        return out
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipMemcpy.restype = C.c_int
    torch.cuda.synchronize()
    rc = hip.hipMemcpy(out.ctypes.data, C.c_void_p(ptr), out.nbytes, 2)  # hipMemcpyDeviceToHost
    if rc != 0:
        raise L.NrfError(f"hipMemcpy D2H failed ({rc})")
    return out


def copy_numpy_to_device(ptr: int, arr: np.ndarray) -> None:
    """Synchronous H2D copy of a contiguous array to a raw device address."""
    arr = np.ascontiguousarray(arr)
    if arr.nbytes == 0:
        return
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipMemcpy.restype = C.c_int
    torch.cuda.synchronize()
    rc = hip.hipMemcpy(C.c_void_p(ptr), arr.ctypes.data, arr.nbytes, 1)  # hipMemcpyHostToDevice
    if rc != 0:
        raise L.NrfError(f"hipMemcpy H2D failed ({rc})")
