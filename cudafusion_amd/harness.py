"""Workload driver: synthetic scene -> device inputs -> the non-rigid step.

Used by bench.py, smoke() and the GPU tests.  Frame 0 is integrated
rigidly, the node graph is sampled on the canonical surface and the voxel
kNN snapshot taken (the reference's initialize_node_graph,
include/fusion_app.h:994-1003); every later frame is one call of
nrf_pipeline_step on inputs that are already resident in device memory.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L
from . import scene as S


def _t(a: np.ndarray, dev) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def identity_pose() -> np.ndarray:
    return np.eye(4, dtype=np.float32).reshape(16)


def dq_to_rt(q: np.ndarray):
    """Rotation matrix and translation of a unit dual quaternion (x, y, z, w)."""
    q = np.asarray(q, np.float64)
    r, d = q[:4] / np.linalg.norm(q[:4]), q[4:] / np.linalg.norm(q[:4])
    x, y, z, w = r
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    # t = 2 d r*
    cx, cy, cz, cw = -x, -y, -z, w
    dx, dy, dz, dw = d
    tx = dw * cx + dx * cw + dy * cz - dz * cy
    ty = dw * cy - dx * cz + dy * cw + dz * cx
    tz = dw * cz + dx * cy - dy * cx + dz * cw
    return R, 2 * np.array([tx, ty, tz])


def transform_errors(Ta, Tb):
    """Max rotation angle (rad) and translation (m) differences between two
    transform sets (DUALQUAT records or (N, 8) arrays)."""
    A = np.ascontiguousarray(Ta).view(np.float32).reshape(-1, 8) if np.asarray(Ta).dtype == L.DUALQUAT \
        else np.asarray(Ta, np.float64).reshape(-1, 8)
    B = np.ascontiguousarray(Tb).view(np.float32).reshape(-1, 8) if np.asarray(Tb).dtype == L.DUALQUAT \
        else np.asarray(Tb, np.float64).reshape(-1, 8)
    er, et = 0.0, 0.0
    for a, b in zip(A, B):
        Ra, ta = dq_to_rt(a)
        Rb, tb = dq_to_rt(b)
        c = np.clip((np.trace(Ra.T @ Rb) - 1) / 2, -1, 1)
        er = max(er, float(np.arccos(c)))
        et = max(et, float(np.linalg.norm(ta - tb)))
    return er, et


class Harness:
    def __init__(self, kind: str = "body", frames: int = 40, use_graph: bool = False, seed: int = 0,
                 pcg_tolerance: float = 0.0, relative_tolerance: bool = False, precond_fp64: bool = False,
                 arap_consistent_sign: bool = False, validate: bool = True, pcg_pipelined: int = 0,
                 jtj_precision: int = 0, refine_iterations: int = 0, storage16: bool = False,
                 refine_jacobi: bool = False, true_residual_rounds: int = 0, psd_dropped_pairs: bool = False,
                 plain_form: bool = False, workload: S.Workload | None = None):
        self.kind = kind
        self.scene = S.scene_for(workload) if workload is not None else S.make_scene(kind, frames, seed)
        self.w = self.scene.w
        w = self.w
        self.frames = frames
        self.vp = L.volume_params(w.dims, w.voxel_size, w.truncation, w.origin, w.decay_alpha, w.max_weight)
        self.cam = L.camera(w.fx, w.fy, w.cx, w.cy, w.width, w.height)
        self.sp = L.solver_params(w.gn_iterations, w.pcg_iterations, pcg_tolerance, w.lambda_smooth, 1.0,
                                  w.lambda_damping, 0, int(relative_tolerance), int(precond_fp64),
                                  int(arap_consistent_sign), int(pcg_pipelined), int(jtj_precision),
                                  int(refine_iterations), int(refine_jacobi), int(true_residual_rounds),
                                  int(psd_dropped_pairs), int(plain_form))
        self.gn = L.gn_params(w.gn_iterations, w.update_scale, w.max_update_rot, w.max_update_trans,
                              int(validate), w.max_dx_mean)
        self.use_graph = use_graph
        self.storage16 = storage16
        self.pose = identity_pose()
        self.pipe = None

    # ------------------------------------------------------------ setup
    def setup(self, max_triangles: int | None = None, host: "Harness | None" = None) -> None:
        """Creates the pipeline, integrates frame 0, samples the nodes and
        packs every frame's correspondences on the device.  ``host``: take
        the rendered frames and the nodes from another harness of the same
        workload instead of computing them again."""
        lib = L.load()
        dev = torch.device("cuda", torch.cuda.current_device())
        self.dev = dev
        w = self.w
        npx = w.width * w.height
        cfg = L.PipelineConfig()
        cfg.volume = self.vp
        cfg.camera = self.cam
        cfg.solver = self.sp
        cfg.gn = self.gn
        cfg.max_nodes = 8192
        cfg.max_corrs = npx
        cfg.max_triangles = max_triangles or (2_000_000 if w.dims[0] >= 256 else 400_000)
        cfg.min_opt_count = w.min_opt_count
        cfg.use_graph = int(self.use_graph)
        cfg.storage16 = int(self.storage16)
        self.cfg = cfg
        self.pipe = lib.nrf_pipeline_create(C.byref(cfg))
        if not self.pipe:
            raise L.NrfError("nrf_pipeline_create failed: " + lib.nrf_last_error().decode())
        self.lib = lib
        self.frames_host = (host.frames_host if host is not None
                            else [self.scene.render(f) for f in range(self.frames + 1)])
        self.depth = [_t(fr.depth.reshape(-1), dev) for fr in self.frames_host]
        stream = torch.cuda.current_stream().cuda_stream
        L.check(lib.nrf_pipeline_integrate_rigid(self.pipe, self.depth[0].data_ptr(), self.pose.ctypes.data,
                                                 stream), "integrate_rigid")
        torch.cuda.synchronize()
        self.nodes = host.nodes if host is not None else S.scene_nodes(self.scene)
        self.N = len(self.nodes)
        L.check(lib.nrf_pipeline_set_nodes(self.pipe, self.nodes.ctypes.data, self.N), "set_nodes")
        # association inputs for frames 1..F, resident on the device
        self.canonical = [None]
        self.corrs = [None]
        knn = lib.nrf_pipeline_knn(self.pipe)
        knn_w = lib.nrf_pipeline_knn_w(self.pipe)
        for f in range(1, self.frames + 1):
            fr = self.frames_host[f]
            can = _t(fr.canonical.reshape(-1), dev)
            dst = _t(fr.dst.reshape(-1), dev)
            nrm = _t(fr.normal.reshape(-1), dev)
            val = _t(fr.valid.reshape(-1), dev)
            out = torch.empty(npx * 76, dtype=torch.uint8, device=dev)
            L.check(lib.nrf_corr_pack(C.byref(self.vp), knn, knn_w, can.data_ptr(), dst.data_ptr(),
                                      nrm.data_ptr(), None, val.data_ptr(), npx, out.data_ptr(), stream),
                    "corr_pack")
            self.canonical.append(can)
            self.corrs.append(out)
        torch.cuda.synchronize()

    # ------------------------------------------------------------- step
    def step(self, f: int) -> None:
        stream = torch.cuda.current_stream().cuda_stream
        L.check(self.lib.nrf_pipeline_step(self.pipe, self.depth[f].data_ptr(), self.pose.ctypes.data,
                                           self.corrs[f].data_ptr(), self.canonical[f].data_ptr(),
                                           self.w.width * self.w.height, stream), "pipeline_step")

    def track(self, f: int, params: L.TrackParams, frame_id: int | None = None) -> None:
        """Frame f with the device association (nrf_pipeline_track): the
        prediction of the current mesh, projective correspondences against
        frame f's depth, the GN loop, marking, warped integration, MC."""
        stream = torch.cuda.current_stream().cuda_stream
        L.check(self.lib.nrf_pipeline_track(self.pipe, self.depth[f].data_ptr(), self.pose.ctypes.data,
                                            C.byref(params), f if frame_id is None else frame_id, stream),
                "pipeline_track")

    def track_stats(self) -> L.TrackStats:
        out = L.TrackStats()
        L.check(self.lib.nrf_pipeline_track_stats(self.pipe, C.byref(out)), "track_stats")
        return out

    def track_maps(self) -> dict:
        """Host copies of the last tracking frame's device maps."""
        ptrs = [C.c_void_p() for _ in range(6)]
        L.check(self.lib.nrf_pipeline_track_maps(self.pipe, *[C.byref(q) for q in ptrs]), "track_maps")
        npx = self.w.width * self.w.height
        V = int(np.prod(self.w.dims))
        return dict(live_v=self._d2h(ptrs[0].value, 3 * npx, np.float32),
                    live_n=self._d2h(ptrs[1].value, 3 * npx, np.float32),
                    hit=self._d2h(ptrs[2].value, npx, np.int32),
                    corrs=self._d2h(ptrs[3].value, npx, L.CORR),
                    opt=self._d2h(ptrs[4].value, V, np.int32),
                    dnorm=self._d2h(ptrs[5].value, 3 * npx, np.float32))

    def solver_result(self) -> L.PcgResult:
        """The device-side PCG / GN record of the last step (syncs)."""
        solver = self.lib.nrf_pipeline_solver(self.pipe)
        ptr = self.lib.nrf_solver_device_result(solver)
        raw = self._d2h(ptr, C.sizeof(L.PcgResult), np.uint8)
        return L.PcgResult.from_buffer_copy(raw.tobytes())

    def solver_totals(self) -> L.SolverTotals:
        """Running GN / PCG totals over every step so far (syncs)."""
        out = L.SolverTotals()
        solver = self.lib.nrf_pipeline_solver(self.pipe)
        L.check(self.lib.nrf_solver_get_totals(solver, C.byref(out), torch.cuda.current_stream().cuda_stream),
                "solver_totals")
        return out

    def num_blocks(self) -> int:
        solver = self.lib.nrf_pipeline_solver(self.pipe)
        nb = C.c_int()
        L.check(self.lib.nrf_solver_system(solver, None, None, None, C.byref(nb), None, None), "system")
        return nb.value

    def stage_ms(self):
        out = (C.c_float * 3)()
        L.check(self.lib.nrf_pipeline_stage_ms(self.pipe, out), "stage_ms")
        return list(out)

    # ---------------------------------------------------------- readback
    def _d2h(self, ptr: int, count: int, dtype) -> np.ndarray:
        from .fusion import copy_device_to_numpy
        return copy_device_to_numpy(ptr, count, dtype)

    def download_transforms(self) -> np.ndarray:
        return self._d2h(self.lib.nrf_pipeline_transforms(self.pipe), self.N, L.DUALQUAT)

    def download_voxels(self) -> np.ndarray:
        """The volume as nrf_voxel records (the compact volume decoded in the
        precision form)."""
        V = int(np.prod(self.w.dims))
        if not self.storage16:
            return self._d2h(self.lib.nrf_pipeline_voxels(self.pipe), V, L.VOXEL)
        t16, w8 = C.c_void_p(), C.c_void_p()
        L.check(self.lib.nrf_pipeline_volume16(self.pipe, C.byref(t16), C.byref(w8), None, None), "volume16")
        out = np.zeros(V, L.VOXEL)
        out["tsdf"] = self._d2h(t16.value, V, np.float16).astype(np.float32)
        out["weight"] = self._d2h(w8.value, V, np.uint8).astype(np.float32)
        return out

    def download_knn(self):
        V = int(np.prod(self.w.dims))
        return (self._d2h(self.lib.nrf_pipeline_knn(self.pipe), 4 * V, np.int32),
                self._d2h(self.lib.nrf_pipeline_knn_w(self.pipe), 4 * V, np.float32))

    def triangle_count(self) -> int:
        cnt = C.c_void_p()
        self.lib.nrf_pipeline_mesh(self.pipe, C.byref(cnt))
        return int(self._d2h(cnt.value, 2, np.int32)[1])

    def mesh(self):
        cnt = C.c_void_p()
        v = self.lib.nrf_pipeline_mesh(self.pipe, C.byref(cnt))
        T = int(self._d2h(cnt.value, 2, np.int32)[1])
        T = min(T, self.cfg.max_triangles)
        return self._d2h(v, T * 9, np.float32).reshape(-1, 3)

    def save_mesh_ply(self, path: str, warped: bool = False) -> int:
        """The pipeline's current canonical mesh as the reference's ASCII PLY
        (save_mesh_ply, include/fusion_app.h:921-943) with per-face normals;
        warped=True writes it deformed into the live frame
        (save_warped_mesh_ply, :901-919), skinned on the device through the
        voxel kNN as the live surface prediction does (the reference skins on
        the host through each vertex's own 4 nearest nodes) with the
        canonical normals kept, as there.  Returns the triangle count."""
        from . import formats
        v = self.mesh()
        n = formats.mesh_normals(v)
        if warped and self.N > 0 and len(v):
            dev = self.dev
            cen = torch.empty(self.N * 32, dtype=torch.uint8, device=dev)
            L.check(self.lib.nrf_warp_centered_table(self.lib.nrf_pipeline_nodes(self.pipe),
                                                     self.lib.nrf_pipeline_transforms(self.pipe), self.N,
                                                     cen.data_ptr(), None), "centered")
            dv = _t(v.reshape(-1), dev)
            out = torch.empty_like(dv)
            L.check(self.lib.nrf_skin_vertices(dv.data_ptr(), len(v), C.byref(self.vp), cen.data_ptr(), self.N,
                                               self.lib.nrf_pipeline_knn(self.pipe),
                                               self.lib.nrf_pipeline_knn_w(self.pipe), out.data_ptr(), None), "skin")
            torch.cuda.synchronize()
            v = out.cpu().numpy().reshape(-1, 3)
        formats.save_mesh_ply(path, v, n, np.arange(len(v), dtype=np.int32).reshape(-1, 3))
        return len(v) // 3

    def frame_corrs_host(self, f: int):
        npx = self.w.width * self.w.height
        corrs = self.corrs[f].cpu().numpy().view(L.CORR)
        can = self.canonical[f].cpu().numpy().astype(np.float32)
        assert len(corrs) == npx
        return corrs, can

    def close(self) -> None:
        if self.pipe:
            self.lib.nrf_pipeline_destroy(self.pipe)
            self.pipe = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchHarness:
    """Several harnesses (already set up) stepped together through the
    nrf_batch_* entry points."""

    def __init__(self, harnesses: list, use_graph: bool = False, tuning: L.BatchTuning | None = None):
        self.h = list(harnesses)
        self.lib = L.load()
        arr = (C.c_void_p * len(self.h))(*[x.pipe for x in self.h])
        self.batch = self.lib.nrf_batch_create(arr, len(self.h), C.byref(tuning) if tuning is not None else None,
                                               int(use_graph))
        if not self.batch:
            raise L.NrfError("nrf_batch_create failed: " + self.lib.nrf_last_error().decode())
        self._frames = {}

    def frames(self, f: int):
        """The BatchFrame array of frame f (kept alive, same address per f)."""
        if f not in self._frames:
            arr = (L.BatchFrame * len(self.h))()
            for k, x in enumerate(self.h):
                arr[k].depth = x.depth[f].data_ptr()
                arr[k].camera_pose[:] = [float(v) for v in x.pose.reshape(-1)]
                arr[k].corrs = x.corrs[f].data_ptr()
                arr[k].canonical = x.canonical[f].data_ptr()
                arr[k].num_corrs = x.w.width * x.w.height
            self._frames[f] = arr
        return self._frames[f]

    def step(self, f: int) -> None:
        L.check(self.lib.nrf_batch_step(self.batch, self.frames(f), torch.cuda.current_stream().cuda_stream),
                "batch_step")

    def autotune(self, f: int, use_cache: bool = True) -> L.BatchTuning:
        out = L.BatchTuning()
        L.check(self.lib.nrf_batch_autotune(self.batch, self.frames(f), int(use_cache), C.byref(out),
                                            torch.cuda.current_stream().cuda_stream), "batch_autotune")
        return out

    def tuning(self) -> L.BatchTuning:
        out = L.BatchTuning()
        L.check(self.lib.nrf_batch_get_tuning(self.batch, C.byref(out)), "batch_get_tuning")
        return out

    def set_tuning(self, t: L.BatchTuning) -> None:
        L.check(self.lib.nrf_batch_set_tuning(self.batch, C.byref(t)), "batch_set_tuning")

    def stage_ms(self):
        out = (C.c_float * 3)()
        L.check(self.lib.nrf_batch_stage_ms(self.batch, out), "batch_stage_ms")
        return list(out)

    def close(self) -> None:
        if self.batch:
            self.lib.nrf_batch_destroy(self.batch)
            self.batch = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
