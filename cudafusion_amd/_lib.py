"""ctypes binding of ``libnrf.so`` (the C ABI declared in ``include/nrf.h``).

The library is the only compute path: if it is missing, or if it cannot be
loaded, every call raises.  ``torch`` is imported first so that the HIP
runtime torch ships is the one the library binds to (both use the soname
``libamdhip64.so.7``); device buffers handed to the library are torch
tensors, passed by address.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnrf.so")

# ------------------------------------------------------------ record dtypes
# Byte layouts of the reference's POD structs (include/types.h:86-206).
VOXEL = np.dtype([("tsdf", "<f4"), ("weight", "<f4"), ("color", "u1", 3), ("pad", "u1")])
NODE = np.dtype([("pos", "<f4", 3), ("normal", "<f4", 3), ("radius", "<f4"), ("neighbors", "<i4", 8),
                 ("neighbor_w", "<f4", 8), ("num_neighbors", "<i4")])
DUALQUAT = np.dtype([("real", "<f4", 4), ("dual", "<f4", 4)])
CORR = np.dtype([("src", "<f4", 3), ("dst", "<f4", 3), ("normal", "<f4", 3), ("node_ids", "<i4", 4),
                 ("node_ws", "<f4", 4), ("weight", "<f4"), ("valid", "u1"), ("pad", "u1", 3)])
assert VOXEL.itemsize == 12 and NODE.itemsize == 96 and DUALQUAT.itemsize == 32 and CORR.itemsize == 76


class VolumeParams(C.Structure):
    _fields_ = [("dims", C.c_int32 * 3), ("voxel_size", C.c_float), ("truncation", C.c_float),
                ("origin", C.c_float * 3), ("decay_alpha", C.c_float), ("max_weight", C.c_float)]


class Camera(C.Structure):
    _fields_ = [("fx", C.c_float), ("fy", C.c_float), ("cx", C.c_float), ("cy", C.c_float),
                ("width", C.c_int32), ("height", C.c_int32)]


class AssocParams(C.Structure):
    _fields_ = [("dist_threshold", C.c_float), ("angle_threshold", C.c_float), ("view_threshold", C.c_float),
                ("search_radius_px", C.c_int32)]


def assoc_params(dist_threshold=0.05, angle_threshold=0.7, view_threshold=0.2, search_radius_px=3) -> "AssocParams":
    """Defaults are the reference pipeline's (include/fusion_app.h:79-82)."""
    return AssocParams(dist_threshold, angle_threshold, view_threshold, int(search_radius_px))


TRACK_MAX_GN = 16


class TrackParams(C.Structure):
    _fields_ = [("assoc", AssocParams), ("min_valid_corrs", C.c_int32), ("line_search", C.c_int32)]


def track_params(assoc=None, min_valid_corrs=2000, line_search=True) -> "TrackParams":
    """Defaults are the reference pipeline's (include/fusion_app.h:79-83;
    the line search of :494-641 on)."""
    return TrackParams(assoc if assoc is not None else assoc_params(), int(min_valid_corrs), int(bool(line_search)))


class TrackStats(C.Structure):
    _fields_ = [("gn_run", C.c_int32), ("gn_applied", C.c_int32), ("stop_reason", C.c_int32),
                ("last_valid", C.c_int32), ("num_valid", C.c_int32 * TRACK_MAX_GN),
                ("best_scale", C.c_float * TRACK_MAX_GN), ("base_rmse", C.c_double * TRACK_MAX_GN),
                ("best_rmse", C.c_double * TRACK_MAX_GN)]


class PruneParams(C.Structure):
    _fields_ = [("enabled", C.c_int32), ("remove_disconnected", C.c_int32), ("min_observed_voxels", C.c_int32),
                ("min_surface_voxels", C.c_int32), ("min_component_size", C.c_int32),
                ("support_radius_factor", C.c_float), ("surface_tsdf_abs", C.c_float), ("empty_tsdf", C.c_float)]


def prune_params(enabled=True, remove_disconnected=True, min_observed_voxels=8, min_surface_voxels=2,
                 min_component_size=4, support_radius_factor=1.0, surface_tsdf_abs=0.45,
                 empty_tsdf=0.75) -> "PruneParams":
    """Defaults are WarpField::PruneParams' (include/warp_field.h:44-53)."""
    return PruneParams(int(enabled), int(remove_disconnected), min_observed_voxels, min_surface_voxels,
                       min_component_size, support_radius_factor, surface_tsdf_abs, empty_tsdf)


class SolverParams(C.Structure):
    _fields_ = [("gn_iterations", C.c_int32), ("pcg_iterations", C.c_int32), ("pcg_tolerance", C.c_float),
                ("lambda_smooth", C.c_float), ("lambda_rot", C.c_float), ("lambda_damping", C.c_float),
                ("debug", C.c_int32), ("relative_tolerance", C.c_int32), ("precond_fp64", C.c_int32),
                ("pcg_pipelined", C.c_int32), ("arap_consistent_sign", C.c_int32), ("jtj_precision", C.c_int32),
                ("refine_iterations", C.c_int32), ("refine_jacobi", C.c_int32),
                ("true_residual_rounds", C.c_int32), ("psd_dropped_pairs", C.c_int32), ("plain_form", C.c_int32)]


class GNParams(C.Structure):
    _fields_ = [("gn_iterations", C.c_int32), ("update_scale", C.c_float), ("max_update_rot", C.c_float),
                ("max_update_trans", C.c_float), ("validate", C.c_int32), ("max_dx_mean", C.c_float)]


class PcgResult(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("total_iterations", C.c_int32), ("flags", C.c_int32),
                ("gn_done", C.c_int32), ("final_residual", C.c_double), ("rhs_norm", C.c_double),
                ("converged", C.c_int32), ("breakdowns", C.c_int32), ("runs", C.c_int32), ("pad", C.c_int32)]


class SolverTotals(C.Structure):
    _fields_ = [("solves", C.c_int64), ("gn_done", C.c_int64), ("pcg_iterations", C.c_int64),
                ("rejected", C.c_int64), ("pcg_runs", C.c_int64), ("pcg_converged", C.c_int64),
                ("pcg_breakdowns", C.c_int64)]


MAX_BATCH = 32


class BatchFrame(C.Structure):
    _fields_ = [("depth", C.c_void_p), ("camera_pose", C.c_float * 16), ("corrs", C.c_void_p),
                ("canonical", C.c_void_p), ("num_corrs", C.c_int32), ("pad", C.c_int32)]


class BatchTuning(C.Structure):
    _fields_ = [("lanes_per_row", C.c_int32), ("wg_nodes", C.c_int32), ("z_chunk", C.c_int32),
                ("launch_groups", C.c_int32), ("pcg_ms", C.c_float), ("integrate_ms", C.c_float)]


class PipelineConfig(C.Structure):
    _fields_ = [("volume", VolumeParams), ("camera", Camera), ("solver", SolverParams), ("gn", GNParams),
                ("max_nodes", C.c_int32), ("max_corrs", C.c_int32), ("max_triangles", C.c_int32),
                ("min_opt_count", C.c_int32), ("use_graph", C.c_int32), ("storage16", C.c_int32)]


P = C.c_void_p
I = C.c_int
F = C.c_float

# name: (restype, argtypes)
_SIGS = {
    "nrf_status_string": (C.c_char_p, [I]),
    "nrf_last_error": (C.c_char_p, []),
    "nrf_build_info": (C.c_char_p, []),
    "nrf_tsdf_integrate": (I, [C.POINTER(VolumeParams), P, P, C.POINTER(Camera), P, P, P, I, P, P, P, P, I, P, P]),
    "nrf_tsdf_reset": (I, [C.POINTER(VolumeParams), P, P]),
    "nrf_mc_workspace_size": (C.c_size_t, [C.POINTER(VolumeParams)]),
    "nrf_mc_extract": (I, [C.POINTER(VolumeParams), P, P, C.c_size_t, P, P, P, I, P, P]),
    "nrf_mc_cases": (P, [P]),
    "nrf_mc_extract_noalloc": (I, [C.POINTER(VolumeParams), P, P, P, P, I, P, P]),
    "nrf_depth_normals": (I, [P, C.POINTER(Camera), P, P]),
    "nrf_pixel_knn": (I, [P, I, P, P, P, P, P]),
    "nrf_find_correspondences": (I, [P, P, P, P, C.POINTER(Camera), P, P, P, C.POINTER(AssocParams), P, P, P, P]),
    "nrf_mark_optimized_voxels": (I, [P, P, I, C.POINTER(VolumeParams), I, P, P, P]),
    "nrf_skin_vertices": (I, [P, I, C.POINTER(VolumeParams), P, I, P, P, P, P]),
    "nrf_raster_workspace_size": (C.c_size_t, [C.POINTER(Camera)]),
    "nrf_rasterize_mesh": (I, [P, P, P, I, I, C.POINTER(Camera), P, C.POINTER(VolumeParams), P, C.c_size_t, P, P,
                               P, P, P]),
    "nrf_tsdf16_reset": (I, [C.POINTER(VolumeParams), P, P, P]),
    "nrf_tsdf16_from_voxels": (I, [C.POINTER(VolumeParams), P, P, P, P]),
    "nrf_knn_compact": (I, [C.POINTER(VolumeParams), P, P, P, P, P]),
    "nrf_tsdf16_integrate": (I, [C.POINTER(VolumeParams), P, P, P, C.POINTER(Camera), P, P, I, P, P, P, P]),
    "nrf_mc16_extract": (I, [C.POINTER(VolumeParams), P, P, P, C.c_size_t, P, P, I, P, P]),
    "nrf_voxel_knn": (I, [C.POINTER(VolumeParams), P, P, I, P, P, P]),
    "nrf_warp_apply_twist": (I, [P, P, I, F, F, F, P, P, P]),
    "nrf_warp_centered_table": (I, [P, P, I, P, P]),
    "nrf_warp_points": (I, [P, P, P, I, P, P, I, P, P]),
    "nrf_warp_normals": (I, [P, P, P, I, P, I, P, P]),
    "nrf_corr_pack": (I, [C.POINTER(VolumeParams), P, P, P, P, P, P, P, I, P, P]),
    "nrf_solver_create": (P, [C.POINTER(SolverParams), I]),
    "nrf_solver_destroy": (None, [P]),
    "nrf_solver_set_params": (I, [P, C.POINTER(SolverParams)]),
    "nrf_solver_set_pattern": (I, [P, P, I]),
    "nrf_solver_solve": (I, [P, P, I, P, P, I, P, C.POINTER(PcgResult), P]),
    "nrf_solver_solve_gn": (I, [P, P, P, I, P, P, P, I, C.POINTER(GNParams), P, C.POINTER(PcgResult), P]),
    "nrf_solver_system": (I, [P, C.POINTER(P), C.POINTER(P), C.POINTER(P), C.POINTER(I), C.POINTER(P),
                              C.POINTER(P)]),
    "nrf_solver_device_result": (P, [P]),
    "nrf_solver_get_totals": (I, [P, C.POINTER(SolverTotals), P]),
    "nrf_solver_true_residual": (I, [P, P, C.POINTER(C.c_double), P]),
    "nrf_solver_resolve": (I, [P, P, C.POINTER(PcgResult), P]),
    "nrf_solver_set_pcg_layout": (I, [P, I, I]),
    "nrf_pipeline_create": (P, [C.POINTER(PipelineConfig)]),
    "nrf_pipeline_destroy": (None, [P]),
    "nrf_pipeline_voxels": (P, [P]),
    "nrf_pipeline_transforms": (P, [P]),
    "nrf_pipeline_nodes": (P, [P]),
    "nrf_pipeline_knn": (P, [P]),
    "nrf_pipeline_knn_w": (P, [P]),
    "nrf_pipeline_mesh": (P, [P, C.POINTER(P)]),
    "nrf_pipeline_solver": (P, [P]),
    "nrf_pipeline_volume16": (I, [P, C.POINTER(P), C.POINTER(P), C.POINTER(P), C.POINTER(P)]),
    "nrf_pipeline_integrate_rigid": (I, [P, P, P, P]),
    "nrf_pipeline_set_nodes": (I, [P, P, I]),
    "nrf_pipeline_step": (I, [P, P, P, P, P, I, P]),
    "nrf_pipeline_stage_ms": (I, [P, C.POINTER(F)]),
    "nrf_pipeline_profile": (I, [P, I]),
    "nrf_pipeline_profile_pcg": (I, [P, I]),
    "nrf_pipeline_profile_every": (I, [P, I]),
    "nrf_pipeline_profile_read": (I, [P, I, C.POINTER(F)]),
    "nrf_pipeline_updated_voxels": (C.c_longlong, [P]),
    "nrf_pipeline_graph_captures": (C.c_longlong, [P]),
    "nrf_pipeline_set_overlap": (C.c_int, [P, C.c_int]),
    "nrf_pipeline_join": (C.c_int, [P, P]),
    "nrf_raycast": (I, [C.POINTER(VolumeParams), P, C.POINTER(Camera), P, P, P, I, P, P, P, P, P, P]),
    "nrf_graph_add_nodes": (I, [P, P, C.POINTER(I), I, P, P, I, C.c_float, C.c_float, C.POINTER(I)]),
    "nrf_graph_add_nodes_geodesic": (I, [P, P, C.POINTER(I), I, P, P, I, P, I, C.c_float, C.c_float, C.POINTER(I),
                                         P]),
    "nrf_graph_prune": (I, [C.POINTER(VolumeParams), P, P, P, C.POINTER(I), C.POINTER(PruneParams), C.POINTER(I), P]),
    "nrf_pipeline_track": (I, [P, P, P, C.POINTER(TrackParams), I, P]),
    "nrf_pipeline_track_stats": (I, [P, C.POINTER(TrackStats)]),
    "nrf_pipeline_track_maps": (I, [P, C.POINTER(P), C.POINTER(P), C.POINTER(P), C.POINTER(P), C.POINTER(P),
                                    C.POINTER(P)]),
    "nrf_batch_create": (P, [C.POINTER(P), I, C.POINTER(BatchTuning), I]),
    "nrf_batch_destroy": (None, [P]),
    "nrf_batch_set_tuning": (I, [P, C.POINTER(BatchTuning)]),
    "nrf_batch_get_tuning": (I, [P, C.POINTER(BatchTuning)]),
    "nrf_batch_autotune": (I, [P, C.POINTER(BatchFrame), I, C.POINTER(BatchTuning), P]),
    "nrf_batch_step": (I, [P, C.POINTER(BatchFrame), P]),
    "nrf_batch_stage_ms": (I, [P, C.POINTER(F)]),
    "nrf_lp_create": (P, [I, I, I]),
    "nrf_lp_destroy": (None, [P]),
    "nrf_lp_param_count": (I, []),
    "nrf_lp_params": (P, [P]),
    "nrf_lp_grads": (P, [P]),
    "nrf_lp_reset_optimizer": (I, [P]),
    "nrf_lp_set_graph": (I, [P, P, I]),
    "nrf_lp_forward": (I, [P, P, P, I, P]),
    "nrf_lp_loss_grad": (I, [P, P, P, P, I, C.POINTER(C.c_double), P]),
    "nrf_lp_adam": (I, [P, F, I, P]),
    "nrf_lp_fold": (I, [P, P, P]),
    "nrf_lp_fallbacks": (I, [P, C.POINTER(C.c_int), P]),
    "nrf_lp_blocks": (I, [P, C.POINTER(P), C.POINTER(P), C.POINTER(P), C.POINTER(P)]),
}

_lib = None


class NrfError(RuntimeError):
    pass


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load (once) and return the library; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise NrfError(f"{path} is missing: build it with __graft_entry__.build() (hipcc, gfx950)")
    import torch  # noqa: F401  (binds the process to torch's HIP runtime first)
    lib = C.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols() -> list[str]:
    return sorted(_SIGS)


def check(status: int, what: str = "") -> None:
    if status != 0:
        lib = load()
        msg = lib.nrf_status_string(status).decode()
        last = lib.nrf_last_error().decode()
        raise NrfError(f"{what}: {msg}" + (f" ({last})" if last else ""))


def volume_params(dims, voxel_size, truncation, origin, decay_alpha=1.0, max_weight=50.0) -> VolumeParams:
    vp = VolumeParams()
    vp.dims[:] = [int(d) for d in dims]
    vp.voxel_size = voxel_size
    vp.truncation = truncation
    vp.origin[:] = [float(o) for o in origin]
    vp.decay_alpha = decay_alpha
    vp.max_weight = max_weight
    return vp


def camera(fx, fy, cx, cy, width, height) -> Camera:
    return Camera(fx, fy, cx, cy, int(width), int(height))


def solver_params(gn_iterations=3, pcg_iterations=10, pcg_tolerance=1e-4, lambda_smooth=1.0, lambda_rot=1.0,
                  lambda_damping=1e-2, debug=0, relative_tolerance=0, precond_fp64=0,
                  arap_consistent_sign=0, pcg_pipelined=0, jtj_precision=0, refine_iterations=0,
                  refine_jacobi=0, true_residual_rounds=0, psd_dropped_pairs=0, plain_form=0) -> SolverParams:
    return SolverParams(gn_iterations, pcg_iterations, pcg_tolerance, lambda_smooth, lambda_rot, lambda_damping,
                        debug, relative_tolerance, precond_fp64, pcg_pipelined, arap_consistent_sign, jtj_precision,
                        refine_iterations, refine_jacobi, true_residual_rounds,
                        psd_dropped_pairs, plain_form)


def gn_params(gn_iterations=3, update_scale=0.25, max_update_rot=0.05, max_update_trans=0.03, validate=1,
              max_dx_mean=0.03) -> GNParams:
    """Defaults are the reference pipeline's (include/fusion_app.h:72-119)."""
    return GNParams(gn_iterations, update_scale, max_update_rot, max_update_trans, validate, max_dx_mean)
