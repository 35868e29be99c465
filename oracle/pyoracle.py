"""ctypes wrapper of the CPU oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
TABLE = os.path.join(HERE, "mc_cases.txt")

P = C.c_void_p
I = C.c_int
L = C.c_long
F = C.c_float
D = C.c_double

_SIGS = {
    "orc_threads": (I, [I]),
    "orc_load_mc_table": (I, [C.c_char_p]),
    "orc_exp_se3_f64": (None, [P, P]),
    "orc_exp_se3_f32": (None, [P, P]),
    "orc_mat_to_dq_f64": (None, [P, P]),
    "orc_dq_to_mat_f64": (None, [P, P]),
    "orc_dq_mul_f64": (None, [P, P, P]),
    "orc_dq_centered_f64": (None, [P, P, P]),
    "orc_dq_transform_f64": (None, [P, P, P]),
    "orc_twist_update_f64": (None, [P, P, D, D, D, P]),
    "orc_twist_update_f32": (None, [P, P, F, F, F, P]),
    "orc_warp_points_f64": (None, [P, P, P, I, P, P, I, P]),
    "orc_warp_points_f32": (None, [P, P, P, I, P, P, I, P]),
    "orc_warp_normals_f32": (None, [P, P, P, I, P, I, P]),
    "orc_warp_normals_f64": (None, [P, P, P, I, P, I, P]),
    "orc_integrate_margin_f64": (None, [P, P, P, P, P, P, I, P, P, P, P]),
    "orc_integrate_f32": (L, [P, P, P, P, P, P, P, I, P, P, P, I]),
    "orc_integrate_f64": (L, [P, P, P, P, P, P, P, P, I, P, P, P, I]),
    "orc_mc_f32": (L, [P, P, P, P, P, L]),
    "orc_mc_f64": (L, [P, P, P, P, P, P, L]),
    "orc_knn_f32": (None, [P, P, P, I, P, P]),
    "orc_knn_f64": (None, [P, P, P, I, P, P]),
    "orc_pattern": (I, [P, I, P, P, I]),
    "orc_assemble_f32": (I, [P, I, P, P, P, I, F, F, I, P, P, P, I, I]),
    "orc_assemble_f64": (I, [P, I, P, P, P, I, D, D, P, P, P, I, I]),
    "orc_pcg_f32": (D, [P, P, P, P, P, I, I, F, I, P, P, P, P]),
    "orc_pcg_f64": (D, [P, P, P, P, P, I, I, D, I, P, P, P, P]),
    "orc_spmv_f64": (None, [P, P, P, P, P, I]),
    "orc_validate_f32": (I, [P, I, P]),
    "orc_depth_normals_f32": (None, [P, P, P]),
    "orc_pixel_knn": (None, [P, I, P, P, P, P]),
    "orc_find_correspondences_f32": (I, [P, P, P, P, P, P, P, P, P, P, P]),
    "orc_mark_optimized_voxels": (None, [P, P, I, P, I, P, P]),
    "orc_skin_vertices_f32": (None, [P, I, P, P, P, I, P, P, P]),
    "orc_rasterize_f32": (None, [P, P, P, I, I, P, P, P, P, P, P, P]),
    "orc_validate_f64": (I, [P, I, P]),
    "orc_apply_step_f32": (None, [P, P, I, P, F]),
    "orc_assemble_pattern_f32": (None, [P, I, P, P, P, I, F, F, P, P, P, P, P, I]),
    "orc_corr_rmse": (D, [P, I]),
    "orc_raycast_f32": (None, [P, P, P, P, P, P, I, P, P, P, P, P]),
    "orc_graph_add_nodes": (I, [P, P, P, I, P, P, I, F, F]),
    "orc_graph_add_nodes_geodesic": (I, [P, P, P, I, P, P, I, P, I, F, F, P, P]),
    "orc_graph_prune": (I, [P, P, P, P, P, P, P]),
    "orc_solve_gn_f32": (I, [P, P, I, P, I, P, P, P, P, P, P, P]),
    "orc_solve_gn_f64": (I, [P, P, I, P, I, P, P, P, P, P, P, P]),
    "orc_lp_param_count": (I, []),
    "orc_lp_offsets": (I, [P, I]),
    "orc_lp_features_f64": (None, [P, I, P, P, P, P]),
    "orc_lp_forward_f64": (None, [P, I, P, I, P, P, P]),
    "orc_lp_apply_f64": (None, [I, P, I, P, P, P, P, P, I, P]),
    "orc_lp_loss_grad_f64": (D, [P, I, P, I, P, P, P, P, P, P, P, I, P]),
    "orc_lp_pcg_f64": (I, [I, P, I, P, P, P, P, P, P, P, P, I, D, P, P]),
}

_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    lib = C.CDLL(LIB)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.orc_load_mc_table(TABLE.encode()) != 0:
        raise RuntimeError("oracle: cannot read " + TABLE)
    _lib = lib
    return lib


def ptr(a):
    """Address of a numpy array or ctypes object.  The caller must keep the
    object referenced until the C call returns (no temporaries)."""
    if a is None:
        return None
    if isinstance(a, (C.Structure, C.Array)):
        return C.addressof(a)
    return a.ctypes.data


def threads(n: int = 0) -> int:
    return load().orc_threads(n)


def identity_dq64(n: int) -> np.ndarray:
    t = np.zeros((n, 8))
    t[:, 3] = 1.0
    return t


def dq_records_to64(T) -> np.ndarray:
    a = np.ascontiguousarray(T).view(np.float32).reshape(-1, 8)
    return a.astype(np.float64)


# ------------------------------------------------------------ components

def integrate_f32(vp, vox, depth, cam, pose, nodes=None, T=None, knn=None, knn_w=None, opt=None, min_opt=0):
    lib = load()
    n = 0 if nodes is None else len(nodes)
    return lib.orc_integrate_f32(C.byref(vp), ptr(vox), ptr(depth), C.byref(cam), ptr(pose), ptr(nodes), ptr(T), n,
                                 ptr(knn), ptr(knn_w), ptr(opt), min_opt)


def integrate_f64(vp, tsdf, weight, depth, cam, pose, nodes=None, T64=None, knn=None, knn_w=None, opt=None,
                  min_opt=0):
    lib = load()
    n = 0 if nodes is None else len(nodes)
    return lib.orc_integrate_f64(C.byref(vp), ptr(tsdf), ptr(weight), ptr(depth), C.byref(cam), ptr(pose),
                                 ptr(nodes), ptr(T64), n, ptr(knn), ptr(knn_w), ptr(opt), min_opt)


def integrate_margin_f64(vp, depth, cam, pose, nodes=None, T64=None, knn=None, knn_w=None):
    """Per-voxel distance of the FP64 projection from the pixel-rounding
    boundary (pixels) and of the sdf from the truncation cut (units of the
    truncation)."""
    V = int(np.prod(vp.dims[:]))
    m_px = np.empty(V)
    m_sdf = np.empty(V)
    n = 0 if nodes is None else len(nodes)
    T64 = None if T64 is None else np.ascontiguousarray(T64, np.float64)
    load().orc_integrate_margin_f64(C.byref(vp), ptr(depth), C.byref(cam), ptr(pose), ptr(nodes), ptr(T64), n,
                                    ptr(knn), ptr(knn_w), ptr(m_px), ptr(m_sdf))
    return m_px, m_sdf


def tsdf_fp64_errors(gpu_vox, t64, w64, margins, px_eps=2e-3, sdf_eps=1e-4):
    """Compares an FP32 device TSDF with the FP64 oracle on the voxels that
    are not within a rounding margin of a projection or truncation boundary.
    Returns (max |tsdf error| on unambiguous voxels, fraction of ambiguous
    voxels among those projecting into the image, count of unambiguous
    weight mismatches).  The voxel grid puts a few percent of projections
    exactly on half-pixels (rational coincidences such as 262.5 * 0.075 /
    0.625 = 31.5), where FP32 and FP64 may round to different pixels."""
    m_px, m_sdf = margins
    ambiguous = (m_px < px_eps) | (m_sdf < sdf_eps)
    in_image = np.isfinite(m_px)
    ok = ~ambiguous
    wmis = int(((gpu_vox["weight"] != w64.astype(np.float32)) & ok).sum())
    both = ok & (gpu_vox["weight"] == w64.astype(np.float32))
    err = float(np.abs(gpu_vox["tsdf"][both] - t64[both]).max()) if both.any() else 0.0
    amb = float((ambiguous & in_image).sum() / max(1, int(in_image.sum())))
    return err, amb, wmis


def mc_f32(vp, vox, normals=True):
    lib = load()
    V = int(np.prod(vp.dims[:]))
    cases = np.zeros(V, np.uint8)
    T = lib.orc_mc_f32(C.byref(vp), ptr(vox), ptr(cases), None, None, 0)
    verts = np.zeros(max(T, 1) * 9, np.float32)
    nrm = np.zeros(max(T, 1) * 9, np.float32) if normals else None
    lib.orc_mc_f32(C.byref(vp), ptr(vox), None, ptr(verts), ptr(nrm), T)
    return cases, verts[: T * 9].reshape(-1, 3), (nrm[: T * 9].reshape(-1, 3) if normals else None)


def mc_f64(vp, tsdf, weight, normals=True):
    lib = load()
    V = int(np.prod(vp.dims[:]))
    cases = np.zeros(V, np.uint8)
    T = lib.orc_mc_f64(C.byref(vp), ptr(tsdf), ptr(weight), ptr(cases), None, None, 0)
    verts = np.zeros(max(T, 1) * 9, np.float64)
    nrm = np.zeros(max(T, 1) * 9, np.float64) if normals else None
    lib.orc_mc_f64(C.byref(vp), ptr(tsdf), ptr(weight), None, ptr(verts), ptr(nrm), T)
    return cases, verts[: T * 9].reshape(-1, 3), (nrm[: T * 9].reshape(-1, 3) if normals else None)


def knn_f32(vp, vox, nodes):
    V = int(np.prod(vp.dims[:]))
    knn = np.zeros(V * 4, np.int32)
    w = np.zeros(V * 4, np.float32)
    load().orc_knn_f32(C.byref(vp), ptr(vox), ptr(nodes), len(nodes), ptr(knn), ptr(w))
    return knn, w


def validate(dx, gn, fp64=False) -> bool:
    """The caller's step validation (fusion_app.h:409-491) on dx [6N]."""
    lib = load()
    d = np.ascontiguousarray(dx, dtype=np.float64 if fp64 else np.float32)
    fn = lib.orc_validate_f64 if fp64 else lib.orc_validate_f32
    return bool(fn(ptr(d), d.size // 6, C.byref(gn)))


# ---------------------------------------------------- projective association

def depth_normals_f32(depth, cam):
    n = np.zeros(cam.width * cam.height * 3, np.float32)
    d = np.ascontiguousarray(depth, np.float32)
    load().orc_depth_normals_f32(ptr(d), C.byref(cam), ptr(n))
    return n


def pixel_knn(hit, voxel_knn, voxel_knn_w):
    hit = np.ascontiguousarray(hit, np.int32)
    ids = np.zeros(len(hit) * 4, np.int32)
    ws = np.zeros(len(hit) * 4, np.float32)
    load().orc_pixel_knn(ptr(hit), len(hit), ptr(voxel_knn), ptr(voxel_knn_w), ptr(ids), ptr(ws))
    return ids, ws


def find_correspondences_f32(live_v, live_n, depth, dnorm, cam, pose, pknn, pknn_w, ap):
    """Returns (corrs, valid count, stats[6])."""
    from cudafusion_amd import _lib as L
    arrs = [np.ascontiguousarray(a, np.float32) for a in (live_v, live_n, depth, dnorm, pose)]
    corrs = np.zeros(cam.width * cam.height, L.CORR)
    stats = np.zeros(6, np.int32)
    pk = None if pknn is None else np.ascontiguousarray(pknn, np.int32)
    pw = None if pknn_w is None else np.ascontiguousarray(pknn_w, np.float32)
    n = load().orc_find_correspondences_f32(ptr(arrs[0]), ptr(arrs[1]), ptr(arrs[2]), ptr(arrs[3]), C.byref(cam),
                                            ptr(arrs[4]), ptr(pk), ptr(pw), C.byref(ap), ptr(corrs), ptr(stats))
    return corrs, n, stats


def mark_optimized_voxels(corrs, hit, dims, frame_id, counts, last_frame):
    """Updates counts / last_frame (int32 arrays of the volume size) in place."""
    hit = np.ascontiguousarray(hit, np.int32)
    d = np.ascontiguousarray(dims, np.int32)
    load().orc_mark_optimized_voxels(ptr(corrs), ptr(hit), len(hit), ptr(d), int(frame_id), ptr(counts),
                                     ptr(last_frame))


def skin_vertices_f32(vp, canon, nodes, T, knn, knn_w):
    c = np.ascontiguousarray(canon, np.float32)
    out = np.zeros_like(c)
    N = 0 if nodes is None else len(nodes)
    load().orc_skin_vertices_f32(ptr(c), c.size // 3, C.byref(vp), ptr(nodes), ptr(T), N, ptr(knn), ptr(knn_w),
                                 ptr(out))
    return out


def rasterize_f32(canon, warped, tris, cam, pose, vp):
    """Returns (live vertices, live normals, hit voxel, canonical points) per pixel."""
    c = np.ascontiguousarray(canon, np.float32)
    w = np.ascontiguousarray(warped, np.float32)
    t = np.ascontiguousarray(tris, np.int32)
    P_ = np.ascontiguousarray(pose, np.float32)
    npx = cam.width * cam.height
    lv, ln, cp = (np.zeros(npx * 3, np.float32) for _ in range(3))
    hit = np.zeros(npx, np.int32)
    load().orc_rasterize_f32(ptr(c), ptr(w), ptr(t), c.size // 3, t.size // 3, C.byref(cam), ptr(P_), C.byref(vp),
                             ptr(lv), ptr(ln), ptr(hit), ptr(cp))
    return lv, ln, hit, cp


def raycast_f32(vp, vox, cam, pose, nodes=None, T=None, knn=None, knn_w=None):
    """TSDFVolume::raycast (src/tsdf_volume.cu:65-92): (vertices, normals,
    canonical voxel index) per pixel."""
    npx = cam.width * cam.height
    v = np.zeros(npx * 3, np.float32)
    n = np.zeros(npx * 3, np.float32)
    idx = np.zeros(npx, np.int32)
    P_ = np.ascontiguousarray(pose, np.float32)
    N = 0 if nodes is None else len(nodes)
    load().orc_raycast_f32(C.byref(vp), ptr(vox), C.byref(cam), ptr(P_), ptr(nodes), ptr(T), N, ptr(knn), ptr(knn_w),
                           ptr(v), ptr(n), ptr(idx))
    return v, n, idx


def graph_add_nodes(nodes, T, num_nodes, verts, normals, min_dist, radius, tris=None):
    """add_nodes_from_surface (tris None) or add_nodes_from_mesh_geodesic on
    host arrays of capacity len(nodes) (updated in place); returns (added,
    num_nodes, geodesic ids, geodesic distances) (the last two None without
    tris)."""
    v = np.ascontiguousarray(verts, np.float32).reshape(-1)
    nrm = None if normals is None else np.ascontiguousarray(normals, np.float32).reshape(-1)
    n = C.c_int(num_nodes)
    if tris is None:
        a = load().orc_graph_add_nodes(ptr(nodes), ptr(T), C.byref(n), len(nodes), ptr(v), ptr(nrm), v.size // 3,
                                       min_dist, radius)
        return a, n.value, None, None
    t = np.ascontiguousarray(tris, np.int32).reshape(-1)
    gi = np.zeros(len(nodes) * 8, np.int32)
    gd = np.zeros(len(nodes) * 8, np.float32)
    a = load().orc_graph_add_nodes_geodesic(ptr(nodes), ptr(T), C.byref(n), len(nodes), ptr(v), ptr(nrm),
                                            v.size // 3, ptr(t), t.size // 3, min_dist, radius, ptr(gi), ptr(gd))
    return a, n.value, gi.reshape(-1, 8)[: n.value], gd.reshape(-1, 8)[: n.value]


def graph_prune(vp, vox, nodes, T, num_nodes, pp):
    """prune_nodes (warp_field.cu:986-1104) in place; returns (removed,
    num_nodes, keep flags before compaction)."""
    n = C.c_int(num_nodes)
    keep = np.zeros(max(num_nodes, 1), np.uint8)
    r = load().orc_graph_prune(C.byref(vp), ptr(vox), ptr(nodes), ptr(T), C.byref(n), C.byref(pp), ptr(keep))
    return r, n.value, keep[:num_nodes]


def pattern(nodes):
    lib = load()
    N = len(nodes)
    rp = np.zeros(N + 1, np.int32)
    nb = lib.orc_pattern(ptr(nodes), N, ptr(rp), None, 0)
    ci = np.zeros(max(nb, 1), np.int32)
    lib.orc_pattern(ptr(nodes), N, ptr(rp), ptr(ci), nb)
    return rp, ci[:nb]


def assemble_f32(corrs, nodes, T, lam_s, lam_d, canonical=None, precond_fp64=0, consistent_sign=0, psd_pairs=0):
    N = len(nodes)
    rp, ci = pattern(nodes)
    nb = len(ci)
    A = np.zeros(nb * 36, np.float32)
    b = np.zeros(N * 6, np.float32)
    M = np.zeros(N * 36, np.float32)
    load().orc_assemble_f32(ptr(corrs), len(corrs), ptr(canonical), ptr(nodes), ptr(T), N, lam_s, lam_d,
                            precond_fp64, ptr(A), ptr(b), ptr(M), consistent_sign, psd_pairs)
    return rp, ci, A.reshape(nb, 6, 6), b, M.reshape(N, 6, 6)


def assemble_pattern_f32(corrs, nodes, T, lam_s, lam_d, rp, ci, canonical=None, consistent_sign=0):
    """The FP32 system on a given pattern (rp, ci) instead of the nodes' own."""
    N = len(nodes)
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    nb = len(ci)
    A = np.zeros(nb * 36, np.float32)
    b = np.zeros(N * 6, np.float32)
    M = np.zeros(N * 36, np.float32)
    load().orc_assemble_pattern_f32(ptr(corrs), len(corrs), ptr(canonical), ptr(nodes), ptr(T), N, lam_s, lam_d,
                                    ptr(rp), ptr(ci), ptr(A), ptr(b), ptr(M), consistent_sign)
    return A.reshape(nb, 6, 6), b, M.reshape(N, 6, 6)


def assemble_f64(corrs, nodes, T64, lam_s, lam_d, canonical=None, consistent_sign=0, psd_pairs=0):
    N = len(nodes)
    rp, ci = pattern(nodes)
    nb = len(ci)
    A = np.zeros(nb * 36)
    b = np.zeros(N * 6)
    M = np.zeros(N * 36)
    load().orc_assemble_f64(ptr(corrs), len(corrs), ptr(canonical), ptr(nodes), ptr(T64), N, lam_s, lam_d, ptr(A),
                            ptr(b), ptr(M), consistent_sign, psd_pairs)
    return rp, ci, A.reshape(nb, 6, 6), b, M.reshape(N, 6, 6)


def pcg_f64(rp, ci, A, b, M, maxit, tol, rel=0, trace=False):
    N = len(rp) - 1
    x = np.zeros(6 * N)
    it = C.c_int()
    fl = C.c_int()
    tr = np.zeros(max(maxit, 1)) if trace else None
    A = np.ascontiguousarray(A, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    M = np.ascontiguousarray(M, np.float64)
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    r = load().orc_pcg_f64(ptr(rp), ptr(ci), ptr(A), ptr(b), ptr(M), N, maxit, tol, rel, ptr(x), C.byref(it),
                           C.byref(fl), ptr(tr))
    return x, it.value, fl.value, r, (tr[: it.value] if trace else None)


def pcg_f32(rp, ci, A, b, M, maxit, tol, rel=0):
    N = len(rp) - 1
    x = np.zeros(6 * N, np.float32)
    it = C.c_int()
    fl = C.c_int()
    A = np.ascontiguousarray(A, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    M = np.ascontiguousarray(M, np.float32)
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    r = load().orc_pcg_f32(ptr(rp), ptr(ci), ptr(A), ptr(b), ptr(M), N, maxit, tol, rel, ptr(x), C.byref(it),
                           C.byref(fl), None)
    return x, it.value, fl.value, r


def solve_gn_f32(corrs, canonical, nodes, T, gn, sp):
    T = np.array(T, copy=True)
    N = len(nodes)
    dx = np.zeros(6 * N, np.float32)
    tot, done, last = C.c_int(), C.c_int(), C.c_int()
    load().orc_solve_gn_f32(ptr(corrs), ptr(canonical), len(corrs), ptr(nodes), N, ptr(T), C.byref(gn),
                            C.byref(sp), ptr(dx), C.byref(tot), C.byref(done), C.byref(last))
    return T, dx, tot.value, done.value, last.value


def solve_gn_f64(corrs, canonical, nodes, T64, gn, sp):
    T64 = np.array(T64, dtype=np.float64, copy=True)
    N = len(nodes)
    dx = np.zeros(6 * N)
    tot, done, last = C.c_int(), C.c_int(), C.c_int()
    load().orc_solve_gn_f64(ptr(corrs), ptr(canonical), len(corrs), ptr(nodes), N, ptr(T64), C.byref(gn),
                            C.byref(sp), ptr(dx), C.byref(tot), C.byref(done), C.byref(last))
    return T64, dx, tot.value, done.value, last.value


def apply_step_f32(dx, T, gn, scale):
    """upload_scaled_delta_x(scale) + apply_twist_increment on every node
    (fusion_app.h:1353-1375, warp_field.cu:1121-1141); returns new T."""
    T = np.array(T, copy=True)
    d = np.ascontiguousarray(dx, np.float32)
    load().orc_apply_step_f32(ptr(d), ptr(T), len(T), C.byref(gn), float(scale))
    return T


def corr_rmse(corrs) -> float:
    """correspondence_rmse (fusion_app.h:1331-1351)."""
    return float(load().orc_corr_rmse(ptr(corrs), len(corrs)))


def track_frame_f32(vp, cam, pose, depth, verts, nodes, T, knn, knn_w, gn, sp, tp, frame_id, opt, opt_frame,
                    vox=None, min_opt=0):
    """One tracking frame of DynamicFusionPipeline::process_frame
    (include/fusion_app.h:253-669, frames after the first, no rigid
    tracking / SIFT / bounding box), composed of the oracle's stages in the
    reference's order: per GN iteration the live surface prediction of the
    canonical mesh `verts` (a triangle soup, 3 vertices per triangle)
    (:259-263, 1216-1320), the association (:288-291, 1791-1845), the
    min_valid_corrs break (:342-343), then either the step at update_scale
    (tp.line_search 0) or the base RMSE, one linearisation, the validation
    (:409-491) and the ten-scale search (:494-641); then
    mark_optimized_voxels (:1322-1329) and, with vox, the warped integration
    gated by min_opt (:655-666).  opt / opt_frame are updated in place.
    Returns (T, stats dict, corrs, hit) where corrs / hit are the slots the
    marking read."""
    from cudafusion_amd import _lib as Lb
    npx = cam.width * cam.height
    depth = np.ascontiguousarray(depth, np.float32)
    dn = depth_normals_f32(depth, cam)
    verts = np.ascontiguousarray(verts, np.float32).reshape(-1)
    tris = np.arange(verts.size // 3, dtype=np.int32)
    G = max(1, gn.gn_iterations)
    ls = bool(tp.line_search)
    st = dict(num_valid=[], best_scale=[], base_rmse=[], best_rmse=[], stop_reason=0, gn_applied=0)
    g1 = Lb.gn_params(1, gn.update_scale, gn.max_update_rot, gn.max_update_trans, gn.validate, gn.max_dx_mean)

    def associate(Tc):
        warped = skin_vertices_f32(vp, verts, nodes, Tc, knn, knn_w)
        lv, ln, hit, _ = rasterize_f32(verts, warped, tris, cam, pose, vp)
        pk, pw = pixel_knn(hit, knn, knn_w)
        c, nv, _ = find_correspondences_f32(lv, ln, depth, dn, cam, pose, pk, pw, tp.assoc)
        return c, hit, nv

    us = np.float32(gn.update_scale)
    trials = [np.float32(f) * us for f in (1.0, 0.5, 0.25, 0.125, 0.0625, -0.0625, -0.125, -0.25, -0.5)]
    trials[0] = us
    trials.append(-us)
    corrs = hit = None
    for g in range(G):
        corrs, hit, nv = associate(T)
        st["num_valid"].append(nv)
        if nv < tp.min_valid_corrs:
            st["stop_reason"] = 1
            break
        if not ls:
            T, _, _, done, _ = solve_gn_f32(corrs, None, nodes, T, g1, sp)
            if not done:
                st["stop_reason"] = 2
                break
            st["gn_applied"] += 1
            continue
        base = corr_rmse(corrs)
        st["base_rmse"].append(base)
        _, dx, _, _, _ = solve_gn_f32(corrs, None, nodes, T, g1, sp)
        if gn.validate and not validate(dx, gn):
            st["stop_reason"] = 2
            break
        if not gn.update_scale > 0 or not np.isfinite(base):
            st["stop_reason"] = 4
            break
        Tsave = np.array(T, copy=True)
        best, best_scale = base, np.float32(0)
        for s in trials:
            Tt = apply_step_f32(dx, Tsave, gn, s)
            corrs, hit, n2 = associate(Tt)
            r = corr_rmse(corrs)
            if (n2 >= tp.min_valid_corrs and n2 >= int(np.float32(0.75) * np.float32(nv)) and np.isfinite(r)
                    and r < best):
                best, best_scale = r, s
        st["best_scale"].append(float(best_scale))
        st["best_rmse"].append(best)
        T = Tsave
        if best_scale == 0:
            st["stop_reason"] = 3
            break
        T = apply_step_f32(dx, Tsave, gn, best_scale)
        st["gn_applied"] += 1
        corrs, hit, _ = associate(T)
    mark_optimized_voxels(corrs, hit, vp.dims[:], frame_id, opt, opt_frame)
    if vox is not None:
        integrate_f32(vp, vox, depth, cam, pose, nodes, T, knn, knn_w, opt, min_opt)
    st["last_valid"] = int((corrs["valid"] != 0).sum())
    return T, st, corrs, hit


# ------------------------------------------------- learned preconditioner
LP_NAMES = ["W0", "b0", "Wm0", "bm0", "Wu0", "bu0", "Wm1", "bm1", "Wu1", "bu1", "Wm2", "bm2", "Wu2", "bu2", "Wo", "bo",
            "We", "be"]


def lp_offsets() -> dict:
    """Parameter name -> (offset, size) of the flat parameter vector."""
    out = np.zeros(32, np.int32)
    n = load().orc_lp_offsets(ptr(out), 32)
    offs = [int(v) for v in out[:n]]
    return {name: (offs[k], offs[k + 1] - offs[k]) for k, name in enumerate(LP_NAMES)}


def lp_param_count() -> int:
    return int(load().orc_lp_param_count())


def lp_diag_blocks(rp, ci, A):
    """The diagonal 6x6 block of every block row of a BSR system."""
    N = len(rp) - 1
    out = np.zeros((N, 36))
    for i in range(N):
        for q in range(rp[i], rp[i + 1]):
            if ci[q] == i:
                out[i] = np.asarray(A[q], np.float64).reshape(36)
    return out


def lp_features_f64(nodes, Adiag, Minv):
    N = len(nodes)
    x = np.zeros((N, 28))
    Cm = np.zeros((N, 36))
    Adiag = np.ascontiguousarray(Adiag, np.float64).reshape(N, 36)
    Minv = np.ascontiguousarray(Minv, np.float64).reshape(N, 36)
    load().orc_lp_features_f64(ptr(nodes), N, ptr(Adiag), ptr(Minv), ptr(x), ptr(Cm))
    return x, Cm


def lp_forward_f64(params, variant, nodes, x):
    N = len(nodes)
    Td = np.zeros((N, 21))
    Te = np.zeros((N, 8, 36))
    params = np.ascontiguousarray(params, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    load().orc_lp_forward_f64(ptr(params), variant, ptr(nodes), N, ptr(x), ptr(Td), ptr(Te))
    return Td, Te


def lp_loss_grad_f64(params, variant, nodes, x, Cm, Minv, rp, ci, A, V, grad=True):
    N = len(nodes)
    params = np.ascontiguousarray(params, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    Cm = np.ascontiguousarray(Cm, np.float64)
    Minv = np.ascontiguousarray(Minv, np.float64).reshape(N, 36)
    A = np.ascontiguousarray(A, np.float64)
    V = np.ascontiguousarray(V, np.float64)
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    g = np.zeros(len(params)) if grad else None
    loss = load().orc_lp_loss_grad_f64(ptr(params), variant, ptr(nodes), N, ptr(x), ptr(Cm), ptr(Minv), ptr(rp),
                                       ptr(ci), ptr(A), ptr(V), V.shape[0], ptr(g))
    return loss, g


def lp_apply_f64(variant, nodes, Cm, Minv, Td, Te, R):
    N = len(nodes)
    R = np.ascontiguousarray(R, np.float64).reshape(-1, 6 * N)
    Z = np.zeros_like(R)
    args = [np.ascontiguousarray(a, np.float64) for a in (Cm, Minv, Td, Te)]
    load().orc_lp_apply_f64(variant, ptr(nodes), N, *[ptr(a) for a in args], ptr(R), R.shape[0], ptr(Z))
    return Z


def lp_pcg_f64(variant, nodes, rp, ci, A, b, Cm, Minv, Td, Te, maxit, tol):
    """FP64 PCG with M^-1 = BJ^-1 + L L^T (relative tolerance); returns
    (iterations, true relative residual, x)."""
    N = len(nodes)
    x = np.zeros(6 * N)
    tr = C.c_double()
    arrs = [np.ascontiguousarray(a, np.float64) for a in (A, b, Cm, Minv, Td, Te)]
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    it = load().orc_lp_pcg_f64(variant, ptr(nodes), N, ptr(rp), ptr(ci), *[ptr(a) for a in arrs], maxit, tol,
                               ptr(x), C.byref(tr))
    return it, tr.value, x
