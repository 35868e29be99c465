"""Formats and IO on either side of the step.

* Mesh export: the ASCII PLY of DynamicFusionPipeline::save_mesh_ply /
  save_warped_mesh_ply (include/fusion_app.h:901-943): vertices with normals,
  faces as "3 i j k", numbers as C++ streams print a float (6 significant
  digits, %g); the warped form skins the canonical mesh on the device
  (nrf_skin_vertices) instead of the reference's CPU skinning.
* Depth sequences: DepthSequence (include/depth_sequence.h:27-270), TUM
  (associations.txt or depth/ + color/ folders), ICL (depth/*.png) and raw
  directories of PNGs, with the reference's start_frame / max_frames / depth
  scale semantics and its raw-directory range filter; PNG decoded here
  (zlib + the five scanline filters, 8/16-bit grey, grey-alpha, RGB, RGBA;
  no palette, no interlace).  OpenEXR (RAW_EXR) needs a decoder that this
  image does not have and is refused.
* Configuration: load_config (include/yaml_helper.h:39-193) with the same
  keys, the same required / optional split (a missing required key raises
  KeyError) and the reference's defaults (fusion_app.h:72-119,
  solver.h:13-22, tsdf_volume.h:12-20); pipeline_config() turns it into the
  C-ABI structs of nrf_pipeline_create / nrf_pipeline_track.
"""
from __future__ import annotations

import dataclasses
import enum
import os
import struct
import zlib

import numpy as np

# ------------------------------------------------------------------ PLY


def _g(x) -> str:
    return "%g" % float(x)


def ply_text(vertices, normals, triangles) -> str:
    """The reference's ASCII PLY (fusion_app.h:924-942)."""
    v = np.asarray(vertices, np.float32).reshape(-1, 3)
    n = np.asarray(normals, np.float32).reshape(-1, 3)
    t = np.asarray(triangles, np.int64).reshape(-1, 3)
    if len(n) != len(v):
        raise ValueError("one normal per vertex")
    out = ["ply\nformat ascii 1.0\n", f"element vertex {len(v)}\n",
           "property float x\nproperty float y\nproperty float z\n",
           "property float nx\nproperty float ny\nproperty float nz\n", f"element face {len(t)}\n",
           "property list uchar int vertex_indices\n", "end_header\n"]
    out += [" ".join(_g(c) for c in (*p, *q)) + "\n" for p, q in zip(v, n)]
    out += [f"3 {a} {b} {c}\n" for a, b, c in t]
    return "".join(out)


def save_mesh_ply(path: str, vertices, normals, triangles) -> None:
    with open(path, "w") as f:
        f.write(ply_text(vertices, normals, triangles))


def read_ply(path: str):
    """(vertices, normals, triangles) of an ASCII PLY written by save_mesh_ply."""
    with open(path) as f:
        lines = f.read().splitlines()
    nv = nf = 0
    k = 0
    while lines[k] != "end_header":
        parts = lines[k].split()
        if parts[:2] == ["element", "vertex"]:
            nv = int(parts[2])
        elif parts[:2] == ["element", "face"]:
            nf = int(parts[2])
        k += 1
    k += 1
    vn = np.array([[float(x) for x in lines[k + i].split()] for i in range(nv)], np.float32).reshape(-1, 6)
    faces = np.array([[int(x) for x in lines[k + nv + i].split()[1:4]] for i in range(nf)], np.int32).reshape(-1, 3)
    return vn[:, :3], vn[:, 3:], faces


def mesh_normals(vertices) -> np.ndarray:
    """Per-face normals repeated on the face's 3 vertices of a triangle soup
    (the extract_surface convention, src/marching_cubes.cu:391-407):
    normalize(cross(v1 - v0, v2 - v0))."""
    v = np.asarray(vertices, np.float32).reshape(-1, 3, 3)
    n = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    ln = np.linalg.norm(n, axis=1, keepdims=True)
    n = np.where(ln > 1e-9, n / np.maximum(ln, 1e-30), n).astype(np.float32)
    return np.repeat(n, 3, axis=0)


# ------------------------------------------------------------------ PNG
_PNG_SIG = b"\x89PNG\r\n\x1a\n"
_CHANNELS = {0: 1, 2: 3, 4: 2, 6: 4}


def read_png(path: str) -> np.ndarray:
    """Decoded image: (H, W) or (H, W, C), uint8 or uint16."""
    with open(path, "rb") as f:
        data = f.read()
    if data[:8] != _PNG_SIG:
        raise ValueError(f"{path}: not a PNG file")
    pos, idat, hdr = 8, [], None
    while pos + 8 <= len(data):
        n, typ = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        pos += 12 + n
        if typ == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", body)
        elif typ == b"IDAT":
            idat.append(body)
        elif typ == b"IEND":
            break
    if hdr is None:
        raise ValueError(f"{path}: no IHDR chunk")
    w, h, depth, ctype, _, _, interlace = hdr
    if ctype not in _CHANNELS or depth not in (8, 16) or interlace != 0:
        raise NotImplementedError(f"{path}: PNG colour type {ctype}, depth {depth}, interlace {interlace}")
    ch = _CHANNELS[ctype]
    bpp = ch * depth // 8
    stride = w * bpp
    raw = np.frombuffer(zlib.decompress(b"".join(idat)), np.uint8)
    if raw.size != h * (stride + 1):
        raise ValueError(f"{path}: truncated image data")
    rows = raw.reshape(h, stride + 1)
    ftype = rows[:, 0].astype(np.int64)
    if (ftype > 4).any():
        raise ValueError(f"{path}: unknown scanline filter")
    F = rows[:, 1:].reshape(h, w, bpp).astype(np.int64)
    # Filters reconstructed along anti-diagonals: pixel (y, x) needs its left,
    # upper and upper-left neighbours only, so every pixel of a diagonal is
    # independent of the others.
    O = np.zeros((h + 1, w + 1, bpp), np.int64)
    for d in range(h + w - 1):
        y = np.arange(max(0, d - w + 1), min(h - 1, d) + 1)
        x = d - y
        a, b, c, f = O[y + 1, x], O[y, x + 1], O[y, x], F[y, x]
        p = a + b - c
        pa, pb, pc = np.abs(p - a), np.abs(p - b), np.abs(p - c)
        paeth = np.where((pa <= pb) & (pa <= pc), a, np.where(pb <= pc, b, c))
        t = ftype[y][:, None]
        pred = np.select([t == 0, t == 1, t == 2, t == 3], [0, a, b, (a + b) >> 1], paeth)
        O[y + 1, x + 1] = (f + pred) & 255
    img = O[1:, 1:].astype(np.uint8).reshape(h, stride)
    if depth == 16:
        img = img.view(">u2").astype(np.uint16)
    img = img.reshape(h, w, ch)
    return img[:, :, 0] if ch == 1 else img


def to_gray(img: np.ndarray) -> np.ndarray:
    """cv::IMREAD_GRAYSCALE of a decoded image, 8-bit: 0.299 R + 0.587 G +
    0.114 B rounded (alpha dropped; 16-bit scaled down by 256)."""
    a = np.asarray(img)
    if a.dtype == np.uint16:
        a = (a >> 8).astype(np.uint8)
    if a.ndim == 2:
        return a.astype(np.uint8)
    if a.shape[2] in (1, 2):
        return a[:, :, 0].astype(np.uint8)
    g = 0.299 * a[:, :, 0] + 0.587 * a[:, :, 1] + 0.114 * a[:, :, 2]
    return np.clip(np.rint(g), 0, 255).astype(np.uint8)


def resize_linear(img: np.ndarray, shape) -> np.ndarray:
    """cv::resize(..., INTER_LINEAR) to (H, W): pixel centres aligned, edge
    samples clamped, rounded to the input's integer type."""
    H, W = shape
    h, w = img.shape[:2]
    if (h, w) == (H, W):
        return img
    ys = np.clip((np.arange(H) + 0.5) * h / H - 0.5, 0, h - 1)
    xs = np.clip((np.arange(W) + 0.5) * w / W - 0.5, 0, w - 1)
    y0, x0 = np.floor(ys).astype(int), np.floor(xs).astype(int)
    y1, x1 = np.minimum(y0 + 1, h - 1), np.minimum(x0 + 1, w - 1)
    fy, fx = (ys - y0)[:, None], (xs - x0)[None, :]
    a = img.astype(np.float64)
    if a.ndim == 3:
        fy, fx = fy[..., None], fx[..., None]
    top = a[y0][:, x0] * (1 - fx) + a[y0][:, x1] * fx
    bot = a[y1][:, x0] * (1 - fx) + a[y1][:, x1] * fx
    out = top * (1 - fy) + bot * fy
    return np.clip(np.rint(out), 0, np.iinfo(img.dtype).max).astype(img.dtype)


# --------------------------------------------------------- depth sequences
class Format(enum.Enum):
    TUM = "tum"
    ICL = "icl"
    RAW_PNG = "raw"
    RAW_EXR = "exr"


@dataclasses.dataclass
class SequenceConfig:
    """DepthSequence::Config (include/depth_sequence.h:38-46)."""
    path: str = ""
    format: Format = Format.TUM
    depth_scale: float = 0.001
    start_frame: int = 0
    max_frames: int = -1
    cam: object = None  # _lib.Camera


@dataclasses.dataclass
class DepthFrame:
    depth_m: np.ndarray        # (H, W) float32 metres
    color_gray: np.ndarray     # (H, W) uint8
    timestamp: float = 0.0
    index: int = 0


def _to_metres(raw: np.ndarray, scale: float) -> np.ndarray:
    """raw.convertTo(depth_m, CV_32F, depth_scale): the product in double of
    the float scale, rounded to float."""
    if raw.dtype == np.float32:
        return raw
    return (raw.astype(np.float64) * float(np.float32(scale))).astype(np.float32)


class DepthSequence:
    """include/depth_sequence.h:27-270."""

    def __init__(self, cfg: SequenceConfig):
        self.cfg = cfg
        self.frames: list[DepthFrame] = []
        self._i = 0
        if cfg.format == Format.TUM:
            self._load_tum()
        elif cfg.format == Format.ICL:
            self._load_dir(os.path.join(cfg.path, "depth"), ".png")
        elif cfg.format == Format.RAW_PNG:
            self._load_dir(cfg.path, ".png")
        else:
            raise NotImplementedError("OpenEXR depth (RAW_EXR) needs a decoder that is not available here")

    def has_next(self) -> bool:
        return self._i < len(self.frames)

    def next(self) -> DepthFrame | None:
        if not self.has_next():
            return None
        self._i += 1
        return self.frames[self._i - 1]

    def total(self) -> int:
        return len(self.frames)

    def camera(self):
        return self.cfg.cam

    @staticmethod
    def depth_to_gray(depth_m: np.ndarray) -> np.ndarray:
        """depth_sequence.h:82-98: near = bright over the valid (> 0.01 m)
        range, 0 elsewhere."""
        valid = depth_m > 0.01
        if not valid.any():
            return np.zeros(depth_m.shape, np.uint8)
        lo, hi = float(depth_m[valid].min()), float(depth_m[valid].max())
        if hi <= lo:
            return np.zeros(depth_m.shape, np.uint8)
        g = depth_m.astype(np.float64) * (-255.0 / (hi - lo)) + 255.0 * hi / (hi - lo)
        g = np.clip(np.rint(g), 0, 255).astype(np.uint8)
        g[~valid] = 0
        return g

    def _too_many(self) -> bool:
        return self.cfg.max_frames > 0 and len(self.frames) >= self.cfg.max_frames

    def _load_tum(self) -> None:
        cfg = self.cfg
        assoc = os.path.join(cfg.path, "associations.txt")
        if not os.path.exists(assoc):
            ddir, cdir = os.path.join(cfg.path, "depth"), os.path.join(cfg.path, "color")
            if not (os.path.isdir(ddir) and os.path.isdir(cdir)):
                return
            dp = sorted(os.path.join(ddir, f) for f in os.listdir(ddir) if f.endswith(".png"))
            cp = sorted(os.path.join(cdir, f) for f in os.listdir(cdir) if f.endswith(".png"))
            added = 0
            for i in range(min(len(dp), len(cp))):
                if added < cfg.start_frame:
                    added += 1
                    continue
                if self._too_many():
                    break
                try:
                    raw, color = read_png(dp[i]), to_gray(read_png(cp[i]))
                except (OSError, ValueError):
                    continue
                depth = _to_metres(raw, cfg.depth_scale)
                self.frames.append(DepthFrame(depth, resize_linear(color, depth.shape), float(i), i))
            return
        count = 0
        with open(assoc) as f:
            for line in f:
                line = line.rstrip("\n")
                if not line or line[0] == "#":
                    continue
                if count < cfg.start_frame:
                    count += 1
                    continue
                if self._too_many():
                    break
                parts = line.split()
                ts_rgb, rgb_file, ts_depth, depth_file = float(parts[0]), parts[1], float(parts[2]), parts[3]
                try:
                    raw = read_png(os.path.join(cfg.path, depth_file))
                except (OSError, ValueError):
                    count += 1
                    continue
                depth = _to_metres(raw, cfg.depth_scale)
                try:
                    color = resize_linear(to_gray(read_png(os.path.join(cfg.path, rgb_file))), depth.shape)
                except (OSError, ValueError):
                    color = np.zeros((0, 0), np.uint8)
                self.frames.append(DepthFrame(depth, color, ts_depth, count))
                count += 1

    def _load_dir(self, d: str, ext: str) -> None:
        cfg = self.cfg
        if not os.path.isdir(d):
            return
        paths = sorted(os.path.join(d, f) for f in os.listdir(d) if f.endswith(ext))
        count = 0
        for p in paths:
            if count < cfg.start_frame:
                count += 1
                continue
            if self._too_many():
                break
            try:
                raw = read_png(p)
            except (OSError, ValueError):
                count += 1
                continue
            if raw.ndim == 3:
                raw = raw[:, :, 0]
            depth = _to_metres(raw, cfg.depth_scale)
            # cv::threshold TOZERO at 0.1 m, then TOZERO_INV at 6 m
            depth = np.where(depth > np.float32(0.1), depth, np.float32(0))
            depth = np.where(depth > np.float32(6.0), np.float32(0), depth).astype(np.float32)
            self.frames.append(DepthFrame(depth, self.depth_to_gray(depth), float(count), count))
            count += 1


# ------------------------------------------------------------ YAML config
@dataclasses.dataclass
class TsdfParams:
    """TSDFVolume::Params (include/tsdf_volume.h:12-20)."""
    dims: tuple = (256, 256, 256)
    voxel_size: float = 0.006
    truncation: float = 0.03
    origin: tuple = (-0.768, -0.768, 0.5)
    decay_alpha: float = 0.90
    max_weight: float = 50.0


@dataclasses.dataclass
class SolverParamsCfg:
    """GaussNewtonSolver::Params (include/solver.h:13-22)."""
    gn_iterations: int = 3
    pcg_iterations: int = 10
    pcg_tolerance: float = 1e-4
    lambda_smooth: float = 1.0
    lambda_rot: float = 1.0
    lambda_damping: float = 1e-2
    debug: bool = False


@dataclasses.dataclass
class BBox:
    enabled: bool = False
    min_pt: tuple = (-1.0, -1.0, 0.0)
    max_pt: tuple = (1.0, 1.0, 2.0)


@dataclasses.dataclass
class PipelineParams:
    """DynamicFusionPipeline::Params (include/fusion_app.h:72-119)."""
    tsdf: TsdfParams = dataclasses.field(default_factory=TsdfParams)
    solver: SolverParamsCfg = dataclasses.field(default_factory=SolverParamsCfg)
    cam: object = None
    node_radius: float = 0.04
    node_min_dist: float = 0.025
    dist_threshold: float = 0.05
    angle_threshold: float = 0.7
    view_threshold: float = 0.2
    search_radius_px: int = 3
    min_valid_corrs: int = 2000
    max_nodes: int = 4096
    node_update_every_n: int = 5
    prune_nodes: bool = True
    prune_disconnected: bool = True
    prune_min_observed_voxels: int = 8
    prune_min_surface_voxels: int = 2
    prune_min_component_size: int = 4
    prune_support_radius_factor: float = 1.0
    prune_surface_tsdf_abs: float = 0.45
    prune_empty_tsdf: float = 0.75
    max_dx_mean: float = 0.03
    max_update_rot: float = 0.05
    max_update_trans: float = 0.03
    update_scale: float = 0.25
    bbox: BBox = dataclasses.field(default_factory=BBox)
    integrate_warped: bool = False
    integrate_unwarped_fallback: bool = False
    integrate_min_optimized_count: int = 3
    rigid_tracking: bool = True
    use_sift: bool = False
    sift_max_features: int = 8192
    sift_max_history: int = 30
    sift_max_matches_per_frame: int = 64
    sift_octaves: int = 5
    sift_threshold: float = 3.0
    sift_max_match_error: float = 5.0
    sift_max_ambiguity: float = 0.85
    sift_max_3d_dist: float = 0.20
    sift_max_pixel_dist: float = 0.0
    sift_weight: float = 1.0
    debug_vis: bool = False
    debug_every_n: int = 1
    profile_timings: bool = False
    profile_every_n: int = 1
    quiet: bool = False


@dataclasses.dataclass
class AppConfig:
    seq: SequenceConfig
    df: PipelineParams
    use_vis: bool = False
    quiet: bool = False


def _opt(node: dict, key: str, cast, obj, attr: str | None = None) -> None:
    if node is not None and key in node and node[key] is not None:
        setattr(obj, attr or key, cast(node[key]))


def _f3(v) -> tuple:
    return (float(v[0]), float(v[1]), float(v[2]))


def load_config(path: str) -> AppConfig:
    """load_config (include/yaml_helper.h:39-193)."""
    import yaml

    from . import _lib as L
    with open(path) as f:
        cfg = yaml.safe_load(f)
    s = cfg["sequence"]
    fmt = {"tum": Format.TUM, "icl": Format.ICL}.get(str(s["format"]), Format.RAW_PNG)
    c = cfg["camera"]
    cam = L.camera(float(c["fx"]), float(c["fy"]), float(c["cx"]), float(c["cy"]), int(c["width"]),
                   int(c["height"]))
    seq = SequenceConfig(path=str(s["path"]), format=fmt, depth_scale=float(s["depth_scale"]),
                         start_frame=int(s["start_frame"]), max_frames=int(s["max_frames"]), cam=cam)
    df = PipelineParams(cam=cam)
    t = cfg["tsdf"]
    df.tsdf.dims = tuple(int(x) for x in t["dims"][:3])
    df.tsdf.voxel_size = float(t["voxel_size"])
    df.tsdf.truncation = float(t["truncation"])
    df.tsdf.origin = _f3(t["origin"])
    _opt(t, "decay_alpha", float, df.tsdf)
    _opt(t, "max_weight", float, df.tsdf)
    sol = cfg["solver"]
    df.solver.gn_iterations = int(sol["gn_iterations"])
    df.solver.pcg_iterations = int(sol["pcg_iterations"])
    df.solver.lambda_smooth = float(sol["lambda_smooth"])
    _opt(sol, "lambda_damping", float, df.solver)
    _opt(sol, "debug", bool, df.solver)
    w = cfg["warp"]
    df.node_radius = float(w["node_radius"])
    df.node_min_dist = float(w["node_min_dist"])
    df.max_nodes = int(w["max_nodes"])
    for key, cast in (("node_update_every_n", int), ("prune_nodes", bool), ("prune_disconnected", bool),
                      ("prune_min_observed_voxels", int), ("prune_min_surface_voxels", int),
                      ("prune_min_component_size", int), ("prune_support_radius_factor", float),
                      ("prune_surface_tsdf_abs", float), ("prune_empty_tsdf", float), ("max_dx_mean", float),
                      ("max_update_rot", float), ("max_update_trans", float), ("update_scale", float),
                      ("integrate_warped", bool), ("integrate_unwarped_fallback", bool),
                      ("integrate_min_optimized_count", int)):
        _opt(w, key, cast, df)
    icp = cfg["icp"]
    df.dist_threshold = float(icp["dist_threshold"])
    df.angle_threshold = float(icp["angle_threshold"])
    for key, cast in (("view_threshold", float), ("search_radius_px", int), ("min_valid_corrs", int),
                      ("rigid_tracking", bool)):
        _opt(icp, key, cast, df)
    if cfg.get("sift") is not None:
        sift = cfg["sift"]
        _opt(sift, "enabled", bool, df, "use_sift")
        for key, cast in (("max_features", int), ("max_history", int), ("max_matches_per_frame", int),
                          ("octaves", int), ("threshold", float), ("max_match_error", float),
                          ("max_ambiguity", float), ("max_3d_dist", float), ("max_pixel_dist", float),
                          ("weight", float)):
            _opt(sift, key, cast, df, "sift_" + key)
    quiet = False
    if cfg.get("profiler") is not None:
        prof = cfg["profiler"]
        _opt(prof, "enabled", bool, df, "profile_timings")
        if prof.get("every_n") is not None:
            df.profile_every_n = max(1, int(prof["every_n"]))
        if prof.get("quiet") is not None:
            quiet = bool(prof["quiet"])
    b = cfg["bbox"]
    df.bbox = BBox(bool(b["enabled"]), _f3(b["min_pt"]), _f3(b["max_pt"]))
    use_vis = bool(cfg["visualizer"]["enabled"])
    df.quiet = quiet
    return AppConfig(seq=seq, df=df, use_vis=use_vis, quiet=quiet)


def pipeline_config(app: AppConfig, max_triangles: int = 2_000_000):
    """(nrf_pipeline_config, nrf_track_params) of a loaded configuration: the
    volume, camera, solver (the reference's absolute PCG tolerance), GN step
    clamps with the step validation on, graph size and the association
    thresholds; the line search on (the reference always runs it)."""
    from . import _lib as L
    df = app.df
    t = df.tsdf
    cfg = L.PipelineConfig()
    cfg.volume = L.volume_params(t.dims, t.voxel_size, t.truncation, t.origin, t.decay_alpha, t.max_weight)
    cfg.camera = df.cam
    s = df.solver
    cfg.solver = L.solver_params(s.gn_iterations, s.pcg_iterations, s.pcg_tolerance, s.lambda_smooth, s.lambda_rot,
                                 s.lambda_damping)
    cfg.gn = L.gn_params(s.gn_iterations, df.update_scale, df.max_update_rot, df.max_update_trans, 1, df.max_dx_mean)
    cfg.max_nodes = df.max_nodes
    cfg.max_corrs = df.cam.width * df.cam.height
    cfg.max_triangles = max_triangles
    cfg.min_opt_count = df.integrate_min_optimized_count
    tp = L.track_params(L.assoc_params(df.dist_threshold, df.angle_threshold, df.view_threshold,
                                       df.search_radius_px), df.min_valid_corrs, True)
    return cfg, tp
