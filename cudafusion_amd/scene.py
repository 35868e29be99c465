"""Seeded synthetic inputs for the non-rigid step (harness, host side).

Two scenes, one per workload:

* ``SphereScene`` - 128^3 volume at 6 mm (truncation 3 voxels), 320x240
  depth (fx = fy = 262.5, cx = 160, cy = 120) of a sphere of radius 0.25 m
  at z = 1.0 m bent each frame by a sinusoidal x-shear (1 cm amplitude,
  0.3 m wavelength, phase advancing per frame), depth noise
  N(0, (0.0012 z^2)^2), nodes every 25 mm, 3 GN x 20 PCG, ARAP weight 10.
* ``BodyScene`` - 256^3 volume at 4 mm (truncation 5 voxels = 20 mm),
  640x480 depth (fx = fy = 525, cx = 319.5, cy = 239.5) of an articulated
  capsule body (torso r = 0.15 m / length 0.5 m, two 2-segment arms
  r = 0.04 m / 0.3 m) at z ~ 1.2 m whose joint angles follow seeded random
  walks of at most 3 degrees per frame, same noise model, nodes every 20 mm,
  5 GN x 30 PCG.

Depth is rendered by exact ray casting through pixel coordinates (u, v), as
CameraIntrinsics::unproject (include/types.h:157-161) expects.  The data
association is the ground truth: the correspondence of a pixel is the
canonical point that the true motion maps onto the pixel's surface point
(dst = the noisy back-projected depth, normal = the true surface normal,
weight 1); the 4-NN of each canonical point is copied from the voxel kNN
field on the device (nrf_corr_pack).  Node sampling and the node graph
follow WarpField::add_nodes_from_surface / add_nodes_from_mesh_geodesic
(src/warp_field.cu:709-955) with Euclidean instead of geodesic distance;
they are inputs of the step, built once.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from ._lib import NODE

# ------------------------------------------------------------------ configs


@dataclass
class Workload:
    name: str
    dims: tuple
    voxel_size: float
    truncation: float
    origin: tuple
    width: int
    height: int
    fx: float
    fy: float
    cx: float
    cy: float
    frames: int
    node_spacing: float
    gn_iterations: int
    pcg_iterations: int
    lambda_smooth: float
    lambda_damping: float = 1e-2
    decay_alpha: float = 1.0
    max_weight: float = 50.0
    # caller step policy: the reference pipeline's defaults (include/fusion_app.h:72-119)
    update_scale: float = 0.25
    max_update_rot: float = 0.05
    max_update_trans: float = 0.03
    max_dx_mean: float = 0.03
    min_opt_count: int = 0
    seed: int = 0
    extra: dict = field(default_factory=dict)


def sphere_workload(frames: int = 10, seed: int = 0) -> Workload:
    return Workload("sphere-128^3-320x240", (128, 128, 128), 0.006, 0.018, (-0.384, -0.384, 0.616), 320, 240,
                    262.5, 262.5, 160.0, 120.0, frames, 0.025, 3, 20, 10.0, seed=seed)


def body_workload(frames: int = 40, seed: int = 0) -> Workload:
    return Workload("capsule-body-256^3-640x480", (256, 256, 256), 0.004, 0.020, (-0.512, -0.512, 0.70), 640,
                    480, 525.0, 525.0, 319.5, 239.5, frames, 0.020, 5, 30, 10.0, seed=seed)


def body_instance_workload(nodes: int, dims: int, voxel_size: float, width: int, frames: int = 40,
                           seed: int = 0, gn_iterations: int = 5, pcg_iterations: int = 30) -> Workload:
    """A capsule-body sequence of another size: node spacing chosen for about
    ``nodes`` nodes (2382 at 20 mm, count ~ spacing^-2), a dims^3 volume of
    ``voxel_size`` centred on the body (truncation 5 voxels), and a 640x480 or
    320x240 camera with the intrinsics of the full-scale workload scaled."""
    ext = dims * voxel_size
    sc = width / 640.0
    spacing = 0.020 * math.sqrt(2382.0 / nodes)
    return Workload(f"capsule-body-{dims}^3-{width}x{int(480 * sc)}-n{nodes}", (dims, dims, dims), voxel_size,
                    5 * voxel_size, (-ext / 2, -ext / 2, 1.21 - ext / 2), width, int(480 * sc), 525.0 * sc,
                    525.0 * sc, 319.5 * sc + 0.5 * (sc - 1.0), 239.5 * sc + 0.5 * (sc - 1.0), frames, spacing,
                    gn_iterations, pcg_iterations, 10.0, seed=seed)


# (dims, voxel size) pairs of the batched form: 128^3 to 320^3 at 4-8 mm,
# each about 1.0-1.3 m across so that the body fits.
BATCH_VOLUMES = ((128, 0.008), (160, 0.0065), (192, 0.0055), (224, 0.005), (256, 0.004), (288, 0.004),
                 (320, 0.004))


def batch_workloads(count: int = 16, frames: int = 40, seed: int = 0, nodes_lo: int = 500, nodes_hi: int = 8000,
                    gn_iterations: int = 5, pcg_iterations: int = 30) -> list:
    """Uneven capsule-body sequences for the batched form: node counts
    log-uniform in [nodes_lo, nodes_hi], volumes from BATCH_VOLUMES, 640x480
    or 320x240 depth; instance k's sequence (joint walk, noise) uses seed k."""
    rng = np.random.default_rng(seed + 4242)
    out = []
    for k in range(count):
        n = int(round(math.exp(rng.uniform(math.log(nodes_lo), math.log(nodes_hi)))))
        dims, vs = BATCH_VOLUMES[int(rng.integers(len(BATCH_VOLUMES)))]
        width = 640 if rng.uniform() < 0.5 else 320
        out.append(body_instance_workload(n, dims, vs, width, frames, seed=k, gn_iterations=gn_iterations,
                                          pcg_iterations=pcg_iterations))
    return out


# ------------------------------------------------------------- primitives

def pixel_rays(w: Workload) -> np.ndarray:
    """Unnormalised ray directions through integer pixel coordinates, z = 1."""
    u, v = np.meshgrid(np.arange(w.width, dtype=np.float64), np.arange(w.height, dtype=np.float64))
    d = np.stack([(u - w.cx) / w.fx, (v - w.cy) / w.fy, np.ones_like(u)], axis=-1)
    return d.reshape(-1, 3)


def add_depth_noise(z: np.ndarray, rng: np.random.Generator) -> np.ndarray:
    out = z.copy()
    m = z > 0
    out[m] = z[m] + rng.normal(0.0, 1.0, size=int(m.sum())) * (0.0012 * z[m] ** 2)
    out[out < 0] = 0
    return out


def backproject(w: Workload, depth: np.ndarray) -> np.ndarray:
    d = pixel_rays(w)
    return d * depth.reshape(-1, 1)


@dataclass
class Frame:
    depth: np.ndarray       # (H, W) float32, noisy, 0 = no measurement
    canonical: np.ndarray   # (H*W, 3) float32 canonical point per pixel
    dst: np.ndarray         # (H*W, 3) float32 back-projected noisy depth
    normal: np.ndarray      # (H*W, 3) float32 true live surface normal
    valid: np.ndarray       # (H*W,) uint8


# ---------------------------------------------------------- sphere scene

class SphereScene:
    def __init__(self, w: Workload | None = None, radius=0.25, center=(0.0, 0.0, 1.0), amp=0.01,
                 wavelength=0.3, phase_step=0.5):
        self.w = w or sphere_workload()
        self.r = radius
        self.c = np.array(center, np.float64)
        self.amp = amp
        self.k = 2 * math.pi / wavelength
        self.phase_step = phase_step
        self.rng = np.random.default_rng(self.w.seed)

    def shift(self, y, f):
        return self.amp * np.sin(self.k * y + self.phase_step * f)

    def dshift(self, y, f):
        return self.amp * self.k * np.cos(self.k * y + self.phase_step * f)

    def _g(self, q, f):
        u = q.copy()
        u[:, 0] -= self.shift(q[:, 1], f)
        return np.linalg.norm(u - self.c, axis=1) - self.r

    def render(self, f: int) -> Frame:
        w = self.w
        d = pixel_rays(w)
        # initial guess: plain sphere intersection (camera at the origin)
        a = (d * d).sum(1)
        b = -2 * (d @ self.c)
        cc = self.c @ self.c - self.r ** 2
        disc = b * b - 4 * a * cc
        hit = disc > 0
        t = np.where(hit, (-b - np.sqrt(np.maximum(disc, 0))) / (2 * a), 0.0)
        for _ in range(12):  # Newton on g(t d) = 0
            q = d * t[:, None]
            g = self._g(q, f)
            eps = 1e-6
            g2 = self._g(d * (t + eps)[:, None], f)
            dg = (g2 - g) / eps
            step = np.where(np.abs(dg) > 1e-9, g / np.where(np.abs(dg) > 1e-9, dg, 1.0), 0.0)
            t = t - step
        q = d * t[:, None]
        ok = hit & (np.abs(self._g(q, f)) < 1e-7) & (t > 0)
        u = q.copy()
        u[:, 0] -= self.shift(q[:, 1], f)
        un = (u - self.c) / np.linalg.norm(u - self.c, axis=1, keepdims=True)
        grad = un.copy()
        grad[:, 1] -= un[:, 0] * self.dshift(q[:, 1], f)
        n = grad / np.linalg.norm(grad, axis=1, keepdims=True)
        z = np.where(ok, q[:, 2], 0.0)
        noisy = add_depth_noise(z, self.rng)
        canon = q.copy()  # canonical = frame 0 surface
        canon[:, 0] = q[:, 0] - self.shift(q[:, 1], f) + self.shift(q[:, 1], 0)
        dst = backproject(w, noisy)
        valid = ok & (noisy > 0)
        return Frame(noisy.reshape(w.height, w.width).astype(np.float32), canon.astype(np.float32),
                     dst.astype(np.float32), n.astype(np.float32), valid.astype(np.uint8))

    def surface_samples(self, step: float):
        """Points and outward normals on the canonical (frame 0) surface facing the camera."""
        n_t = max(8, int(math.pi * self.r / step * 2))
        pts = []
        th = np.linspace(0, math.pi, n_t)
        for t in th:
            ring = max(1, int(2 * math.pi * self.r * math.sin(t) / step * 2))
            ph = np.linspace(0, 2 * math.pi, ring, endpoint=False)
            p = np.stack([np.sin(t) * np.cos(ph), np.sin(t) * np.sin(ph), np.full_like(ph, np.cos(t))], 1)
            pts.append(p)
        un = np.concatenate(pts)
        un = un[un[:, 2] <= 0.05]  # the side facing the camera (-z)
        p = self.c + self.r * un
        p[:, 0] += self.shift(p[:, 1], 0)
        grad = un.copy()
        grad[:, 1] -= un[:, 0] * self.dshift(p[:, 1], 0)
        n = grad / np.linalg.norm(grad, axis=1, keepdims=True)
        return p, n


# ------------------------------------------------------------ body scene

def _rot_z(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


def _rot_x(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])


def ray_capsules(d: np.ndarray, caps) -> tuple:
    """Nearest hit t along unit rays d from the origin over capsules (a, b, r)."""
    n = d.shape[0]
    best = np.full(n, np.inf)
    who = np.full(n, -1, np.int32)
    for k, (a, b, r) in enumerate(caps):
        ba = b - a
        oa = -a
        baba = ba @ ba
        bard = d @ ba
        baoa = ba @ oa
        rdoa = d @ oa
        oaoa = oa @ oa
        A = baba - bard * bard
        B = baba * rdoa - baoa * bard
        Cc = baba * oaoa - baoa * baoa - r * r * baba
        h = B * B - A * Cc
        t = np.full(n, np.inf)
        m = (h >= 0) & (np.abs(A) > 1e-12)
        tb = np.where(m, (-B - np.sqrt(np.maximum(h, 0))) / np.where(np.abs(A) > 1e-12, A, 1.0), np.inf)
        y = baoa + tb * bard
        body = m & (y > 0) & (y < baba) & (tb > 0)
        t = np.where(body, tb, t)
        for end in (a, b):  # spherical caps
            oc = -end
            b2 = d @ oc
            c2 = oc @ oc - r * r
            h2 = b2 * b2 - c2
            tc = np.where(h2 > 0, -b2 - np.sqrt(np.maximum(h2, 0)), np.inf)
            tc = np.where(tc > 0, tc, np.inf)
            t = np.minimum(t, tc)
        closer = t < best
        best = np.where(closer, t, best)
        who = np.where(closer, k, who)
    return best, who


class BodyScene:
    """Articulated capsule body; segment 0 = torso, 1/2 = right upper/fore arm, 3/4 = left."""

    def __init__(self, w: Workload | None = None):
        self.w = w or body_workload()
        self.rng = np.random.default_rng(self.w.seed)
        self.torso = (np.array([0.0, -0.25, 1.2]), np.array([0.0, 0.25, 1.2]), 0.15)
        self.arm_r = 0.04
        self.seg = 0.3
        # joint angles per side: theta1 (down from horizontal), psi (about x), theta2 (elbow bend)
        deg = math.pi / 180
        ang0 = np.array([[60 * deg, 0.0, 15 * deg], [60 * deg, 0.0, 15 * deg]])
        lo = np.array([45 * deg, -20 * deg, 0.0])
        hi = np.array([80 * deg, 20 * deg, 40 * deg])
        self.angles = [ang0.copy()]
        a = ang0.copy()
        for _ in range(1, self.w.frames + 1):
            a = np.clip(a + self.rng.uniform(-3 * deg, 3 * deg, size=a.shape), lo, hi)
            self.angles.append(a.copy())
        self.noise_rng = np.random.default_rng(self.w.seed + 7919)

    def _arm(self, side: int, ang):
        s = 1.0 if side == 0 else -1.0
        S = np.array([s * 0.17, -0.2, 1.2])
        R1 = _rot_x(ang[1]) @ _rot_z(s * ang[0])
        R2 = R1 @ _rot_z(s * ang[2])
        ex = np.array([s, 0.0, 0.0])
        E = S + self.seg * (R1 @ ex)
        H = E + self.seg * (R2 @ ex)
        return S, E, H, R1, R2

    def capsules(self, f: int):
        caps = [self.torso]
        for side in (0, 1):
            S, E, H, _, _ = self._arm(side, self.angles[f][side])
            caps.append((S, E, self.arm_r))
            caps.append((E, H, self.arm_r))
        return caps

    def segment_transforms(self, f: int):
        """(R, t) per segment mapping canonical (frame 0) points to frame f."""
        out = [(np.eye(3), np.zeros(3))]
        for side in (0, 1):
            S0, E0, _, R10, R20 = self._arm(side, self.angles[0][side])
            S, E, _, R1, R2 = self._arm(side, self.angles[f][side])
            Ru = R1 @ R10.T
            out.append((Ru, S - Ru @ S0))
            Rf = R2 @ R20.T
            out.append((Rf, E - Rf @ E0))
        return out

    def render(self, f: int) -> Frame:
        w = self.w
        d = pixel_rays(w)
        dn = d / np.linalg.norm(d, axis=1, keepdims=True)
        caps = self.capsules(f)
        t, who = ray_capsules(dn, caps)
        ok = np.isfinite(t)
        q = dn * np.where(ok, t, 0.0)[:, None]
        n = np.zeros_like(q)
        canon = np.zeros_like(q)
        xf = self.segment_transforms(f)
        for k, (a, b, r) in enumerate(caps):
            m = ok & (who == k)
            if not m.any():
                continue
            ba = b - a
            h = np.clip(((q[m] - a) @ ba) / (ba @ ba), 0, 1)
            c = a + h[:, None] * ba
            nn = q[m] - c
            n[m] = nn / np.linalg.norm(nn, axis=1, keepdims=True)
            R, tr = xf[k]
            canon[m] = (q[m] - tr) @ R  # R^T (q - t)
        z = np.where(ok, q[:, 2], 0.0)
        noisy = add_depth_noise(z, self.noise_rng)
        dst = backproject(w, noisy)
        valid = ok & (noisy > 0)
        return Frame(noisy.reshape(w.height, w.width).astype(np.float32), canon.astype(np.float32),
                     dst.astype(np.float32), n.astype(np.float32), valid.astype(np.uint8))

    def surface_samples(self, step: float):
        """Points and outward normals on the whole canonical body surface."""
        caps = self.capsules(0)
        pts, nrm = [], []
        for (a, b, r) in caps:
            ax = (b - a) / np.linalg.norm(b - a)
            tmp = np.array([1.0, 0, 0]) if abs(ax[0]) < 0.9 else np.array([0, 1.0, 0])
            e1 = np.cross(ax, tmp)
            e1 /= np.linalg.norm(e1)
            e2 = np.cross(ax, e1)
            L = np.linalg.norm(b - a)
            nphi = max(8, int(2 * math.pi * r / step))
            phi = np.linspace(0, 2 * math.pi, nphi, endpoint=False)
            for hgt in np.linspace(0, L, max(2, int(L / step) + 1)):
                nn = np.cos(phi)[:, None] * e1 + np.sin(phi)[:, None] * e2
                pts.append(a + hgt * ax + r * nn)
                nrm.append(nn)
            for end, sgn in ((a, -1.0), (b, 1.0)):
                nth = max(2, int(0.5 * math.pi * r / step))
                for th in np.linspace(0, 0.5 * math.pi, nth + 1)[1:]:
                    ring = max(1, int(2 * math.pi * r * math.cos(th) / step))
                    ph = np.linspace(0, 2 * math.pi, ring, endpoint=False)
                    nn = (math.cos(th) * (np.cos(ph)[:, None] * e1 + np.sin(ph)[:, None] * e2)
                          + sgn * math.sin(th) * ax)
                    pts.append(end + r * nn)
                    nrm.append(nn)
        p = np.concatenate(pts)
        n = np.concatenate(nrm)
        inside = np.zeros(len(p), bool)
        for (a, b, r) in caps:  # drop samples buried in another capsule
            ba = b - a
            h = np.clip(((p - a) @ ba) / (ba @ ba), 0, 1)
            dist = np.linalg.norm(p - (a + h[:, None] * ba), axis=1)
            inside |= dist < r - 1e-4
        return p[~inside], n[~inside]


# --------------------------------------------------------- nodes / graph

def sample_nodes(points: np.ndarray, normals: np.ndarray, spacing: float, radius: float, origin, voxel_size,
                 max_nodes: int = 8192, graph_dist: float | None = None, min_normal_dot: float = 0.70):
    """Greedy Poisson-disk node sampling and the K_GRAPH = 8 node graph.

    Candidates are visited in volume scan order (z, y, x of their voxel), a
    candidate becomes a node when no node lies within ``spacing``
    (add_nodes_from_surface, src/warp_field.cu:709-841).  Each node then gets
    up to 8 nearest neighbours within max(3.5 spacing, 1.5 radius) whose
    normals agree (dot >= 0.7), weighted exp(-d^2 / 2 r^2) - the edge rule
    of add_nodes_from_mesh_geodesic (:843-955) with Euclidean distance.
    """
    org = np.asarray(origin, np.float64)
    vox = np.floor((points - org) / voxel_size).astype(np.int64)
    order = np.lexsort((vox[:, 0], vox[:, 1], vox[:, 2]))
    inv = 1.0 / spacing
    grid: dict = {}
    chosen = []
    sp2 = spacing * spacing
    for i in order:
        p = points[i]
        c = tuple(np.floor(p * inv).astype(np.int64))
        close = False
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for j in grid.get((c[0] + dx, c[1] + dy, c[2] + dz), ()):
                        q = points[j]
                        if (p[0] - q[0]) ** 2 + (p[1] - q[1]) ** 2 + (p[2] - q[2]) ** 2 < sp2:
                            close = True
                            break
                    if close:
                        break
                if close:
                    break
            if close:
                break
        if close:
            continue
        grid.setdefault(c, []).append(i)
        chosen.append(i)
        if len(chosen) >= max_nodes:
            break
    P = points[chosen]
    Nn = normals[chosen]
    N = len(chosen)
    nodes = np.zeros(N, NODE)
    nodes["pos"] = P.astype(np.float32)
    nodes["normal"] = Nn.astype(np.float32)
    nodes["radius"] = np.float32(radius)
    nodes["neighbors"] = -1
    gd = graph_dist if graph_dist is not None else max(3.5 * spacing, 1.5 * radius)
    d2 = ((P[:, None, :] - P[None, :, :]) ** 2).sum(-1) if N <= 4096 else None
    for i in range(N):
        if d2 is not None:
            di = d2[i]
        else:
            di = ((P - P[i]) ** 2).sum(1)
        cand = np.nonzero(di <= gd * gd)[0]
        cand = cand[cand != i]
        cand = cand[(Nn[cand] @ Nn[i]) >= min_normal_dot]
        cand = cand[np.argsort(di[cand], kind="stable")][:8]
        k = len(cand)
        nodes["neighbors"][i, :k] = cand
        dist = np.sqrt(di[cand]).astype(np.float32)
        r = max(radius, 1e-6)
        nodes["neighbor_w"][i, :k] = np.exp(-(dist * dist) / (2 * r * r))
        nodes["num_neighbors"][i] = k
    return nodes


def scene_for(w: Workload):
    """The scene of a workload: capsule body for "capsule-body*", else sphere."""
    return BodyScene(w) if w.name.startswith("capsule-body") else SphereScene(w)


def make_scene(kind: str, frames: int | None = None, seed: int = 0):
    if kind == "sphere":
        w = sphere_workload(frames or 10, seed)
        return SphereScene(w)
    if kind == "body":
        w = body_workload(frames or 40, seed)
        return BodyScene(w)
    raise ValueError(kind)


def scene_nodes(scene) -> np.ndarray:
    w = scene.w
    p, n = scene.surface_samples(w.node_spacing / 4)
    return sample_nodes(p, n, w.node_spacing, w.node_spacing, w.origin, w.voxel_size)
