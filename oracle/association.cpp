// association.cpp - CPU restatement of the reference's projective data
// association (the step that produces the solver's Correspondence array and
// the integration's optimised-voxel counts).
//
// TEST INFRASTRUCTURE ONLY (linked into liboracle.so; never into the
// product library).  FP32 throughout, in the reference's operation order
// (one CUDA thread run sequentially, atomics replaced by in-order updates
// whose result does not depend on the order):
//   depth normals        kernels/tsdf_kernels.cu:180-230 (K15)
//   pixel kNN            kernels/tsdf_kernels.cu:573-604 (K18)
//   correspondences      kernels/tsdf_kernels.cu:386-563 (K17)
//   optimised voxels     kernels/tsdf_kernels.cu:606-651 (K19)
// Vectors are float[3] arrays; the camera pose is the row-major 4x4
// camera-to-world matrix, inverted as include/fusion_math.h:80-95 does.
#include <cmath>
#include <cstdint>
#include <limits>
#include <cstring>
#include <vector>
#include <algorithm>

#include "../include/nrf.h"

namespace {

struct F3 {
  float x, y, z;
};
inline F3 f3(float x, float y, float z) { return {x, y, z}; }
inline F3 ld3(const float *p) { return {p[0], p[1], p[2]}; }
inline void st3(float *p, F3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }
inline F3 sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
// se3_math.cuh:12-33
inline F3 cross3(F3 a, F3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float dot3(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float norm3(F3 v) { return std::sqrt(dot3(v, v)); }
inline F3 normalize3(F3 v) {
  const float n = norm3(v);
  return n > 1e-9f ? F3{v.x / n, v.y / n, v.z / n} : v;
}
// CameraIntrinsics::unproject (include/types.h:157-161)
inline F3 unproject(const nrf_camera &c, float u, float v, float z) {
  return {z * (u - c.cx) / c.fx, z * (v - c.cy) / c.fy, z};
}
struct Rigid {
  float m[3][4];
  F3 point(F3 p) const {
    return {m[0][0] * p.x + m[0][1] * p.y + m[0][2] * p.z + m[0][3],
            m[1][0] * p.x + m[1][1] * p.y + m[1][2] * p.z + m[1][3],
            m[2][0] * p.x + m[2][1] * p.y + m[2][2] * p.z + m[2][3]};
  }
  F3 normal(F3 n) const {
    return {m[0][0] * n.x + m[0][1] * n.y + m[0][2] * n.z, m[1][0] * n.x + m[1][1] * n.y + m[1][2] * n.z,
            m[2][0] * n.x + m[2][1] * n.y + m[2][2] * n.z};
  }
};
Rigid from_pose(const float pose[16]) {
  Rigid r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 4; ++j) r.m[i][j] = pose[i * 4 + j];
  return r;
}
Rigid inverse_rigid(const float pose[16]) {
  Rigid r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.m[i][j] = pose[j * 4 + i];
  for (int i = 0; i < 3; ++i) r.m[i][3] = -(r.m[i][0] * pose[3] + r.m[i][1] * pose[7] + r.m[i][2] * pose[11]);
  return r;
}
// round half to even (__float2int_rn)
inline int rn(float v) { return (int)std::nearbyint(v); }

}  // namespace

extern "C" {

// compute_depth_normals_kernel (kernels/tsdf_kernels.cu:180-230)
void orc_depth_normals_f32(const float *depth, const nrf_camera *cam, float *normals) {
  const int w = cam->width, h = cam->height;
  for (int v = 0; v < h; ++v)
    for (int u = 0; u < w; ++u) {
      float *out = normals + 3 * ((long)v * w + u);
      st3(out, f3(0, 0, 0));
      if (u >= w - 1 || v >= h - 1) continue;
      const float d00 = depth[v * w + u], d01 = depth[v * w + u + 1], d10 = depth[(v + 1) * w + u];
      if (d00 <= 0 || d01 <= 0 || d10 <= 0) continue;
      const F3 p00 = unproject(*cam, (float)u, (float)v, d00);
      const F3 p01 = unproject(*cam, (float)(u + 1), (float)v, d01);
      const F3 p10 = unproject(*cam, (float)u, (float)(v + 1), d10);
      const F3 n = normalize3(cross3(sub(p01, p00), sub(p10, p00)));
      st3(out, f3(-n.x, -n.y, -n.z));
    }
}

// compute_pixel_knn_kernel (kernels/tsdf_kernels.cu:573-604)
void orc_pixel_knn(const int32_t *hit_voxel, int num_pixels, const int32_t *voxel_knn, const float *voxel_knn_w,
                   int32_t *pixel_knn, float *pixel_knn_w) {
  for (int px = 0; px < num_pixels; ++px) {
    const int vidx = hit_voxel[px];
    for (int k = 0; k < 4; ++k) {
      pixel_knn[px * 4 + k] = vidx < 0 ? -1 : voxel_knn[(long)vidx * 4 + k];
      pixel_knn_w[px * 4 + k] = vidx < 0 ? 0.f : voxel_knn_w[(long)vidx * 4 + k];
    }
  }
}

// find_correspondences_kernel (kernels/tsdf_kernels.cu:386-563).  Pixels
// without a correspondence get valid = 0, weight 0 and zero vectors, ids -1
// (the reference leaves those fields unset).  stats (6 ints, added to):
// invalid live, out of projection, no depth, none close, angle or view,
// valid.  Returns the valid count.
int orc_find_correspondences_f32(const float *live_v, const float *live_n, const float *depth, const float *dnorm,
                                 const nrf_camera *cam, const float pose[16], const int32_t *pixel_knn,
                                 const float *pixel_knn_w, const nrf_assoc_params *ap, nrf_corr *corrs,
                                 int32_t *stats) {
  const int W = cam->width, H = cam->height;
  const Rigid Tcw = inverse_rigid(pose), Twc = from_pose(pose);
  int valid = 0;
  int st[6] = {0, 0, 0, 0, 0, 0};
  for (int px = 0; px < W * H; ++px) {
    nrf_corr c;
    std::memset(&c, 0, sizeof(c));
    for (int k = 0; k < 4; ++k) c.node_ids[k] = -1;
    const F3 vl = ld3(live_v + 3 * px), nl = ld3(live_n + 3 * px);
    do {
      if (norm3(vl) < 1e-6f || norm3(nl) < 1e-6f) { st[0]++; break; }
      const F3 vc = Tcw.point(vl);
      if (vc.z <= 0) { st[1]++; break; }
      const float fu = cam->fx * vc.x / vc.z + cam->cx, fv = cam->fy * vc.y / vc.z + cam->cy;
      if (!(fu > -1e9f && fu < 1e9f && fv > -1e9f && fv < 1e9f)) { st[1]++; break; }
      const int pu = rn(fu), pv = rn(fv);
      if (pu < 0 || pu >= W || pv < 0 || pv >= H) { st[1]++; break; }
      const F3 nc = Tcw.normal(nl);
      F3 best_dst = f3(0, 0, 0), best_n = f3(0, 0, 0);
      float best_score = ap->dist_threshold, best_weight = 0.f;
      const float euclidean_cap = std::fmax(3.0f * ap->dist_threshold, ap->dist_threshold + 0.03f);
      const float normal_eps = std::fmax(1.0f - ap->angle_threshold, 1e-4f);
      const float view_eps = std::fmax(1.0f - ap->view_threshold, 1e-4f);
      const F3 view_dir = normalize3(f3(-vc.x, -vc.y, -vc.z));
      const float view_cos = std::fabs(dot3(nc, view_dir));
      if (view_cos < ap->view_threshold) { st[4]++; break; }
      bool saw_depth = false, saw_close = false, saw_angle = false;
      const int R = ap->search_radius_px > 0 ? ap->search_radius_px : 0;
      for (int dy = -R; dy <= R; ++dy) {
        const int y = pv + dy;
        if (y < 0 || y >= H) continue;
        for (int dx = -R; dx <= R; ++dx) {
          const int x = pu + dx;
          if (x < 0 || x >= W) continue;
          const float dd = depth[y * W + x];
          if (dd <= 0) continue;
          saw_depth = true;
          const F3 vd = unproject(*cam, (float)x, (float)y, dd);
          const F3 diff = sub(vc, vd);
          const float ed = norm3(diff);
          if (ed > euclidean_cap) continue;
          const F3 nd = ld3(dnorm + 3 * (y * W + x));
          if (norm3(nd) < 1e-6f) continue;
          const float cosine = std::fabs(dot3(nc, nd));
          if (cosine < ap->angle_threshold) continue;
          saw_angle = true;
          const float plane = std::fabs(dot3(nd, diff));
          if (plane > ap->dist_threshold || plane >= best_score) continue;
          saw_close = true;
          const float phi_d = std::fmax(0.0f, 1.0f - ed / euclidean_cap);
          const float phi_n = std::fmax(0.0f, 1.0f - (1.0f - cosine) / normal_eps);
          const float phi_v = std::fmax(0.0f, 1.0f - (1.0f - view_cos) / view_eps);
          float conf = (phi_d + phi_n + phi_v) / 3.0f;
          conf = conf * conf;
          best_score = plane;
          best_weight = conf;
          best_dst = vd;
          best_n = nd;
        }
      }
      if (!saw_depth) { st[2]++; break; }
      if (!saw_close) { st[3]++; break; }
      if (!saw_angle) { st[4]++; break; }
      st3(c.src, vl);
      st3(c.dst, Twc.point(best_dst));
      st3(c.normal, normalize3(Twc.normal(best_n)));
      c.weight = best_weight;
      c.valid = 1;
      for (int k = 0; k < 4; ++k) {
        c.node_ids[k] = pixel_knn ? pixel_knn[px * 4 + k] : -1;
        c.node_ws[k] = pixel_knn_w ? pixel_knn_w[px * 4 + k] : 0.f;
      }
      ++valid;
      st[5]++;
    } while (false);
    corrs[px] = c;
  }
  if (stats)
    for (int k = 0; k < 6; ++k) stats[k] += st[k];
  return valid;
}

// mark_optimized_voxels_kernel (kernels/tsdf_kernels.cu:606-651): each voxel
// of the 3x3x3 neighbourhood of a valid correspondence's hit voxel counts
// once per frame.
void orc_mark_optimized_voxels(const nrf_corr *corrs, const int32_t *hit_voxel, int num_pixels, const int32_t dims[3],
                               int frame_id, int32_t *opt_counts, int32_t *opt_frame) {
  for (int px = 0; px < num_pixels; ++px) {
    if (!corrs[px].valid || corrs[px].weight <= 0.0f) continue;
    const int vidx = hit_voxel[px];
    if (vidx < 0) continue;
    const int vx = vidx % dims[0], vy = (vidx / dims[0]) % dims[1], vz = vidx / (dims[0] * dims[1]);
    for (int dz = -1; dz <= 1; ++dz) {
      const int z = vz + dz;
      if (z < 0 || z >= dims[2]) continue;
      for (int dy = -1; dy <= 1; ++dy) {
        const int y = vy + dy;
        if (y < 0 || y >= dims[1]) continue;
        for (int dx = -1; dx <= 1; ++dx) {
          const int x = vx + dx;
          if (x < 0 || x >= dims[0]) continue;
          const long n = ((long)z * dims[1] + y) * dims[0] + x;
          const int prev = opt_frame[n];
          opt_frame[n] = frame_id;
          if (prev != frame_id) opt_counts[n] += 1;
        }
      }
    }
  }
}

// correspondence_rmse (include/fusion_app.h:1331-1351): over the valid slots
// with weight > 0, r = n . (src - dst), the difference in FP32 and the dot
// in FP64, summed in slot order in FP64; infinity without weight.
double orc_corr_rmse(const nrf_corr *corrs, int n) {
  double sum = 0.0, wsum = 0.0;
  for (int i = 0; i < n; ++i) {
    const nrf_corr &c = corrs[i];
    if (!c.valid || c.weight <= 0.0f) continue;
    const float dx = c.src[0] - c.dst[0], dy = c.src[1] - c.dst[1], dz = c.src[2] - c.dst[2];
    const double r = (double)c.normal[0] * dx + (double)c.normal[1] * dy + (double)c.normal[2] * dz;
    const double w = std::max(0.0f, c.weight);
    sum += w * r * r;
    wsum += w;
  }
  return wsum > 0.0 ? std::sqrt(sum / wsum) : std::numeric_limits<double>::infinity();
}

}  // extern "C"

// ------------------------------------------------- live surface prediction
namespace {
inline float edge_fn(float ax, float ay, float bx, float by, float px, float py) {
  return (px - ax) * (by - ay) - (py - ay) * (bx - ax);
}
inline int sat_int(float v) { return (int)std::fmin(std::fmax(v, -1073741824.f), 1073741824.f); }
inline int voxel_of(F3 p, const nrf_volume_params &vp) {
  const int vx = sat_int(std::floor((p.x - vp.origin[0]) / vp.voxel_size));
  const int vy = sat_int(std::floor((p.y - vp.origin[1]) / vp.voxel_size));
  const int vz = sat_int(std::floor((p.z - vp.origin[2]) / vp.voxel_size));
  if (vx < 0 || vx >= vp.dims[0] || vy < 0 || vy >= vp.dims[1] || vz < 0 || vz >= vp.dims[2]) return -1;
  return (vz * vp.dims[1] + vy) * vp.dims[0] + vx;
}
// project_triangle (src/mesh_raster_kernels.cu:44-72)
bool project_tri(const float *warped, const int ids[3], const nrf_camera &cam, const Rigid &Tcw, F3 wc[3], F3 cc[3],
                 float px[3], float py[3], float &area) {
  for (int k = 0; k < 3; ++k) {
    wc[k] = ld3(warped + 3 * (long)ids[k]);
    cc[k] = Tcw.point(wc[k]);
  }
  if (cc[0].z <= 0.0f || cc[1].z <= 0.0f || cc[2].z <= 0.0f) return false;
  for (int k = 0; k < 3; ++k) {
    px[k] = cam.fx * cc[k].x / cc[k].z + cam.cx;
    py[k] = cam.fy * cc[k].y / cc[k].z + cam.cy;
  }
  area = edge_fn(px[0], py[0], px[1], py[1], px[2], py[2]);
  return std::fabs(area) >= 1e-6f;
}
}  // namespace

extern "C" {

// rasterize_triangles_z_kernel + resolve_rasterized_triangles_kernel
// (src/mesh_raster_kernels.cu:76-200) with the caller's resets: per pixel
// the nearest covering triangle (then the lowest index), its interpolated
// live vertex, face normal, canonical point and that point's voxel.
void orc_rasterize_f32(const float *canon, const float *warped, const int32_t *tris, int nv, int nt,
                       const nrf_camera *cam, const float pose[16], const nrf_volume_params *vp, float *live_v,
                       float *live_n, int32_t *hit, float *canon_pt) {
  const int W = cam->width, H = cam->height;
  const Rigid Tcw = inverse_rigid(pose);
  std::vector<unsigned long long> zt((size_t)W * H, ~0ull);
  auto ids_of = [&](int t, int ids[3]) {
    for (int k = 0; k < 3; ++k) ids[k] = tris[3 * (long)t + k];
    return !(ids[0] < 0 || ids[1] < 0 || ids[2] < 0 || ids[0] >= nv || ids[1] >= nv || ids[2] >= nv);
  };
  for (int t = 0; t < nt; ++t) {
    int ids[3];
    if (!ids_of(t, ids)) continue;
    F3 wc[3], cc[3];
    float px[3], py[3], area = 0;
    if (!project_tri(warped, ids, *cam, Tcw, wc, cc, px, py, area)) continue;
    const int min_u = std::max(0, sat_int(std::floor(std::fmin(px[0], std::fmin(px[1], px[2])))));
    const int max_u = std::min(W - 1, sat_int(std::ceil(std::fmax(px[0], std::fmax(px[1], px[2])))));
    const int min_v = std::max(0, sat_int(std::floor(std::fmin(py[0], std::fmin(py[1], py[2])))));
    const int max_v = std::min(H - 1, sat_int(std::ceil(std::fmax(py[0], std::fmax(py[1], py[2])))));
    if (min_u > max_u || min_v > max_v) continue;
    const F3 fn = normalize3(cross3(sub(wc[1], wc[0]), sub(wc[2], wc[0])));
    if (norm3(fn) < 1e-6f) continue;
    for (int y = min_v; y <= max_v; ++y)
      for (int x = min_u; x <= max_u; ++x) {
        const float qx = (float)x + 0.5f, qy = (float)y + 0.5f;
        const float w0 = edge_fn(px[1], py[1], px[2], py[2], qx, qy) / area;
        const float w1 = edge_fn(px[2], py[2], px[0], py[0], qx, qy) / area;
        const float w2 = edge_fn(px[0], py[0], px[1], py[1], qx, qy) / area;
        if (w0 < -1e-5f || w1 < -1e-5f || w2 < -1e-5f) continue;
        const float z = w0 * cc[0].z + w1 * cc[1].z + w2 * cc[2].z;
        if (z <= 0.0f || !std::isfinite(z)) continue;
        uint32_t zb;
        std::memcpy(&zb, &z, 4);
        const unsigned long long key = ((unsigned long long)zb << 32) | (unsigned)t;
        unsigned long long &slot = zt[(size_t)y * W + x];
        if (key < slot) slot = key;
      }
  }
  for (int p = 0; p < W * H; ++p) {
    st3(live_v + 3 * (long)p, f3(0, 0, 0));
    st3(live_n + 3 * (long)p, f3(0, 0, 0));
    if (canon_pt) st3(canon_pt + 3 * (long)p, f3(0, 0, 0));
    hit[p] = -1;
    if (zt[p] == ~0ull) continue;
    const int t = (int)(zt[p] & 0xffffffffull);
    int ids[3];
    if (!ids_of(t, ids)) continue;
    F3 wc[3], cc[3];
    float px[3], py[3], area = 0;
    if (!project_tri(warped, ids, *cam, Tcw, wc, cc, px, py, area)) continue;
    const float qx = (float)(p % W) + 0.5f, qy = (float)(p / W) + 0.5f;
    const float w0 = edge_fn(px[1], py[1], px[2], py[2], qx, qy) / area;
    const float w1 = edge_fn(px[2], py[2], px[0], py[0], qx, qy) / area;
    const float w2 = edge_fn(px[0], py[0], px[1], py[1], qx, qy) / area;
    if (w0 < -1e-5f || w1 < -1e-5f || w2 < -1e-5f) continue;
    st3(live_v + 3 * (long)p, f3(w0 * wc[0].x + w1 * wc[1].x + w2 * wc[2].x, w0 * wc[0].y + w1 * wc[1].y + w2 * wc[2].y,
                                w0 * wc[0].z + w1 * wc[1].z + w2 * wc[2].z));
    st3(live_n + 3 * (long)p, normalize3(cross3(sub(wc[1], wc[0]), sub(wc[2], wc[0]))));
    const F3 c0 = ld3(canon + 3 * (long)ids[0]), c1 = ld3(canon + 3 * (long)ids[1]), c2 = ld3(canon + 3 * (long)ids[2]);
    const F3 cp = f3(w0 * c0.x + w1 * c1.x + w2 * c2.x, w0 * c0.y + w1 * c1.y + w2 * c2.y,
                     w0 * c0.z + w1 * c1.z + w2 * c2.z);
    if (canon_pt) st3(canon_pt + 3 * (long)p, cp);
    hit[p] = voxel_of(cp, *vp);
  }
}

}  // extern "C"
