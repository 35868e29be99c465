// association.hip - projective data association for gfx950.
//
// The step before the solve: depth-map normals (kernels/tsdf_kernels.cu:
// 180-230), the 4-NN of each pixel's canonical hit voxel (:573-604), the
// projective point-to-plane correspondence search (:386-563) and the
// optimised-voxel counts that gate the warped integration (:606-651).
// FP32 in the reference's operation order; the search visits the window in
// the reference's row-major order with its strict < tie rule, so the chosen
// depth pixel is the reference's on the same arithmetic.
//
// Mapping: one lane per pixel on 64 x 4 tiles (x fastest, coalesced maps).
// The depth map (1.2 MB at 640x480) and its normals stay resident in the
// L2 / Infinity Cache for the window reads; counters are summed per wave
// (ballot + popcount) before one atomic per wave and reason, so the totals
// are exact and order independent.
#include "nrf_internal.h"
#include "nrf_math.h"

namespace nrf {
namespace {

constexpr int kTx = 64, kTy = 4;

__device__ inline float3 ld3(const float *p) { return make_float3(p[0], p[1], p[2]); }
__device__ inline void st3(float *p, float3 v) {
  p[0] = v.x;
  p[1] = v.y;
  p[2] = v.z;
}
__device__ inline float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline float norm3(float3 v) { return sqrtf(dot3(v, v)); }
// include/se3_math.cuh:28-33: left unchanged at (near) zero length
__device__ inline float3 normalize3(float3 v) {
  const float n = norm3(v);
  return n > 1e-9f ? make_float3(v.x / n, v.y / n, v.z / n) : v;
}
// CameraIntrinsics::unproject (include/types.h:157-161)
__device__ inline float3 unproject(const nrf_camera &c, float u, float v, float z) {
  return make_float3(z * (u - c.cx) / c.fx, z * (v - c.cy) / c.fy, z);
}
__device__ inline float3 xf_point(const CamXf &T, float3 p) {
  return make_float3(T.m[0][0] * p.x + T.m[0][1] * p.y + T.m[0][2] * p.z + T.m[0][3],
                     T.m[1][0] * p.x + T.m[1][1] * p.y + T.m[1][2] * p.z + T.m[1][3],
                     T.m[2][0] * p.x + T.m[2][1] * p.y + T.m[2][2] * p.z + T.m[2][3]);
}
__device__ inline float3 xf_normal(const CamXf &T, float3 n) {
  return make_float3(T.m[0][0] * n.x + T.m[0][1] * n.y + T.m[0][2] * n.z,
                     T.m[1][0] * n.x + T.m[1][1] * n.y + T.m[1][2] * n.z,
                     T.m[2][0] * n.x + T.m[2][1] * n.y + T.m[2][2] * n.z);
}
// one atomic per wave: the number of lanes with `flag`
__device__ inline void wave_count(int32_t *dst, bool flag) {
  const unsigned long long m = __ballot(flag);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(dst, (int)__popcll(m));
}

__global__ __launch_bounds__(kTx *kTy) void depth_normals_kernel(const float *__restrict__ depth, nrf_camera cam,
                                                                 float *__restrict__ normals) {
  const int u = blockIdx.x * kTx + (threadIdx.x & (kTx - 1));
  const int v = blockIdx.y * kTy + threadIdx.x / kTx;
  const int w = cam.width, h = cam.height;
  if (u >= w || v >= h) return;
  float3 n = make_float3(0.f, 0.f, 0.f);
  if (u < w - 1 && v < h - 1) {
    const float d00 = depth[v * w + u], d01 = depth[v * w + u + 1], d10 = depth[(v + 1) * w + u];
    if (!(d00 <= 0 || d01 <= 0 || d10 <= 0)) {
      const float3 p00 = unproject(cam, (float)u, (float)v, d00);
      const float3 p01 = unproject(cam, (float)(u + 1), (float)v, d01);
      const float3 p10 = unproject(cam, (float)u, (float)(v + 1), d10);
      const float3 c = normalize3(cross(sub(p01, p00), sub(p10, p00)));
      n = make_float3(-c.x, -c.y, -c.z);
    }
  }
  st3(normals + 3 * ((size_t)v * w + u), n);
}

__global__ __launch_bounds__(256) void pixel_knn_kernel(const int32_t *__restrict__ hit, int n,
                                                        const int4 *__restrict__ vknn, const float4 *__restrict__ vw,
                                                        int4 *__restrict__ pknn, float4 *__restrict__ pw) {
  const int px = blockIdx.x * 256 + threadIdx.x;
  if (px >= n) return;
  const int vidx = hit[px];
  pknn[px] = vidx < 0 ? make_int4(-1, -1, -1, -1) : vknn[vidx];
  pw[px] = vidx < 0 ? make_float4(0.f, 0.f, 0.f, 0.f) : vw[vidx];
}

struct AssocArgs {
  const float *live_v, *live_n, *depth, *dnorm;
  nrf_camera cam;
  CamXf Tcw, Twc;
  const int4 *pknn;
  const float4 *pw;
  nrf_assoc_params ap;
  nrf_corr *corrs;
  int32_t *num_valid, *stats;
  // the pixel 4-NN read through the hit voxel (the pixel kNN step fused;
  // used when pknn is null and hit is set)
  const int32_t *hit;
  const int4 *vknn;
  const float4 *vw;
  const int32_t *skip;  // optional: nonzero = leave every output as it is
};

__global__ __launch_bounds__(kTx *kTy) void find_corrs_kernel(AssocArgs a) {
  if (a.skip && *a.skip) return;
  const int u = blockIdx.x * kTx + (threadIdx.x & (kTx - 1));
  const int v = blockIdx.y * kTy + threadIdx.x / kTx;
  const int W = a.cam.width, H = a.cam.height;
  const bool inside = u < W && v < H;
  // reason: 0 invalid live, 1 out of projection, 2 no depth, 3 none close,
  // 4 angle / view, 5 valid, -1 outside the image
  int reason = -1;
  nrf_corr c;
  for (int k = 0; k < 3; ++k) c.src[k] = c.dst[k] = c.normal[k] = 0.f;
  for (int k = 0; k < 4; ++k) {
    c.node_ids[k] = -1;
    c.node_ws[k] = 0.f;
  }
  c.weight = 0.f;
  c.valid = 0;
  c.pad[0] = c.pad[1] = c.pad[2] = 0;
  const int px = v * W + u;
  if (inside) {
    do {
      const float3 vl = ld3(a.live_v + 3 * (size_t)px), nl = ld3(a.live_n + 3 * (size_t)px);
      if (norm3(vl) < 1e-6f || norm3(nl) < 1e-6f) { reason = 0; break; }
      const float3 vc = xf_point(a.Tcw, vl);
      if (vc.z <= 0) { reason = 1; break; }
      const float fu = a.cam.fx * vc.x / vc.z + a.cam.cx, fv = a.cam.fy * vc.y / vc.z + a.cam.cy;
      // round half to even; the range test only keeps the conversion defined
      if (!(fu > -1.0e9f && fu < 1.0e9f && fv > -1.0e9f && fv < 1.0e9f)) { reason = 1; break; }
      const int pu = (int)rintf(fu), pv = (int)rintf(fv);
      if (pu < 0 || pu >= W || pv < 0 || pv >= H) { reason = 1; break; }
      const float3 nc = xf_normal(a.Tcw, nl);
      float3 best_dst = make_float3(0.f, 0.f, 0.f), best_n = make_float3(0.f, 0.f, 0.f);
      float best_score = a.ap.dist_threshold, best_weight = 0.f;
      const float euclidean_cap = fmaxf(3.0f * a.ap.dist_threshold, a.ap.dist_threshold + 0.03f);
      const float normal_eps = fmaxf(1.0f - a.ap.angle_threshold, 1e-4f);
      const float view_eps = fmaxf(1.0f - a.ap.view_threshold, 1e-4f);
      const float3 view_dir = normalize3(make_float3(-vc.x, -vc.y, -vc.z));
      const float view_cos = fabsf(dot3(nc, view_dir));
      if (view_cos < a.ap.view_threshold) { reason = 4; break; }
      bool saw_depth = false, saw_close = false, saw_angle = false;
      const int R = max(0, a.ap.search_radius_px);
      for (int dy = -R; dy <= R; ++dy) {
        const int y = pv + dy;
        if (y < 0 || y >= H) continue;
        for (int dx = -R; dx <= R; ++dx) {
          const int x = pu + dx;
          if (x < 0 || x >= W) continue;
          const float dd = a.depth[y * W + x];
          if (dd <= 0) continue;
          saw_depth = true;
          const float3 vd = unproject(a.cam, (float)x, (float)y, dd);
          const float3 diff = sub(vc, vd);
          const float ed = norm3(diff);
          if (ed > euclidean_cap) continue;
          const float3 nd = ld3(a.dnorm + 3 * (size_t)(y * W + x));
          if (norm3(nd) < 1e-6f) continue;
          const float cosine = fabsf(dot3(nc, nd));
          if (cosine < a.ap.angle_threshold) continue;
          saw_angle = true;
          const float plane = fabsf(dot3(nd, diff));
          if (plane > a.ap.dist_threshold || plane >= best_score) continue;
          saw_close = true;
          const float phi_d = fmaxf(0.0f, 1.0f - ed / euclidean_cap);
          const float phi_n = fmaxf(0.0f, 1.0f - (1.0f - cosine) / normal_eps);
          const float phi_v = fmaxf(0.0f, 1.0f - (1.0f - view_cos) / view_eps);
          float conf = (phi_d + phi_n + phi_v) / 3.0f;
          conf = conf * conf;
          best_score = plane;
          best_weight = conf;
          best_dst = vd;
          best_n = nd;
        }
      }
      if (!saw_depth) { reason = 2; break; }
      if (!saw_close) { reason = 3; break; }
      if (!saw_angle) { reason = 4; break; }
      reason = 5;
      st3(c.src, vl);
      st3(c.dst, xf_point(a.Twc, best_dst));
      st3(c.normal, normalize3(xf_normal(a.Twc, best_n)));
      c.weight = best_weight;
      c.valid = 1;
      int4 id = make_int4(-1, -1, -1, -1);
      float4 wk = make_float4(0.f, 0.f, 0.f, 0.f);
      bool have = false;
      if (a.pknn) {
        id = a.pknn[px];
        wk = a.pw[px];
        have = true;
      } else if (a.hit) {
        const int vidx = a.hit[px];
        if (vidx >= 0) {
          id = a.vknn[vidx];
          wk = a.vw[vidx];
        }
        have = true;
      }
      if (have) {
        c.node_ids[0] = id.x; c.node_ids[1] = id.y; c.node_ids[2] = id.z; c.node_ids[3] = id.w;
        c.node_ws[0] = wk.x; c.node_ws[1] = wk.y; c.node_ws[2] = wk.z; c.node_ws[3] = wk.w;
      }
    } while (false);
    a.corrs[px] = c;
  }
  wave_count(a.num_valid, reason == 5);
  if (a.stats)
    for (int k = 0; k < 6; ++k) wave_count(a.stats + k, reason == k);
}

__global__ __launch_bounds__(256) void mark_opt_kernel(const nrf_corr *__restrict__ corrs,
                                                       const int32_t *__restrict__ hit, int n, int X, int Y, int Z,
                                                       int frame_id, int32_t *__restrict__ counts,
                                                       int32_t *__restrict__ last_frame) {
  const int px = blockIdx.x * 256 + threadIdx.x;
  if (px >= n) return;
  if (!corrs[px].valid || corrs[px].weight <= 0.0f) return;
  const int vidx = hit[px];
  if (vidx < 0) return;
  const int vx = vidx % X, vy = (vidx / X) % Y, vz = vidx / (X * Y);
  for (int dz = -1; dz <= 1; ++dz) {
    const int z = vz + dz;
    if (z < 0 || z >= Z) continue;
    for (int dy = -1; dy <= 1; ++dy) {
      const int y = vy + dy;
      if (y < 0 || y >= Y) continue;
      for (int dx = -1; dx <= 1; ++dx) {
        const int x = vx + dx;
        if (x < 0 || x >= X) continue;
        const int nidx = (z * Y + y) * X + x;
        if (atomicExch(&last_frame[nidx], frame_id) != frame_id) atomicAdd(&counts[nidx], 1);
      }
    }
  }
}

}  // namespace
}  // namespace nrf

using namespace nrf;

extern "C" int nrf_depth_normals(const float *depth, const nrf_camera *cam, float *normals, void *stream) {
  if (!depth || !cam || !normals || cam->width <= 0 || cam->height <= 0) return NRF_ERR_INVALID_ARGUMENT;
  const dim3 grid((cam->width + kTx - 1) / kTx, (cam->height + kTy - 1) / kTy);
  depth_normals_kernel<<<grid, kTx * kTy, 0, as_stream(stream)>>>(depth, *cam, normals);
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}

extern "C" int nrf_pixel_knn(const int32_t *hit_voxel, int num_pixels, const int32_t *voxel_knn,
                             const float *voxel_knn_w, int32_t *pixel_knn, float *pixel_knn_w, void *stream) {
  if (num_pixels < 0 || (num_pixels > 0 && (!hit_voxel || !voxel_knn || !voxel_knn_w || !pixel_knn || !pixel_knn_w)))
    return NRF_ERR_INVALID_ARGUMENT;
  if (num_pixels == 0) return NRF_OK;
  pixel_knn_kernel<<<(num_pixels + 255) / 256, 256, 0, as_stream(stream)>>>(
      hit_voxel, num_pixels, reinterpret_cast<const int4 *>(voxel_knn), reinterpret_cast<const float4 *>(voxel_knn_w),
      reinterpret_cast<int4 *>(pixel_knn), reinterpret_cast<float4 *>(pixel_knn_w));
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}

namespace nrf {
int find_correspondences(const float *live_vertices, const float *live_normals, const float *depth,
                         const float *depth_normals, const nrf_camera *cam, const float camera_pose[16],
                         const int32_t *pixel_knn, const float *pixel_knn_w, const int32_t *hit_voxel,
                         const int32_t *voxel_knn, const float *voxel_knn_w, const nrf_assoc_params *params,
                         nrf_corr *corrs, int32_t *num_valid, int32_t *stats, const int32_t *skip, hipStream_t st) {
  if (!live_vertices || !live_normals || !depth || !depth_normals || !cam || !camera_pose || !params || !corrs ||
      !num_valid || cam->width <= 0 || cam->height <= 0 || (!pixel_knn != !pixel_knn_w) ||
      (hit_voxel && (!voxel_knn || !voxel_knn_w)))
    return NRF_ERR_INVALID_ARGUMENT;
  AssocArgs a;
  a.live_v = live_vertices;
  a.live_n = live_normals;
  a.depth = depth;
  a.dnorm = depth_normals;
  a.cam = *cam;
  a.Tcw = invert_pose(camera_pose);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 4; ++j) a.Twc.m[i][j] = camera_pose[i * 4 + j];
  a.pknn = reinterpret_cast<const int4 *>(pixel_knn);
  a.pw = reinterpret_cast<const float4 *>(pixel_knn_w);
  a.ap = *params;
  a.corrs = corrs;
  a.num_valid = num_valid;
  a.stats = stats;
  a.hit = pixel_knn ? nullptr : hit_voxel;
  a.vknn = reinterpret_cast<const int4 *>(voxel_knn);
  a.vw = reinterpret_cast<const float4 *>(voxel_knn_w);
  a.skip = skip;
  const dim3 grid((cam->width + kTx - 1) / kTx, (cam->height + kTy - 1) / kTy);
  find_corrs_kernel<<<grid, kTx * kTy, 0, st>>>(a);
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}
}  // namespace nrf

extern "C" int nrf_find_correspondences(const float *live_vertices, const float *live_normals, const float *depth,
                                        const float *depth_normals, const nrf_camera *cam,
                                        const float camera_pose[16], const int32_t *pixel_knn,
                                        const float *pixel_knn_w, const nrf_assoc_params *params, nrf_corr *corrs,
                                        int32_t *num_valid, int32_t *stats, void *stream) {
  return find_correspondences(live_vertices, live_normals, depth, depth_normals, cam, camera_pose, pixel_knn,
                              pixel_knn_w, nullptr, nullptr, nullptr, params, corrs, num_valid, stats, nullptr,
                              as_stream(stream));
}

extern "C" int nrf_mark_optimized_voxels(const nrf_corr *corrs, const int32_t *hit_voxel, int num_pixels,
                                         const nrf_volume_params *vp, int frame_id, int32_t *voxel_opt_counts,
                                         int32_t *voxel_opt_frame, void *stream) {
  if (!volume_ok(vp) || num_pixels < 0 ||
      (num_pixels > 0 && (!corrs || !hit_voxel || !voxel_opt_counts || !voxel_opt_frame)))
    return NRF_ERR_INVALID_ARGUMENT;
  if (num_pixels == 0) return NRF_OK;
  mark_opt_kernel<<<(num_pixels + 255) / 256, 256, 0, as_stream(stream)>>>(
      corrs, hit_voxel, num_pixels, vp->dims[0], vp->dims[1], vp->dims[2], frame_id, voxel_opt_counts,
      voxel_opt_frame);
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}
