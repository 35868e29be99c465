// raycast.hip - TSDF raycast for gfx950 (the reference's visual check and
// its surface prediction before the mesh rasteriser).
//
// raycast_kernel (kernels/tsdf_kernels.cu:236-361) with
// sample_tsdf_trilinear (:14-64): per pixel, a ray from the camera centre
// through (u, v) (integer pixel coordinates, as the reference) marched from
// t = 0.05 m to 3.5 m in steps of 0.8 voxel (0.5 voxel once |tsdf| < 0.5),
// the first + to - crossing with the previous sample inside (0, 1)
// interpolated linearly; the normal from central differences of six
// trilinear samples one voxel apart; optionally warped through the 4-NN of
// the crossing's base voxel.  The trilinear sample reads the voxel at index k
// as the value at origin + k s (no half-voxel offset, as the reference) and
// fails when any of its 8 voxels is unobserved.
//
// Mapping: one lane per pixel on 64 x 4 tiles, so that a wave's rays march
// through neighbouring voxels together and the 8-voxel gathers of a step
// hit the same cache lines.  FP32 in the reference's operation order
// (compiled without contraction), so the oracle's FP32 restatement gives the
// same bits.
#include <algorithm>

#include "nrf_internal.h"
#include "nrf_math.h"

namespace nrf {
namespace {

constexpr int kTx = 64, kTy = 4;

struct RaycastArgs {
  const nrf_voxel *vox;
  int X, Y, Z;
  float ox, oy, oz, s;
  nrf_camera cam;
  float Twc[3][4];  // camera-to-world (camera_pose rows 0..2)
  const nrf_node *nodes;
  const DQ *T;
  int N;
  const int4 *knn;
  const float4 *knn_w;
  float *verts, *normals;
  int32_t *vidx;
};

// floor of a bounded float as an int (the reference's __float2int_rd)
__device__ inline int floor_int(float v) { return (int)fminf(fmaxf(floorf(v), -1073741824.f), 1073741824.f); }

// sample_tsdf_trilinear (kernels/tsdf_kernels.cu:14-64)
__device__ inline bool sample_trilinear(const RaycastArgs &a, float px, float py, float pz, float &tsdf, int &base) {
  const float gx = (px - a.ox) / a.s, gy = (py - a.oy) / a.s, gz = (pz - a.oz) / a.s;
  const int x0 = floor_int(gx), y0 = floor_int(gy), z0 = floor_int(gz);
  if (x0 < 0 || x0 >= a.X - 1 || y0 < 0 || y0 >= a.Y - 1 || z0 < 0 || z0 >= a.Z - 1) return false;
  const float fx = gx - (float)x0, fy = gy - (float)y0, fz = gz - (float)z0;
  const size_t plane = (size_t)a.X * a.Y;
  const size_t b0 = (size_t)z0 * plane + (size_t)y0 * a.X + x0;
  float accum = 0.0f;
#pragma unroll
  for (int dz = 0; dz <= 1; ++dz) {
    const float wz = dz ? fz : (1.0f - fz);
#pragma unroll
    for (int dy = 0; dy <= 1; ++dy) {
      const float wy = dy ? fy : (1.0f - fy);
#pragma unroll
      for (int dx = 0; dx <= 1; ++dx) {
        const float wx = dx ? fx : (1.0f - fx);
        const nrf_voxel &vx = a.vox[b0 + dz * plane + dy * a.X + dx];
        if (vx.weight <= 0.0f) return false;
        accum += wx * wy * wz * vx.tsdf;
      }
    }
  }
  tsdf = accum;
  base = (int)b0;
  return true;
}

__device__ inline float3 norm3v(float3 v) {
  const float n = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
  return n > 1e-9f ? f3(v.x / n, v.y / n, v.z / n) : v;
}

__global__ __launch_bounds__(kTx *kTy) void raycast_kernel(RaycastArgs a) {
  const int u = blockIdx.x * kTx + (threadIdx.x & (kTx - 1));
  const int v = blockIdx.y * kTy + threadIdx.x / kTx;
  const int W = a.cam.width, H = a.cam.height;
  if (u >= W || v >= H) return;
  const size_t px = (size_t)v * W + u;
  float3 out_v = f3(0.f, 0.f, 0.f), out_n = f3(0.f, 0.f, 0.f);
  int out_idx = -1;
  // CameraIntrinsics::unproject(u, v, 1) (include/types.h:157-161)
  const float3 dc = f3(1.0f * ((float)u - a.cam.cx) / a.cam.fx, 1.0f * ((float)v - a.cam.cy) / a.cam.fy, 1.0f);
  const float3 o = f3(a.Twc[0][0] * 0.f + a.Twc[0][1] * 0.f + a.Twc[0][2] * 0.f + a.Twc[0][3],
                      a.Twc[1][0] * 0.f + a.Twc[1][1] * 0.f + a.Twc[1][2] * 0.f + a.Twc[1][3],
                      a.Twc[2][0] * 0.f + a.Twc[2][1] * 0.f + a.Twc[2][2] * 0.f + a.Twc[2][3]);
  const float3 d = norm3v(f3(a.Twc[0][0] * dc.x + a.Twc[0][1] * dc.y + a.Twc[0][2] * dc.z,
                             a.Twc[1][0] * dc.x + a.Twc[1][1] * dc.y + a.Twc[1][2] * dc.z,
                             a.Twc[2][0] * dc.x + a.Twc[2][1] * dc.y + a.Twc[2][2] * dc.z));
  float t = 0.05f;
  const float t_max = 3.5f;
  float step = a.s * 0.8f;
  float tsdf_prev = 0.f, t_prev = t;
  bool have_prev = false;
  for (; t < t_max; t += step) {
    const float pxw = o.x + t * d.x, pyw = o.y + t * d.y, pzw = o.z + t * d.z;
    float cur = 1.0f;
    int vidx = -1;
    if (!sample_trilinear(a, pxw, pyw, pzw, cur, vidx)) {
      have_prev = false;
      continue;
    }
    if (have_prev && tsdf_prev > 0 && cur <= 0 && tsdf_prev < 1.0f) {
      const float ts = t - (t - t_prev) * cur / (cur - tsdf_prev);
      const float sx = o.x + ts * d.x, sy = o.y + ts * d.y, sz = o.z + ts * d.z;
      float g[6];
      int unused = -1;
      const bool ok = sample_trilinear(a, sx + a.s, sy, sz, g[0], unused) &&
                      sample_trilinear(a, sx - a.s, sy, sz, g[1], unused) &&
                      sample_trilinear(a, sx, sy + a.s, sz, g[2], unused) &&
                      sample_trilinear(a, sx, sy - a.s, sz, g[3], unused) &&
                      sample_trilinear(a, sx, sy, sz + a.s, g[4], unused) &&
                      sample_trilinear(a, sx, sy, sz - a.s, g[5], unused);
      if (!ok) break;
      float3 n = norm3v(f3(g[0] - g[1], g[2] - g[3], g[4] - g[5]));
      out_idx = vidx;
      float3 p = f3(sx, sy, sz);
      if (a.N > 0 && a.nodes) {
        const int4 i4 = a.knn[vidx];
        const float4 w4 = a.knn_w[vidx];
        const int ids[4] = {i4.x, i4.y, i4.z, i4.w};
        const float ws[4] = {w4.x, w4.y, w4.z, w4.w};
        // warp_point (include/se3_math.cuh:450-510): the blend of the
        // neighbours' centred transforms
        DQ b;
        b.r = make_float4(0.f, 0.f, 0.f, 0.f);
        b.d = make_float4(0.f, 0.f, 0.f, 0.f);
        float wsum = 0.f;
        for (int k = 0; k < 4; ++k) {
          const int id = ids[k];
          const float w = ws[k];
          if (id < 0 || id >= a.N || w < 1e-8f) continue;
          DQ q = dq_centered(a.T[id], f3(a.nodes[id].pos));
          if (qdot(q.r, b.r) < 0) {
            q.r.x = -q.r.x; q.r.y = -q.r.y; q.r.z = -q.r.z; q.r.w = -q.r.w;
            q.d.x = -q.d.x; q.d.y = -q.d.y; q.d.z = -q.d.z; q.d.w = -q.d.w;
          }
          b.r.x += w * q.r.x; b.r.y += w * q.r.y; b.r.z += w * q.r.z; b.r.w += w * q.r.w;
          b.d.x += w * q.d.x; b.d.y += w * q.d.y; b.d.z += w * q.d.z; b.d.w += w * q.d.w;
          wsum += w;
        }
        if (wsum >= 1e-8f) p = dq_transform(dq_normalize(b), p);
        n = dq_warp_normal(a.T, a.N, ids, ws, n);
      }
      out_v = p;
      out_n = n;
      break;
    }
    tsdf_prev = cur;
    have_prev = true;
    t_prev = t;
    step = fabsf(cur) < 0.5f ? a.s * 0.5f : a.s * 0.8f;
  }
  a.verts[3 * px] = out_v.x;
  a.verts[3 * px + 1] = out_v.y;
  a.verts[3 * px + 2] = out_v.z;
  a.normals[3 * px] = out_n.x;
  a.normals[3 * px + 1] = out_n.y;
  a.normals[3 * px + 2] = out_n.z;
  if (a.vidx) a.vidx[px] = out_idx;
}

}  // namespace
}  // namespace nrf

using namespace nrf;

extern "C" int nrf_raycast(const nrf_volume_params *vp, const nrf_voxel *voxels, const nrf_camera *cam,
                           const float camera_pose[16], const nrf_node *nodes, const nrf_dualquat *transforms,
                           int num_nodes, const int32_t *voxel_knn, const float *voxel_knn_w, float *vertices,
                           float *normals, int32_t *canonical_vidx, void *stream) {
  if (!volume_ok(vp) || !voxels || !cam || !camera_pose || !vertices || !normals || cam->width <= 0 ||
      cam->height <= 0 || num_nodes < 0)
    return NRF_ERR_INVALID_ARGUMENT;
  const bool warp = nodes && num_nodes > 0;
  if (warp && (!transforms || !voxel_knn || !voxel_knn_w)) return NRF_ERR_INVALID_ARGUMENT;
  RaycastArgs a;
  a.vox = voxels;
  a.X = vp->dims[0]; a.Y = vp->dims[1]; a.Z = vp->dims[2];
  a.ox = vp->origin[0]; a.oy = vp->origin[1]; a.oz = vp->origin[2];
  a.s = vp->voxel_size;
  a.cam = *cam;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 4; ++j) a.Twc[i][j] = camera_pose[4 * i + j];
  a.nodes = warp ? nodes : nullptr;
  a.T = reinterpret_cast<const DQ *>(transforms);
  a.N = warp ? num_nodes : 0;
  a.knn = reinterpret_cast<const int4 *>(voxel_knn);
  a.knn_w = reinterpret_cast<const float4 *>(voxel_knn_w);
  a.verts = vertices;
  a.normals = normals;
  a.vidx = canonical_vidx;
  const dim3 grid((cam->width + kTx - 1) / kTx, (cam->height + kTy - 1) / kTy);
  raycast_kernel<<<grid, kTx * kTy, 0, as_stream(stream)>>>(a);
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}
