// warp.hip - warp-field kernels for gfx950: voxel 4-NN snapshot, twist
// update, the per-node centred-DQ table and point warping.
#include "nrf_internal.h"
#include "nrf_math.h"

namespace nrf {
namespace {

constexpr int kKnnBlock = 256;
constexpr int kKnnTile = 1024;  // node positions staged per LDS tile (12 KiB)

// compute_voxel_knn_kernel (src/warp_field.cu:299-405): for voxels with
// weight > 0, the four nearest nodes by squared distance from the voxel
// centre, visited in index order with a strict '<' insertion so that ties
// keep the lower index; weights exp(-d^2 / 2r^2) normalised to sum 1.
// Node positions are staged through LDS in tiles shared by the workgroup;
// workgroups whose 256 voxels are all unobserved skip the node scan.
__global__ __launch_bounds__(kKnnBlock) void knn_kernel(const nrf_voxel *__restrict__ vox, int X, int Y, int Z,
                                                         float3 org, float s, const nrf_node *__restrict__ nodes,
                                                         int N, int4 *__restrict__ knn, float4 *__restrict__ knn_w) {
  __shared__ float sx[kKnnTile], sy[kKnnTile], sz[kKnnTile];
  const int64_t V = (int64_t)X * Y * Z;
  const int64_t idx = (int64_t)blockIdx.x * kKnnBlock + threadIdx.x;
  const bool inb = idx < V;
  const bool obs = inb && N > 0 && vox[idx].weight > 0;
  float bd0 = 1e30f, bd1 = 1e30f, bd2 = 1e30f, bd3 = 1e30f;
  int bi0 = -1, bi1 = -1, bi2 = -1, bi3 = -1;
  float3 p = f3(0.f, 0.f, 0.f);
  if (inb) {
    const int x = (int)(idx % X);
    const int y = (int)((idx / X) % Y);
    const int z = (int)(idx / ((int64_t)X * Y));
    p = f3(org.x + (x + 0.5f) * s, org.y + (y + 0.5f) * s, org.z + (z + 0.5f) * s);
  }
  if (__syncthreads_or(obs)) {
    for (int t0 = 0; t0 < N; t0 += kKnnTile) {
      const int tn = min(kKnnTile, N - t0);
      for (int j = threadIdx.x; j < tn; j += kKnnBlock) {
        const nrf_node &n = nodes[t0 + j];
        sx[j] = n.pos[0];
        sy[j] = n.pos[1];
        sz[j] = n.pos[2];
      }
      __syncthreads();
      if (obs) {
        for (int j = 0; j < tn; ++j) {
          const float dx = p.x - sx[j], dy = p.y - sy[j], dz = p.z - sz[j];
          const float d2 = dx * dx + dy * dy + dz * dz;
          if (d2 < bd3) {
            bd3 = d2; bi3 = t0 + j;
            if (bd3 < bd2) { float td = bd2; bd2 = bd3; bd3 = td; int ti = bi2; bi2 = bi3; bi3 = ti; }
            if (bd2 < bd1) { float td = bd1; bd1 = bd2; bd2 = td; int ti = bi1; bi1 = bi2; bi2 = ti; }
            if (bd1 < bd0) { float td = bd0; bd0 = bd1; bd1 = td; int ti = bi0; bi0 = bi1; bi1 = ti; }
          }
        }
      }
      __syncthreads();
    }
  }
  if (!inb) return;
  if (!obs) {
    knn[idx] = make_int4(-1, -1, -1, -1);
    knn_w[idx] = make_float4(0.f, 0.f, 0.f, 0.f);
    return;
  }
  const float bd[4] = {bd0, bd1, bd2, bd3};
  const int bi[4] = {bi0, bi1, bi2, bi3};
  float w[4];
  float wsum = 0.f;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (bi[k] < 0) {
      w[k] = 0.f;
      continue;
    }
    const float r = fmaxf(nodes[bi[k]].radius, 1e-6f);
    w[k] = expf(-bd[k] / (2.0f * r * r));
    wsum += w[k];
  }
  const float inv = (wsum > 1e-8f) ? 1.0f / wsum : 0.0f;
  knn[idx] = make_int4(bi0, bi1, bi2, bi3);
  knn_w[idx] = make_float4(w[0] * inv, w[1] * inv, w[2] * inv, w[3] * inv);
}

// apply_twist_kernel (src/warp_field.cu:250-291): scale, clamp |omega| and
// |v| separately, left-compose the exponential.  Optionally refreshes the
// centred table used by the blends.
__global__ void twist_kernel(const float *__restrict__ dx, DQ *__restrict__ T, int N, float max_rot,
                             float max_trans, float scale, const nrf_node *__restrict__ nodes,
                             DQ *__restrict__ centered) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  float tw[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) tw[k] = scale * dx[i * 6 + k];
  const float rn = sqrtf(tw[0] * tw[0] + tw[1] * tw[1] + tw[2] * tw[2]);
  if (max_rot > 0.0f && rn > max_rot) {
    const float f = max_rot / rn;
    tw[0] *= f; tw[1] *= f; tw[2] *= f;
  }
  const float tn = sqrtf(tw[3] * tw[3] + tw[4] * tw[4] + tw[5] * tw[5]);
  if (max_trans > 0.0f && tn > max_trans) {
    const float f = max_trans / tn;
    tw[3] *= f; tw[4] *= f; tw[5] *= f;
  }
  const DQ dT = mat_to_dq(exp_se3(tw));
  const DQ Tn = dq_mul(dT, T[i]);
  T[i] = Tn;
  if (centered) centered[i] = dq_centered(Tn, f3(nodes[i].pos));
}

__global__ void centered_kernel(const nrf_node *__restrict__ nodes, const DQ *__restrict__ T, int N,
                                DQ *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  out[i] = dq_centered(T[i], f3(nodes[i].pos));
}

__global__ void warp_points_kernel(const float *__restrict__ pts, const int4 *__restrict__ ids4,
                                   const float4 *__restrict__ ws4, int count, const nrf_node *__restrict__ nodes,
                                   const DQ *__restrict__ T, int N, float *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  float3 p = f3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
  const int4 i4 = ids4[i];
  const float4 w4 = ws4[i];
  const int ids[4] = {i4.x, i4.y, i4.z, i4.w};
  const float ws[4] = {w4.x, w4.y, w4.z, w4.w};
  DQ b;
  b.r = make_float4(0.f, 0.f, 0.f, 0.f);
  b.d = make_float4(0.f, 0.f, 0.f, 0.f);
  float wsum = 0.f;
  for (int k = 0; k < 4; ++k) {
    const int id = ids[k];
    const float w = ws[k];
    if (id < 0 || id >= N || w < 1e-8f) continue;
    DQ q = dq_centered(T[id], f3(nodes[id].pos));
    if (qdot(q.r, b.r) < 0) {
      q.r.x = -q.r.x; q.r.y = -q.r.y; q.r.z = -q.r.z; q.r.w = -q.r.w;
      q.d.x = -q.d.x; q.d.y = -q.d.y; q.d.z = -q.d.z; q.d.w = -q.d.w;
    }
    b.r.x += w * q.r.x; b.r.y += w * q.r.y; b.r.z += w * q.r.z; b.r.w += w * q.r.w;
    b.d.x += w * q.d.x; b.d.y += w * q.d.y; b.d.z += w * q.d.z; b.d.w += w * q.d.w;
    wsum += w;
  }
  if (wsum >= 1e-8f) p = dq_transform(dq_normalize(b), p);
  out[3 * i] = p.x;
  out[3 * i + 1] = p.y;
  out[3 * i + 2] = p.z;
}

// warp_normal (include/se3_math.cuh:512-558), dq_warp_normal per point.
__global__ void warp_normals_kernel(const float *__restrict__ nin, const int4 *__restrict__ ids4,
                                    const float4 *__restrict__ ws4, int count, const DQ *__restrict__ T, int N,
                                    float *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int4 i4 = ids4[i];
  const float4 w4 = ws4[i];
  const int ids[4] = {i4.x, i4.y, i4.z, i4.w};
  const float ws[4] = {w4.x, w4.y, w4.z, w4.w};
  const float3 o = dq_warp_normal(T, N, ids, ws, f3(nin[3 * i], nin[3 * i + 1], nin[3 * i + 2]));
  out[3 * i] = o.x;
  out[3 * i + 1] = o.y;
  out[3 * i + 2] = o.z;
}

// Association input packing: a correspondence per point with src = the
// canonical point and the 4-NN of the voxel containing it, as
// compute_pixel_knn_kernel (kernels/tsdf_kernels.cu:573-604) copies the
// hit voxel's kNN and find_correspondences_kernel (:386-563) fills the
// record.  Points outside the volume or in voxels without nodes are invalid.
__global__ void corr_pack_kernel(const float *__restrict__ canon, const float *__restrict__ dst,
                                 const float *__restrict__ nrm, const float *__restrict__ wgt,
                                 const uint8_t *__restrict__ valid, int count, int X, int Y, int Z, float3 org,
                                 float s, const int4 *__restrict__ knn, const float4 *__restrict__ knn_w,
                                 nrf_corr *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  nrf_corr c;
  for (int k = 0; k < 3; ++k) {
    c.src[k] = canon[3 * i + k];
    c.dst[k] = dst[3 * i + k];
    c.normal[k] = nrm[3 * i + k];
  }
  c.weight = wgt ? wgt[i] : 1.0f;
  c.pad[0] = c.pad[1] = c.pad[2] = 0;
  int4 ids = make_int4(-1, -1, -1, -1);
  float4 ws = make_float4(0.f, 0.f, 0.f, 0.f);
  bool ok = valid[i] != 0;
  if (ok) {
    const float gx = (c.src[0] - org.x) / s, gy = (c.src[1] - org.y) / s, gz = (c.src[2] - org.z) / s;
    const int x = (int)floorf(gx), y = (int)floorf(gy), z = (int)floorf(gz);
    if (x < 0 || y < 0 || z < 0 || x >= X || y >= Y || z >= Z) {
      ok = false;
    } else {
      const int64_t v = ((int64_t)z * Y + y) * X + x;
      ids = knn[v];
      ws = knn_w[v];
      ok = ids.x >= 0;
    }
  }
  c.node_ids[0] = ids.x; c.node_ids[1] = ids.y; c.node_ids[2] = ids.z; c.node_ids[3] = ids.w;
  c.node_ws[0] = ws.x; c.node_ws[1] = ws.y; c.node_ws[2] = ws.z; c.node_ws[3] = ws.w;
  c.valid = ok ? 1 : 0;
  if (!ok) c.weight = 0.0f;
  out[i] = c;
}

}  // namespace
}  // namespace nrf

using namespace nrf;

extern "C" int nrf_voxel_knn(const nrf_volume_params *vp, const nrf_voxel *voxels, const nrf_node *nodes,
                             int num_nodes, int32_t *voxel_knn, float *voxel_knn_w, void *stream) {
  if (!volume_ok(vp) || !voxels || !voxel_knn || !voxel_knn_w || num_nodes < 0) return NRF_ERR_INVALID_ARGUMENT;
  if (num_nodes > 0 && !nodes) return NRF_ERR_INVALID_ARGUMENT;
  const int64_t V = volume_count(vp);
  knn_kernel<<<(unsigned)((V + kKnnBlock - 1) / kKnnBlock), kKnnBlock, 0, as_stream(stream)>>>(
      voxels, vp->dims[0], vp->dims[1], vp->dims[2], make_float3(vp->origin[0], vp->origin[1], vp->origin[2]),
      vp->voxel_size, nodes, num_nodes, reinterpret_cast<int4 *>(voxel_knn), reinterpret_cast<float4 *>(voxel_knn_w));
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}

extern "C" int nrf_warp_apply_twist(const float *delta_x, nrf_dualquat *transforms, int num_nodes, float max_rot,
                                    float max_trans, float update_scale, const nrf_node *nodes,
                                    nrf_dualquat *centered, void *stream) {
  if (num_nodes < 0) return NRF_ERR_INVALID_ARGUMENT;
  if (num_nodes == 0) return NRF_OK;
  if (!delta_x || !transforms || (centered && !nodes)) return NRF_ERR_INVALID_ARGUMENT;
  twist_kernel<<<(num_nodes + 255) / 256, 256, 0, as_stream(stream)>>>(
      delta_x, reinterpret_cast<DQ *>(transforms), num_nodes, max_rot, max_trans, update_scale, nodes,
      reinterpret_cast<DQ *>(centered));
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}

extern "C" int nrf_warp_centered_table(const nrf_node *nodes, const nrf_dualquat *transforms, int num_nodes,
                                       nrf_dualquat *centered, void *stream) {
  if (num_nodes < 0) return NRF_ERR_INVALID_ARGUMENT;
  if (num_nodes == 0) return NRF_OK;
  if (!nodes || !transforms || !centered) return NRF_ERR_INVALID_ARGUMENT;
  centered_kernel<<<(num_nodes + 255) / 256, 256, 0, as_stream(stream)>>>(
      nodes, reinterpret_cast<const DQ *>(transforms), num_nodes, reinterpret_cast<DQ *>(centered));
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}

extern "C" int nrf_warp_points(const float *points, const int32_t *knn_ids, const float *knn_ws, int count,
                               const nrf_node *nodes, const nrf_dualquat *transforms, int num_nodes, float *out,
                               void *stream) {
  if (count < 0 || num_nodes < 0) return NRF_ERR_INVALID_ARGUMENT;
  if (count == 0) return NRF_OK;
  if (!points || !knn_ids || !knn_ws || !out || (num_nodes > 0 && (!nodes || !transforms)))
    return NRF_ERR_INVALID_ARGUMENT;
  warp_points_kernel<<<(count + 255) / 256, 256, 0, as_stream(stream)>>>(
      points, reinterpret_cast<const int4 *>(knn_ids), reinterpret_cast<const float4 *>(knn_ws), count, nodes,
      reinterpret_cast<const DQ *>(transforms), num_nodes, out);
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}

extern "C" int nrf_corr_pack(const nrf_volume_params *vp, const int32_t *voxel_knn, const float *voxel_knn_w,
                             const float *canonical, const float *dst, const float *normals, const float *weights,
                             const uint8_t *valid, int count, nrf_corr *out, void *stream) {
  if (!volume_ok(vp) || count < 0) return NRF_ERR_INVALID_ARGUMENT;
  if (count == 0) return NRF_OK;
  if (!voxel_knn || !voxel_knn_w || !canonical || !dst || !normals || !valid || !out) return NRF_ERR_INVALID_ARGUMENT;
  corr_pack_kernel<<<(count + 255) / 256, 256, 0, as_stream(stream)>>>(
      canonical, dst, normals, weights, valid, count, vp->dims[0], vp->dims[1], vp->dims[2],
      make_float3(vp->origin[0], vp->origin[1], vp->origin[2]), vp->voxel_size,
      reinterpret_cast<const int4 *>(voxel_knn), reinterpret_cast<const float4 *>(voxel_knn_w), out);
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}

extern "C" int nrf_warp_normals(const float *normals, const int32_t *knn_ids, const float *knn_ws, int count,
                                const nrf_dualquat *transforms, int num_nodes, float *out, void *stream) {
  if (count < 0 || num_nodes < 0) return NRF_ERR_INVALID_ARGUMENT;
  if (count == 0) return NRF_OK;
  if (!normals || !knn_ids || !knn_ws || !out || (num_nodes > 0 && !transforms)) return NRF_ERR_INVALID_ARGUMENT;
  warp_normals_kernel<<<(count + 255) / 256, 256, 0, as_stream(stream)>>>(
      normals, reinterpret_cast<const int4 *>(knn_ids), reinterpret_cast<const float4 *>(knn_ws), count,
      reinterpret_cast<const DQ *>(transforms), num_nodes, out);
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}
