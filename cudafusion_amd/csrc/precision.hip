// precision.hip - the compact-storage form of integration and marching cubes.
//
// The TSDF is stored as an FP16 value plus a uint8 weight (3 B/voxel instead
// of the reference's 12 B TSDFVoxel) and the voxel kNN field as uint16 ids
// plus FP16 weights (16 B/voxel instead of 32 B).  Arithmetic is the FP32
// arithmetic of integrate.hip / mc.hip on the decoded values: the state is
// decoded exactly, updated in FP32 and rounded once (round-to-nearest-even
// to FP16; the weight is rounded to the nearest integer, exact for the
// decay_alpha = 1 configurations since every update adds 1).
#include <hip/hip_fp16.h>

#include "mc_front.h"
#include "mc_tables.h"
#include "nrf_internal.h"
#include "nrf_math.h"

namespace nrf {
namespace {

__global__ void compact_knn_kernel(const int4 *__restrict__ knn, const float4 *__restrict__ w, int64_t V,
                                   ushort4 *__restrict__ knn16, uint2 *__restrict__ w16) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V) return;
  const int4 k = knn[i];
  const float4 f = w[i];
  auto id16 = [](int v) -> unsigned short { return v < 0 ? (unsigned short)0xffff : (unsigned short)v; };
  knn16[i] = make_ushort4(id16(k.x), id16(k.y), id16(k.z), id16(k.w));
  const __half2 a = __floats2half2_rn(f.x, f.y), b = __floats2half2_rn(f.z, f.w);
  w16[i] = make_uint2(*reinterpret_cast<const unsigned *>(&a), *reinterpret_cast<const unsigned *>(&b));
}

__global__ void voxels_to_compact_kernel(const nrf_voxel *__restrict__ v, int64_t V, __half *__restrict__ t16,
                                         uint8_t *__restrict__ w8) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V) return;
  t16[i] = __float2half_rn(v[i].tsdf);
  const float w = fminf(fmaxf(rintf(v[i].weight), 0.f), 255.f);
  w8[i] = (uint8_t)w;
}

__global__ void fill16_kernel(__half *t16, uint8_t *w8, int64_t V) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V) return;
  t16[i] = __float2half_rn(1.f);
  w8[i] = 0;
}

constexpr int kMcBlock = 256;

__global__ void zero_i32(int32_t *p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

struct VoxelReaderF16 {
  const __half *t;
  const uint8_t *w;
  __device__ float2 operator()(size_t i) const { return make_float2(__half2float(t[i]), (float)w[i]); }
};



__device__ inline float3 edge_point16(int e, const float *val, int x, int y, int z, float3 org, float s) {
  const int c0 = kMcEdge[e][0], c1 = kMcEdge[e][1];
  const float3 p0 = f3((float)(x + kMcCorner[c0][0]) + 0.5f, (float)(y + kMcCorner[c0][1]) + 0.5f,
                       (float)(z + kMcCorner[c0][2]) + 0.5f);
  const float3 p1 = f3((float)(x + kMcCorner[c1][0]) + 0.5f, (float)(y + kMcCorner[c1][1]) + 0.5f,
                       (float)(z + kMcCorner[c1][2]) + 0.5f);
  const float v0 = val[c0], v1 = val[c1];
  const float den = v1 - v0;
  float t = fabsf(den) > 1e-12f ? -v0 / den : 0.5f;
  t = fminf(1.f, fmaxf(0.f, t));
  const float3 p = f3(p0.x + (p1.x - p0.x) * t, p0.y + (p1.y - p0.y) * t, p0.z + (p1.z - p0.z) * t);
  return f3(org.x + p.x * s, org.y + p.y * s, org.z + p.z * s);
}

__global__ __launch_bounds__(kMcBlock) void emit16_kernel(const __half *__restrict__ t16, int X, int Y, int Z,
                                                          float3 org, float s, const uint8_t *__restrict__ cases,
                                                          const int32_t *__restrict__ rowcnt,
                                                          const int32_t *__restrict__ rowoff, int max_tris,
                                                          float *__restrict__ verts, int32_t *__restrict__ tris,
                                                          const uint8_t *__restrict__ occ) {
  const int y = blockIdx.x, z = blockIdx.y;
  const int row = z * Y + y;
  if (rowcnt[row] == 0) return;
  __shared__ int32_t wsum[kMcBlock / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int run = rowoff[row];
  for (int x0 = 0; x0 < X - 1; x0 += kMcBlock) {
    const int x = x0 + threadIdx.x;
    const int idx = z * X * Y + y * X + x;
    const int cube = (x < X - 1 && (!occ || occ[mc_brick_of(x, y, z, X, Y)])) ? cases[idx] : 0;
    const int nt = kMcNTri[cube];
    int inc = nt;
    for (int d = 1; d < 64; d <<= 1) {
      int v = __shfl_up(inc, d, 64);
      if (lane >= d) inc += v;
    }
    if (lane == 63) wsum[wv] = inc;
    __syncthreads();
    int before = 0, total = 0;
    for (int w = 0; w < kMcBlock / 64; ++w) {
      if (w < wv) before += wsum[w];
      total += wsum[w];
    }
    __syncthreads();
    const int first = run + before + inc - nt;
    run += total;
    if (nt == 0) continue;
    float val[8];
#pragma unroll
    for (int c = 0; c < 8; ++c)
      val[c] = __half2float(t16[idx + kMcCorner[c][0] + kMcCorner[c][1] * X + kMcCorner[c][2] * X * Y]);
    for (int t = 0; t < nt; ++t) {
      const int tri = first + t;
      if (tri >= max_tris) break;
      const float3 a = edge_point16(kMcTri[cube][3 * t + 0], val, x, y, z, org, s);
      const float3 b = edge_point16(kMcTri[cube][3 * t + 1], val, x, y, z, org, s);
      const float3 c = edge_point16(kMcTri[cube][3 * t + 2], val, x, y, z, org, s);
      float *vo = verts + (size_t)tri * 9;
      vo[0] = a.x; vo[1] = a.y; vo[2] = a.z;
      vo[3] = b.x; vo[4] = b.y; vo[5] = b.z;
      vo[6] = c.x; vo[7] = c.y; vo[8] = c.z;
      int32_t *to = tris + (size_t)tri * 3;
      to[0] = 3 * tri; to[1] = 3 * tri + 1; to[2] = 3 * tri + 2;
    }
  }
}

}  // namespace
}  // namespace nrf

using namespace nrf;

extern "C" int nrf_knn_compact(const nrf_volume_params *vp, const int32_t *voxel_knn, const float *voxel_knn_w,
                               uint16_t *knn16, uint16_t *knn_w16, void *stream) {
  if (!volume_ok(vp) || !voxel_knn || !voxel_knn_w || !knn16 || !knn_w16) return NRF_ERR_INVALID_ARGUMENT;
  const int64_t V = volume_count(vp);
  compact_knn_kernel<<<(unsigned)((V + 255) / 256), 256, 0, as_stream(stream)>>>(
      reinterpret_cast<const int4 *>(voxel_knn), reinterpret_cast<const float4 *>(voxel_knn_w), V,
      reinterpret_cast<ushort4 *>(knn16), reinterpret_cast<uint2 *>(knn_w16));
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}

extern "C" int nrf_tsdf16_from_voxels(const nrf_volume_params *vp, const nrf_voxel *voxels, uint16_t *tsdf16,
                                      uint8_t *weight8, void *stream) {
  if (!volume_ok(vp) || !voxels || !tsdf16 || !weight8) return NRF_ERR_INVALID_ARGUMENT;
  const int64_t V = volume_count(vp);
  voxels_to_compact_kernel<<<(unsigned)((V + 255) / 256), 256, 0, as_stream(stream)>>>(
      voxels, V, reinterpret_cast<__half *>(tsdf16), weight8);
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}

extern "C" int nrf_tsdf16_reset(const nrf_volume_params *vp, uint16_t *tsdf16, uint8_t *weight8, void *stream) {
  if (!volume_ok(vp) || !tsdf16 || !weight8) return NRF_ERR_INVALID_ARGUMENT;
  const int64_t V = volume_count(vp);
  fill16_kernel<<<(unsigned)((V + 255) / 256), 256, 0, as_stream(stream)>>>(reinterpret_cast<__half *>(tsdf16),
                                                                            weight8, V);
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}

extern "C" int nrf_tsdf16_integrate(const nrf_volume_params *vp, uint16_t *tsdf16, uint8_t *weight8,
                                    const float *depth, const nrf_camera *cam, const float camera_pose[16],
                                    const nrf_dualquat *centered, int num_nodes, const uint16_t *knn16,
                                    const uint16_t *knn_w16, int32_t *updated_count, void *stream) {
  return tsdf16_integrate(vp, tsdf16, weight8, depth, cam, camera_pose, centered, num_nodes, knn16, knn_w16,
                          nullptr, updated_count, as_stream(stream));
}

int nrf::mc16_extract(const nrf_volume_params *vp, const uint16_t *tsdf16, const uint8_t *weight8,
                      const uint8_t *occ, void *workspace, size_t workspace_bytes, float *vertices,
                      int32_t *triangles, int max_triangles, int32_t *counts, hipStream_t st) {
  if (!volume_ok(vp) || !tsdf16 || !weight8 || !workspace || !counts) return NRF_ERR_INVALID_ARGUMENT;
  if (max_triangles < 0 || (max_triangles > 0 && (!vertices || !triangles))) return NRF_ERR_INVALID_ARGUMENT;
  const int X = vp->dims[0], Y = vp->dims[1], Z = vp->dims[2];
  if (Y > 65535 || Z > 65535) return NRF_ERR_TOO_LARGE;
  const size_t R = (size_t)Y * Z;
  const McLayout L = mc_layout(vp);
  if (workspace_bytes < L.total) return NRF_ERR_WORKSPACE;
  char *ws = reinterpret_cast<char *>(workspace);
  uint8_t *cases = reinterpret_cast<uint8_t *>(ws + L.cases_off);
  int32_t *rowcnt = reinterpret_cast<int32_t *>(ws + L.rowcnt_off);
  int32_t *rowoff = reinterpret_cast<int32_t *>(ws + L.rowoff_off);
  const __half *t16 = reinterpret_cast<const __half *>(tsdf16);
  const int nz = (int)L.zero_ints;  // the counts and the scan's status words
  zero_i32<<<(unsigned)((nz + 255) / 256), 256, 0, st>>>(rowcnt, nz);
  NRF_LAUNCH_CHECK();
  classify_brick_kernel<<<classify_brick_grid(X, Y, Z), kMcBrickThreads, 0, st>>>(VoxelReaderF16{t16, weight8}, X,
                                                                                   Y, Z, cases, rowcnt, occ);
  NRF_LAUNCH_CHECK();
  mc_scan_rows(L, rowcnt, rowoff, (int)R, counts, workspace, st);
  NRF_LAUNCH_CHECK();
  if (X >= 2 && Y >= 2 && Z >= 2 && max_triangles > 0) {
    emit16_kernel<<<dim3(Y, Z), kMcBlock, 0, st>>>(t16, X, Y, Z,
                                                   make_float3(vp->origin[0], vp->origin[1], vp->origin[2]),
                                                   vp->voxel_size, cases, rowcnt, rowoff, max_triangles, vertices,
                                                   triangles, occ);
    NRF_LAUNCH_CHECK();
  }
  return NRF_OK;
}

extern "C" int nrf_mc16_extract(const nrf_volume_params *vp, const uint16_t *tsdf16, const uint8_t *weight8,
                                void *workspace, size_t workspace_bytes, float *vertices, int32_t *triangles,
                                int max_triangles, int32_t *counts, void *stream) {
  return mc16_extract(vp, tsdf16, weight8, nullptr, workspace, workspace_bytes, vertices, triangles, max_triangles,
                      counts, as_stream(stream));
}
