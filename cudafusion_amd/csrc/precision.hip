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

constexpr int kTileX = 64;
constexpr int kTileY = 4;
constexpr int kZChunk = 4;

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

struct Int16Args {
  __half *t16;
  uint8_t *w8;
  int X, Y, Z;
  float ox, oy, oz, s, trunc;
  const float *depth;
  int W, H;
  float fx, fy, cx, cy;
  CamXf T;
  const DQ *centered;
  int N;
  const ushort4 *knn;
  const uint2 *knn_w;
  float decay, wmax;
  int32_t *updated;
};

__global__ __launch_bounds__(256) void integrate16_kernel(Int16Args a) {
  const int x = blockIdx.x * kTileX + (threadIdx.x & (kTileX - 1));
  const int y = blockIdx.y * kTileY + (threadIdx.x / kTileX);
  const int z0 = blockIdx.z * kZChunk;
  int n_upd = 0;
  if (x < a.X && y < a.Y) {
    const float px_w = a.ox + (x + 0.5f) * a.s;
    const float py_w = a.oy + (y + 0.5f) * a.s;
#pragma unroll
    for (int dz = 0; dz < kZChunk; ++dz) {
      const int z = z0 + dz;
      if (z >= a.Z) break;
      const int idx = z * a.X * a.Y + y * a.X + x;
      float3 p = f3(px_w, py_w, a.oz + (z + 0.5f) * a.s);
      if (a.N > 0) {
        const ushort4 k4 = a.knn[idx];
        const uint2 w2 = a.knn_w[idx];
        const __half2 wa = *reinterpret_cast<const __half2 *>(&w2.x);
        const __half2 wb = *reinterpret_cast<const __half2 *>(&w2.y);
        const float2 fa = __half22float2(wa), fb = __half22float2(wb);
        const int ids[4] = {k4.x == 0xffff ? -1 : (int)k4.x, k4.y == 0xffff ? -1 : (int)k4.y,
                            k4.z == 0xffff ? -1 : (int)k4.z, k4.w == 0xffff ? -1 : (int)k4.w};
        const float ws[4] = {fa.x, fa.y, fb.x, fb.y};
        DQ b;
        b.r = make_float4(0.f, 0.f, 0.f, 0.f);
        b.d = make_float4(0.f, 0.f, 0.f, 0.f);
        float wsum = 0.f;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int id = ids[k];
          const float w = ws[k];
          if (id < 0 || id >= a.N || w < 1e-8f) continue;
          DQ q = a.centered[id];
          if (qdot(q.r, b.r) < 0) {
            q.r.x = -q.r.x; q.r.y = -q.r.y; q.r.z = -q.r.z; q.r.w = -q.r.w;
            q.d.x = -q.d.x; q.d.y = -q.d.y; q.d.z = -q.d.z; q.d.w = -q.d.w;
          }
          b.r.x += w * q.r.x; b.r.y += w * q.r.y; b.r.z += w * q.r.z; b.r.w += w * q.r.w;
          b.d.x += w * q.d.x; b.d.y += w * q.d.y; b.d.z += w * q.d.z; b.d.w += w * q.d.w;
          wsum += w;
        }
        if (wsum >= 1e-8f) p = dq_transform(dq_normalize(b), p);
      }
      const float cxp = a.T.m[0][0] * p.x + a.T.m[0][1] * p.y + a.T.m[0][2] * p.z + a.T.m[0][3];
      const float cyp = a.T.m[1][0] * p.x + a.T.m[1][1] * p.y + a.T.m[1][2] * p.z + a.T.m[1][3];
      const float czp = a.T.m[2][0] * p.x + a.T.m[2][1] * p.y + a.T.m[2][2] * p.z + a.T.m[2][3];
      if (czp <= 0) continue;
      const float fu = a.fx * cxp / czp + a.cx;
      const float fv = a.fy * cyp / czp + a.cy;
      if (!(fu > -1.0e9f && fu < 1.0e9f && fv > -1.0e9f && fv < 1.0e9f)) continue;
      const int u = (int)rintf(fu);
      const int v = (int)rintf(fv);
      if (u < 0 || u >= a.W || v < 0 || v >= a.H) continue;
      const float d = a.depth[v * a.W + u];
      if (d <= 0) continue;
      const float sdf = d - czp;
      if (sdf < -a.trunc) continue;
      const float t = fminf(1.f, sdf / a.trunc);
      const float w_old = (float)a.w8[idx];
      const float dw = w_old * a.decay;
      const float nw = dw + 1.f;
      a.t16[idx] = __float2half_rn((__half2float(a.t16[idx]) * dw + t * 1.f) / nw);
      a.w8[idx] = (uint8_t)fminf(rintf(fminf(nw, a.wmax)), 255.f);
      ++n_upd;
    }
  }
  if (a.updated) {
    for (int off = 32; off > 0; off >>= 1) n_upd += __shfl_xor(n_upd, off, 64);
    if ((threadIdx.x & 63) == 0 && n_upd) atomicAdd(a.updated, n_upd);
  }
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
                                                          float *__restrict__ verts, int32_t *__restrict__ tris) {
  const int y = blockIdx.x, z = blockIdx.y;
  const int row = z * Y + y;
  if (rowcnt[row] == 0) return;
  __shared__ int32_t wsum[kMcBlock / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int run = rowoff[row];
  for (int x0 = 0; x0 < X - 1; x0 += kMcBlock) {
    const int x = x0 + threadIdx.x;
    const int idx = z * X * Y + y * X + x;
    const int cube = (x < X - 1) ? cases[idx] : 0;
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
  if (!volume_ok(vp) || !tsdf16 || !weight8 || !depth || !cam || !camera_pose) return NRF_ERR_INVALID_ARGUMENT;
  if (num_nodes > 0 && (!centered || !knn16 || !knn_w16)) return NRF_ERR_INVALID_ARGUMENT;
  if (num_nodes > 65535) return NRF_ERR_TOO_LARGE;
  if (vp->max_weight > 255.f) return NRF_ERR_TOO_LARGE;
  Int16Args a;
  a.t16 = reinterpret_cast<__half *>(tsdf16);
  a.w8 = weight8;
  a.X = vp->dims[0]; a.Y = vp->dims[1]; a.Z = vp->dims[2];
  a.ox = vp->origin[0]; a.oy = vp->origin[1]; a.oz = vp->origin[2];
  a.s = vp->voxel_size;
  a.trunc = vp->truncation;
  a.depth = depth;
  a.W = cam->width; a.H = cam->height;
  a.fx = cam->fx; a.fy = cam->fy; a.cx = cam->cx; a.cy = cam->cy;
  a.T = invert_pose(camera_pose);
  a.centered = reinterpret_cast<const DQ *>(centered);
  a.N = num_nodes;
  a.knn = reinterpret_cast<const ushort4 *>(knn16);
  a.knn_w = reinterpret_cast<const uint2 *>(knn_w16);
  a.decay = vp->decay_alpha;
  a.wmax = vp->max_weight;
  a.updated = updated_count;
  dim3 grid((a.X + kTileX - 1) / kTileX, (a.Y + kTileY - 1) / kTileY, (a.Z + kZChunk - 1) / kZChunk);
  integrate16_kernel<<<grid, 256, 0, as_stream(stream)>>>(a);
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}

extern "C" int nrf_mc16_extract(const nrf_volume_params *vp, const uint16_t *tsdf16, const uint8_t *weight8,
                                void *workspace, size_t workspace_bytes, float *vertices, int32_t *triangles,
                                int max_triangles, int32_t *counts, void *stream) {
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
  hipStream_t st = as_stream(stream);
  const __half *t16 = reinterpret_cast<const __half *>(tsdf16);
  const int nz = (int)L.zero_ints;  // the counts and the scan's status words
  zero_i32<<<(unsigned)((nz + 255) / 256), 256, 0, st>>>(rowcnt, nz);
  NRF_LAUNCH_CHECK();
  classify_brick_kernel<<<classify_brick_grid(X, Y, Z), kMcBrickThreads, 0, st>>>(VoxelReaderF16{t16, weight8}, X,
                                                                                   Y, Z, cases, rowcnt);
  NRF_LAUNCH_CHECK();
  mc_scan_rows(L, rowcnt, rowoff, (int)R, counts, workspace, st);
  NRF_LAUNCH_CHECK();
  if (X >= 2 && Y >= 2 && Z >= 2 && max_triangles > 0) {
    emit16_kernel<<<dim3(Y, Z), kMcBlock, 0, st>>>(t16, X, Y, Z,
                                                   make_float3(vp->origin[0], vp->origin[1], vp->origin[2]),
                                                   vp->voxel_size, cases, rowcnt, rowoff, max_triangles, vertices,
                                                   triangles);
    NRF_LAUNCH_CHECK();
  }
  return NRF_OK;
}
