// mc.hip - deterministic marching cubes for gfx950.
//
// Semantics follow src/marching_cubes.cu:315-451 of the reference: a cube is
// skipped when any corner weight is <= 0 or its case is 0/255; bit c of the
// case is set when corner c has tsdf > 0; edge points interpolate at
// t = -v0/(v1-v0) clamped to [0,1] (0.5 for |v1-v0| <= 1e-12) between voxel
// centres; each triangle gets three fresh vertices and its face normal.
//
// Instead of the reference's two passes with global atomic output counters
// (nondeterministic order, a host round trip between them) the extraction is
// three stream-ordered kernels with no host synchronisation:
//   1. classify (mc_front.h): one workgroup per 64 x 4 x 4 brick of cubes,
//      voxels staged once in LDS; writes the case byte and adds the row's
//      triangle count (x-rows of cubes, integer atomics, so deterministic);
//   2. scan: exclusive prefix sum of the row counts (one workgroup);
//   3. emit: one workgroup per non-empty row; a block scan of per-cube
//      triangle counts places every triangle at its rank in cube order.
// Output order is therefore increasing cube index, reproducible run to run.
#include "mc_front.h"
#include "mc_tables.h"
#include "nrf_internal.h"
#include "nrf_math.h"

#include <mutex>

namespace nrf {
namespace {

constexpr int kMcBlock = 256;

// Scratch of the workspace-free extraction (nrf_mc_extract_noalloc): row
// counts, the scan's status and flag words and the row offsets of up to
// kMcNoAllocMaxRows rows (dims[1] * dims[2]), laid out as mc_layout_for(0, R)
// (row counts at 0).  One copy per device, held until the process ends.
constexpr size_t kMcNoAllocMaxRows = 512 * 512;
constexpr size_t cx_align(size_t v, size_t a) { return (v + a - 1) / a * a; }
constexpr size_t kMcNoAllocBytes =
    cx_align(cx_align(kMcNoAllocMaxRows * 4, 8) + ((kMcNoAllocMaxRows + kScanBlock - 1) / kScanBlock + 1) * 8, 256) +
    cx_align((kMcNoAllocMaxRows + 1) * 4, 256);
__device__ __align__(256) unsigned char scratch[kMcNoAllocBytes];


struct VoxelReaderF32 {
  const nrf_voxel *v;
  __device__ float2 operator()(size_t i) const { return make_float2(v[i].tsdf, v[i].weight); }
};

__global__ void fill_zero_kernel(int32_t *p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}


__device__ inline float3 edge_point(int e, const float *val, int x, int y, int z, float3 org, float s) {
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

// Case index of the cube at voxel index idx from its eight corners, by the
// rule of classify_brick_kernel (a corner with weight <= 0 voids the cube;
// bit c set when tsdf_c > 0; 0 and 255 emit nothing).
__device__ inline int cube_case(const nrf_voxel *__restrict__ vox, int idx, int X, int Y) {
  int cube = 0;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const nrf_voxel &v = vox[idx + kMcCorner[c][0] + kMcCorner[c][1] * X + kMcCorner[c][2] * X * Y];
    if (v.weight <= 0.f) cube = -1;
    if (cube >= 0 && v.tsdf > 0.f) cube |= (1 << c);
  }
  return (cube <= 0 || cube >= 255) ? 0 : cube;
}

// kCases: the case bytes come from the classify pass; otherwise (the
// workspace-free form, which keeps no case array) each cube's case is
// recomputed from its corners.
template <bool kCases>
__global__ __launch_bounds__(kMcBlock) void emit_kernel(const nrf_voxel *__restrict__ vox, int X, int Y, int Z,
                                                        float3 org, float s, const uint8_t *__restrict__ cases,
                                                        const int32_t *__restrict__ rowcnt,
                                                        const int32_t *__restrict__ rowoff, int max_tris,
                                                        float *__restrict__ verts, float *__restrict__ normals,
                                                        int32_t *__restrict__ tris, const uint8_t *__restrict__ occ) {
  const int y = blockIdx.x, z = blockIdx.y;
  const int row = z * Y + y;
  if (rowcnt[row] == 0) return;
  __shared__ int32_t wsum[kMcBlock / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int run = rowoff[row];
  for (int x0 = 0; x0 < X - 1; x0 += kMcBlock) {
    const int x = x0 + threadIdx.x;
    const int idx = z * X * Y + y * X + x;
    int cube = 0;
    if (x < X - 1 && (!occ || occ[mc_brick_of(x, y, z, X, Y)]))
      cube = kCases ? (int)cases[idx] : cube_case(vox, idx, X, Y);
    const int nt = kMcNTri[cube];
    int inc = nt;  // inclusive wave scan
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
      val[c] = vox[idx + kMcCorner[c][0] + kMcCorner[c][1] * X + kMcCorner[c][2] * X * Y].tsdf;
    for (int t = 0; t < nt; ++t) {
      const int tri = first + t;
      if (tri >= max_tris) break;
      const float3 a = edge_point(kMcTri[cube][3 * t + 0], val, x, y, z, org, s);
      const float3 b = edge_point(kMcTri[cube][3 * t + 1], val, x, y, z, org, s);
      const float3 c = edge_point(kMcTri[cube][3 * t + 2], val, x, y, z, org, s);
      float *vo = verts + (size_t)tri * 9;
      vo[0] = a.x; vo[1] = a.y; vo[2] = a.z;
      vo[3] = b.x; vo[4] = b.y; vo[5] = b.z;
      vo[6] = c.x; vo[7] = c.y; vo[8] = c.z;
      if (normals) {
        const float3 n = cross(sub(b, a), sub(c, a));
        const float len = sqrtf(n.x * n.x + n.y * n.y + n.z * n.z);
        float3 nn = f3(0.f, 0.f, 0.f);
        if (!(len <= 1e-12f)) {
          const float inv = 1.f / len;
          nn = f3(n.x * inv, n.y * inv, n.z * inv);
        }
        float *no = normals + (size_t)tri * 9;
        no[0] = nn.x; no[1] = nn.y; no[2] = nn.z;
        no[3] = nn.x; no[4] = nn.y; no[5] = nn.z;
        no[6] = nn.x; no[7] = nn.y; no[8] = nn.z;
      }
      int32_t *to = tris + (size_t)tri * 3;
      to[0] = 3 * tri; to[1] = 3 * tri + 1; to[2] = 3 * tri + 2;
    }
  }
}

}  // namespace
}  // namespace nrf

using namespace nrf;

extern "C" size_t nrf_mc_workspace_size(const nrf_volume_params *vp) {
  if (!volume_ok(vp)) return 0;
  return mc_layout(vp).total;
}

extern "C" const uint8_t *nrf_mc_cases(const void *workspace) {
  return reinterpret_cast<const uint8_t *>(workspace);
}

int nrf::mc_extract(const nrf_volume_params *vp, const nrf_voxel *voxels, const uint8_t *occ, void *workspace,
                    size_t workspace_bytes, float *vertices, float *normals, int32_t *triangles, int max_triangles,
                    int32_t *counts, hipStream_t st) {
  if (!volume_ok(vp) || !voxels || !workspace || !counts) return NRF_ERR_INVALID_ARGUMENT;
  if (max_triangles < 0 || (max_triangles > 0 && (!vertices || !triangles))) return NRF_ERR_INVALID_ARGUMENT;
  const int X = vp->dims[0], Y = vp->dims[1], Z = vp->dims[2];
  if (Y > 65535 || Z > 65535) return NRF_ERR_TOO_LARGE;
  McLayout L = mc_layout(vp);
  if (workspace_bytes < L.total) return NRF_ERR_WORKSPACE;
  char *ws = reinterpret_cast<char *>(workspace);
  uint8_t *cases = reinterpret_cast<uint8_t *>(ws + L.cases_off);
  int32_t *rowcnt = reinterpret_cast<int32_t *>(ws + L.rowcnt_off);
  int32_t *rowoff = reinterpret_cast<int32_t *>(ws + L.rowoff_off);
  const int R = Y * Z;
  const int nz = (int)L.zero_ints;  // the counts and the scan's status words
  fill_zero_kernel<<<(nz + 255) / 256, 256, 0, st>>>(rowcnt, nz);
  NRF_LAUNCH_CHECK();
  classify_brick_kernel<<<classify_brick_grid(X, Y, Z), kMcBrickThreads, 0, st>>>(VoxelReaderF32{voxels}, X, Y, Z,
                                                                                   cases, rowcnt, occ);
  NRF_LAUNCH_CHECK();
  mc_scan_rows(L, rowcnt, rowoff, R, counts, workspace, st);
  NRF_LAUNCH_CHECK();
  if (X >= 2 && Y >= 2 && Z >= 2 && max_triangles > 0) {
    emit_kernel<true><<<dim3(Y, Z), kMcBlock, 0, st>>>(voxels, X, Y, Z,
                                                 make_float3(vp->origin[0], vp->origin[1], vp->origin[2]),
                                                 vp->voxel_size, cases, rowcnt, rowoff, max_triangles, vertices,
                                                 normals, triangles, occ);
    NRF_LAUNCH_CHECK();
  }
  return NRF_OK;
}

extern "C" int nrf_mc_extract(const nrf_volume_params *vp, const nrf_voxel *voxels, void *workspace,
                              size_t workspace_bytes, float *vertices, float *normals, int32_t *triangles,
                              int max_triangles, int32_t *counts, void *stream) {
  return mc_extract(vp, voxels, nullptr, workspace, workspace_bytes, vertices, normals, triangles, max_triangles,
                    counts, as_stream(stream));
}

size_t nrf::mc_occupancy_bytes(const nrf_volume_params *vp) {
  return mc_brick_count(vp->dims[0], vp->dims[1], vp->dims[2]);
}

// Device address of this device's copy of `scratch`, looked up once per
// device (the host-side name of the array is not a device address).
static int mc_scratch_ptr(unsigned char **out) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static unsigned char *ptrs[kMaxDevices] = {};
  int dev = 0;
  NRF_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return NRF_ERR_TOO_LARGE;
  std::lock_guard<std::mutex> lock(mu);
  if (!ptrs[dev]) {
    void *p = nullptr;
    NRF_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    ptrs[dev] = reinterpret_cast<unsigned char *>(p);
  }
  *out = ptrs[dev];
  return NRF_OK;
}

extern "C" int nrf_mc_extract_noalloc(const nrf_volume_params *vp, const nrf_voxel *voxels, float *vertices,
                                      float *normals, int32_t *triangles, int max_triangles, int32_t *counts,
                                      void *stream) {
  if (!volume_ok(vp) || !voxels || !counts) return NRF_ERR_INVALID_ARGUMENT;
  if (max_triangles < 0 || (max_triangles > 0 && (!vertices || !triangles))) return NRF_ERR_INVALID_ARGUMENT;
  const int X = vp->dims[0], Y = vp->dims[1], Z = vp->dims[2];
  const size_t R = (size_t)Y * Z;
  if (R > kMcNoAllocMaxRows) {
    set_last_error_msg("nrf_mc_extract_noalloc: dims[1] * dims[2] = " + std::to_string(R) +
                       " rows exceeds the scratch capacity of " + std::to_string(kMcNoAllocMaxRows) + " rows");
    return NRF_ERR_TOO_LARGE;
  }
  unsigned char *ws = nullptr;
  int rc = mc_scratch_ptr(&ws);
  if (rc) return rc;
  const McLayout L = mc_layout_for(0, R);
  if (L.total > kMcNoAllocBytes) return NRF_ERR_TOO_LARGE;  // cannot happen for R <= kMcNoAllocMaxRows
  int32_t *rowcnt = reinterpret_cast<int32_t *>(ws + L.rowcnt_off);
  int32_t *rowoff = reinterpret_cast<int32_t *>(ws + L.rowoff_off);
  hipStream_t st = as_stream(stream);
  const int nz = (int)L.zero_ints;  // the counts and the scan's status words (a kernel, no memset node)
  fill_zero_kernel<<<(nz + 255) / 256, 256, 0, st>>>(rowcnt, nz);
  NRF_LAUNCH_CHECK();
  classify_brick_kernel<<<classify_brick_grid(X, Y, Z), kMcBrickThreads, 0, st>>>(VoxelReaderF32{voxels}, X, Y, Z,
                                                                                   nullptr, rowcnt, nullptr);
  NRF_LAUNCH_CHECK();
  mc_scan_rows(L, rowcnt, rowoff, (int)R, counts, ws, st);
  NRF_LAUNCH_CHECK();
  if (X >= 2 && Y >= 2 && Z >= 2 && max_triangles > 0) {
    emit_kernel<false><<<dim3(Y, Z), kMcBlock, 0, st>>>(voxels, X, Y, Z,
                                                        make_float3(vp->origin[0], vp->origin[1], vp->origin[2]),
                                                        vp->voxel_size, nullptr, rowcnt, rowoff, max_triangles,
                                                        vertices, normals, triangles, nullptr);
    NRF_LAUNCH_CHECK();
  }
  return NRF_OK;
}
