// raster.hip - live surface prediction for gfx950: the canonical mesh
// skinned through the voxel kNN (src/skinning_kernels.cu:29-63) and
// rasterised into per-pixel live vertices, normals and canonical hit voxels
// (src/mesh_raster_kernels.cu:76-200), the input of the projective
// association.
//
// Rasterisation keeps the reference's two passes: every triangle writes
// atomicMin(depth bits << 32 | triangle) into the pixels its barycentric
// test accepts (a deterministic result: the nearest, then the lowest
// triangle index), then one lane per pixel resolves the winner's
// interpolated live vertex, face normal and canonical point.  One lane per
// triangle in the first pass (the full-scale mesh has ~180k triangles of a
// few pixels each); FP32 in the reference's operation order.
#include <algorithm>

#include "nrf_internal.h"
#include "nrf_math.h"

namespace nrf {
namespace {

__device__ inline float3 ld3(const float *p) { return make_float3(p[0], p[1], p[2]); }
__device__ inline void st3(float *p, float3 v) {
  p[0] = v.x;
  p[1] = v.y;
  p[2] = v.z;
}
__device__ inline float norm3(float3 v) { return sqrtf(v.x * v.x + v.y * v.y + v.z * v.z); }
__device__ inline float3 normalize3(float3 v) {
  const float n = norm3(v);
  return n > 1e-9f ? make_float3(v.x / n, v.y / n, v.z / n) : v;
}
__device__ inline float edge_fn(float2 a, float2 b, float2 p) {
  return (p.x - a.x) * (b.y - a.y) - (p.y - a.y) * (b.x - a.x);
}
// float -> int of a rounded value, saturated (the reference's __float2int_rd /
// __float2int_ru saturate; a plain conversion of an out-of-range value is
// undefined)
__device__ inline int sat_int(float v) { return (int)fminf(fmaxf(v, -1073741824.f), 1073741824.f); }
// the voxel containing p, or -1 outside the volume (floor of the division)
__device__ inline int voxel_of(float3 p, float3 org, float s, int X, int Y, int Z) {
  const int vx = sat_int(floorf((p.x - org.x) / s)), vy = sat_int(floorf((p.y - org.y) / s)),
            vz = sat_int(floorf((p.z - org.z) / s));
  if (vx < 0 || vx >= X || vy < 0 || vy >= Y || vz < 0 || vz >= Z) return -1;
  return (vz * Y + vy) * X + vx;
}

struct RasterArgs {
  const float *canon, *warped;
  const int32_t *tris;
  int nv, nt;
  const int32_t *d_nt;  // optional: the triangle count on the device (at most nt; vertices 3 per triangle)
  nrf_camera cam;
  CamXf Tcw;
  float3 org;
  float s;
  int X, Y, Z;
  unsigned long long *ztri;
  float *live_v, *live_n, *canon_pt;
  int32_t *hit;
  const int32_t *skip;  // optional: nonzero = leave every output as it is
};

// project_triangle (src/mesh_raster_kernels.cu:44-72)
__device__ inline bool project_tri(const RasterArgs &a, const int ids[3], float3 wc[3], float3 cc[3], float2 p[3],
                                   float &area) {
  for (int k = 0; k < 3; ++k) {
    wc[k] = ld3(a.warped + 3 * (size_t)ids[k]);
    const CamXf &T = a.Tcw;
    cc[k] = make_float3(T.m[0][0] * wc[k].x + T.m[0][1] * wc[k].y + T.m[0][2] * wc[k].z + T.m[0][3],
                        T.m[1][0] * wc[k].x + T.m[1][1] * wc[k].y + T.m[1][2] * wc[k].z + T.m[1][3],
                        T.m[2][0] * wc[k].x + T.m[2][1] * wc[k].y + T.m[2][2] * wc[k].z + T.m[2][3]);
  }
  if (cc[0].z <= 0.0f || cc[1].z <= 0.0f || cc[2].z <= 0.0f) return false;
  for (int k = 0; k < 3; ++k)
    p[k] = make_float2(a.cam.fx * cc[k].x / cc[k].z + a.cam.cx, a.cam.fy * cc[k].y / cc[k].z + a.cam.cy);
  area = edge_fn(p[0], p[1], p[2]);
  return fabsf(area) >= 1e-6f;
}

__device__ inline int tri_count(const RasterArgs &a) { return a.d_nt ? min(max(*a.d_nt, 0), a.nt) : a.nt; }
__device__ inline int vert_count(const RasterArgs &a) { return a.d_nt ? 3 * tri_count(a) : a.nv; }
__device__ inline bool tri_ids(const RasterArgs &a, int t, int ids[3]) {
  const int nv = vert_count(a);
  for (int k = 0; k < 3; ++k) ids[k] = a.tris[3 * (size_t)t + k];
  return !(ids[0] < 0 || ids[1] < 0 || ids[2] < 0 || ids[0] >= nv || ids[1] >= nv || ids[2] >= nv);
}

// skin_vertices_from_voxel_knn_kernel (src/skinning_kernels.cu:29-63): the
// vertex warped through the 4-NN of the voxel containing it (the centred
// node table gives the same blend as warp_point), unchanged outside the
// volume or without nodes.
// Grid-stride over the vertices, so that a launch sized for a capacity
// serves a count held on the device.
__global__ __launch_bounds__(256) void skin_kernel(const float *__restrict__ canon, int n, float3 org, float s,
                                                   int X, int Y, int Z, const DQ *__restrict__ centered, int N,
                                                   const int4 *__restrict__ knn, const float4 *__restrict__ knn_w,
                                                   float *__restrict__ out, const int32_t *__restrict__ d_ntri,
                                                   const int32_t *__restrict__ skip) {
  if (skip && *skip) return;
  const int cnt = d_ntri ? 3 * min(max(*d_ntri, 0), n / 3) : n;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < cnt; i += gridDim.x * 256) {
    float3 p = ld3(canon + 3 * (size_t)i);
    if (N > 0 && centered) {
      const int v = voxel_of(p, org, s, X, Y, Z);
      if (v >= 0) {
        const int4 i4 = knn[v];
        const float4 w4 = knn_w[v];
        int ids[4] = {i4.x, i4.y, i4.z, i4.w};
        const float ws[4] = {w4.x, w4.y, w4.z, w4.w};
        for (int k = 0; k < 4; ++k)
          if (ids[k] >= N) ids[k] = -1;
        DQ b;
        if (dq_blend4(centered, ids, ws, b)) p = dq_transform(b, p);
      }
    }
    st3(out + 3 * (size_t)i, p);
  }
}

__global__ __launch_bounds__(256) void raster_reset_kernel(RasterArgs a) {
  if (a.skip && *a.skip) return;
  const int px = blockIdx.x * 256 + threadIdx.x;
  if (px >= a.cam.width * a.cam.height) return;
  a.ztri[px] = ~0ull;
  st3(a.live_v + 3 * (size_t)px, make_float3(0.f, 0.f, 0.f));
  st3(a.live_n + 3 * (size_t)px, make_float3(0.f, 0.f, 0.f));
  if (a.canon_pt) st3(a.canon_pt + 3 * (size_t)px, make_float3(0.f, 0.f, 0.f));
  a.hit[px] = -1;
}

// rasterize_triangles_z_kernel (src/mesh_raster_kernels.cu:76-137)
__device__ inline void raster_z_tri(const RasterArgs &a, int t) {
  int ids[3];
  if (!tri_ids(a, t, ids)) return;
  float3 wc[3], cc[3];
  float2 p[3];
  float area = 0.f;
  if (!project_tri(a, ids, wc, cc, p, area)) return;
  const int min_u = max(0, sat_int(floorf(fminf(p[0].x, fminf(p[1].x, p[2].x)))));
  const int max_u = min(a.cam.width - 1, sat_int(ceilf(fmaxf(p[0].x, fmaxf(p[1].x, p[2].x)))));
  const int min_v = max(0, sat_int(floorf(fminf(p[0].y, fminf(p[1].y, p[2].y)))));
  const int max_v = min(a.cam.height - 1, sat_int(ceilf(fmaxf(p[0].y, fmaxf(p[1].y, p[2].y)))));
  if (min_u > max_u || min_v > max_v) return;
  const float3 fn = normalize3(cross(sub(wc[1], wc[0]), sub(wc[2], wc[0])));
  if (norm3(fn) < 1e-6f) return;
  for (int y = min_v; y <= max_v; ++y)
    for (int x = min_u; x <= max_u; ++x) {
      const float2 q = make_float2((float)x + 0.5f, (float)y + 0.5f);
      const float w0 = edge_fn(p[1], p[2], q) / area;
      const float w1 = edge_fn(p[2], p[0], q) / area;
      const float w2 = edge_fn(p[0], p[1], q) / area;
      if (w0 < -1e-5f || w1 < -1e-5f || w2 < -1e-5f) continue;
      const float z = w0 * cc[0].z + w1 * cc[1].z + w2 * cc[2].z;
      if (z <= 0.0f || !isfinite(z)) continue;
      const unsigned long long key = ((unsigned long long)__float_as_uint(z) << 32) | (unsigned)t;
      atomicMin(&a.ztri[(size_t)y * a.cam.width + x], key);
    }
}
// grid-stride over the triangles (see skin_kernel)
__global__ __launch_bounds__(256) void raster_z_kernel(RasterArgs a) {
  if (a.skip && *a.skip) return;
  const int nt = tri_count(a);
  for (int t = blockIdx.x * 256 + threadIdx.x; t < nt; t += gridDim.x * 256) raster_z_tri(a, t);
}

// resolve_rasterized_triangles_kernel (src/mesh_raster_kernels.cu:139-200),
// and the interpolated canonical point itself when canon_pt is set.
__global__ __launch_bounds__(256) void raster_resolve_kernel(RasterArgs a) {
  if (a.skip && *a.skip) return;
  const int px = blockIdx.x * 256 + threadIdx.x;
  if (px >= a.cam.width * a.cam.height) return;
  const unsigned long long key = a.ztri[px];
  if (key == ~0ull) return;
  const int t = (int)(key & 0xffffffffull);
  int ids[3];
  if (!tri_ids(a, t, ids)) return;
  float3 wc[3], cc[3];
  float2 p[3];
  float area = 0.f;
  if (!project_tri(a, ids, wc, cc, p, area)) return;
  const int x = px % a.cam.width, y = px / a.cam.width;
  const float2 q = make_float2((float)x + 0.5f, (float)y + 0.5f);
  const float w0 = edge_fn(p[1], p[2], q) / area;
  const float w1 = edge_fn(p[2], p[0], q) / area;
  const float w2 = edge_fn(p[0], p[1], q) / area;
  if (w0 < -1e-5f || w1 < -1e-5f || w2 < -1e-5f) return;
  st3(a.live_v + 3 * (size_t)px,
      make_float3(w0 * wc[0].x + w1 * wc[1].x + w2 * wc[2].x, w0 * wc[0].y + w1 * wc[1].y + w2 * wc[2].y,
                  w0 * wc[0].z + w1 * wc[1].z + w2 * wc[2].z));
  st3(a.live_n + 3 * (size_t)px, normalize3(cross(sub(wc[1], wc[0]), sub(wc[2], wc[0]))));
  const float3 c0 = ld3(a.canon + 3 * (size_t)ids[0]), c1 = ld3(a.canon + 3 * (size_t)ids[1]),
               c2 = ld3(a.canon + 3 * (size_t)ids[2]);
  const float3 cp = make_float3(w0 * c0.x + w1 * c1.x + w2 * c2.x, w0 * c0.y + w1 * c1.y + w2 * c2.y,
                                w0 * c0.z + w1 * c1.z + w2 * c2.z);
  if (a.canon_pt) st3(a.canon_pt + 3 * (size_t)px, cp);
  a.hit[px] = voxel_of(cp, a.org, a.s, a.X, a.Y, a.Z);
}

// blocks of a grid-stride launch over n items: one item per lane up to
// 4096 workgroups (16 per CU)
inline int grid_for(int n) { return std::max(1, std::min((n + 255) / 256, 4096)); }

}  // namespace

int skin_vertices(const float *canonical, int n, const nrf_volume_params *vp, const nrf_dualquat *centered,
                  int num_nodes, const int32_t *voxel_knn, const float *voxel_knn_w, float *out, hipStream_t st,
                  const int32_t *d_ntri, const int32_t *skip) {
  if (!volume_ok(vp) || n < 0 || num_nodes < 0) return NRF_ERR_INVALID_ARGUMENT;
  if (n == 0) return NRF_OK;
  if (!canonical || !out || (num_nodes > 0 && (!centered || !voxel_knn || !voxel_knn_w)))
    return NRF_ERR_INVALID_ARGUMENT;
  skin_kernel<<<grid_for(n), 256, 0, st>>>(canonical, n, make_float3(vp->origin[0], vp->origin[1], vp->origin[2]),
                                               vp->voxel_size, vp->dims[0], vp->dims[1], vp->dims[2],
                                               reinterpret_cast<const DQ *>(centered), num_nodes,
                                               reinterpret_cast<const int4 *>(voxel_knn),
                                               reinterpret_cast<const float4 *>(voxel_knn_w), out, d_ntri, skip);
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}

int rasterize_mesh(const float *canonical, const float *warped, const int32_t *triangles, int num_vertices,
                   int num_triangles, const nrf_camera *cam, const float camera_pose[16], const nrf_volume_params *vp,
                   unsigned long long *ztri, float *live_vertices, float *live_normals, int32_t *hit_voxel,
                   float *canonical_points, hipStream_t st, const int32_t *d_ntri, const int32_t *skip) {
  if (!cam || !camera_pose || !volume_ok(vp) || !ztri || !live_vertices || !live_normals || !hit_voxel ||
      num_vertices < 0 || num_triangles < 0 || cam->width <= 0 || cam->height <= 0)
    return NRF_ERR_INVALID_ARGUMENT;
  if (num_triangles > 0 && (!canonical || !warped || !triangles)) return NRF_ERR_INVALID_ARGUMENT;
  RasterArgs a;
  a.canon = canonical;
  a.warped = warped;
  a.tris = triangles;
  a.nv = num_vertices;
  a.nt = num_triangles;
  a.d_nt = d_ntri;
  a.cam = *cam;
  a.Tcw = invert_pose(camera_pose);
  a.org = make_float3(vp->origin[0], vp->origin[1], vp->origin[2]);
  a.s = vp->voxel_size;
  a.X = vp->dims[0]; a.Y = vp->dims[1]; a.Z = vp->dims[2];
  a.ztri = ztri;
  a.live_v = live_vertices;
  a.live_n = live_normals;
  a.canon_pt = canonical_points;
  a.hit = hit_voxel;
  a.skip = skip;
  const int npx = cam->width * cam->height;
  raster_reset_kernel<<<(npx + 255) / 256, 256, 0, st>>>(a);
  NRF_LAUNCH_CHECK();
  if (num_triangles == 0 || num_vertices == 0) return NRF_OK;
  raster_z_kernel<<<grid_for(num_triangles), 256, 0, st>>>(a);
  NRF_LAUNCH_CHECK();
  raster_resolve_kernel<<<(npx + 255) / 256, 256, 0, st>>>(a);
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}

}  // namespace nrf

using namespace nrf;

extern "C" int nrf_skin_vertices(const float *canonical_vertices, int num_vertices, const nrf_volume_params *vp,
                                 const nrf_dualquat *centered, int num_nodes, const int32_t *voxel_knn,
                                 const float *voxel_knn_w, float *warped_vertices, void *stream) {
  return skin_vertices(canonical_vertices, num_vertices, vp, centered, num_nodes, voxel_knn, voxel_knn_w,
                       warped_vertices, as_stream(stream));
}

extern "C" size_t nrf_raster_workspace_size(const nrf_camera *cam) {
  if (!cam || cam->width <= 0 || cam->height <= 0) return 0;
  return (size_t)cam->width * cam->height * sizeof(unsigned long long);
}

extern "C" int nrf_rasterize_mesh(const float *canonical_vertices, const float *warped_vertices,
                                  const int32_t *triangles, int num_vertices, int num_triangles,
                                  const nrf_camera *cam, const float camera_pose[16], const nrf_volume_params *vp,
                                  void *workspace, size_t workspace_bytes, float *live_vertices, float *live_normals,
                                  int32_t *hit_voxel, float *canonical_points, void *stream) {
  if (!workspace) return NRF_ERR_INVALID_ARGUMENT;
  if (workspace_bytes < nrf_raster_workspace_size(cam)) return NRF_ERR_WORKSPACE;
  return rasterize_mesh(canonical_vertices, warped_vertices, triangles, num_vertices, num_triangles, cam, camera_pose,
                        vp, reinterpret_cast<unsigned long long *>(workspace), live_vertices, live_normals, hit_voxel,
                        canonical_points, as_stream(stream));
}
