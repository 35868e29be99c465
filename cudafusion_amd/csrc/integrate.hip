// integrate.hip - DQ-warped TSDF integration for gfx950.
//
// Semantics follow tsdf_integrate_kernel (kernels/tsdf_kernels.cu:85-174)
// and TSDFVolume::integrate (src/tsdf_volume.cu:23-63): voxel centre at
// (v + 0.5) s, optional DQ warp through the voxel's 4-NN, projection with
// round-half-to-even, running average with weight decay and a weight cap.
//
// Layout and mapping: the reference's 12-byte AoS voxels and [V][4] kNN
// arrays are kept (they are the API).  A workgroup of 256 lanes covers a
// 64 (x) x 4 (y) tile and walks ZC voxels down z, so each wave issues
// contiguous 768-byte voxel rows and 1 KiB kNN rows (int4/float4 per lane)
// and every lane has ZC independent voxels in flight.  The per-voxel DQ
// blend reads a per-node table of centred transforms (N x 32 B, resident in
// L2) instead of recomputing dq_centered four times per voxel.
#include <hip/hip_fp16.h>

#include <vector>

#include "nrf_internal.h"
#include "nrf_math.h"

namespace nrf {

namespace {

constexpr int kTileX = 64;
constexpr int kTileY = 4;
constexpr int kZChunk = 16;  // 8 and 32 measured slower on the 256^3 body volume

struct IntegrateArgs {
  nrf_voxel *vox;
  int X, Y, Z;
  float ox, oy, oz, s, trunc;
  const float *depth;
  int W, H;
  float fx, fy, cx, cy;
  CamXf T;
  const CamXf *Tdev;  // optional: the world-to-camera transform read from device memory instead of T
  const nrf_node *nodes;
  const DQ *transforms;
  const DQ *centered;
  int N;
  const int4 *knn;
  const float4 *knn_w;
  const uint32_t *knn_rows;  // row occupancy of knn, or null
  int nsx;                   // runs of 64 voxels per x row
  const int32_t *opt;
  int min_opt;
  float decay, wmax;
  int32_t *updated;
  uint8_t *occ;  // marching-cubes brick occupancy: set for every brick holding an updated voxel, or null
  const float *dmax;  // depth block maxima (depth_block_max_kernel) for the wave cull, or null
  int dmw;            // blocks per image row
};
static_assert(kTileX == kMcBrickX && kTileY == kMcBrickY, "a tile's waves lie in one column of bricks");

// Storage of the volume: the reference's 12-byte AoS voxels, or the compact
// form (FP16 value + uint8 weight, the update rounded once when stored:
// FP16 round-to-nearest-even, weight to the nearest integer, capped at 255).
struct VoxAoS {
  nrf_voxel *v;
  __device__ float2 load(int idx) const {
    const nrf_voxel &x = v[idx];
    return make_float2(x.tsdf, x.weight);
  }
  __device__ void store(int idx, float tsdf, float weight) const {
    nrf_voxel &x = v[idx];
    x.tsdf = tsdf;
    x.weight = weight;
  }
};
struct VoxCompact {
  __half *t;
  uint8_t *w;
  __device__ float2 load(int idx) const { return make_float2(__half2float(t[idx]), (float)w[idx]); }
  __device__ void store(int idx, float tsdf, float weight) const {
    t[idx] = __float2half_rn(tsdf);
    w[idx] = (uint8_t)fminf(rintf(weight), 255.f);
  }
};
// Storage of the voxel kNN: int32 ids + FP32 weights, or uint16 ids
// (0xffff = none) + FP16 weights, decoded exactly.
struct KnnF32 {
  const int4 *id;
  const float4 *w;
  __device__ int4 ids(int idx) const { return id[idx]; }
  __device__ float4 ws(int idx) const { return w[idx]; }
};
struct Knn16 {
  const ushort4 *id;
  const uint2 *w;
  __device__ int4 ids(int idx) const {
    const ushort4 k = id[idx];
    auto d = [](unsigned short v) { return v == 0xffff ? -1 : (int)v; };
    return make_int4(d(k.x), d(k.y), d(k.z), d(k.w));
  }
  __device__ float4 ws(int idx) const {
    const uint2 u = w[idx];
    const float2 a = __half22float2(*reinterpret_cast<const __half2 *>(&u.x));
    const float2 b = __half22float2(*reinterpret_cast<const __half2 *>(&u.y));
    return make_float4(a.x, a.y, b.x, b.y);
  }
};

// Depth block maxima for the wave cull: one float per kDepthBlock x
// kDepthBlock pixels, the largest depth the integration can use there (d > 0,
// or NaN, which passes the integration's tests and counts as +inf), -1 when
// the block holds none.
constexpr int kDepthBlock = 8;
// One lane per pixel row of a block (8 pixels), the block's 8 rows on 8
// adjacent lanes, reduced by lane moves.
__global__ __launch_bounds__(256) void depth_block_max_kernel(const float *__restrict__ depth, int W, int H, int bw,
                                                              int bh, float *__restrict__ dmax) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = t / kDepthBlock, r = t % kDepthBlock;
  float m = -1.f;
  if (b < bw * bh) {
    const int u0 = (b % bw) * kDepthBlock, v = (b / bw) * kDepthBlock + r;
    if (v < H) {
      const float *row = depth + (size_t)v * W;
      float d[kDepthBlock];
#pragma unroll
      for (int du = 0; du < kDepthBlock; ++du) d[du] = (u0 + du < W) ? row[u0 + du] : 0.f;
#pragma unroll
      for (int du = 0; du < kDepthBlock; ++du)
        if (!(d[du] <= 0.f)) m = fmaxf(m, d[du] != d[du] ? INFINITY : d[du]);
    }
  }
#pragma unroll
  for (int off = 1; off < kDepthBlock; off <<= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
  if (b < bw * bh && r == 0) dmax[b] = m;
}

// True when no voxel of the wave's set (the 64 voxels along x of row y in
// z slices z0 .. z0 + ZC - 1) can be updated: all of them integrate unwarped
// (no kNN id in their runs), the rectangle of their centres lies in front of
// the camera, and every pixel a centre can round to (the image bounds of the
// rectangle's corners, which bound every centre's projection, widened by
// 2 px) holds no usable depth or one nearer than the rectangle by more than
// the truncation (plus 0.1 mm, far above the rounding of the per-voxel
// arithmetic).  Every centre then fails the image, d > 0 or sdf >= -trunc
// test.  Wave-uniform (every lane takes the reduced values).
template <bool WARP, int ZC>
__device__ bool wave_culled(const IntegrateArgs &a, const CamXf &T, int bx, int y, int z0) {
  const int lane = threadIdx.x & 63;
  if (y >= a.Y) return true;
  if (WARP) {
    if (!a.knn_rows) return false;
    const bool obs = lane < ZC && z0 + lane < a.Z && a.knn_rows[((z0 + lane) * a.Y + y) * a.nsx + bx] != 0u;
    if (__ballot(obs) != 0ull) return false;
  }
  const int xa = bx * kTileX, xb = min(a.X, xa + kTileX) - 1, zb = min(a.Z, z0 + ZC) - 1;
  float czmin = INFINITY, umin = INFINITY, umax = -INFINITY, vmin = INFINITY, vmax = -INFINITY;
  if (lane < 4) {
    const float3 p = f3(a.ox + (((lane & 1) ? xb : xa) + 0.5f) * a.s, a.oy + (y + 0.5f) * a.s,
                        a.oz + (((lane & 2) ? zb : z0) + 0.5f) * a.s);
    const float cxp = T.m[0][0] * p.x + T.m[0][1] * p.y + T.m[0][2] * p.z + T.m[0][3];
    const float cyp = T.m[1][0] * p.x + T.m[1][1] * p.y + T.m[1][2] * p.z + T.m[1][3];
    const float czp = T.m[2][0] * p.x + T.m[2][1] * p.y + T.m[2][2] * p.z + T.m[2][3];
    czmin = czp;
    umin = umax = a.fx * cxp / czp + a.cx;
    vmin = vmax = a.fy * cyp / czp + a.cy;
  }
#pragma unroll
  for (int off = 1; off < 4; off <<= 1) {  // lanes 0..3, then broadcast
    czmin = fminf(czmin, __shfl_xor(czmin, off, 64));
    umin = fminf(umin, __shfl_xor(umin, off, 64));
    umax = fmaxf(umax, __shfl_xor(umax, off, 64));
    vmin = fminf(vmin, __shfl_xor(vmin, off, 64));
    vmax = fmaxf(vmax, __shfl_xor(vmax, off, 64));
  }
  czmin = __shfl(czmin, 0, 64);
  umin = __shfl(umin, 0, 64);
  umax = __shfl(umax, 0, 64);
  vmin = __shfl(vmin, 0, 64);
  vmax = __shfl(vmax, 0, 64);
  // near the camera plane, or projections far outside any image: no cull
  if (!(czmin > 1e-3f) || !(umin > -1e6f && umax < 1e6f && vmin > -1e6f && vmax < 1e6f)) return false;
  const int u0 = max(0, (int)floorf(umin) - 2), u1 = min(a.W - 1, (int)ceilf(umax) + 2);
  const int v0 = max(0, (int)floorf(vmin) - 2), v1 = min(a.H - 1, (int)ceilf(vmax) + 2);
  if (u0 > u1 || v0 > v1) return true;  // every centre projects outside the image
  const int bu0 = u0 / kDepthBlock, bv0 = v0 / kDepthBlock;
  const int nbu = u1 / kDepthBlock - bu0 + 1, nb = nbu * (v1 / kDepthBlock - bv0 + 1);
  float m = -1.f;
  for (int j = lane; j < nb; j += 64) m = fmaxf(m, a.dmax[(bv0 + j / nbu) * a.dmw + bu0 + j % nbu]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
  return m < 0.f || m < czmin - a.trunc - 1e-4f;
}

__device__ inline DQ fetch_centered(const IntegrateArgs &a, int id, bool table) {
  if (table) return a.centered[id];
  const nrf_node &n = a.nodes[id];
  return dq_centered(a.transforms[id], f3(n.pos));
}

// A lane's ZC voxels (one column along z) go through three stages so that
// each stage's loads are in flight together: (1) warp and projection, one
// voxel after another (the kNN of a run without an observed voxel is not
// read: those voxels integrate unwarped, as with all-(-1) ids), leaving the
// pixel and camera depth of each voxel; (2) the depth pixels; (3) the
// voxels that pass the tests, then their updates.  Same arithmetic per
// voxel as one voxel at a time.  Vox / Knn: the storage of the volume and of
// the kNN (the same arithmetic on the decoded values).
template <bool WARP, bool TABLE, int ZC, class Vox, class Knn>
__device__ void integrate_tile(const IntegrateArgs &a, const Vox &vs, const Knn &kn, int bx, int by, int bz) {
  const int x = bx * kTileX + (threadIdx.x & (kTileX - 1));
  const int y = by * kTileY + (threadIdx.x / kTileX);
  const int z0 = bz * ZC;
  int n_upd = 0;
  unsigned occ_bits = 0;  // bit k: a voxel of brick z0 / 4 + k was updated
  const CamXf T = a.Tdev ? *a.Tdev : a.T;
  if (a.dmax && wave_culled<WARP, ZC>(a, T, bx, y, z0)) return;
  if (x < a.X && y < a.Y) {
    const float px_w = a.ox + (x + 0.5f) * a.s;
    const float py_w = a.oy + (y + 0.5f) * a.s;
    const int plane = a.X * a.Y, col = y * a.X + x;
    int pix[ZC];
    float czs[ZC];
#pragma unroll
    for (int dz = 0; dz < ZC; ++dz) {
      pix[dz] = -1;
      czs[dz] = 0.f;
    }
#pragma unroll
    for (int dz = 0; dz < ZC; ++dz) {
      const int z = z0 + dz;
      if (z >= a.Z) break;
      const int idx = z * plane + col;
      if (WARP && a.opt && a.opt[idx] < a.min_opt) continue;
      float3 p = f3(px_w, py_w, a.oz + (z + 0.5f) * a.s);
      if (WARP) {
        // a run without any id >= 0 (its word is zero) is read as all -1:
        // the same unwarped path without the 16-byte read per voxel
        int4 id4 = make_int4(-1, -1, -1, -1);
        const int run = __builtin_amdgcn_readfirstlane((z * a.Y + y) * a.nsx + bx);  // one per wave
        if (!a.knn_rows || a.knn_rows[run] != 0u) id4 = kn.ids(idx);
        if ((id4.x & id4.y & id4.z & id4.w) >= 0) {  // sign bit clear: some id >= 0
          const float4 w4 = kn.ws(idx);
          const int ids[4] = {id4.x, id4.y, id4.z, id4.w};
          const float ws[4] = {w4.x, w4.y, w4.z, w4.w};
          DQ b;
          b.r = make_float4(0.f, 0.f, 0.f, 0.f);
          b.d = make_float4(0.f, 0.f, 0.f, 0.f);
          float wsum = 0.f;
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const int id = ids[k];
            const float w = ws[k];
            if (id < 0 || id >= a.N || w < 1e-8f) continue;
            DQ q = fetch_centered(a, id, TABLE);
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
      }
      const float cxp = T.m[0][0] * p.x + T.m[0][1] * p.y + T.m[0][2] * p.z + T.m[0][3];
      const float cyp = T.m[1][0] * p.x + T.m[1][1] * p.y + T.m[1][2] * p.z + T.m[1][3];
      const float czp = T.m[2][0] * p.x + T.m[2][1] * p.y + T.m[2][2] * p.z + T.m[2][3];
      if (czp <= 0) continue;
      const float fu = a.fx * cxp / czp + a.cx;
      const float fv = a.fy * cyp / czp + a.cy;
      // round half to even; the range test only keeps the conversion defined
      if (!(fu > -1.0e9f && fu < 1.0e9f && fv > -1.0e9f && fv < 1.0e9f)) continue;
      const int u = (int)rintf(fu);
      const int v = (int)rintf(fv);
      if (u < 0 || u >= a.W || v < 0 || v >= a.H) continue;
      pix[dz] = v * a.W + u;
      czs[dz] = czp;
    }
    float tv[ZC];
#pragma unroll
    for (int dz = 0; dz < ZC; ++dz) tv[dz] = pix[dz] >= 0 ? a.depth[pix[dz]] : 0.f;
#pragma unroll
    for (int dz = 0; dz < ZC; ++dz) {  // depth -> t, or NaN when the voxel is not updated
      const float d = tv[dz];
      const float sdf = d - czs[dz];
      const bool upd = pix[dz] >= 0 && !(d <= 0) && !(sdf < -a.trunc);
      tv[dz] = upd ? fminf(1.f, sdf / a.trunc) : __int_as_float(0x7fc00000);
      pix[dz] = upd ? 1 : 0;
    }
    float2 tw[ZC];
#pragma unroll
    for (int dz = 0; dz < ZC; ++dz) {
      tw[dz] = make_float2(0.f, 0.f);
      if (pix[dz]) tw[dz] = vs.load((z0 + dz) * plane + col);
    }
#pragma unroll
    for (int dz = 0; dz < ZC; ++dz) {
      if (!pix[dz]) continue;
      const float dw = tw[dz].y * a.decay;
      const float nw = dw + 1.f;
      vs.store((z0 + dz) * plane + col, (tw[dz].x * dw + tv[dz] * 1.f) / nw, fminf(nw, a.wmax));
      ++n_upd;
      occ_bits |= 1u << (((z0 + dz) / kMcBrickZ) - z0 / kMcBrickZ);
    }
  }
  if (a.occ) {
    // the wave's 64 voxels along x at one y share their bricks
    for (int off = 32; off > 0; off >>= 1) occ_bits |= __shfl_xor(occ_bits, off, 64);
    if ((threadIdx.x & 63) == 0 && occ_bits && y < a.Y) {
      for (int k = 0; occ_bits >> k; ++k)
        if ((occ_bits >> k) & 1u) a.occ[mc_brick_of(bx * kTileX, y, (z0 / kMcBrickZ + k) * kMcBrickZ, a.X, a.Y)] = 1;
    }
  }
  if (a.updated) {
    // wave-level sum, one atomic per wave
    for (int off = 32; off > 0; off >>= 1) n_upd += __shfl_xor(n_upd, off, 64);
    if ((threadIdx.x & 63) == 0 && n_upd) atomicAdd(a.updated, n_upd);
  }
}
template <bool WARP, bool TABLE>
__global__ __launch_bounds__(256) void integrate_kernel(IntegrateArgs a) {
  integrate_tile<WARP, TABLE, kZChunk>(a, VoxAoS{a.vox}, KnnF32{a.knn, a.knn_w}, blockIdx.x, blockIdx.y,
                                       blockIdx.z);
}
// The compact form (warped through the centred table when a.N > 0).
template <bool WARP>
__global__ __launch_bounds__(256) void integrate16_kernel(IntegrateArgs a, VoxCompact vs, Knn16 kn) {
  integrate_tile<WARP, true, kZChunk>(a, vs, kn, blockIdx.x, blockIdx.y, blockIdx.z);
}

// All instances of a batch in one launch.  Instance k owns the tiles
// [tile_base[k], tile_base[k + 1]) of the 1-D grid, numbered x fastest, then
// y, then z as in the single-volume grid; depth and camera transform are the
// frame's (kernel arguments), everything else the instance's record.
struct IntBatchDev {
  IntegrateArgs a;
  int tiles_x, tiles_y;
};
template <int ZC>
__global__ __launch_bounds__(256) void integrate_batch_kernel(const IntBatchDev *__restrict__ inst,
                                                              const int32_t *__restrict__ tile_base, int count,
                                                              BatchDepth fr) {
  const int t = blockIdx.x;
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tile_base[mid] <= t) lo = mid; else hi = mid - 1;
  }
  const IntBatchDev &d = inst[lo];
  IntegrateArgs a = d.a;
  a.depth = fr.depth[lo];
  a.T = fr.xf[lo];
  a.Tdev = nullptr;
  const int l = t - tile_base[lo];
  const int bx = l % d.tiles_x, r = l / d.tiles_x;
  integrate_tile<true, true, ZC>(a, VoxAoS{a.vox}, KnnF32{a.knn, a.knn_w}, bx, r % d.tiles_y, r / d.tiles_y);
}

// Row occupancy of a kNN: four waves per workgroup, one run of 64 voxels per
// wave.
__global__ __launch_bounds__(256) void knn_rows_kernel(const int4 *__restrict__ knn, int X, int nsx, int64_t nrows,
                                                       uint32_t *__restrict__ rows) {
  const int64_t r = (int64_t)blockIdx.x * 4 + threadIdx.x / 64;
  if (r >= nrows) return;  // whole waves
  const int lane = threadIdx.x & 63;
  const int x = (int)(r % nsx) * kTileX + lane;
  const int64_t yz = r / nsx;
  bool obs = false;
  if (x < X) {
    const int4 id = knn[yz * X + x];
    obs = (id.x & id.y & id.z & id.w) >= 0;
  }
  const unsigned long long m = __ballot(obs);
  if (lane == 0) rows[r] = m != 0ull ? 1u : 0u;
}

__global__ __launch_bounds__(256) void reset_kernel(nrf_voxel *v, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  nrf_voxel o;
  o.tsdf = 1.f;
  o.weight = 0.f;
  o.color[0] = o.color[1] = o.color[2] = 128;
  o.pad = 0;
  v[i] = o;
}

}  // namespace

}  // namespace nrf

using namespace nrf;

struct nrf::IntegrateBatch {
  int count = 0, z_chunk = 4, tiles = 0;
  IntBatchDev *d_inst = nullptr;
  int32_t *tile_base = nullptr;
};

namespace nrf {

void integrate_batch_destroy(IntegrateBatch *b) {
  if (!b) return;
  if (b->d_inst) (void)hipFree(b->d_inst);
  if (b->tile_base) (void)hipFree(b->tile_base);
  delete b;
}

int integrate_batch_create(const IntegrateBatchInstance *inst, int count, int z_chunk, IntegrateBatch **out) {
  if (!inst || !out || count < 1 || count > kMaxBatch) return NRF_ERR_INVALID_ARGUMENT;
  if (z_chunk != 1 && z_chunk != 2 && z_chunk != 4 && z_chunk != 8 && z_chunk != 16) return NRF_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  std::vector<IntBatchDev> h(count);
  std::vector<int32_t> base(count + 1, 0);
  long tiles = 0;
  for (int k = 0; k < count; ++k) {
    const IntegrateBatchInstance &q = inst[k];
    if (!volume_ok(&q.vp) || !q.vox || q.cam.width <= 0 || q.cam.height <= 0 || q.N <= 0 || !q.nodes || !q.T ||
        !q.centered || !q.knn || !q.knn_w)
      return NRF_ERR_INVALID_ARGUMENT;
    IntegrateArgs &a = h[k].a;
    a.vox = q.vox;
    a.X = q.vp.dims[0]; a.Y = q.vp.dims[1]; a.Z = q.vp.dims[2];
    a.ox = q.vp.origin[0]; a.oy = q.vp.origin[1]; a.oz = q.vp.origin[2];
    a.s = q.vp.voxel_size;
    a.trunc = q.vp.truncation;
    a.depth = nullptr;
    a.W = q.cam.width; a.H = q.cam.height;
    a.fx = q.cam.fx; a.fy = q.cam.fy; a.cx = q.cam.cx; a.cy = q.cam.cy;
    a.nodes = q.nodes;
    a.transforms = reinterpret_cast<const DQ *>(q.T);
    a.centered = reinterpret_cast<const DQ *>(q.centered);
    a.N = q.N;
    a.knn = reinterpret_cast<const int4 *>(q.knn);
    a.knn_w = reinterpret_cast<const float4 *>(q.knn_w);
    a.knn_rows = q.knn_rows;
    a.nsx = (a.X + kTileX - 1) / kTileX;
    a.opt = nullptr;
    a.min_opt = 0;
    a.decay = q.vp.decay_alpha;
    a.wmax = q.vp.max_weight;
    a.updated = q.updated;
    a.occ = q.occ;
    a.dmax = nullptr;  // the wave cull needs each frame's block maxima: not in the batched form
    a.dmw = 0;
    h[k].tiles_x = (a.X + kTileX - 1) / kTileX;
    h[k].tiles_y = (a.Y + kTileY - 1) / kTileY;
    base[k] = (int32_t)tiles;
    tiles += (long)h[k].tiles_x * h[k].tiles_y * ((a.Z + z_chunk - 1) / z_chunk);
    if (tiles > INT32_MAX) return NRF_ERR_TOO_LARGE;
  }
  base[count] = (int32_t)tiles;
  IntegrateBatch *b = new IntegrateBatch();
  b->count = count;
  b->z_chunk = z_chunk;
  b->tiles = (int)tiles;
  bool ok = hipMalloc(&b->d_inst, sizeof(IntBatchDev) * count) == hipSuccess &&
            hipMalloc(&b->tile_base, sizeof(int32_t) * (count + 1)) == hipSuccess &&
            hipMemcpy(b->d_inst, h.data(), sizeof(IntBatchDev) * count, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(b->tile_base, base.data(), sizeof(int32_t) * (count + 1), hipMemcpyHostToDevice) ==
                hipSuccess;
  if (!ok) {
    set_last_error("integrate_batch_create", hipGetLastError());
    integrate_batch_destroy(b);
    return NRF_ERR_HIP;
  }
  *out = b;
  return NRF_OK;
}

int integrate_batch_run(IntegrateBatch *b, const BatchDepth &fr, hipStream_t st) {
  if (!b) return NRF_ERR_INVALID_ARGUMENT;
  switch (b->z_chunk) {
    case 1: integrate_batch_kernel<1><<<b->tiles, 256, 0, st>>>(b->d_inst, b->tile_base, b->count, fr); break;
    case 2: integrate_batch_kernel<2><<<b->tiles, 256, 0, st>>>(b->d_inst, b->tile_base, b->count, fr); break;
    case 8: integrate_batch_kernel<8><<<b->tiles, 256, 0, st>>>(b->d_inst, b->tile_base, b->count, fr); break;
    case 16: integrate_batch_kernel<16><<<b->tiles, 256, 0, st>>>(b->d_inst, b->tile_base, b->count, fr); break;
    default: integrate_batch_kernel<4><<<b->tiles, 256, 0, st>>>(b->d_inst, b->tile_base, b->count, fr); break;
  }
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}

size_t depth_block_count(const nrf_camera *cam) {
  if (cam->width <= 0 || cam->height <= 0) return 0;
  return (size_t)((cam->width + kDepthBlock - 1) / kDepthBlock) * ((cam->height + kDepthBlock - 1) / kDepthBlock);
}

int64_t knn_row_words(const nrf_volume_params *vp) {
  return (int64_t)((vp->dims[0] + kTileX - 1) / kTileX) * vp->dims[1] * vp->dims[2];
}

int knn_row_flags(const nrf_volume_params *vp, const int32_t *knn, uint32_t *rows, hipStream_t st) {
  if (!volume_ok(vp) || !knn || !rows) return NRF_ERR_INVALID_ARGUMENT;
  const int64_t nrows = knn_row_words(vp);
  const int nsx = (vp->dims[0] + kTileX - 1) / kTileX;
  knn_rows_kernel<<<(unsigned)((nrows + 3) / 4), 256, 0, st>>>(reinterpret_cast<const int4 *>(knn), vp->dims[0],
                                                                 nsx, nrows, rows);
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}

__global__ void set_pose_kernel(CamXf *dev, CamXf T) {
  if (threadIdx.x == 0) *dev = T;
}
int set_device_pose(CamXf *dev, const float camera_pose[16], hipStream_t st) {
  if (!dev || !camera_pose) return NRF_ERR_INVALID_ARGUMENT;
  set_pose_kernel<<<1, 64, 0, st>>>(dev, invert_pose(camera_pose));
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}

int tsdf_integrate(const nrf_volume_params *vp, nrf_voxel *voxels, const float *depth, const nrf_camera *cam,
                   const float camera_pose[16], const nrf_node *nodes, const nrf_dualquat *transforms, int num_nodes,
                   const nrf_dualquat *centered, const int32_t *voxel_knn, const float *voxel_knn_w,
                   const uint32_t *knn_rows, const int32_t *voxel_opt_counts, int min_opt_count,
                   int32_t *updated_count, hipStream_t st, const void *dev_pose, uint8_t *occ,
                   float *depth_blocks) {
  if (!volume_ok(vp) || !voxels || !depth || !cam || !camera_pose) return NRF_ERR_INVALID_ARGUMENT;
  if (cam->width <= 0 || cam->height <= 0) return NRF_ERR_INVALID_ARGUMENT;
  const bool warp = nodes != nullptr && num_nodes > 0;
  if (warp && (!transforms || !voxel_knn || !voxel_knn_w)) return NRF_ERR_INVALID_ARGUMENT;
  IntegrateArgs a;
  a.vox = voxels;
  a.X = vp->dims[0]; a.Y = vp->dims[1]; a.Z = vp->dims[2];
  a.ox = vp->origin[0]; a.oy = vp->origin[1]; a.oz = vp->origin[2];
  a.s = vp->voxel_size;
  a.trunc = vp->truncation;
  a.depth = depth;
  a.W = cam->width; a.H = cam->height;
  a.fx = cam->fx; a.fy = cam->fy; a.cx = cam->cx; a.cy = cam->cy;
  a.T = invert_pose(camera_pose);
  a.Tdev = static_cast<const CamXf *>(dev_pose);
  a.nodes = nodes;
  a.transforms = reinterpret_cast<const DQ *>(transforms);
  a.centered = reinterpret_cast<const DQ *>(centered);
  a.N = num_nodes;
  a.knn = reinterpret_cast<const int4 *>(voxel_knn);
  a.knn_w = reinterpret_cast<const float4 *>(voxel_knn_w);
  a.knn_rows = knn_rows;
  a.nsx = (a.X + kTileX - 1) / kTileX;
  a.opt = voxel_opt_counts;
  a.min_opt = min_opt_count;
  a.decay = vp->decay_alpha;
  a.wmax = vp->max_weight;
  a.updated = updated_count;
  a.occ = occ;
  a.dmax = depth_blocks;
  a.dmw = (a.W + kDepthBlock - 1) / kDepthBlock;
  if (depth_blocks) {
    const int bh = (a.H + kDepthBlock - 1) / kDepthBlock;
    depth_block_max_kernel<<<(a.dmw * bh * kDepthBlock + 255) / 256, 256, 0, st>>>(depth, a.W, a.H, a.dmw, bh,
                                                                                   depth_blocks);
    NRF_LAUNCH_CHECK();
  }
  dim3 grid((a.X + kTileX - 1) / kTileX, (a.Y + kTileY - 1) / kTileY, (a.Z + kZChunk - 1) / kZChunk);
  if (!warp)
    integrate_kernel<false, false><<<grid, 256, 0, st>>>(a);
  else if (centered)
    integrate_kernel<true, true><<<grid, 256, 0, st>>>(a);
  else
    integrate_kernel<true, false><<<grid, 256, 0, st>>>(a);
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}

int tsdf16_integrate(const nrf_volume_params *vp, uint16_t *tsdf16, uint8_t *weight8, const float *depth,
                     const nrf_camera *cam, const float camera_pose[16], const nrf_dualquat *centered, int num_nodes,
                     const uint16_t *knn16, const uint16_t *knn_w16, const uint32_t *knn_rows,
                     int32_t *updated_count, hipStream_t st, const int32_t *voxel_opt_counts, int min_opt_count,
                     const void *dev_pose, uint8_t *occ, float *depth_blocks) {
  if (!volume_ok(vp) || !tsdf16 || !weight8 || !depth || !cam || !camera_pose) return NRF_ERR_INVALID_ARGUMENT;
  if (cam->width <= 0 || cam->height <= 0) return NRF_ERR_INVALID_ARGUMENT;
  if (num_nodes > 0 && (!centered || !knn16 || !knn_w16)) return NRF_ERR_INVALID_ARGUMENT;
  if (num_nodes > 65535) return NRF_ERR_TOO_LARGE;
  if (vp->max_weight > 255.f) return NRF_ERR_TOO_LARGE;
  IntegrateArgs a;
  a.vox = nullptr;
  a.X = vp->dims[0]; a.Y = vp->dims[1]; a.Z = vp->dims[2];
  a.ox = vp->origin[0]; a.oy = vp->origin[1]; a.oz = vp->origin[2];
  a.s = vp->voxel_size;
  a.trunc = vp->truncation;
  a.depth = depth;
  a.W = cam->width; a.H = cam->height;
  a.fx = cam->fx; a.fy = cam->fy; a.cx = cam->cx; a.cy = cam->cy;
  a.T = invert_pose(camera_pose);
  a.Tdev = static_cast<const CamXf *>(dev_pose);
  a.nodes = nullptr;
  a.transforms = nullptr;
  a.centered = reinterpret_cast<const DQ *>(centered);
  a.N = num_nodes;
  a.knn = nullptr;
  a.knn_w = nullptr;
  a.knn_rows = knn_rows;
  a.nsx = (a.X + kTileX - 1) / kTileX;
  a.opt = voxel_opt_counts;
  a.min_opt = min_opt_count;
  a.decay = vp->decay_alpha;
  a.wmax = vp->max_weight;
  a.updated = updated_count;
  a.occ = occ;
  a.dmax = depth_blocks;
  a.dmw = (a.W + kDepthBlock - 1) / kDepthBlock;
  if (depth_blocks) {
    const int bh = (a.H + kDepthBlock - 1) / kDepthBlock;
    depth_block_max_kernel<<<(a.dmw * bh * kDepthBlock + 255) / 256, 256, 0, st>>>(depth, a.W, a.H, a.dmw, bh,
                                                                                   depth_blocks);
    NRF_LAUNCH_CHECK();
  }
  const VoxCompact vs{reinterpret_cast<__half *>(tsdf16), weight8};
  const Knn16 kn{reinterpret_cast<const ushort4 *>(knn16), reinterpret_cast<const uint2 *>(knn_w16)};
  dim3 grid((a.X + kTileX - 1) / kTileX, (a.Y + kTileY - 1) / kTileY, (a.Z + kZChunk - 1) / kZChunk);
  if (num_nodes > 0)
    integrate16_kernel<true><<<grid, 256, 0, st>>>(a, vs, kn);
  else
    integrate16_kernel<false><<<grid, 256, 0, st>>>(a, vs, kn);
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}

}  // namespace nrf

extern "C" int nrf_tsdf_integrate(const nrf_volume_params *vp, nrf_voxel *voxels, const float *depth,
                                  const nrf_camera *cam, const float camera_pose[16], const nrf_node *nodes,
                                  const nrf_dualquat *transforms, int num_nodes, const nrf_dualquat *centered,
                                  const int32_t *voxel_knn, const float *voxel_knn_w,
                                  const int32_t *voxel_opt_counts, int min_opt_count, int32_t *updated_count,
                                  void *stream) {
  return tsdf_integrate(vp, voxels, depth, cam, camera_pose, nodes, transforms, num_nodes, centered, voxel_knn,
                        voxel_knn_w, nullptr, voxel_opt_counts, min_opt_count, updated_count, as_stream(stream));
}

extern "C" int nrf_tsdf_reset(const nrf_volume_params *vp, nrf_voxel *voxels, void *stream) {
  if (!volume_ok(vp) || !voxels) return NRF_ERR_INVALID_ARGUMENT;
  int64_t n = volume_count(vp);
  reset_kernel<<<(unsigned)((n + 255) / 256), 256, 0, as_stream(stream)>>>(voxels, n);
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}
