// mc_front.h - the two passes of marching cubes before emission, shared by
// the FP32 volume (mc.hip) and the compact FP16 + u8 volume (precision.hip):
// the case index of every cube with per-row triangle counts, and the
// exclusive scan of the row counts.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "mc_tables.h"

namespace nrf {
namespace {

// Case index of every cube (marching_cubes.cu:323-355 semantics: a corner
// with weight <= 0 voids the cube, bit c is set when tsdf_c > 0 (Q9), cases
// 0 and 255 emit nothing) and the triangle count of every x-row of cubes.
// One workgroup per brick of 64 x 4 x 4 cubes: the brick's 65 x 5 x 5 voxels
// are read once, coalesced along x, into LDS as (tsdf, weight) pairs; wave w
// then classifies cube row y0 + w for the four z-slices.  Reader maps a voxel
// index to its (tsdf, weight).
constexpr int kMcBx = 64, kMcBy = 4, kMcBz = 4;
constexpr int kMcVx = kMcBx + 1, kMcVy = kMcBy + 1, kMcVz = kMcBz + 1;
constexpr int kMcBrickThreads = 64 * kMcBy;

template <class Reader>
__global__ __launch_bounds__(kMcBrickThreads) void classify_brick_kernel(Reader rd, int X, int Y, int Z,
                                                                         uint8_t *__restrict__ cases,
                                                                         int32_t *__restrict__ rowcnt) {
  __shared__ float2 sv[kMcVz][kMcVy][kMcVx];
  const int x0 = blockIdx.x * kMcBx, y0 = blockIdx.y * kMcBy, z0 = blockIdx.z * kMcBz;
  const int t = threadIdx.x;
  for (int j = t; j < kMcVx * kMcVy * kMcVz; j += kMcBrickThreads) {
    const int vx = j % kMcVx, vy = (j / kMcVx) % kMcVy, vz = j / (kMcVx * kMcVy);
    const int x = x0 + vx, y = y0 + vy, z = z0 + vz;
    float2 v = make_float2(0.f, 0.f);  // outside the volume: weight 0
    if (x < X && y < Y && z < Z) v = rd(((size_t)z * Y + y) * X + x);
    sv[vz][vy][vx] = v;
  }
  __syncthreads();
  const int lx = t & 63, ly = t >> 6;
  const int x = x0 + lx, y = y0 + ly;
#pragma unroll
  for (int lz = 0; lz < kMcBz; ++lz) {
    const int z = z0 + lz;
    int cube = 0;
    if (x < X - 1 && y < Y - 1 && z < Z - 1) {
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const float2 v = sv[lz + kMcCorner[c][2]][ly + kMcCorner[c][1]][lx + kMcCorner[c][0]];
        if (v.y <= 0.f) cube = -1;
        if (cube >= 0 && v.x > 0.f) cube |= (1 << c);
      }
      if (cube <= 0 || cube >= 255) cube = 0;
    }
    if (x < X && y < Y && z < Z) cases[((size_t)z * Y + y) * X + x] = (uint8_t)cube;
    int nt = kMcNTri[cube];
    for (int off = 32; off > 0; off >>= 1) nt += __shfl_xor(nt, off, 64);
    if (lx == 0 && nt && y < Y && z < Z) atomicAdd(&rowcnt[z * Y + y], nt);
  }
}

inline dim3 classify_brick_grid(int X, int Y, int Z) {
  return dim3((X + kMcBx - 1) / kMcBx, (Y + kMcBy - 1) / kMcBy, (Z + kMcBz - 1) / kMcBz);
}

// Exclusive scan of n row counts by one workgroup of 1024 lanes.  Lane t
// owns the contiguous chunk [t * per, (t + 1) * per) with per a multiple of
// 4, read and written as 16-byte vectors (the chunk's loads are issued
// together); the chunk sums are scanned in LDS.  cnt and off must be 16-byte
// aligned.  counts (optional) receives {3 T, T} with T the total.
__global__ __launch_bounds__(1024) void scan_rows_kernel(const int32_t *__restrict__ cnt, int32_t *__restrict__ off,
                                                         int n, int32_t *__restrict__ counts) {
  __shared__ int32_t part[1024];
  const int t = threadIdx.x;
  const int per = (((n + 1023) / 1024) + 3) & ~3;
  const int b = min(n, t * per), e = min(n, b + per);
  const int nv = (e - b) >> 2;  // whole vectors; the tail (< 4) only in the last non-empty chunk
  const int4 *c4 = reinterpret_cast<const int4 *>(cnt + b);
  int s = 0;
  for (int k0 = 0; k0 < nv; k0 += 16) {
    int4 v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = (k0 + k < nv) ? c4[k0 + k] : make_int4(0, 0, 0, 0);
#pragma unroll
    for (int k = 0; k < 16; ++k) s += v[k].x + v[k].y + v[k].z + v[k].w;
  }
  for (int i = b + 4 * nv; i < e; ++i) s += cnt[i];
  part[t] = s;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {  // Hillis-Steele inclusive scan
    int v = (t >= d) ? part[t - d] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = part[t] - s;
  int4 *o4 = reinterpret_cast<int4 *>(off + b);
  for (int k0 = 0; k0 < nv; k0 += 16) {
    int4 v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = (k0 + k < nv) ? c4[k0 + k] : make_int4(0, 0, 0, 0);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (k0 + k < nv) {
        int4 w;
        w.x = run; run += v[k].x;
        w.y = run; run += v[k].y;
        w.z = run; run += v[k].z;
        w.w = run; run += v[k].w;
        o4[k0 + k] = w;
      }
    }
  }
  for (int i = b + 4 * nv; i < e; ++i) {
    off[i] = run;
    run += cnt[i];
  }
  if (t == 1023) {
    off[n] = part[1023];
    if (counts) {
      counts[0] = 3 * part[1023];
      counts[1] = part[1023];
    }
  }
}

}  // namespace
}  // namespace nrf
