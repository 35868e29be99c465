// mc_front.h - the two passes of marching cubes before emission, shared by
// the FP32 volume (mc.hip) and the compact FP16 + u8 volume (precision.hip):
// the case index of every cube with per-row triangle counts, and the
// exclusive scan of the row counts.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "mc_tables.h"
#include "nrf_internal.h"

namespace nrf {
namespace {

// Case index of every cube (marching_cubes.cu:323-355 semantics: a corner
// with weight <= 0 voids the cube, bit c is set when tsdf_c > 0 (Q9), cases
// 0 and 255 emit nothing) and the triangle count of every x-row of cubes.
// One workgroup per brick of 64 x 4 x 4 cubes: the brick's 65 x 5 x 5 voxels
// are read once, coalesced along x, into LDS as (tsdf, weight) pairs; wave w
// then classifies cube row y0 + w for the four z-slices.  Reader maps a voxel
// index to its (tsdf, weight).  cases may be null (only the row counts).
//
// Brick occupancy (occ, optional): one byte per brick, brick (bx, by, bz) at
// (bz * nby + by) * nbx + bx, nonzero when some voxel of the brick's own
// 64 x 4 x 4 region may have weight > 0.  Corner 0 of every cube of a brick
// lies in that region, so a zero byte means that every cube of the brick has
// a zero-weight corner: case 0, no triangle.  Such a brick reads no voxel and
// writes no case byte (readers of the case array take occ into account).
constexpr int kMcBx = kMcBrickX, kMcBy = kMcBrickY, kMcBz = kMcBrickZ;
constexpr int kMcVx = kMcBx + 1, kMcVy = kMcBy + 1, kMcVz = kMcBz + 1;
constexpr int kMcBrickThreads = 64 * kMcBy;
static_assert(kMcBx == 64, "one wave per x-row of a brick");

template <class Reader>
__global__ __launch_bounds__(kMcBrickThreads) void classify_brick_kernel(Reader rd, int X, int Y, int Z,
                                                                         uint8_t *__restrict__ cases,
                                                                         int32_t *__restrict__ rowcnt,
                                                                         const uint8_t *__restrict__ occ) {
  __shared__ float2 sv[kMcVz][kMcVy][kMcVx];
  if (occ && occ[((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] == 0) return;
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
    if (cases && x < X && y < Y && z < Z) cases[((size_t)z * Y + y) * X + x] = (uint8_t)cube;
    int nt = kMcNTri[cube];
    for (int off = 32; off > 0; off >>= 1) nt += __shfl_xor(nt, off, 64);
    if (lx == 0 && nt && y < Y && z < Z) atomicAdd(&rowcnt[z * Y + y], nt);
  }
}

inline dim3 classify_brick_grid(int X, int Y, int Z) {
  return dim3(mc_bricks_x(X), mc_bricks_y(Y), (Z + kMcBz - 1) / kMcBz);
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

// Exclusive scan of n row counts in blocks of kScanBlock rows, one
// workgroup of kScanThreads lanes per block (four rows per lane, one 16-byte
// vector): the block scans its rows, publishes its total in status[b] (flag
// in the high word; the status words start zeroed), and adds the totals of
// all earlier blocks, which one wave reads.  Workgroups are dispatched in
// index order and publish before they wait, so every wait ends; the wait is
// bounded all the same: a block that gives up raises the flag word
// status[gridDim.x] and writes counts[1] = -1, and the last block, after
// writing T, reads the flag and writes -1 if it is raised (sequentially
// consistent atomics, so whichever store comes last, -1 stays).  Integer
// sums: the result does not depend on the order of the additions.  cnt and
// off 16-byte aligned; status holds gridDim.x + 1 zeroed words; counts
// (optional) receives {3 T, T}.
constexpr int kScanBlock = 1024, kScanThreads = 256, kScanMaxBlocks = 4096;
constexpr unsigned kScanSpinLimit = 1u << 24;
__global__ __launch_bounds__(kScanThreads) void scan_rows_lb_kernel(const int32_t *__restrict__ cnt,
                                                                    int32_t *__restrict__ off, int n,
                                                                    int32_t *__restrict__ counts,
                                                                    unsigned long long *__restrict__ status) {
  __shared__ int32_t wsum[kScanThreads / 64];
  __shared__ int32_t s_prefix;
  __shared__ int s_bad;
  const int b = blockIdx.x, t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int r0 = b * kScanBlock + 4 * t;
  int4 v = make_int4(0, 0, 0, 0);
  if (r0 + 3 < n) {
    v = *reinterpret_cast<const int4 *>(cnt + r0);
  } else {
    if (r0 < n) v.x = cnt[r0];
    if (r0 + 1 < n) v.y = cnt[r0 + 1];
    if (r0 + 2 < n) v.z = cnt[r0 + 2];
  }
  const int s = v.x + v.y + v.z + v.w;
  int incl = s;  // inclusive scan of the wave
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int u = __shfl_up(incl, d, 64);
    if (lane >= d) incl += u;
  }
  if (lane == 63) wsum[w] = incl;
  if (t == 0) s_bad = 0;
  __syncthreads();
  int wpre = 0, total = 0;
#pragma unroll
  for (int k = 0; k < kScanThreads / 64; ++k) {
    wpre += k < w ? wsum[k] : 0;
    total += wsum[k];
  }
  if (t == 0)
    __hip_atomic_store(&status[b], (1ull << 32) | (unsigned)total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (w == 0) {
    int acc = 0;
    for (int l0 = 0; l0 < b; l0 += 64) {
      const int l = l0 + lane;
      if (l < b) {
        unsigned long long sv = 0;
        unsigned spins = 0;
        for (;;) {
          sv = __hip_atomic_load(&status[l], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (sv >> 32) break;
          if (++spins > kScanSpinLimit) {
            s_bad = 1;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
        acc += (int)(unsigned)sv;
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (lane == 0) s_prefix = acc;
  }
  __syncthreads();
  int run = s_prefix + wpre + incl - s;
  int4 o4;
  o4.x = run; run += v.x;
  o4.y = run; run += v.y;
  o4.z = run; run += v.z;
  o4.w = run;
  if (r0 + 3 < n) {
    *reinterpret_cast<int4 *>(off + r0) = o4;
  } else {
    if (r0 < n) off[r0] = o4.x;
    if (r0 + 1 < n) off[r0 + 1] = o4.y;
    if (r0 + 2 < n) off[r0 + 2] = o4.z;
  }
  if (t == 0 && s_bad) {
    __hip_atomic_store(&status[gridDim.x], 1ull, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
    if (counts) __hip_atomic_store(&counts[1], -1, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (b == (int)gridDim.x - 1 && t == 0) {
    const int T = s_prefix + total;
    off[n] = T;
    if (counts) {
      counts[0] = 3 * T;
      __hip_atomic_store(&counts[1], s_bad ? -1 : T, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
      if (__hip_atomic_load(&status[gridDim.x], __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT) != 0)
        __hip_atomic_store(&counts[1], -1, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// Workspace of marching cubes: the case byte of every cube, the row counts
// followed by the scan's block status words and its timeout flag word
// (zeroed together before the classify pass: zero_ints words from
// rowcnt_off), the row offsets.
struct McLayout {
  size_t cases_off, rowcnt_off, status_off, rowoff_off, total, zero_ints;
  int nb;
};
// case_bytes = 0 leaves out the case array (the workspace-free form).
inline McLayout mc_layout_for(size_t case_bytes, size_t R) {
  McLayout L;
  L.nb = (int)((R + kScanBlock - 1) / kScanBlock);
  L.cases_off = 0;
  L.rowcnt_off = align_up(case_bytes, 256);
  L.status_off = L.rowcnt_off + align_up(R * 4, 8);
  L.rowoff_off = align_up(L.status_off + (size_t)(L.nb + 1) * 8, 256);
  L.total = L.rowoff_off + align_up((R + 1) * 4, 256);
  L.zero_ints = (L.status_off - L.rowcnt_off) / 4 + (size_t)(L.nb + 1) * 2;
  return L;
}
inline McLayout mc_layout(const nrf_volume_params *vp) {
  return mc_layout_for((size_t)volume_count(vp), (size_t)vp->dims[1] * vp->dims[2]);
}

// The scan of the row counts: the multi-workgroup form, or one workgroup
// beyond kScanMaxBlocks blocks.
inline void mc_scan_rows(const McLayout &L, const int32_t *rowcnt, int32_t *rowoff, int R, int32_t *counts,
                         void *workspace, hipStream_t st) {
  if (L.nb <= kScanMaxBlocks) {
    auto *status = reinterpret_cast<unsigned long long *>(reinterpret_cast<char *>(workspace) + L.status_off);
    scan_rows_lb_kernel<<<L.nb, kScanThreads, 0, st>>>(rowcnt, rowoff, R, counts, status);
  } else {
    scan_rows_kernel<<<1, 1024, 0, st>>>(rowcnt, rowoff, R, counts);
  }
}

}  // namespace
}  // namespace nrf
