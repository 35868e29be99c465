// solver.hip - Gauss-Newton warp-field solve for gfx950.
//
// Reference: GaussNewtonSolver (include/solver.h, kernels/solver_kernels.cu).
// The normal equations are the reference's: point-to-plane data term with
// Huber weights (delta 0.2) over DQ-skinned correspondences, ARAP smoothness
// over graph edges i->j with j > i (Huber 0.05, strain weight), Levenberg
// damping of the diagonal, block-Jacobi PCG.  The implementation differs:
//
//  * Assembly is atomic-free and deterministic.  Per frame, the
//    (correspondence, neighbour-slot) pairs are bucketed by node and each
//    bucket sorted by key corr * 4 + slot ("memberships").  Per GN
//    iteration a prep launch writes one 128-byte record per valid
//    correspondence (warped source, residual, Huber weight, the four slot
//    Jacobians) and one record per ARAP edge-side.  One wavefront then owns
//    a chunk of at most 128 memberships of a block row and forms its 6 x 6
//    blocks and RHS as sums of outer products on the matrix cores
//    (v_mfma_f32_16x16x4f32, or BF16 / FP16 products in the precision
//    form), with the membership's dense row staged in LDS.  Rows longer than
//    one chunk leave partial sums that a tail launch adds in chunk order,
//    together with the ARAP edge-sides, damping and the block-Jacobi
//    Cholesky inverse of the diagonal block.  These launches replace the
//    reference's data / smooth / damping / preconditioner kernels and their
//    float atomics; every sum has a fixed order.
//  * PCG is one persistent kernel (ordinary launch, grid sized to be
//    co-resident): each workgroup owns a contiguous range of nodes and keeps
//    their matrix blocks in registers and LDS for all iterations.  Values
//    that cross workgroups (halo entries of the direction vector, per-
//    workgroup FP64 partial dot products) travel as tagged 64-bit words,
//    tag = launch epoch << 16 | synchronisation point, written and polled
//    with agent-scope relaxed atomics, so no grid barrier, fence or cache
//    maintenance is needed and every workgroup combines the partial sums in
//    the same order (bit-identical alpha / beta, the same branch).  Four
//    forms share the layout: textbook, overlapped (A p by recurrence),
//    single-reduction (Chronopoulos-Gear) and pipelined (Ghysels-Vanroose,
//    FP64 recurrences).  Every wait is bounded: on timeout the kernel
//    raises a flag, every workgroup exits, and the call reports
//    NRF_ERR_BARRIER_TIMEOUT.
#include <atomic>
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "nrf_internal.h"
#include "nrf_math.h"

namespace nrf {
namespace {

constexpr int kPcgThreads = 1024;
constexpr unsigned kBarrierSpinLimit = 1u << 22;

// ---------------------------------------------------------------- helpers
__global__ void fill_i32(int32_t *p, int n, int32_t v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Memberships (node -> its correspondence slots) are built once per frame:
// count per node, exclusive scan, scatter, then each node's bucket is
// ordered by key (corr * 4 + slot), which fixes the order in which the
// assembly sums a row's contributions.

// The node of slot k of a valid corr, or -1 when the slot takes no part
// (kernels/solver_kernels.cu:94-96: id in range and weight > 1e-8).
__device__ inline int member_node(const nrf_corr &cr, int k, int N) {
  const int id = cr.node_ids[k];
  return (id >= 0 && id < N && cr.node_ws[k] > 1e-8f) ? id : -1;
}

// Count and scatter aggregate a 256-lane workgroup's claims in an LDS hash
// table first (one entry per distinct node, at most 1024: four slots per
// lane), so a node seen by many of the workgroup's pixels costs one global
// atomic per workgroup, and the workgroup's global atomics are issued in
// parallel.  Positions within a bucket depend on the atomics' order; the
// per-bucket sort that follows fixes the order.
constexpr int kMemThreads = 256, kMemHash = 4 * kMemThreads;
struct MemHash {
  int32_t key[kMemHash], cnt[kMemHash], base[kMemHash];
};
__device__ inline void mem_hash_clear(MemHash &hs) {
  for (int j = threadIdx.x; j < kMemHash; j += kMemThreads) {
    hs.key[j] = -1;
    hs.cnt[j] = 0;
  }
  __syncthreads();
}
// slot of id in the table (inserted if absent); linear probing always ends
// since the table holds more entries than the workgroup can insert
__device__ inline int mem_hash_slot(MemHash &hs, int id) {
  int h = (int)(((unsigned)id * 2654435761u) >> 22) & (kMemHash - 1);
  for (;;) {
    const int prev = atomicCAS(&hs.key[h], -1, id);
    if (prev == -1 || prev == id) return h;
    h = (h + 1) & (kMemHash - 1);
  }
}
__device__ inline void mem_count_body(const nrf_corr *__restrict__ corrs, int n, int N, int32_t *__restrict__ cnt,
                                      int c) {
  __shared__ MemHash hs;
  mem_hash_clear(hs);
  const bool ok = c < n && corrs[c].valid;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int id = ok ? member_node(corrs[c], k, N) : -1;
    if (id >= 0) atomicAdd(&hs.cnt[mem_hash_slot(hs, id)], 1);
  }
  __syncthreads();
  for (int j = threadIdx.x; j < kMemHash; j += kMemThreads)
    if (hs.key[j] >= 0) atomicAdd(&cnt[hs.key[j]], hs.cnt[j]);
}
// A copy of the node transforms and their centred table taken by a solve's
// first kernel (solver_set_snapshot); dst_T null: none.
struct SnapArgs {
  const DQ *T = nullptr, *C = nullptr;
  DQ *dst_T = nullptr, *dst_C = nullptr;
  int N = 0;
};
// vmask (optional): 1 per valid slot, 0 per other slot (the record
// launches of the frame's GN iterations read it instead of the slots).
__global__ __launch_bounds__(256) void mem_count_kernel(const nrf_corr *__restrict__ corrs, int n, int N,
                                                        int32_t *__restrict__ cnt, SnapArgs sn,
                                                        uint8_t *__restrict__ vmask) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (sn.dst_T && c < sn.N) {
    sn.dst_T[c] = sn.T[c];
    sn.dst_C[c] = sn.C[c];
  }
  mem_count_body(corrs, n, N, cnt, c);
  if (vmask && c < n) vmask[c] = corrs[c].valid ? 1 : 0;
}

// The memberships' scan of one frame: off / cur from the per-node counts,
// and in the same pass the assembly's chunk offsets (max(1, ceil(count /
// chunk)) chunks per row); the counts are cleared behind the reads, so the
// next frame's count pass starts from zero without a fill launch.
constexpr int kAsmChunkMem = kAsmChunk;
__device__ inline void member_scan_body(int32_t *__restrict__ cnt, int32_t *__restrict__ off,
                                        int32_t *__restrict__ cur, int32_t *__restrict__ chunk_off, int n) {
  __shared__ int32_t part[1024], partc[1024];
  const int t = threadIdx.x;
  const int per = (n + 1023) / 1024;
  const int b = t * per, e = min(n, b + per);
  int s = 0, sc = 0;
  for (int i = b; i < e; ++i) {
    const int c = cnt[i];
    s += c;
    sc += max(1, (c + kAsmChunkMem - 1) / kAsmChunkMem);
  }
  part[t] = s;
  partc[t] = sc;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {
    const int v = (t >= d) ? part[t - d] : 0, vc = (t >= d) ? partc[t - d] : 0;
    __syncthreads();
    part[t] += v;
    partc[t] += vc;
    __syncthreads();
  }
  int run = part[t] - s, runc = partc[t] - sc;
  for (int i = b; i < e; ++i) {
    const int c = cnt[i];
    off[i] = run;
    cur[i] = run;
    chunk_off[i] = runc;
    run += c;
    runc += max(1, (c + kAsmChunkMem - 1) / kAsmChunkMem);
    cnt[i] = 0;
  }
  if (t == 1023) {
    off[n] = part[1023];
    chunk_off[n] = partc[1023];
  }
}
__device__ inline void gn_begin_body(int32_t *stop, nrf_pcg_result *res, nrf_solver_totals *tot);
// stop / res / tot (optional): the solve's GN begin (gn_begin_body) by one lane.
__global__ __launch_bounds__(1024) void member_scan_kernel(int32_t *__restrict__ cnt, int32_t *__restrict__ off,
                                                           int32_t *__restrict__ cur, int32_t *__restrict__ chunk_off,
                                                           int n, int32_t *stop, nrf_pcg_result *res,
                                                           nrf_solver_totals *tot) {
  if (stop && threadIdx.x == 0) gn_begin_body(stop, res, tot);
  member_scan_body(cnt, off, cur, chunk_off, n);
}

// vmask (optional): the count kernel's valid mask, read instead of the slots'
// valid bytes so that invalid slots are not fetched.
__device__ inline void mem_scatter_body(const nrf_corr *__restrict__ corrs, int n, int N,
                                        int32_t *__restrict__ cur, int32_t *__restrict__ keys, int c,
                                        const uint8_t *__restrict__ vmask = nullptr) {
  __shared__ MemHash hs;
  mem_hash_clear(hs);
  const bool ok = c < n && (vmask ? vmask[c] != 0 : corrs[c].valid != 0);
  int slot[4], rank[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int id = ok ? member_node(corrs[c], k, N) : -1;
    slot[k] = -1;
    rank[k] = 0;
    if (id >= 0) {
      slot[k] = mem_hash_slot(hs, id);
      rank[k] = atomicAdd(&hs.cnt[slot[k]], 1);
    }
  }
  __syncthreads();
  for (int j = threadIdx.x; j < kMemHash; j += kMemThreads)
    if (hs.key[j] >= 0) hs.base[j] = atomicAdd(&cur[hs.key[j]], hs.cnt[j]);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (slot[k] >= 0) keys[hs.base[slot[k]] + rank[k]] = c * 4 + k;
}
__global__ __launch_bounds__(256) void mem_scatter_kernel(const nrf_corr *__restrict__ corrs, int n, int N,
                                                          int32_t *__restrict__ cur, int32_t *__restrict__ keys,
                                                          const uint8_t *__restrict__ vmask) {
  mem_scatter_body(corrs, n, N, cur, keys, blockIdx.x * blockDim.x + threadIdx.x, vmask);
}

// Orders each node's bucket by key.  Buckets of up to kSortLds keys are
// bitonic-sorted in LDS (keys are unique, padding is INT_MAX); longer ones
// fall back to ranking every key against the whole bucket.
constexpr int kSortLds = 2048;
constexpr int kSortThreads = 256;
// chunk_off (optional): the row's chunks get chunk_row[c] = i.
__device__ void mem_sort_body(const int32_t *__restrict__ off, const int32_t *__restrict__ keys_in,
                              int32_t *__restrict__ keys_out, int i, const int32_t *__restrict__ chunk_off = nullptr,
                              int32_t *__restrict__ chunk_row = nullptr) {
  __shared__ int32_t sk[kSortLds];
  const int t = threadIdx.x;
  if (chunk_off)
    for (int c = chunk_off[i] + t; c < chunk_off[i + 1]; c += kSortThreads) chunk_row[c] = i;
  const int b = off[i], n = off[i + 1] - b;
  if (n <= 0) return;  // uniform in the workgroup
  if (n > kSortLds) {
    for (int j = t; j < n; j += kSortThreads) {
      const int kj = keys_in[b + j];
      int rank = 0;
      for (int m = 0; m < n; ++m) rank += (keys_in[b + m] < kj);
      keys_out[b + rank] = kj;
    }
    return;
  }
  int P = 1;
  while (P < n) P <<= 1;
  for (int j = t; j < P; j += kSortThreads) sk[j] = j < n ? keys_in[b + j] : INT_MAX;
  __syncthreads();
  for (int k = 2; k <= P; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int x = t; x < P; x += kSortThreads) {
        const int y = x ^ j;
        if (y > x) {
          const bool up = (x & k) == 0;
          const int u = sk[x], v = sk[y];
          if ((u > v) == up) {
            sk[x] = v;
            sk[y] = u;
          }
        }
      }
      __syncthreads();
    }
  for (int j = t; j < n; j += kSortThreads) keys_out[b + j] = sk[j];
}
__global__ __launch_bounds__(kSortThreads) void mem_sort_kernel(const int32_t *__restrict__ off,
                                                                const int32_t *__restrict__ keys_in,
                                                                int32_t *__restrict__ keys_out,
                                                                const int32_t *__restrict__ chunk_off,
                                                                int32_t *__restrict__ chunk_row) {
  mem_sort_body(off, keys_in, keys_out, (int)blockIdx.x, chunk_off, chunk_row);
}

// ------------------------------------------------------------ assembly
struct PcgCtl;
struct AsmArgs {
  const nrf_corr *corrs;
  const float *canonical;  // NULL: use stored src
  const nrf_node *nodes;
  const DQ *T;
  const DQ *centered;
  int N;
  const int32_t *row_ptr, *col_idx, *diag_rel;
  const int32_t *m_off, *m_keys;
  const int32_t *es_off, *es_entry, *es_offrel;
  const int32_t *e_i, *e_j;
  const float *e_w;
  float lambda_smooth, lambda_damping;
  float arap_sign;  // +1 reference rotation rows, -1 consistent sign
  int psd_pairs;    // 1: a data-term pair outside the pattern adds f J_k^T J_k to the diagonal block
  int precond_fp64;
  int Rmax;
  float *values, *b, *Minv;
  const int32_t *stop;
  const float *crec;  // per-corr records (corr_prep_kernel), MFMA form only
  // per-edge-side ARAP records (corr_prep_kernel, kArapRec floats each), or
  // null: the row tail then computes them
  float *arec;
  int n_es;
  // chunks of a row's memberships (MFMA form): chunk_off[i] .. chunk_off[i+1]
  // are row i's chunks, chunk_row maps a chunk to its row; rows of several
  // chunks leave per-chunk partial sums in part
  const int32_t *chunk_off, *chunk_row;
  float *part;
  // when set, the assembly's last launch opens the next PCG launch (advances
  // its epoch and clears its abort flag), which then needs no launch of its own
  PcgCtl *pcg_open;
  // per correspondence slot, 1 if valid (made by the frame's membership
  // count): the record launch reads a slot's body only when it is set
  const uint8_t *vmask = nullptr;
};
__device__ void pcg_open(PcgCtl *ctl);

// Memberships per assembly chunk: a row with more is split, so that the
// longest serial chain of 32-membership batches is 4 batches long (128
// measured faster than 256 and 64 on the full-scale body workload).
constexpr int kChunk = kAsmChunk;
constexpr int kAsmBatch = 32;     // memberships per batch of the MFMA assembly
constexpr int kAsmBcStride = 34;  // column stride of its dense B operand (2 mod 32 banks)
// floats of one chunk's partial sums: the accumulators of lanes 0..31 (rows
// 0..5 of every column strip live there), NS strips + the FP32 RHS column
__host__ __device__ constexpr int part_floats(int ns) { return (ns + 1) * 32 * 4; }
// chunk slots for N rows and n correspondences (4 memberships each at most)
inline size_t asm_max_chunks(int N, int n) {
  return (size_t)N + ((size_t)4 * (n > 0 ? n : 0) + kChunk - 1) / kChunk;
}

// staging layout per entry (floats): fw, r, kslot, pos[4] (as float bits), J[4][6]
constexpr int kStageData = 3 + 4 + 24;
// staging per ARAP edge-side: lam, side, offrel, r[3], Js[6][3], Jo[6][3]
constexpr int kStageArap = 6 + 18 + 18;
constexpr int kStage = kStageArap > kStageData ? kStageArap : kStageData;
constexpr int kArapRec = 44;  // kStageArap padded to whole float4

__device__ inline void arap_jac(float3 q, float sgn, float J[6][3]) {
  // compute_arap_jacobian (kernels/solver_kernels.cu:140-165).  sgn = 1 gives
  // the reference's rotation rows; sgn = -1 negates them (exact in FP32), which
  // makes them d(omega x q)/d(omega) like the data term's.
  const float3 s = make_float3(sgn * q.x, sgn * q.y, sgn * q.z);
  J[0][0] = 0.0f; J[0][1] = s.z; J[0][2] = -s.y;
  J[1][0] = -s.z; J[1][1] = 0.0f; J[1][2] = s.x;
  J[2][0] = s.y; J[2][1] = -s.x; J[2][2] = 0.0f;
  J[3][0] = 1.0f; J[3][1] = 0.0f; J[3][2] = 0.0f;
  J[4][0] = 0.0f; J[4][1] = 1.0f; J[4][2] = 0.0f;
  J[5][0] = 0.0f; J[5][1] = 0.0f; J[5][2] = 1.0f;
}

template <typename R>
__device__ inline bool chol_inv6(const R *A, R *Ainv_col, int col, bool &ok_out) {
  // invert6x6_cholesky (include/se3_math.cuh:152-202), one column per call.
  R L[6][6];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) L[i][j] = 0;
  for (int i = 0; i < 6; ++i) {
    for (int j = 0; j <= i; ++j) {
      R s = A[i * 6 + j];
      for (int k = 0; k < j; ++k) s -= L[i][k] * L[j][k];
      if (i == j) {
        if (s <= 0) {
          ok_out = false;
          return false;
        }
        L[i][j] = sqrt(s);
      } else {
        L[i][j] = s / L[j][j];
      }
    }
  }
  R y[6], x[6];
  for (int i = 0; i < 6; ++i) {
    R s = (i == col) ? (R)1 : (R)0;
    for (int k = 0; k < i; ++k) s -= L[i][k] * y[k];
    y[i] = s / L[i][i];
  }
  for (int i = 5; i >= 0; --i) {
    R s = y[i];
    for (int k = i + 1; k < 6; ++k) s -= L[k][i] * x[k];
    x[i] = s / L[i][i];
  }
  for (int i = 0; i < 6; ++i) Ainv_col[i] = x[i];
  ok_out = true;
  return true;
}

// Staged residual and Jacobians of one membership (corr c, slot kslot) of
// row i: the DQ-warped source point (re-warped from the canonical point when
// one is given), the Huber-weighted residual and the four skinning-slot
// Jacobians w_l [ (src - x_l) x n, n ] (kernels/solver_kernels.cu:52-138,
// include/se3_math.cuh:121-145).  pos[l] is the slot's block position in the
// row, -1 for an invalid slot, -2 for a valid neighbour whose block is not in
// the pattern (dropped, Q2).  Returns fw (0 for an inactive lane).
__device__ inline float stage_membership(const AsmArgs &a, int key, const int32_t *cols, int rlen, float &r,
                                         int &kslot, float J[4][6], int pos[4]) {
  const int c = key >> 2;
  kslot = key & 3;
  const nrf_corr &cr = a.corrs[c];
  const float3 dst = f3(cr.dst), nrm = f3(cr.normal);
  const int ids[4] = {cr.node_ids[0], cr.node_ids[1], cr.node_ids[2], cr.node_ids[3]};
  const float ws[4] = {cr.node_ws[0], cr.node_ws[1], cr.node_ws[2], cr.node_ws[3]};
  float3 src;
  if (a.canonical) {
    src = f3(a.canonical[3 * c], a.canonical[3 * c + 1], a.canonical[3 * c + 2]);
    DQ bl;
    if (dq_blend4(a.centered, ids, ws, bl)) src = dq_transform(bl, src);
  } else {
    src = f3(cr.src);
  }
  r = dot(nrm, sub(src, dst));
  const float fw = fmaxf(cr.weight, 0.0f) * huber_w(r, 0.2f);
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    pos[l] = -1;
    for (int q = 0; q < 6; ++q) J[l][q] = 0.f;
    const int id = ids[l];
    if (id >= 0 && id < a.N && ws[l] > 1e-8f) {
      const float3 xp = f3(a.nodes[id].pos);
      const float3 lp = sub(src, xp);
      const float3 cp = cross(lp, nrm);
      const float w = ws[l];
      J[l][0] = w * cp.x; J[l][1] = w * cp.y; J[l][2] = w * cp.z;
      J[l][3] = w * nrm.x; J[l][4] = w * nrm.y; J[l][5] = w * nrm.z;
      int p = -2;
      for (int q = 0; q < rlen; ++q)  // rows are short; linear scan
        if (cols[q] == id) { p = q; break; }
      pos[l] = p;
    }
  }
  return fw;
}

// ARAP term of the edge (ni, nj) with graph weight nw (kernels/solver_kernels.cu:
// 236-290): residual r = T_i x_j - T_j x_j, the Jacobians of both ends, and
// the weight lambda * nw * Huber * strain, or 0 when it fails the
// threshold of :271 (Ji, Jj untouched then).
__device__ inline float arap_edge(const AsmArgs &a, int ni, int nj, float nw, float3 &r, float Ji[6][3],
                                  float Jj[6][3]) {
  const float3 xi = f3(a.nodes[ni].pos), xj = f3(a.nodes[nj].pos);
  const DQ Ti = a.T[ni], Tj = a.T[nj];
  const float3 Ti_xj = dq_transform_centered(Ti, xj, xi);
  const float3 Ti_xi = dq_transform_centered(Ti, xi, xi);
  const float3 Tj_xj = dq_transform_centered(Tj, xj, xj);
  const float3 rest = sub(xi, xj);
  const float3 curr = sub(Ti_xi, Tj_xj);
  const float dr = sqrtf(rest.x * rest.x + rest.y * rest.y + rest.z * rest.z);
  const float dc = sqrtf(curr.x * curr.x + curr.y * curr.y + curr.z * curr.z);
  const float strain = fabsf(dc - dr) / fmaxf(dr, 1e-6f);
  const float ratio = strain / 0.25f;
  const float sw = expf(-(ratio * ratio));
  r = sub(Ti_xj, Tj_xj);
  const float rn = sqrtf(r.x * r.x + r.y * r.y + r.z * r.z);
  const float hw = huber_w(rn, 0.05f);
  const float lam = a.lambda_smooth * nw * hw * sw;
  if (lam <= 1e-8f) return 0.f;
  arap_jac(sub(Ti_xj, xi), a.arap_sign, Ji);
  arap_jac(sub(Tj_xj, xj), a.arap_sign, Jj);
  return lam;
}

// One ARAP edge-side e: stages side, offrel, r and the two Jacobians in
// st[1..41] when its weight passes the threshold and returns the weight
// (0 otherwise: st[1..] untouched).
__device__ inline float arap_side(const AsmArgs &a, int e, float *st) {
  const int ent = a.es_entry[e];
  const int edge = ent >> 1, side = ent & 1;
  float3 r;
  float Ji[6][3], Jj[6][3];
  const float lam = arap_edge(a, a.e_i[edge], a.e_j[edge], fmaxf(a.e_w[edge], 0.05f), r, Ji, Jj);
  if (lam == 0.f) return 0.f;
  st[1] = __int_as_float(side);
  st[2] = __int_as_float(a.es_offrel[e]);
  st[3] = r.x; st[4] = r.y; st[5] = r.z;
  for (int p = 0; p < 6; ++p)
    for (int c = 0; c < 3; ++c) {
      st[6 + p * 3 + c] = side ? Jj[p][c] : Ji[p][c];
      st[24 + p * 3 + c] = side ? Ji[p][c] : Jj[p][c];
    }
  return lam;
}

// ARAP edge-sides of row i (kernels/solver_kernels.cu:202-352), damping
// (:167-194), write-out of the row's blocks and RHS, then the block-Jacobi
// inverse of the diagonal block (:358-399, include/se3_math.cuh:152-202).
// acc holds the row's blocks with the data term already in, bacc (lanes
// 36..41) the data term's RHS.
__device__ inline void row_tail(const AsmArgs &a, int i, int rb, int rlen, float *acc, float *stage, float bacc) {
  const int lane = threadIdx.x;
  const int ab = lane, ea = ab / 6, eb = ab % 6, ra = lane - 36;
  const int diag = a.diag_rel[i];
  const int sb = a.es_off[i], se = a.es_off[i + 1];
  for (int base = sb; base < se; base += 64) {
    const int e = base + lane;
    float *st = stage + lane * kStage;
    if (e < se && a.arec) {  // the record corr_prep_kernel made
      const float4 *src = reinterpret_cast<const float4 *>(a.arec + (size_t)e * kArapRec);
#pragma unroll
      for (int k = 0; k < kArapRec / 4; ++k) {
        const float4 v = src[k];
        if (4 * k < kStageArap) st[4 * k] = v.x;
        if (4 * k + 1 < kStageArap) st[4 * k + 1] = v.y;
        if (4 * k + 2 < kStageArap) st[4 * k + 2] = v.z;
        if (4 * k + 3 < kStageArap) st[4 * k + 3] = v.w;
      }
    } else {
      st[0] = e < se ? arap_side(a, e, st) : 0.f;
    }
    __syncthreads();
    const int cnt = min(64, se - base);
    // the diagonal entry of lane ab accumulates in a register (same order of
    // additions; the off-diagonal block differs per edge-side)
    float dacc = lane < 36 ? acc[diag * 36 + ab] : 0.f;
    for (int q = 0; q < cnt; ++q) {
      const float *sq = stage + q * kStage;
      const float lm = sq[0];
      if (lm <= 1e-8f) continue;
      const int side = __float_as_int(sq[1]);
      const int offrel = __float_as_int(sq[2]);
      const float *Js = sq + 6, *Jo = sq + 24;
      if (lane < 36) {
        float vss = 0, vso = 0;
        for (int c = 0; c < 3; ++c) vss += Js[ea * 3 + c] * Js[eb * 3 + c];
        for (int c = 0; c < 3; ++c) vso += Js[ea * 3 + c] * Jo[eb * 3 + c];
        dacc += lm * vss;
        // an edge whose block is absent from the pattern (a pattern kept over
        // a graph change, Q4) drops its off-diagonal term only, as the
        // reference's find_block < 0 test does (kernels/solver_kernels.cu:286-330)
        if (offrel >= 0) acc[offrel * 36 + ab] += -lm * vso;
      } else if (ra >= 0 && ra < 6) {
        float v = 0;
        for (int c = 0; c < 3; ++c) v += Js[ra * 3 + c] * sq[3 + c];
        bacc += side ? lm * v : -lm * v;
      }
    }
    if (lane < 36) acc[diag * 36 + ab] = dacc;
    __syncthreads();
  }

  // ---- damping (kernels/solver_kernels.cu:167-194)
  if (lane < 36 && ea == eb) {
    const float cur = acc[diag * 36 + ab];
    const float fl = (ea < 3) ? 1e-3f : 1e-2f;
    acc[diag * 36 + ab] = cur + a.lambda_damping * fmaxf(fabsf(cur), fl);
  }
  __syncthreads();
  for (int j = lane; j < rlen * 36; j += 64) a.values[(size_t)rb * 36 + j] = acc[j];
  if (ra >= 0 && ra < 6) a.b[i * 6 + ra] = bacc;

  // ---- block-Jacobi inverse (kernels/solver_kernels.cu:358-399)
  if (lane < 6) {
    const float *D = acc + diag * 36;
    float outc[6];
    bool ok = true;
    if (a.precond_fp64) {
      double d[36], x[6];
      for (int k = 0; k < 36; ++k) d[k] = D[k];
      for (int k = 0; k < 6; ++k) d[k * 7] += (double)1e-4f;
      chol_inv6<double>(d, x, lane, ok);
      for (int k = 0; k < 6; ++k) outc[k] = (float)x[k];
      if (!ok)
        for (int k = 0; k < 6; ++k) outc[k] = (k == lane) ? (float)(1.0 / (d[k * 7] + (double)1e-6f)) : 0.f;
    } else {
      float d[36];
      for (int k = 0; k < 36; ++k) d[k] = D[k];
      for (int k = 0; k < 6; ++k) d[k * 7] += 1e-4f;
      chol_inv6<float>(d, outc, lane, ok);
      if (!ok)
        for (int k = 0; k < 6; ++k) outc[k] = (k == lane) ? 1.f / (d[k * 7] + 1e-6f) : 0.f;
    }
    for (int k = 0; k < 6; ++k) a.Minv[(size_t)i * 36 + k * 6 + lane] = outc[k];
  }
}

// One wavefront per block row i: the data term with every membership's
// contribution accumulated in LDS by 36 lanes (one per block entry) in
// membership order, then row_tail.  Used for rows wider than the MFMA form
// takes.
__global__ __launch_bounds__(64) void assemble_rows_kernel(AsmArgs a) {
  extern __shared__ float lds[];
  const int i = blockIdx.x;
  if (a.pcg_open && i == 0 && threadIdx.x == 0) pcg_open(a.pcg_open);
  if (a.stop && *a.stop) return;
  const int lane = threadIdx.x;
  const int rb = a.row_ptr[i], rlen = a.row_ptr[i + 1] - rb;
  float *acc = lds;                                      // [Rmax][36]
  int32_t *cols = reinterpret_cast<int32_t *>(acc + a.Rmax * 36);  // [Rmax]
  float *stage = reinterpret_cast<float *>(cols + a.Rmax);         // [64][kStage]
  for (int j = lane; j < rlen * 36; j += 64) acc[j] = 0.f;
  for (int j = lane; j < rlen; j += 64) cols[j] = a.col_idx[rb + j];
  __syncthreads();
  const int ab = lane;  // lanes 0..35: entry (a, b) = (ab / 6, ab % 6)
  const int ea = ab / 6, eb = ab % 6;
  const int ra = lane - 36;  // lanes 36..41: RHS row ra
  float bacc = 0.f;
  const int mb = a.m_off[i], me = a.m_off[i + 1];
  for (int base = mb; base < me; base += 64) {
    const int e = base + lane;
    float *st = stage + lane * kStage;
    float fw = 0.f;
    if (e < me) {
      float r, J[4][6];
      int kslot, pos[4];
      fw = stage_membership(a, a.m_keys[e], cols, rlen, r, kslot, J, pos);
      st[1] = r;
      st[2] = __int_as_float(kslot);
      for (int l = 0; l < 4; ++l) {
        st[3 + l] = __int_as_float(pos[l]);
        for (int q = 0; q < 6; ++q) st[7 + 6 * l + q] = J[l][q];
      }
    }
    st[0] = fw;  // inactive lanes hold 0; NaN stays active as in the reference
    __syncthreads();
    const int cnt = min(64, me - base);
    for (int q = 0; q < cnt; ++q) {
      const float *sq = stage + q * kStage;
      const float f = sq[0];
      if (f <= 1e-8f) continue;  // kernels/solver_kernels.cu:84
      const int ks = __float_as_int(sq[2]);
      const float *Jk = sq + 7 + 6 * ks;
      if (lane < 36) {
        const float fa = f * Jk[ea];
#pragma unroll
        for (int l = 0; l < 4; ++l) {
          const int pos = __float_as_int(sq[3 + l]);
          if (pos >= 0)
            acc[pos * 36 + ab] += fa * sq[7 + 6 * l + eb];
          else if (pos == -2 && a.psd_pairs)  // outside the pattern: f J_k^T J_k on the diagonal
            acc[a.diag_rel[i] * 36 + ab] += fa * Jk[eb];
        }
      } else if (ra >= 0 && ra < 6) {
        bacc += -f * Jk[ra] * sq[1];
      }
    }
    __syncthreads();
  }
  row_tail(a, i, rb, rlen, acc, stage, bacc);
}

// Per-correspondence record of one GN iteration, computed once per valid
// corr instead of once per membership (each corr belongs to up to four
// rows): [fw, r, id0..id3 (int bits, -1 = slot takes no part), J[4][6],
// pad] = 32 floats (kRec).  Same expressions as stage_membership.
constexpr int kRec = 32;
typedef float floatx4_t __attribute__((ext_vector_type(4)));
__device__ void corr_prep_one(const AsmArgs &a, int c, float *__restrict__ rec) {
  if (a.vmask && !a.vmask[c]) return;
  const nrf_corr &cr = a.corrs[c];
  if (!cr.valid) return;
  const float3 dst = f3(cr.dst), nrm = f3(cr.normal);
  const int ids[4] = {cr.node_ids[0], cr.node_ids[1], cr.node_ids[2], cr.node_ids[3]};
  const float ws[4] = {cr.node_ws[0], cr.node_ws[1], cr.node_ws[2], cr.node_ws[3]};
  float3 src;
  if (a.canonical) {
    src = f3(a.canonical[3 * c], a.canonical[3 * c + 1], a.canonical[3 * c + 2]);
    DQ bl;
    if (dq_blend4(a.centered, ids, ws, bl)) src = dq_transform(bl, src);
  } else {
    src = f3(cr.src);
  }
  const float r = dot(nrm, sub(src, dst));
  const float fw = fmaxf(cr.weight, 0.0f) * huber_w(r, 0.2f);
  float o[kRec];
  o[0] = fw;
  o[1] = r;
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    const int id = ids[l];
    const bool ok = id >= 0 && id < a.N && ws[l] > 1e-8f;
    o[2 + l] = __int_as_float(ok ? id : -1);
    float J[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (ok) {
      const float3 xp = f3(a.nodes[id].pos);
      const float3 lp = sub(src, xp);
      const float3 cp = cross(lp, nrm);
      const float w = ws[l];
      J[0] = w * cp.x; J[1] = w * cp.y; J[2] = w * cp.z;
      J[3] = w * nrm.x; J[4] = w * nrm.y; J[5] = w * nrm.z;
    }
    for (int q = 0; q < 6; ++q) o[6 + 6 * l + q] = J[q];
  }
  o[30] = o[31] = 0.f;
  float4 *dst4 = reinterpret_cast<float4 *>(rec + (size_t)c * kRec);
#pragma unroll
  for (int k = 0; k < kRec / 4; ++k) dst4[k] = make_float4(o[4 * k], o[4 * k + 1], o[4 * k + 2], o[4 * k + 3]);
}
// Threads [0, n): the per-corr records; [n, n + n_es): the ARAP edge-side
// records (when a.arec is set), so that the row tails read them instead of
// walking edge -> node -> transform loads themselves.
__global__ __launch_bounds__(256) void corr_prep_kernel(AsmArgs a, int n, float *__restrict__ rec) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) {
    corr_prep_one(a, c, rec);
  } else if (a.arec && c - n < a.n_es) {
    const int e = c - n;
    float o[kArapRec];
#pragma unroll
    for (int k = 0; k < kArapRec; ++k) o[k] = 0.f;
    o[0] = arap_side(a, e, o);
    float4 *dst = reinterpret_cast<float4 *>(a.arec + (size_t)e * kArapRec);
#pragma unroll
    for (int k = 0; k < kArapRec / 4; ++k) dst[k] = make_float4(o[4 * k], o[4 * k + 1], o[4 * k + 2], o[4 * k + 3]);
  }
}

// MFMA form of the data term (NS = column slabs of 16 covering the 6 R
// block columns plus one RHS column).  For a batch of 64 memberships e of
// row i, with a_e = f_e J_k(e) (the row node's Jacobian times the Huber
// weight) and the dense row b_e = [J_l(e) in the columns of block pos_l(e),
// r_e in column 6 R], the batch adds
//   C (6 x (6 R + 1)) += sum_e a_e b_e^T
// as v_mfma_f32_16x16x4f32 products (rows 6..15 of A are zero).  Then the
// blocks are C[:, 0:6R] and the RHS is -C[:, 6R].  Same products as the
// scalar form, FP32, summed in the matrix unit's order.
// Reduced-precision operands of the JtJ products: bf16 or fp16, both by the
// hardware conversion (round to nearest even).
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));
template <int PREC> struct LpType;
template <> struct LpType<1> { typedef short elem; typedef bf16x4 vec; };
template <> struct LpType<2> { typedef _Float16 elem; typedef f16x4 vec; };
template <> struct LpType<0> { typedef short elem; typedef bf16x4 vec; };  // unused
// PREC 3 / 4: BF16 / FP16 operands split in two (hi = rounded value, lo =
// the rounded remainder), three products per pair (lo hi + hi lo + hi hi).
template <> struct LpType<3> : LpType<1> {};
template <> struct LpType<4> : LpType<2> {};
template <int PREC> constexpr int lp_base() { return PREC >= 3 ? PREC - 2 : PREC; }
template <int PREC>
__device__ inline typename LpType<PREC>::elem to_lp(float f) {
  if constexpr (PREC == 2)
    return (_Float16)f;
  else
    return __builtin_bit_cast(short, (__bf16)f);  // v_cvt_pk_bf16_f32: round to nearest even
}
template <int PREC>
__device__ inline float from_lp(typename LpType<PREC>::elem v) {
  if constexpr (PREC == 2)
    return (float)v;
  else
    return __uint_as_float((unsigned)(unsigned short)v << 16);
}
template <int PREC>
__device__ inline floatx4_t mfma_lp(typename LpType<PREC>::vec a, typename LpType<PREC>::vec b, floatx4_t c) {
  if constexpr (PREC == 2)
    return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
  else
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Hand-off between the lanes of a one-wave workgroup through LDS: a
// compiler barrier only (LDS accesses of one wave execute in program order).
__device__ inline void wave_sync() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_wave_barrier();
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// mode 0: the whole row; mode 1: memberships [pb, pe) of the row, partial
// sums to chunk slot `slot`; mode 2: the sums of the row's nslot chunk slots
// from `slot` on, added in slot order, then the rest of the row.
template <int NS, int PREC>
__device__ void assemble_row_mfma(const AsmArgs &a, int i, float *lds, float *s_rhs, int mode = 0, int pb = 0,
                                  int pe = 0, int slot = 0, int nslot = 0) {
  if (a.stop && *a.stop) return;
  const int lane = threadIdx.x;
  const int rb = a.row_ptr[i], rlen = a.row_ptr[i + 1] - rb;
  float *acc = lds;                                      // [Rmax][36]
  int32_t *cols = reinterpret_cast<int32_t *>(acc + a.Rmax * 36);  // [Rmax]
  float *stage = reinterpret_cast<float *>(cols + a.Rmax);         // dense operands, then the tail's scratch
  for (int j = lane; j < rlen; j += 64) cols[j] = a.col_idx[rb + j];
  __syncthreads();
  typedef floatx4_t floatx4;
  typedef typename LpType<PREC>::vec lp4;
  floatx4 C[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) C[s] = floatx4{0.f, 0.f, 0.f, 0.f};
  floatx4 Crhs = {0.f, 0.f, 0.f, 0.f};  // FP32 RHS column of the reduced-precision forms
  const int ncol = 6 * rlen + 1;  // block columns + RHS column
  const int am = lane % 16;       // A row / B, D column within a slab
  const int kq = lane / 16;       // k within a 4-step
  const int mb = mode == 1 ? pb : a.m_off[i], me = mode == 1 ? pe : (mode == 2 ? mb : a.m_off[i + 1]);
  if (mode == 2 && lane < 32) {
    for (int k = 0; k < nslot; ++k) {
      const float4 *src = reinterpret_cast<const float4 *>(a.part + (size_t)(slot + k) * part_floats(NS));
#pragma unroll
      for (int s2 = 0; s2 < NS; ++s2) {
        const float4 v = src[s2 * 32 + lane];
        C[s2][0] += v.x; C[s2][1] += v.y; C[s2][2] += v.z; C[s2][3] += v.w;
      }
      const float4 v = src[NS * 32 + lane];
      Crhs[0] += v.x; Crhs[1] += v.y; Crhs[2] += v.z; Crhs[3] += v.w;
    }
  }
  // Dense operands of the products, per batch of kAsmBatch = 32 memberships:
  // Ab[m][0..7] = a_m (six entries, then two zeros) and Bc[j][m] = b_m[j],
  // column j of membership m's dense row (J_l in the columns of block
  // pos_l, r in column 6 R, zero elsewhere), stored by column with stride
  // kAsmBcStride so that the 16 x 2 lanes of a half wave reading one
  // k-step of a slab hit 32 distinct banks.  Lane (m, h) stages membership
  // m: half h scatters the Jacobians of slots 2h and 2h + 1 after clearing
  // the two blocks it scattered for the previous batch; half 0 also writes
  // a_m and r.  The operands are the scalar form's, so are the products.
  //
  // Software pipeline over the batches: the keys of the chunk (at most
  // kChunk / kAsmBatch batches) are loaded up front, and the records of
  // batch k + 1 are requested once batch k is staged, so they travel while
  // batch k's products run.  The workgroup is one wave: its LDS accesses
  // are ordered by program order, and the hand-offs use wave barriers,
  // which (unlike a workgroup barrier's fence) do not wait for the
  // outstanding record loads.
  constexpr int kBat = kAsmBatch, kBS = kAsmBcStride, kKeys = kChunk / kAsmBatch;
  float *Ab = stage;            // [kBat][8]
  float *Bc = stage + kBat * 8;  // [16 NS][kBS]
  const int m = lane & (kBat - 1), h = lane / kBat;
  if (mb < me) {
    for (int j = lane; j < 16 * NS * kBS; j += 64) Bc[j] = 0.f;
    if (lane < kBat) {
      Ab[lane * 8 + 6] = 0.f;
      Ab[lane * 8 + 7] = 0.f;
    }
  }
  // four named registers (an array here ended up in scratch memory, whose
  // loads wait for every record load in flight)
  static_assert(kKeys == 4, "the chunk's keys live in four registers");
  int k0 = 0, k1 = 0, k2 = 0, k3 = 0;
  if (mb < me) {  // indices clamped into [mb, me): the loads go out together
    k0 = a.m_keys[min(mb + m, me - 1)];
    k1 = a.m_keys[min(mb + kBat + m, me - 1)];
    k2 = a.m_keys[min(mb + 2 * kBat + m, me - 1)];
    k3 = a.m_keys[min(mb + 3 * kBat + m, me - 1)];
  }
  const int32_t *m_keys = a.m_keys;
  auto key_of = [k0, k1, k2, k3, m_keys](int bi, int e) -> int {
    if (bi >= kKeys) return m_keys[e];
    return bi == 0 ? k0 : bi == 1 ? k1 : bi == 2 ? k2 : k3;
  };
  // Every loaded word is used, including the two zero padding words (30 and
  // 31): a load in flight whose register looked dead would be waited for as
  // soon as the products reuse that register.
  float R[kRec];
  auto fetch = [&](int bi) {
    const int e = mb + kBat * bi + m;
    if (e < me) {
      const float4 *src4 = reinterpret_cast<const float4 *>(a.crec + (size_t)(key_of(bi, e) >> 2) * kRec);
#pragma unroll
      for (int k = 0; k < kRec / 4; ++k) {
        const float4 v = src4[k];
        R[4 * k] = v.x; R[4 * k + 1] = v.y; R[4 * k + 2] = v.z; R[4 * k + 3] = v.w;
      }
    } else {
#pragma unroll
      for (int k = 0; k < kRec; ++k) R[k] = 0.f;
    }
  };
  int prev[2] = {-1, -1};  // blocks this lane scattered for the previous batch
  // the row's columns in scalar registers (uniform), for the position search
  constexpr int kMaxBlk = (16 * NS - 1) / 6;  // 6 R + 1 <= 16 NS
  int cr[kMaxBlk];
#pragma unroll
  for (int q = 0; q < kMaxBlk; ++q) cr[q] = q < rlen ? a.col_idx[rb + q] : -2;
  wave_sync();
  if (mb < me) fetch(0);
#pragma nounroll
  for (int base = mb, bi = 0; base < me; base += kBat, ++bi) {
    const int e = base + m;
    const int kslot = e < me ? (key_of(bi, e) & 3) : 0;
    const float fw = R[0];
    // f <= 1e-8 skips the membership (kernels/solver_kernels.cu:84); NaN stays
    const bool act = e < me && !(fw <= 1e-8f);
    // this half's two slots: their blocks in the row (absent: dropped, Q2)
    int pos[2];
    float Js[2][6];
#pragma unroll
    for (int l2 = 0; l2 < 2; ++l2) {
      const int id = __float_as_int(h ? R[4 + l2] : R[2 + l2]);
#pragma unroll
      for (int c = 0; c < 6; ++c) Js[l2][c] = h ? R[18 + 6 * l2 + c] : R[6 + 6 * l2 + c];
      int p = -1;
      if (act && id >= 0) {
#pragma unroll
        for (int q = 0; q < kMaxBlk; ++q) p = cr[q] == id ? q : p;  // columns are distinct
      }
      pos[l2] = p;
    }
    if (a.psd_pairs) {
      // psd_dropped_pairs: each active slot outside the row's pattern adds
      // f J_k^T J_k to the diagonal block, i.e. the row node's slot k enters
      // the dense row (1 + m) times, m = the membership's dropped slots
      int dropped = 0;
#pragma unroll
      for (int l2 = 0; l2 < 2; ++l2) {
        const int id = __float_as_int(h ? R[4 + l2] : R[2 + l2]);
        dropped += (act && id >= 0 && pos[l2] < 0) ? 1 : 0;
      }
      const int mtot = dropped + __shfl_xor(dropped, 32, 64);
      if ((kslot >> 1) == h) {
        const float sc = 1.f + (float)mtot;
        const bool s1 = (kslot & 1) != 0;  // static indices: Js stays in registers
#pragma unroll
        for (int c = 0; c < 6; ++c) {
          Js[0][c] = s1 ? Js[0][c] : Js[0][c] * sc;
          Js[1][c] = s1 ? Js[1][c] * sc : Js[1][c];
        }
      }
    }
#pragma unroll
    for (int l2 = 0; l2 < 2; ++l2)
      if (prev[l2] >= 0)
        for (int c = 0; c < 6; ++c) Bc[(prev[l2] * 6 + c) * kBS + m] = 0.f;
    wave_sync();  // every clear before any scatter (a lane's new block may be another's old one)
#pragma unroll
    for (int l2 = 0; l2 < 2; ++l2) {
      if (pos[l2] >= 0)
        for (int c = 0; c < 6; ++c) Bc[(pos[l2] * 6 + c) * kBS + m] = Js[l2][c];
      prev[l2] = pos[l2];
    }
    if (h == 0) {
#pragma unroll
      for (int q = 0; q < 6; ++q) {  // the row node's slot, selected without dynamic indexing
        const float jk = kslot == 0 ? R[6 + q] : kslot == 1 ? R[12 + q] : kslot == 2 ? R[18 + q] : R[24 + q];
        Ab[m * 8 + q] = act ? fw * jk : 0.f;
      }
      // r; adding the record's two padding words (zero) keeps their
      // registers live, so no load in flight lands on a reused register
      Bc[(ncol - 1) * kBS + m] = act ? R[1] + (R[30] + R[31]) : 0.f;
    }
    if (base + kBat < me) fetch(bi + 1);  // R is staged: the next batch's records
    wave_sync();
    const int ar = am < 6 ? am : 6;  // A rows 6..15 read the zero entry
    if constexpr (PREC == 0) {
#pragma unroll
      for (int t = 0; t < kBat / 4; ++t) {
        const float av = Ab[(4 * t + kq) * 8 + ar];
        const float *bq = Bc + am * kBS + 4 * t + kq;
#pragma unroll
        for (int s = 0; s < NS; ++s)
          C[s] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bq[16 * s * kBS], C[s], 0, 0, 0);
      }
    } else {
      // JtJ blocks in BF16 (PREC 1, 3) or FP16 (PREC 2, 4) with FP32
      // accumulation: 16 memberships per product, lane (am, kq) supplying
      // memberships 16 t + 4 kq + q, q = 0..3.  The split forms (3, 4) carry
      // each operand as hi + lo, both in the type, and add lo_a hi_b +
      // hi_a lo_b before hi_a hi_b: about 16 (BF16) / 22 (FP16) significant
      // bits per operand, only lo_a lo_b dropped.  The FP16 split scales
      // both operands by 2^6 (undone exactly on the FP32 sums) so that the
      // low parts stay normal numbers.  The RHS column stays FP32 (below).
      constexpr int B = lp_base<PREC>();
      constexpr bool kSplit = PREC >= 3;
      constexpr float kSc = PREC == 4 ? 64.f : 1.f;
#pragma unroll
      for (int t = 0; t < kBat / 16; ++t) {
        lp4 ah, al;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const float v = Ab[(16 * t + 4 * kq + q) * 8 + ar] * kSc;
          ah[q] = to_lp<B>(v);
          if constexpr (kSplit) al[q] = to_lp<B>(v - from_lp<B>(ah[q]));
        }
        const float *bq = Bc + am * kBS + 16 * t + 4 * kq;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          lp4 bh, bl;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const float v = bq[16 * s * kBS + q] * kSc;
            bh[q] = to_lp<B>(v);
            if constexpr (kSplit) bl[q] = to_lp<B>(v - from_lp<B>(bh[q]));
          }
          if constexpr (kSplit) {
            C[s] = mfma_lp<B>(al, bh, C[s]);
            C[s] = mfma_lp<B>(ah, bl, C[s]);
          }
          C[s] = mfma_lp<B>(ah, bh, C[s]);
        }
      }
#pragma unroll
      for (int t = 0; t < kBat / 4; ++t)
        Crhs = __builtin_amdgcn_mfma_f32_16x16x4f32(Ab[(4 * t + kq) * 8 + ar],
                                                    am == 0 ? Bc[(ncol - 1) * kBS + 4 * t + kq] : 0.f, Crhs, 0, 0, 0);
    }
    wave_sync();
  }
  if constexpr (PREC == 4) {  // undo the operand scaling (exact: a power of two)
    if (mode != 2) {
#pragma unroll
      for (int s = 0; s < NS; ++s) C[s] = C[s] * (1.f / 4096.f);
    }
  }
  if (mode == 1) {  // partial sums of this chunk
    if (lane < 32) {
      float4 *dst = reinterpret_cast<float4 *>(a.part + (size_t)slot * part_floats(NS));
#pragma unroll
      for (int s2 = 0; s2 < NS; ++s2) dst[s2 * 32 + lane] = make_float4(C[s2][0], C[s2][1], C[s2][2], C[s2][3]);
      dst[NS * 32 + lane] = make_float4(Crhs[0], Crhs[1], Crhs[2], Crhs[3]);
    }
    return;
  }
  // D[m][j] sits in lane j % 16 + 16 (m / 4), element m % 4: rows 0..5 are in
  // lanes 0..31.
  for (int j = lane; j < rlen * 36; j += 64) acc[j] = 0.f;
  __syncthreads();
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const int j = 16 * s + am;
    if (lane < 32 && j < ncol) {
      for (int v = 0; v < 4; ++v) {
        const int m = 4 * kq + v;
        if (m < 6) {
          if (j < ncol - 1)
            acc[(j / 6) * 36 + m * 6 + j % 6] = C[s][v];
          else if (PREC == 0)
            s_rhs[m] = -C[s][v];
        }
      }
    }
  }
  if (PREC != 0 && lane < 32 && am == 0)  // column 0 of the FP32 RHS product
    for (int v = 0; v < 4; ++v)
      if (4 * kq + v < 6) s_rhs[4 * kq + v] = -Crhs[v];
  __syncthreads();
  const int ra = lane - 36;
  const float bacc = (ra >= 0 && ra < 6) ? s_rhs[ra] : 0.f;
  row_tail(a, i, rb, rlen, acc, stage, bacc);
}
// One workgroup per chunk b (grid: an upper bound on the chunk count, the
// rest leave): a row of one chunk is assembled whole, a chunk of a longer row
// leaves its partial sums for assemble_rows_tail_kernel.
template <int NS, int PREC>
__device__ inline void assemble_chunk(const AsmArgs &a, int N, int b, float *lds, float *s_rhs) {
  if (b >= a.chunk_off[N]) return;
  const int i = a.chunk_row[b];
  const int c0 = a.chunk_off[i], np = a.chunk_off[i + 1] - c0;
  if (np == 1) {
    assemble_row_mfma<NS, PREC>(a, i, lds, s_rhs);
  } else {
    const int mb = a.m_off[i] + (b - c0) * kChunk;
    assemble_row_mfma<NS, PREC>(a, i, lds, s_rhs, 1, mb, min(a.m_off[i + 1], mb + kChunk), b);
  }
}
template <int NS, int PREC>
__device__ inline void assemble_tail(const AsmArgs &a, int i, float *lds, float *s_rhs) {
  const int c0 = a.chunk_off[i], np = a.chunk_off[i + 1] - c0;
  if (np > 1) assemble_row_mfma<NS, PREC>(a, i, lds, s_rhs, 2, 0, 0, c0, np);
}
template <int NS, int PREC>
__global__ __launch_bounds__(64) void assemble_rows_mfma_kernel(AsmArgs a) {
  extern __shared__ float lds[];
  __shared__ float s_rhs[6];
  assemble_chunk<NS, PREC>(a, a.N, (int)blockIdx.x, lds, s_rhs);
}
template <int NS, int PREC>
__global__ __launch_bounds__(64) void assemble_rows_tail_kernel(AsmArgs a) {
  extern __shared__ float lds[];
  __shared__ float s_rhs[6];
  if (a.pcg_open && blockIdx.x == 0 && threadIdx.x == 0) pcg_open(a.pcg_open);
  assemble_tail<NS, PREC>(a, (int)blockIdx.x, lds, s_rhs);
}


// ----------------------------------------------------------------- PCG
// Cross-workgroup traffic of the persistent PCG uses a flag-in-data
// protocol and never relies on the L2 (each XCD has its own, and an
// agent-scope fence costs a write-back and an invalidation of a whole L2):
// every value another workgroup reads travels in a 64-bit word
// (tag << 32 | 32-bit payload) written and read with agent-scope relaxed
// atomic accesses, which gfx950 performs coherently at the device level.  A
// 64-bit access is single-copy atomic, so a reader that sees the expected tag
// also sees the payload written with it: no store ordering, no fences and no
// separate generation counters are needed, and a reader waits only for the
// words it uses.  The tag of a word is the synchronisation point it belongs
// to, (launch epoch << 16) + point, compared for equality.
typedef unsigned long long ll_word;

// One workgroup's all-reduce record: per parity the hi / lo halves of up to
// three partial sums, each parity in a 64-byte line of its own.  (Records of
// different workgroups in shared lines, word-major, measured slower: the
// writers of one line contend.)
struct alignas(128) SyncRec {
  ll_word w[2][8];
};
struct PcgCtl {
  unsigned long long epoch;  // advanced once per launch by pcg_prep_kernel
  unsigned abort;
  unsigned started;  // workgroups of the launch that have started (cleared with the epoch's advance)
  unsigned ticket;   // the last launch whose workgroups all started, by its caller's ticket
  unsigned pad;
};
constexpr int kPcgSub = 4;                        // lanes per scalar row
constexpr int kPcgRows = kPcgThreads / kPcgSub;   // scalar rows per workgroup
constexpr int kPcgMaxNodesPerWg = kPcgRows / 6;   // 42 nodes = 252 rows
constexpr int kPcgMaxWg = 256;                    // records (one per workgroup)
constexpr int kPcgMaxPoints = 65535;              // synchronisation points per launch (16 tag bits)
constexpr int kPcgWaves = kPcgThreads / 64;
// workgroup reduction scratch: first stage 3 per wave, second stage 3 per
// wave of record-holding lanes (at most kPcgMaxWg / 64 = 4 waves)
constexpr int kRedSize = 3 * kPcgWaves + 3 * 4;

// Wave-level sum of a double in a fixed order (pairs, quads, half rows,
// rows through DPP lane moves, then rows 0+1 and 2+3 and the two halves via
// the row broadcasts); the total is in lane 63.  Every lane of the wave must
// take part.
template <int CTRL, int ROWMASK>
__device__ inline double dpp_add(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROWMASK, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWMASK, 0xF, false);
  return v + __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
// A double that is equal in every lane, moved to scalar registers (read
// from the first active lane).
__device__ inline double uniform_d(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readfirstlane((int)b);
  const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
// A float from another lane of the same quad (DPP quad_perm; 0xB1 reads
// lane ^ 1, 0x4E lane ^ 2).  Every lane of the wave must take part.
template <int CTRL>
__device__ inline float quad_perm_f(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
// wave_sum_last for values that are equal in each aligned group of SUB
// lanes (SUB = 1, 2, 4), counted once per group: the quad steps that would
// add a group's copies are skipped.  Same bits as wave_sum_last over the
// values kept in the group's first lane and zeros elsewhere.
template <int SUB>
__device__ inline double wave_sum_groups(double v) {
  if constexpr (SUB == 1) v = dpp_add<0xB1, 0xF>(v);  // quad_perm [1,0,3,2]
  if constexpr (SUB <= 2) v = dpp_add<0x4E, 0xF>(v);  // quad_perm [2,3,0,1]
  v = dpp_add<0x141, 0xF>(v);  // row_half_mirror
  v = dpp_add<0x140, 0xF>(v);  // row_mirror
  v = dpp_add<0x142, 0xA>(v);  // row_bcast:15 into rows 1, 3
  v = dpp_add<0x143, 0xC>(v);  // row_bcast:31 into rows 2, 3
  return v;
}
__device__ inline double wave_sum_last(double v) {
  v = dpp_add<0xB1, 0xF>(v);   // quad_perm [1,0,3,2]
  v = dpp_add<0x4E, 0xF>(v);   // quad_perm [2,3,0,1]
  v = dpp_add<0x141, 0xF>(v);  // row_half_mirror
  v = dpp_add<0x140, 0xF>(v);  // row_mirror
  v = dpp_add<0x142, 0xA>(v);  // row_bcast:15 into rows 1, 3
  v = dpp_add<0x143, 0xC>(v);  // row_bcast:31 into rows 2, 3
  return v;
}

// Workgroup sums of a pair of doubles.  red holds 2 * kPcgWaves (first
// stage) + 8 (second stage) doubles; the two stages use separate parts so
// that one barrier each suffices between consecutive reductions.
//
// First stage (this workgroup's partials): every wave reduces its lanes, the
// wave sums are added in wave order by lanes 0..3, which are the lanes that
// publish them.
__device__ inline void wg_sum2_lanes03(double &v0, double &v1, double *red) {
  const double s0 = wave_sum_last(v0), s1 = wave_sum_last(v1);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 63) {
    red[2 * w] = s0;
    red[2 * w + 1] = s1;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double t0 = 0, t1 = 0;
    for (int k = 0; k < kPcgWaves; ++k) {
      t0 += red[2 * k];
      t1 += red[2 * k + 1];
    }
    v0 = t0;
    v1 = t1;
  }
}
// Second stage (the G records held by lanes 0..G-1): the waves that hold
// records reduce them, every lane adds those wave sums in order.
__device__ inline void wg_sum2_records(double &v0, double &v1, int G, double *red) {
  double *r2 = red + 2 * kPcgWaves;
  const int w = threadIdx.x >> 6, nw = (G + 63) >> 6;
  if (w < nw) {
    const double s0 = wave_sum_last(v0), s1 = wave_sum_last(v1);
    if ((threadIdx.x & 63) == 63) {
      r2[2 * w] = s0;
      r2[2 * w + 1] = s1;
    }
  }
  __syncthreads();
  double t0 = 0, t1 = 0;
  for (int k = 0; k < nw; ++k) {
    t0 += r2[2 * k];
    t1 += r2[2 * k + 1];
  }
  v0 = t0;
  v1 = t1;
}

// State of the true-residual rounds of one PCG run (solver option
// true_residual_rounds): after the PCG, r = b - A x is recomputed in FP64;
// while ||r|| >= tol ||b|| a correction PCG solves A d = r to the absolute
// tolerance tol ||b|| and x += d.  The run's outcome (converged, breakdown,
// final residual) is the true residual's, recorded by tr_decide_kernel.
struct TrState {
  int32_t done;        // the run has its outcome (correction rounds are no-ops)
  int32_t last_flags;  // stop flags of the run's last PCG
  double tol_abs;      // tol * ||b|| of the run's system
  float *r, *d;        // rhs and solution of the correction system (6N)
  double *partial;     // per-workgroup partial sums of the residual kernel
};

struct PcgArgs {
  const int32_t *row_ptr, *lcol;
  const float *values, *b, *Minv;
  float *x;
  ll_word *zpub, *ppub0, *ppub1;
  SyncRec *recs;
  PcgCtl *ctl;
  const int32_t *rows, *halo_off, *halo;
  int N, maxit, G, lds_a;
  float tol;
  int rel;
  nrf_pcg_result *res;
  const int32_t *stop;
  TrState *tr;  // true-residual rounds (null: off)
  int round;    // 0: the run's PCG; > 0: a correction round (A d = r, to tr->tol_abs)
  // 1: the pipelined form ends with one refinement step, x += M^-1 (b - A x)
  // with the residual in FP64 (x's halo through zpub); 2 and 4 lanes per row
  int refine;
  // non-zero: the last workgroup to start records it in ctl->ticket
  // (pcg_started_wait_kernel)
  unsigned ticket = 0;
};

// A correction round of a run that already has its outcome does nothing.
__device__ inline bool pcg_skip(const PcgArgs &a) { return a.tr && a.round > 0 && a.tr->done; }
// The stopping tolerance on ||r|| (bb = ||rhs||^2).
__device__ inline double pcg_tol(const PcgArgs &a, double bb) {
  if (a.tr && a.round > 0) return a.tr->tol_abs;
  return a.rel ? (double)a.tol * sqrt(bb) : (double)a.tol;
}
// The run's record (one lane): with true-residual rounds the outcome is left
// to tr_decide_kernel.
__device__ inline void pcg_tally(const PcgArgs &a, int iters, int flags, double rr, double bb) {
  a.res->total_iterations += iters;
  if (!a.tr) {
    a.res->iterations = iters;
    a.res->flags |= flags;
    a.res->final_residual = sqrt(rr);
    a.res->rhs_norm = sqrt(bb);
    a.res->converged += (flags & 2) ? 1 : 0;
    a.res->breakdowns += (flags & 1) ? 1 : 0;
    a.res->runs += 1;
    return;
  }
  a.res->iterations = a.round == 0 ? iters : a.res->iterations + iters;
  a.res->flags |= flags & 8;
  if (a.round == 0) a.tr->done = 0;
  a.tr->last_flags = flags;
}

// The words live in global memory: the casts give global (not flat)
// instructions also where the pointer came from memory (batch kernels).
typedef __attribute__((address_space(1))) ll_word global_ll_word;
__device__ inline void ll_put(ll_word *p, unsigned tag, unsigned payload) {
  __hip_atomic_store((global_ll_word *)p, ((ll_word)tag << 32) | payload, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline ll_word ll_load(const ll_word *p) {
  return __hip_atomic_load((global_ll_word *)const_cast<ll_word *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline bool ll_has(ll_word w, unsigned tag) { return (unsigned)(w >> 32) == tag; }

// Bounded wait of the polling loops: every 256 rounds the launch-wide abort
// flag is checked, and raised when the spin limit ran out.  Returns true when
// the caller has to give up (and then sets *bad).
__device__ inline bool ll_spin(unsigned &spins, PcgCtl *ctl, int *bad) {
  if ((++spins & 255u) == 0) {
    if (__hip_atomic_load(&ctl->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) || spins > kBarrierSpinLimit) {
      __hip_atomic_store(&ctl->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      *bad = 1;
      return true;
    }
  }
  __builtin_amdgcn_s_sleep(1);
  return false;
}

// Polling a word that was not ready at the first load: two loads in flight,
// the second issued kPollStagger sleep units (64 clocks each) after the
// first, re-issued alternately, so that a value landing just after one
// stream's load passed is seen by the other stream's next load.
constexpr int kPollStagger = 6;
__device__ inline ll_word ll_wait2(const ll_word *src, unsigned tag, PcgCtl *ctl, int *bad) {
  ll_word u = ll_load(src);
  __builtin_amdgcn_s_sleep(kPollStagger);
  ll_word w = ll_load(src);
  unsigned spins = 0;
  for (;;) {
    if (ll_has(u, tag) || ll_spin(spins, ctl, bad)) return u;
    u = ll_load(src);
    if (ll_has(w, tag) || ll_spin(spins, ctl, bad)) return w;
    w = ll_load(src);
  }
}

// All-reduce of two doubles across the G workgroups, in two halves so that
// work can be placed between them.  ar2_publish: this workgroup's sums go
// out under `tag` (lanes 0..3 write the hi / lo words).  ar2_collect: lanes
// 0..G-1 each read one workgroup's record (its four words in one round trip
// per poll) and the records are summed in lane order, so every workgroup
// holds the same bits.  Records are double-buffered by the parity of the
// point: a workgroup is at most one point ahead of the slowest one.
// ar2_collect returns true when the solve has to be abandoned.
__device__ inline void ar2_publish(const PcgArgs &a, int g, unsigned tag, int slot, double v0, double v1,
                                   double *red) {
  wg_sum2_lanes03(v0, v1, red);
  const int t = threadIdx.x;
  if (t < 4) {
    const ll_word bits = (ll_word)__double_as_longlong(t < 2 ? v0 : v1);
    ll_put(&a.recs[g].w[slot][t], tag, (t & 1) ? (unsigned)bits : (unsigned)(bits >> 32));
  }
}
__device__ inline bool ar2_collect(const PcgArgs &a, unsigned tag, int slot, double &v0, double &v1,
                                   double *red, int *s_bad) {
  const int t = threadIdx.x;
  double g0 = 0, g1 = 0;
  if (t < a.G) {
    const ll_word *w = a.recs[t].w[slot];
    ll_word w0, w1, w2, w3;
    unsigned spins = 0;
    for (;;) {
      w0 = ll_load(w + 0);
      w1 = ll_load(w + 1);
      w2 = ll_load(w + 2);
      w3 = ll_load(w + 3);
      if (ll_has(w0, tag) && ll_has(w1, tag) && ll_has(w2, tag) && ll_has(w3, tag)) break;
      if (ll_spin(spins, a.ctl, s_bad)) break;
    }
    g0 = __longlong_as_double((long long)((w0 << 32) | (w1 & 0xffffffffull)));
    g1 = __longlong_as_double((long long)((w2 << 32) | (w3 & 0xffffffffull)));
  }
  wg_sum2_records(g0, g1, a.G, red);  // its barrier also publishes *s_bad
  v0 = g0;
  v1 = g1;
  return *s_bad != 0;
}
__device__ inline bool allreduce2(const PcgArgs &a, unsigned tag, int slot, double &v0, double &v1, double *red,
                                  int *s_bad) {
  ar2_publish(a, (int)blockIdx.x, tag, slot, v0, v1, red);
  return ar2_collect(a, tag, slot, v0, v1, red, s_bad);
}

// The same all-reduce for three doubles (lanes 0..5 publish the hi / lo
// words; the sums follow the fixed orders of the two-value form).  The
// partials are equal in each group of SUB lanes (a row's lanes) and counted
// once per group; only the first nwa waves hold rows (the others would add
// zeros and are skipped).
template <int SUB>
__device__ inline void ar3_publish(const PcgArgs &a, int g, unsigned tag, int slot, double (&v)[3], double *red,
                                   int nwa) {
  const int t = threadIdx.x, w = t >> 6;
  if (w < nwa) {
    double s[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) s[k] = wave_sum_groups<SUB>(v[k]);
    if ((t & 63) == 63)
      for (int k = 0; k < 3; ++k) red[3 * w + k] = s[k];
  }
  __syncthreads();
  if (t < 6) {
    double acc = 0;
    for (int kk = 0; kk < nwa; ++kk) acc += red[3 * kk + (t >> 1)];
    const ll_word bits = (ll_word)__double_as_longlong(acc);
    ll_put(&a.recs[g].w[slot][t], tag, (t & 1) ? (unsigned)bits : (unsigned)(bits >> 32));
  }
}
__device__ inline bool ar3_collect(const PcgArgs &a, unsigned tag, int slot, double (&v)[3], double *red,
                                   int *s_bad) {
  const int t = threadIdx.x;
  double g[3] = {0, 0, 0};
  if (t < a.G) {
    const ll_word *w = a.recs[t].w[slot];
    ll_word x[6], y[6];
    bool all = true;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      x[k] = ll_load(w + k);
      all = all && ll_has(x[k], tag);
    }
    if (!all) {  // two streams of polls, as in ll_wait2
#pragma unroll
      for (int k = 0; k < 6; ++k) x[k] = ll_load(w + k);
      __builtin_amdgcn_s_sleep(kPollStagger);
#pragma unroll
      for (int k = 0; k < 6; ++k) y[k] = ll_load(w + k);
      unsigned spins = 0;
      for (;;) {
        all = true;
#pragma unroll
        for (int k = 0; k < 6; ++k) all = all && ll_has(x[k], tag);
        if (all || ll_spin(spins, a.ctl, s_bad)) break;
#pragma unroll
        for (int k = 0; k < 6; ++k) x[k] = ll_load(w + k);
        all = true;
#pragma unroll
        for (int k = 0; k < 6; ++k) all = all && ll_has(y[k], tag);
        if (all || ll_spin(spins, a.ctl, s_bad)) {
#pragma unroll
          for (int k = 0; k < 6; ++k) x[k] = y[k];
          break;
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) y[k] = ll_load(w + k);
      }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k)
      g[k] = __longlong_as_double((long long)((x[2 * k] << 32) | (x[2 * k + 1] & 0xffffffffull)));
  }
  double *r2 = red + 3 * kPcgWaves;
  const int w = t >> 6, nw = (a.G + 63) >> 6;
  if (w < nw) {
    double sm[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) sm[k] = wave_sum_last(g[k]);
    if ((t & 63) == 63)
      for (int k = 0; k < 3; ++k) r2[3 * w + k] = sm[k];
  }
  __syncthreads();  // also publishes *s_bad
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double acc = 0;
    for (int kk = 0; kk < nw; ++kk) acc += r2[3 * kk + k];
    v[k] = acc;
  }
  return *s_bad != 0;
}

// Block-Jacobi PCG (kernels/solver_kernels.cu:616-687, Saad Alg. 9.1, x0 = 0)
// as one persistent kernel.  Workgroup g owns nodes [rows[g], rows[g+1]),
// at most 42, i.e. at most 252 scalar rows; four adjacent lanes share scalar
// row q = t / 4, hold identical copies of its x, r, z, p, M^-1 row in
// registers for the whole solve and split the row's blocks among themselves
// in the product.  The blocks of the owned rows (and their columns,
// renumbered into the workgroup's halo list) are copied to LDS once.  Per
// iteration:
//   A: p = z + beta p; the halo's p_j = z_j + beta p_j: own nodes from LDS,
//      the others from the published words; Ap from LDS; all-reduce p.Ap;
//   B: x += alpha p, r -= alpha Ap, z = M^-1 r, publish z; all-reduce
//      (r.r, r.z).
// Every workgroup evaluates the same expressions on the same published
// bits, so the owner's p and its readers' copies are identical.
__global__ __launch_bounds__(kPcgThreads) void pcg_kernel(PcgArgs a) {
  extern __shared__ float lds[];
  if (threadIdx.x == 0) {  // resident (pcg_started_wait_kernel)
    const unsigned n = __hip_atomic_fetch_add(&a.ctl->started, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1;
    if (a.ticket && n == (unsigned)a.G)
      __hip_atomic_store(&a.ctl->ticket, a.ticket, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __shared__ double red[kRedSize];
  __shared__ int s_bad;
  __shared__ unsigned s_base;
  if ((a.stop && *a.stop) || pcg_skip(a)) return;
  const int g = blockIdx.x, t = threadIdx.x;
  const int n0 = a.rows[g], n1 = a.rows[g + 1];
  const int S = 6 * (n1 - n0);
  const int blk0 = a.row_ptr[n0], nblk = a.row_ptr[n1] - blk0;
  const int h0 = a.halo_off[g], nh = a.halo_off[g + 1] - h0;
  float *Al = lds;
  int32_t *Ll = reinterpret_cast<int32_t *>(Al + (a.lds_a ? (size_t)nblk * 36 : 0));
  float *ph = reinterpret_cast<float *>(Ll + (a.lds_a ? nblk : 0));
  float *rl = ph + (size_t)nh * 6;  // own rows' r, then own rows' p (kPcgRows each)
  float *pl = rl + kPcgRows;
  if (a.lds_a) {
    const float4 *src = reinterpret_cast<const float4 *>(a.values + (size_t)blk0 * 36);
    float4 *dst = reinterpret_cast<float4 *>(Al);
    for (int j = t; j < nblk * 9; j += kPcgThreads) dst[j] = src[j];
    for (int j = t; j < nblk; j += kPcgThreads) Ll[j] = a.lcol[blk0 + j];
  }
  if (t == 0) {
    s_base = (unsigned)(__hip_atomic_load(&a.ctl->epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) << 16);
    s_bad = 0;
  }
  const int q = t / kPcgSub, sub = t % kPcgSub;
  const bool own = q < S, lead = own && sub == 0;
  const int li = q / 6, row = q % 6, i = n0 + li, s = 6 * n0 + q;
  float M[6] = {0, 0, 0, 0, 0, 0};
  float bv = 0.f;
  int rb = 0, re = 0;
  if (own) {
    for (int c = 0; c < 6; ++c) M[c] = a.Minv[(size_t)i * 36 + row * 6 + c];
    bv = a.b[s];
    rb = a.row_ptr[i];
    re = a.row_ptr[i + 1];
  }
  // init: x = 0, r = b, z = M^-1 r, p = z (kernels/solver_kernels.cu:620-632)
  float x = 0.f, r = bv, z = 0.f, p = 0.f;
  if (sub == 0) rl[q] = r;
  __syncthreads();
  const unsigned base = s_base;
  unsigned pt = 1;
  if (own) {
    for (int c = 0; c < 6; ++c) z += M[c] * rl[li * 6 + c];
    if (lead) ll_put(a.zpub + s, base + pt, __float_as_uint(z));
  }
  unsigned z_tag = base + pt, pold_tag = 0;
  double rz = lead ? (double)r * (double)z : 0.0, bb = lead ? (double)bv * (double)bv : 0.0;
  bool aborted = allreduce2(a, base + pt, (int)(pt & 1), rz, bb, red, &s_bad);
  double rr = bb;
  const double tol = pcg_tol(a, bb);
  int iters = 0, flags = 0;
  float betaf = 0.f;
  for (int it = 0; it < a.maxit && !aborted; ++it) {
    const ll_word *Pold = (it & 1) ? a.ppub1 : a.ppub0;
    ll_word *Pnew = (it & 1) ? a.ppub0 : a.ppub1;
    ++pt;  // the p.Ap point
    // A: own p (registers -> LDS), the halo's p, Ap, p.Ap
    p = (it == 0) ? z : z + betaf * p;
    if (lead) {
      ll_put(Pnew + s, base + pt, __float_as_uint(p));
      pl[q] = p;
    }
    __syncthreads();
    for (int e = t; e < nh * 6; e += kPcgThreads) {
      const int nd = a.halo[h0 + e / 6], c = e % 6;
      if (nd >= n0 && nd < n1) {
        ph[e] = pl[(nd - n0) * 6 + c];
      } else {
        const int idx = 6 * nd + c;
        ll_word wz, wp = 0;
        unsigned spins = 0;
        for (;;) {  // both words in one round trip
          wz = ll_load(a.zpub + idx);
          if (it > 0) wp = ll_load(Pold + idx);
          if (ll_has(wz, z_tag) && (it == 0 || ll_has(wp, pold_tag))) break;
          if (ll_spin(spins, a.ctl, &s_bad)) break;
        }
        const float zv = __uint_as_float((unsigned)wz);
        ph[e] = (it == 0) ? zv : zv + betaf * __uint_as_float((unsigned)wp);
      }
    }
    __syncthreads();
    float y = 0.f;
    if (own) {
      for (int bk = rb + sub; bk < re; bk += kPcgSub) {
        const int lc = a.lds_a ? Ll[bk - blk0] : a.lcol[bk];
        const float *A = a.lds_a ? (Al + (size_t)(bk - blk0) * 36) : (a.values + (size_t)bk * 36);
        const float *pj = ph + lc * 6;
        for (int c = 0; c < 6; ++c) y += A[row * 6 + c] * pj[c];
      }
    }
    // the row's four partial products, combined identically in all four lanes
    y += __shfl_xor(y, 1, 64);
    y += __shfl_xor(y, 2, 64);
    pold_tag = base + pt;
    double pAp = lead ? (double)p * (double)y : 0.0, unused = 0.0;
    aborted = allreduce2(a, base + pt, (int)(pt & 1), pAp, unused, red, &s_bad);
    if (aborted) break;
    // the reference's absolute guard (|p.Ap| < 1e-12); with the relative
    // stopping rule only a true breakdown (p.Ap not positive) stops the run
    if (a.rel ? !(pAp > 0.0) : fabs(pAp) < 1e-12) {
      flags |= 1;
      break;
    }
    const float alf = (float)(rz / pAp);
    // B: x += alpha p, r -= alpha Ap, z = M^-1 r, (r.r, r.z)
    ++pt;  // the (r.r, r.z) point
    x += alf * p;
    r += -alf * y;
    if (sub == 0) rl[q] = r;
    __syncthreads();
    z = 0.f;
    if (own) {
      for (int c = 0; c < 6; ++c) z += M[c] * rl[li * 6 + c];
      if (lead) ll_put(a.zpub + s, base + pt, __float_as_uint(z));
    }
    z_tag = base + pt;
    double rr_l = lead ? (double)r * (double)r : 0.0, rz_l = lead ? (double)r * (double)z : 0.0;
    aborted = allreduce2(a, base + pt, (int)(pt & 1), rr_l, rz_l, red, &s_bad);
    if (aborted) break;
    rr = rr_l;
    iters = it + 1;
    if (sqrt(rr) < tol) {
      flags |= 2;
      break;
    }
    betaf = (float)(rz_l / rz);
    rz = rz_l;
  }
  if (aborted) flags |= 8;
  if (lead) a.x[s] = x;
  if (g == 0 && t == 0) {
    pcg_tally(a, iters, flags, rr, bb);
  }
}

// Block-Jacobi PCG with the product overlapping the second all-reduce: the
// recurrences of kernels/solver_kernels.cu:616-687 (Saad Alg. 9.1, x0 = 0)
// with A p_{k+1} = A z_{k+1} + beta A p_k, so that the only operand another
// workgroup's rows contribute is z, published before the (r.r, r.z)
// all-reduce.  Per iteration:
//   A: all-reduce p.Ap; alpha = r.z / p.Ap (breakdown |p.Ap| < 1e-12);
//   B: x += alpha p, r -= alpha Ap, z = M^-1 r, publish z and the (r.r, r.z)
//      partials; while they travel, gather the halo's z and form A z; then
//      collect, stop on ||r|| < tol, beta = r.z / (r.z)_prev,
//      p = z + beta p, Ap = A z + beta Ap.
// Same iterates as pcg_kernel in exact arithmetic; in FP32 the one extra
// recurrence (for Ap) keeps the iterates within a few ulps of the textbook
// ones.  Layout, publication and the lane-per-row split are pcg_kernel's.
template <int SUB, int FORM>
__device__ void pcg_pipe_body(const PcgArgs &a, int g, float *lds, double *red, int &s_bad, unsigned &s_base) {
  constexpr int ROWS = kPcgThreads / SUB;
  // this workgroup is resident (pcg_started_wait_kernel)
  if (threadIdx.x == 0) {
    const unsigned n = __hip_atomic_fetch_add(&a.ctl->started, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1;
    if (a.ticket && n == (unsigned)a.G)
      __hip_atomic_store(&a.ctl->ticket, a.ticket, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if ((a.stop && *a.stop) || pcg_skip(a)) return;
  const int t = threadIdx.x;
  const int n0 = a.rows[g], n1 = a.rows[g + 1];
  const int S = 6 * (n1 - n0);
  const int blk0 = a.row_ptr[n0], nblk = a.row_ptr[n1] - blk0;
  const int h0 = a.halo_off[g], nh = a.halo_off[g + 1] - h0;
  float *Al = lds;
  int32_t *Ll = reinterpret_cast<int32_t *>(Al + (a.lds_a ? (size_t)nblk * 36 : 0));
  // the halo's values, then one zero entry (index nh) that the register
  // blocks of short rows point at
  float *ph = reinterpret_cast<float *>(Ll + (a.lds_a ? nblk : 0));
  float *rl = ph + (size_t)(nh + 1) * 6;  // own rows' r (ROWS entries)
  if (a.lds_a) {
    const float4 *src = reinterpret_cast<const float4 *>(a.values + (size_t)blk0 * 36);
    float4 *dst = reinterpret_cast<float4 *>(Al);
    for (int j = t; j < nblk * 9; j += kPcgThreads) dst[j] = src[j];
    for (int j = t; j < nblk; j += kPcgThreads) Ll[j] = a.lcol[blk0 + j];
  }
  if (t < 6) ph[nh * 6 + t] = 0.f;
  // the own nodes are a contiguous run of the sorted halo list, from s_o0
  __shared__ int s_o0;
  for (int e = t; e < nh; e += kPcgThreads)
    if (a.halo[h0 + e] == n0) s_o0 = e;
  if (t == 0) {
    s_base = (unsigned)(__hip_atomic_load(&a.ctl->epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) << 16);
    s_bad = 0;
  }
  // Lanes (SUB = 2, 4): each wave holds NPW whole nodes (6 SUB lanes each,
  // the rest of the wave idle), so that a node's six r values are exchanged
  // inside one wave, through LDS without a barrier.  Node li of the
  // workgroup sits in wave li / NPW; its row `row` in the SUB lanes from
  // (li % NPW) 6 SUB + row SUB.  SUB = 1 packs the rows contiguously (170
  // nodes in 1020 lanes, nodes straddle waves) and exchanges r behind a
  // barrier.
  constexpr bool WN = SUB != 1;
  constexpr int NPW = WN ? 64 / (6 * SUB) : 64;
  const int ln = WN ? (t & 63) / (6 * SUB) : 0, lr = WN ? (t & 63) % (6 * SUB) : t % SUB;
  const int q = WN ? ((t >> 6) * NPW + ln) * 6 + lr / SUB : t / SUB;
  const int li = q / 6, row = q % 6, sub = lr % SUB;
  const bool own = ln < NPW && q < S, lead = own && sub == 0;
  const int i = n0 + li, s = 6 * n0 + q;
  float M[6] = {0, 0, 0, 0, 0, 0};
  float bv = 0.f;
  int rb = 0, re = 0;
  if (own) {
    for (int c = 0; c < 6; ++c) M[c] = a.Minv[(size_t)i * 36 + row * 6 + c];
    bv = a.b[s];
    rb = a.row_ptr[i];
    re = a.row_ptr[i + 1];
  }
  // The lane's first NR blocks of its row (bk = rb + sub + k SUB) stay in
  // registers for the whole solve: their six coefficients and their halo
  // column.  Missing blocks are zero and point at the zero entry, so the
  // product reads every register block without a branch (adding the exact
  // zeros leaves the sum's bits unchanged).  Blocks past NR (rows longer
  // than NR SUB blocks) are read from LDS or global memory.
  constexpr int NR = FORM == 2 ? (SUB == 4 ? 2 : (SUB == 2 ? 3 : 3)) : SUB == 4 ? 3 : (SUB == 2 ? 5 : 4);
  float Ar[NR][6];
  int Lr[NR];
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    const int bk = rb + sub + k * SUB;
    const bool have = own && bk < re;
    Lr[k] = have ? a.lcol[bk] * 6 : nh * 6;
#pragma unroll
    for (int c = 0; c < 6; ++c) Ar[k][c] = have ? a.values[(size_t)bk * 36 + row * 6 + c] : 0.f;
  }
  const int bk_rest = rb + sub + NR * SUB;
  // the node of this lane's first halo entry (e = t), read once
  const int hn0 = t < nh * 6 ? a.halo[h0 + t / 6] : -1;
  int zo = 0;  // ph index of this lane's row among the own nodes' entries
  // z = M^-1 r for this lane's row, published under `tag` and stored in the
  // own nodes' entries of ph for the product.  A node's six r values come
  // from lanes of the same wave: LDS store and loads in program order, no
  // barrier.
  auto precond_publish = [&](float r, unsigned tag, ll_word *pub) -> float {
    if (lead) rl[q] = r;
    if constexpr (WN)
      __builtin_amdgcn_wave_barrier();  // keeps the loads after the store
    else
      __syncthreads();
    float z = 0.f;
    if (own) {
      for (int c = 0; c < 6; ++c) z += M[c] * rl[li * 6 + c];
      if (lead) {
        ll_put(pub + s, tag, __float_as_uint(z));
        ph[zo] = z;
      }
    }
    return z;
  };
  // (A z) for this lane's row: the halo's z from the words published under
  // `tag` (own nodes' entries are already in ph).  The callers order the
  // gather after every read of ph by the previous product (a barrier lies
  // between).
  // ACC = float: products and sums in FP32.  ACC = double: each product is
  // exact in FP64 (two FP32 factors), the sums FP64.
  auto product = [&](unsigned tag, const ll_word *pub, auto acc0) {
    using ACC = decltype(acc0);
    for (int e = t; e < nh * 6; e += kPcgThreads) {
      const int nd = e == t ? hn0 : a.halo[h0 + e / 6], c = e % 6;
      if (nd < n0 || nd >= n1) {
        const ll_word *src = pub + 6 * nd + c;
        ll_word wv = ll_load(src);
        if (!ll_has(wv, tag)) wv = ll_wait2(src, tag, a.ctl, &s_bad);
        ph[e] = __uint_as_float((unsigned)wv);
      }
    }
    __syncthreads();
    ACC y = 0;
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      const float *pj = ph + Lr[k];
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        y += (ACC)Ar[k][c] * (ACC)pj[c];
      }
    }
    if (own && bk_rest < re) {
      if (a.lds_a) {
        for (int bk = bk_rest; bk < re; bk += SUB) {
          const float *A = Al + (size_t)(bk - blk0) * 36 + row * 6;
          const float *pj = ph + Ll[bk - blk0] * 6;
          for (int c = 0; c < 6; ++c) y += (ACC)A[c] * (ACC)pj[c];
        }
      } else {
        for (int bk = bk_rest; bk < re; bk += SUB) {
          const float *A = a.values + (size_t)bk * 36 + row * 6;
          const float *pj = ph + a.lcol[bk] * 6;
          for (int c = 0; c < 6; ++c) y += (ACC)A[c] * (ACC)pj[c];
        }
      }
    }
    // the row's SUB partial products, combined identically in all its lanes
    if constexpr (std::is_same<ACC, double>::value) {
      if constexpr (SUB >= 2) y = dpp_add<0xB1, 0xF>(y);  // lane ^ 1
      if constexpr (SUB >= 4) y = dpp_add<0x4E, 0xF>(y);  // lane ^ 2
    } else {
      if constexpr (SUB >= 2) y += quad_perm_f<0xB1>(y);  // lane ^ 1
      if constexpr (SUB >= 4) y += quad_perm_f<0x4E>(y);  // lane ^ 2
    }
    return y;
  };

  __syncthreads();
  const unsigned base = s_base;
  zo = (s_o0 + li) * 6 + row;
  // the reduction partials are held by all SUB lanes of a row (equal copies)
  const int nwa = WN ? (n1 - n0 + NPW - 1) / NPW : (S * SUB + 63) >> 6;
  if constexpr (FORM == 2) {
    // Pipelined PCG (Ghysels and Vanroose, 2014, Alg. 4, with the block-Jacobi
    // M): besides x, r, u = M^-1 r and p, each row carries w = A u, s = A p,
    // q = M^-1 s and zz = A q as recurrences.  The all-reduce of
    // (r.r, r.u, w.u) starts an iteration and travels while m = M^-1 w is
    // published, its halo gathered and n = A m formed, so the iteration's one
    // halo exchange and one all-reduce overlap.  p.Ap = w.u - beta^2 (p.Ap)_old
    // as in the single-reduction form; the same iterates as the textbook form
    // in exact arithmetic.
    // Rounding: these coupled recurrences stay close to the textbook iterates
    // only while n is the product of A with exactly the m that enters q, so m
    // is one FP32 value (published, in the own entries of ph and in q alike),
    // n = A m is summed in FP64 from exact products, and the coupled vectors
    // r, u, w, s, q, zz are FP64 (FP32 vectors or an FP32 sum, even per block,
    // leave x 1e-4..1e-3 from the FP64 solve after 30 iterations and stall the
    // residual near 1e-3).  x and p feed nothing else and stay FP32.
    // The published m alternates between two arrays by point parity: a
    // workgroup publishes its record of a point before it gathers that
    // point's halo, so a neighbour can already publish the next point's m, but
    // not the one after (that waits for this workgroup's next record).  After
    // maxit updates one more all-reduce yields ||r||.
    unsigned pt = 1;
    float x = 0.f, p = 0.f;
    double r = bv;
    double u = precond_publish(bv, base + pt, a.ppub1);
    double w = product(base + pt, a.ppub1, 0.0);
    // every read of ph by this product ends before the first iteration
    // stores the own nodes' m there (in the loop, the collection's barrier
    // does this)
    __syncthreads();
    double sv = 0.0, qv = 0.0, zv = 0.0;
    double v[3], gamma_old = 0.0, pAp = 0.0, rr = 0.0, bb = 0.0, tol2 = 0.0;
    int iters = 0, flags = 0;
    bool aborted = false;
    for (int it = 0;; ++it) {
      ++pt;
      ll_word *pub = (pt & 1) ? a.ppub1 : a.ppub0;
      const bool last = it >= a.maxit;
      float m = 0.f;
      double n = 0.0;
      if (!last) m = precond_publish((float)w, base + pt, pub);
      v[0] = own ? r * r : 0.0;
      v[1] = own ? r * u : 0.0;
      v[2] = own ? w * u : 0.0;
      ar3_publish<SUB>(a, g, base + pt, (int)(pt & 1), v, red, nwa);
      if (!last) n = product(base + pt, pub, 0.0);
      aborted = ar3_collect(a, base + pt, (int)(pt & 1), v, red, &s_bad);
      if (aborted) break;
      // the sums are equal in all lanes: scalar registers for the loop state
      rr = uniform_d(v[0]);
      iters = it;
      if (it == 0) {
        bb = rr;
        const double tol = pcg_tol(a, bb);
        tol2 = uniform_d(tol * tol);
      } else if (rr < tol2) {
        flags |= 2;
        break;
      }
      if (last) break;
      const double gamma = uniform_d(v[1]);
      double beta = 0.0;
      if (it == 0) {
        pAp = uniform_d(v[2]);
      } else {
        beta = uniform_d(gamma / gamma_old);
        pAp = uniform_d(v[2] - beta * beta * pAp);
      }
      if (a.rel ? !(pAp > 0.0) : fabs(pAp) < 1e-12) {
        flags |= 1;
        break;
      }
      gamma_old = gamma;
      const double alpha = uniform_d(gamma / pAp);
      zv = n + beta * zv;
      qv = (double)m + beta * qv;
      sv = w + beta * sv;
      p = (float)u + (float)beta * p;
      x += (float)alpha * p;
      r -= alpha * sv;
      u -= alpha * qv;
      w -= alpha * zv;
    }
    if (aborted) flags |= 8;
    if constexpr (WN) {
      if (a.refine && !aborted) {
        // One refinement step inside the launch (the precision form's):
        // x's halo travels as published words like m's; r = b - A x sums
        // exact FP64 products of the FP32 matrix and x; then x += M^-1 r with
        // the node's six residuals from the lanes of its rows (one wave).
        // Every read of ph by the last product ended at the collection's
        // barrier, so the own entries can take x now.
        ++pt;
        if (lead) {
          ll_put(a.zpub + s, base + pt, __float_as_uint(x));
          ph[zo] = x;
        }
        const double res = (double)bv - product(base + pt, a.zpub, 0.0);
        const int l0 = (t & 63) - lr;  // the lane of row 0 of this lane's node
        double z = 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c) z += (double)M[c] * __shfl(res, l0 + c * SUB, 64);
        x = (float)((double)x + z);
        if (s_bad) flags |= 8;
      }
    }
    if (lead) a.x[s] = x;
    if (g == 0 && t == 0) {
      pcg_tally(a, iters, flags, rr, bb);
    }
    return;
  }
  if constexpr (FORM == 1) {
    // One all-reduce per iteration (Chronopoulos and Gear, 1989): with
    // gamma = r.z and delta = z.Az reduced together with r.r, p.Ap follows
    // from delta - beta gamma / alpha_old (equal to p.Ap in exact
    // arithmetic, since z_{k+1}.r_k = 0 and p_k.Ap_k = gamma_k / alpha_k),
    // so A z no longer waits for a p.Ap reduction.  With beta = gamma /
    // gamma_old the correction is beta^2 (p.Ap)_old, which saves a division.
    // A p is formed as A z + beta A p as in the overlapped form.
    unsigned pt = 1;
    float x = 0.f, r = bv;
    float z = precond_publish(r, base + pt, a.zpub);
    float az = product(base + pt, a.zpub, 0.f);
    double v[3] = {own ? (double)r * (double)z : 0.0, own ? (double)z * (double)az : 0.0,
                   own ? (double)bv * (double)bv : 0.0};
    ar3_publish<SUB>(a, g, base + pt, (int)(pt & 1), v, red, nwa);
    bool aborted = ar3_collect(a, base + pt, (int)(pt & 1), v, red, &s_bad);
    double gamma = v[0], pAp = v[1];
    const double bb = v[2];
    double rr = bb;
    const double tol = pcg_tol(a, bb), tol2 = tol * tol;
    int iters = 0, flags = 0;
    float p = z, y = az;
    for (int it = 0; it < a.maxit && !aborted; ++it) {
      if (a.rel ? !(pAp > 0.0) : fabs(pAp) < 1e-12) {
        flags |= 1;
        break;
      }
      const double alpha = gamma / pAp;
      const float alf = (float)alpha;
      x += alf * p;
      r += -alf * y;
      ++pt;
      z = precond_publish(r, base + pt, a.zpub);
      az = product(base + pt, a.zpub, 0.f);
      v[0] = own ? (double)r * (double)r : 0.0;
      v[1] = own ? (double)r * (double)z : 0.0;
      v[2] = own ? (double)z * (double)az : 0.0;
      ar3_publish<SUB>(a, g, base + pt, (int)(pt & 1), v, red, nwa);
      aborted = ar3_collect(a, base + pt, (int)(pt & 1), v, red, &s_bad);
      if (aborted) break;
      rr = v[0];
      iters = it + 1;
      if (rr < tol2) {
        flags |= 2;
        break;
      }
      const double beta = v[1] / gamma;
      gamma = v[1];
      pAp = v[2] - beta * beta * pAp;  // beta gamma / alpha_old = beta^2 (p.Ap)_old
      const float betaf = (float)beta;
      p = z + betaf * p;
      y = az + betaf * y;
    }
    if (aborted) flags |= 8;
    if (lead) a.x[s] = x;
    if (g == 0 && t == 0) {
      pcg_tally(a, iters, flags, rr, bb);
    }
    return;
  }
  // init: x = 0, r = b, z = M^-1 r, p = z, Ap = A z (kernels/solver_kernels.cu:620-632)
  unsigned pt = 1;
  float x = 0.f, r = bv;
  float z = precond_publish(r, base + pt, a.zpub);
  double rz = lead ? (double)r * (double)z : 0.0, bb = lead ? (double)bv * (double)bv : 0.0;
  ar2_publish(a, g, base + pt, (int)(pt & 1), rz, bb, red);
  float p = z, y = product(base + pt, a.zpub, 0.f);
  bool aborted = ar2_collect(a, base + pt, (int)(pt & 1), rz, bb, red, &s_bad);
  double rr = bb;
  const double tol = pcg_tol(a, bb);
  int iters = 0, flags = 0;
  for (int it = 0; it < a.maxit && !aborted; ++it) {
    ++pt;  // the p.Ap point
    double pAp = lead ? (double)p * (double)y : 0.0, unused = 0.0;
    ar2_publish(a, g, base + pt, (int)(pt & 1), pAp, unused, red);
    aborted = ar2_collect(a, base + pt, (int)(pt & 1), pAp, unused, red, &s_bad);
    if (aborted) break;
    // the reference's absolute guard (|p.Ap| < 1e-12); with the relative
    // stopping rule only a true breakdown (p.Ap not positive) stops the run
    if (a.rel ? !(pAp > 0.0) : fabs(pAp) < 1e-12) {
      flags |= 1;
      break;
    }
    const float alf = (float)(rz / pAp);
    ++pt;  // the (r.r, r.z) point
    x += alf * p;
    r += -alf * y;
    z = precond_publish(r, base + pt, a.zpub);
    double rr_l = lead ? (double)r * (double)r : 0.0, rz_l = lead ? (double)r * (double)z : 0.0;
    ar2_publish(a, g, base + pt, (int)(pt & 1), rr_l, rz_l, red);
    const float az = product(base + pt, a.zpub, 0.f);
    aborted = ar2_collect(a, base + pt, (int)(pt & 1), rr_l, rz_l, red, &s_bad);
    if (aborted) break;
    rr = rr_l;
    iters = it + 1;
    if (sqrt(rr) < tol) {
      flags |= 2;
      break;
    }
    const float betaf = (float)(rz_l / rz);
    rz = rz_l;
    p = z + betaf * p;
    y = az + betaf * y;
  }
  if (aborted) flags |= 8;
  if (lead) a.x[s] = x;
  if (g == 0 && t == 0) {
    pcg_tally(a, iters, flags, rr, bb);
  }
}

template <int SUB, int FORM>
__global__ __launch_bounds__(kPcgThreads) void pcg_pipe_kernel(PcgArgs a) {
  extern __shared__ float lds[];
  __shared__ double red[kRedSize];
  __shared__ int s_bad;
  __shared__ unsigned s_base;
  pcg_pipe_body<SUB, FORM>(a, (int)blockIdx.x, lds, red, s_bad, s_base);
}


// r = b - A x for every scalar row, products and sums in FP64 (the matrix
// and x are FP32); optionally stores r (rounded to FP32) and per-workgroup
// partial sums of |r|^2 and |b|^2 for the true relative residual.
constexpr int kResThreads = 256;
__device__ inline void residual_fp64_body(const int32_t *__restrict__ row_ptr, const int32_t *__restrict__ col_idx,
                                          const float *__restrict__ values, const float *__restrict__ b,
                                          const float *__restrict__ x, int N, float *__restrict__ r_out,
                                          double *__restrict__ partial) {
  __shared__ double red[2 * kResThreads / 64];
  const int q = blockIdx.x * kResThreads + threadIdx.x;
  double rr = 0.0, bb = 0.0;
  if (q < 6 * N) {
    const int i = q / 6, row = q % 6;
    double acc = 0.0;
    for (int bk = row_ptr[i]; bk < row_ptr[i + 1]; ++bk) {
      const float *A = values + (size_t)bk * 36 + row * 6;
      const float *xj = x + 6 * col_idx[bk];
      for (int c = 0; c < 6; ++c) acc += (double)A[c] * (double)xj[c];
    }
    const double r = (double)b[q] - acc;
    if (r_out) r_out[q] = (float)r;
    rr = r * r;
    bb = (double)b[q] * (double)b[q];
  }
  rr = wave_sum_last(rr);
  bb = wave_sum_last(bb);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 63) {
    red[2 * w] = rr;
    red[2 * w + 1] = bb;
  }
  __syncthreads();
  if (threadIdx.x == 0 && partial) {
    double s0 = 0, s1 = 0;
    for (int k = 0; k < kResThreads / 64; ++k) {
      s0 += red[2 * k];
      s1 += red[2 * k + 1];
    }
    partial[2 * blockIdx.x] = s0;
    partial[2 * blockIdx.x + 1] = s1;
  }
}
__global__ __launch_bounds__(kResThreads) void residual_fp64_kernel(const int32_t *__restrict__ row_ptr,
                                                                    const int32_t *__restrict__ col_idx,
                                                                    const float *__restrict__ values,
                                                                    const float *__restrict__ b,
                                                                    const float *__restrict__ x, int N,
                                                                    float *__restrict__ r_out,
                                                                    double *__restrict__ partial) {
  residual_fp64_body(row_ptr, col_idx, values, b, x, N, r_out, partial);
}

// ---- true-residual rounds (TrState): residual, decision, correction add.
// A raised GN stop flag, or a run that has its outcome, makes each a no-op.
__global__ __launch_bounds__(kResThreads) void tr_residual_kernel(const int32_t *__restrict__ row_ptr,
                                                                  const int32_t *__restrict__ col_idx,
                                                                  const float *__restrict__ values,
                                                                  const float *__restrict__ b,
                                                                  const float *__restrict__ x, int N,
                                                                  TrState *tr, const int32_t *stop) {
  if ((stop && *stop) || tr->done) return;
  residual_fp64_body(row_ptr, col_idx, values, b, x, N, tr->r, tr->partial);
}
// Sums the residual kernel's partials in index order and decides: converged
// when ||b - A x|| < tol ||b|| (or b = 0); the correction rounds then stop.
// In the final check (or once converged) the run's outcome is recorded:
// runs + 1, converged or (when the last PCG stopped on its breakdown test)
// breakdown, and the true residual as final_residual.
__device__ inline void tr_decide_body(TrState *tr, int nblocks, float tol, int final_check, nrf_pcg_result *res,
                                      const int32_t *stop) {
  if ((stop && *stop) || tr->done) return;
  double rr = 0.0, bb = 0.0;
  for (int k = 0; k < nblocks; ++k) {
    rr += tr->partial[2 * k];
    bb += tr->partial[2 * k + 1];
  }
  const double tl = (double)tol * sqrt(bb);
  const bool conv = sqrt(rr) < tl || bb == 0.0;
  tr->tol_abs = tl;
  if (conv || final_check) {
    tr->done = 1;
    res->final_residual = sqrt(rr);
    res->rhs_norm = sqrt(bb);
    res->runs += 1;
    if (conv) {
      res->converged += 1;
      res->flags |= 2;
    } else if (tr->last_flags & 1) {
      res->breakdowns += 1;
      res->flags |= 1;
    }
  }
}
__global__ void tr_decide_kernel(TrState *tr, int nblocks, float tol, int final_check, nrf_pcg_result *res,
                                 const int32_t *stop) {
  if (threadIdx.x == 0) tr_decide_body(tr, nblocks, tol, final_check, res, stop);
}
__global__ __launch_bounds__(256) void tr_axpy_kernel(float *__restrict__ x, int n, const TrState *tr,
                                                      const int32_t *stop) {
  if ((stop && *stop) || tr->done) return;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) x[i] += tr->d[i];
}

// One iterative-refinement step whose correction solve is the block-Jacobi
// inverse: x_out = x + M^-1 (b - A x), the residual (products and sums) and
// the correction in FP64.  One wave per node (4 nodes per workgroup): lane
// 6 s + r takes row r of the row's blocks s, s + kRefSlots, ..., the six row
// lanes sum the slots in slot order and form the correction by lane moves
// (the order of gn_refine_update_kernel, so both give the same bits).  x and
// x_out must not alias (a row reads its neighbours' x).  A raised GN stop
// flag makes it a no-op.
constexpr int kRefSlots = 10;
__device__ inline double refine_row_residual(const int32_t *__restrict__ row_ptr, const int32_t *__restrict__ col_idx,
                                             const float *__restrict__ values, const float *__restrict__ b,
                                             const float *__restrict__ x, int i, int lane) {
  const int slot = lane / 6, row = lane % 6;
  double acc = 0.0;
  if (lane < 6 * kRefSlots) {
    const int b1 = row_ptr[i + 1];
    for (int bk = row_ptr[i] + slot; bk < b1; bk += kRefSlots) {
      const float *A = values + (size_t)bk * 36 + row * 6;
      const float *xj = x + 6 * col_idx[bk];
      for (int c = 0; c < 6; ++c) acc += (double)A[c] * (double)xj[c];
    }
  }
  double sum = 0.0;
  for (int s = 0; s < kRefSlots; ++s) sum += __shfl(acc, 6 * s + row, 64);
  return (double)b[6 * i + row] - sum;
}
// x + M^-1 res for lanes 0-5 (row = lane), from the six row residuals
__device__ inline float refine_row_step(const float *__restrict__ Minv, const float *__restrict__ x, double res, int i,
                                        int lane) {
  double z = 0.0;
  for (int c = 0; c < 6; ++c) {
    const double rc = __shfl(res, c, 64);
    if (lane < 6) z += (double)Minv[(size_t)i * 36 + lane * 6 + c] * rc;
  }
  return lane < 6 ? (float)((double)x[6 * i + lane] + z) : 0.f;
}
__global__ __launch_bounds__(256) void refine_jacobi_kernel(const int32_t *__restrict__ row_ptr,
                                                           const int32_t *__restrict__ col_idx,
                                                           const float *__restrict__ values,
                                                           const float *__restrict__ b,
                                                           const float *__restrict__ Minv,
                                                           const float *__restrict__ x, int N,
                                                           const int32_t *__restrict__ stop,
                                                           float *__restrict__ x_out) {
  if (stop && *stop) return;
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= N) return;  // the whole wave
  const double res = refine_row_residual(row_ptr, col_idx, values, b, x, i, lane);
  const float xo = refine_row_step(Minv, x, res, i, lane);
  if (lane < 6) x_out[6 * i + lane] = xo;
}

// Sums the per-workgroup partials in index order: out = {|b - Ax|, |b|}.
__global__ void residual_norms_kernel(const double *__restrict__ partial, int nblocks, double *__restrict__ out) {
  if (threadIdx.x == 0) {
    double s0 = 0, s1 = 0;
    for (int k = 0; k < nblocks; ++k) {
      s0 += partial[2 * k];
      s1 += partial[2 * k + 1];
    }
    out[0] = sqrt(s0);
    out[1] = sqrt(s1);
  }
}
__global__ void axpy_kernel(float *__restrict__ x, const float *__restrict__ d, int n) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) x[q] += d[q];
}

// Opens a PCG launch: a new epoch for its tags, the abort flag cleared.
__device__ void pcg_open(PcgCtl *ctl) {
  ctl->epoch += 1;
  ctl->abort = 0;
  ctl->started = 0;
}
// Work on another stream that should not take the device before a PCG
// launch's workgroups are resident waits behind this one-wave kernel: it
// polls until the launch that carries `ticket` has all its workgroups
// started (the stream work before that launch has then ended), bounded:
// after about 20 ms it lets the stream go on regardless.
__global__ void pcg_started_wait_kernel(const PcgCtl *ctl, unsigned ticket) {
  if (threadIdx.x != 0) return;
  unsigned *c = const_cast<unsigned *>(&ctl->ticket);
  for (unsigned spins = 0; spins < (1u << 14); ++spins) {
    if (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == ticket) return;
    __builtin_amdgcn_s_sleep(8);
  }
}
__global__ void pcg_prep_kernel(PcgCtl *ctl) {
  if (threadIdx.x == 0) pcg_open(ctl);
}

__device__ inline void gn_begin_body(int32_t *stop, nrf_pcg_result *res, nrf_solver_totals *tot) {
  {
    // fold the previous solve into the running totals before clearing it
    tot->solves += 1;
    tot->gn_done += res->gn_done;
    tot->pcg_iterations += res->total_iterations;
    tot->rejected += (res->flags & 4) ? 1 : 0;
    tot->pcg_runs += res->runs;
    tot->pcg_converged += res->converged;
    tot->pcg_breakdowns += res->breakdowns;
    *stop = 0;
    res->runs = 0;
    res->converged = 0;
    res->breakdowns = 0;
    res->iterations = 0;
    res->total_iterations = 0;
    res->flags = 0;
    res->gn_done = 0;
    res->final_residual = 0;
    res->rhs_norm = 0;
  }
}
__global__ void gn_begin_kernel(int32_t *stop, nrf_pcg_result *res, nrf_solver_totals *tot) {
  if (threadIdx.x == 0) gn_begin_body(stop, res, tot);
}

// Step validation (include/fusion_app.h:409-491) by one workgroup: a
// rejected step raises the stop flag, which turns the update and every
// later kernel of the solve into no-ops.
struct GnArgs {
  const float *dx;
  DQ *T;
  DQ *centered;
  const nrf_node *nodes;
  int N;
  float scale, max_rot, max_trans, max_dx_mean;
  int validate;
  int32_t *stop;
  nrf_pcg_result *res;
};

__device__ void gn_validate_body(const GnArgs &a) {
  if (*a.stop) return;
  __shared__ double s_sum[4][1024 / 64];
  __shared__ float s_max[4][1024 / 64];
  __shared__ int s_fin[1024 / 64];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  {
    double rs = 0, ts = 0, ers = 0, ets = 0;
    float rm = 0, tm = 0, erm = 0, etm = 0;
    int fin = 1;
    for (int i = t; i < a.N; i += 1024) {
      const float *d = a.dx + i * 6;
      const float rn = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      const float tn = sqrtf(d[3] * d[3] + d[4] * d[4] + d[5] * d[5]);
      fin &= (isfinite(rn) && isfinite(tn)) ? 1 : 0;
      rs += rn; ts += tn;
      rm = fmaxf(rm, rn); tm = fmaxf(tm, tn);
      float er = a.scale * rn, et = a.scale * tn;
      if (a.max_rot > 0.0f) er = fminf(er, a.max_rot);
      if (a.max_trans > 0.0f) et = fminf(et, a.max_trans);
      ers += er; ets += et;
      erm = fmaxf(erm, er); etm = fmaxf(etm, et);
    }
    for (int off = 32; off > 0; off >>= 1) {
      rs += __shfl_xor(rs, off, 64); ts += __shfl_xor(ts, off, 64);
      ers += __shfl_xor(ers, off, 64); ets += __shfl_xor(ets, off, 64);
      rm = fmaxf(rm, __shfl_xor(rm, off, 64)); tm = fmaxf(tm, __shfl_xor(tm, off, 64));
      erm = fmaxf(erm, __shfl_xor(erm, off, 64)); etm = fmaxf(etm, __shfl_xor(etm, off, 64));
      fin &= __shfl_xor(fin, off, 64);
    }
    if (lane == 0) {
      s_sum[0][w] = rs; s_sum[1][w] = ts; s_sum[2][w] = ers; s_sum[3][w] = ets;
      s_max[0][w] = rm; s_max[1][w] = tm; s_max[2][w] = erm; s_max[3][w] = etm;
      s_fin[w] = fin;
    }
    __syncthreads();
    if (t == 0) {
      double S[4] = {0, 0, 0, 0};
      float M[4] = {0, 0, 0, 0};
      int F = 1;
      for (int k = 0; k < 1024 / 64; ++k) {
        for (int q = 0; q < 4; ++q) {
          S[q] += s_sum[q][k];
          M[q] = fmaxf(M[q], s_max[q][k]);
        }
        F &= s_fin[k];
      }
      const float n = (float)max(1, a.N);
      const float eff_rot_mean = (float)S[2] / n, eff_trans_mean = (float)S[3] / n;
      const float reject_rot = a.max_rot > 0.0f ? 5.0f * a.max_rot : INFINITY;
      const float reject_trans = a.max_trans > 0.0f ? 5.0f * a.max_trans : INFINITY;
      const bool bad_spike = (a.scale * M[0] > reject_rot) || (a.scale * M[1] > reject_trans);
      const bool mean_ok = (a.max_dx_mean <= 0.0f) ||
                           (eff_rot_mean <= a.max_dx_mean && eff_trans_mean <= a.max_dx_mean);
      const bool ok = F && isfinite(eff_rot_mean) && isfinite(eff_trans_mean) && isfinite(M[2]) &&
                      isfinite(M[3]) && (M[2] < reject_rot) && (M[3] < reject_trans) && mean_ok && !bad_spike;
      if (!ok) {
        *a.stop = 1;
        a.res->flags |= 4;
      }
    }
  }
}
__global__ __launch_bounds__(1024) void gn_validate_kernel(GnArgs a) { gn_validate_body(a); }

// Application of the validated step, one node per lane: per-node clamp and
// scale (upload_scaled_delta_x, include/fusion_app.h:1353-1375), then
// apply_twist_increment(..., 1.0) (src/warp_field.cu:1121-1141) and the
// centred-table refresh.
__device__ inline void gn_update_node(const GnArgs &a, int i, const float d[6]) {
  const DQ Tn = scaled_twist_update(d, a.scale, a.max_rot, a.max_trans, a.T[i]);
  a.T[i] = Tn;
  if (a.centered) a.centered[i] = dq_centered(Tn, f3(a.nodes[i].pos));
}
__device__ void gn_update_body(const GnArgs &a) {
  if (*a.stop) return;
  const int t = threadIdx.x;
  for (int i = blockIdx.x * 256 + t; i < a.N; i += gridDim.x * 256) gn_update_node(a, i, a.dx + i * 6);
  if (blockIdx.x == 0 && t == 0) a.res->gn_done += 1;
}
__global__ __launch_bounds__(256) void gn_update_kernel(GnArgs a) { gn_update_body(a); }

// The refinement step x = d + M^-1 (b - A d) (residual and correction in
// FP64, as refine_jacobi_kernel) fused with the update when no validation
// sits between them.  One wave per node (refine_row_residual: all of a
// row's blocks in one round for rows of up to 10 blocks, so the loads of a
// node are in flight together); lane 0 applies the refined step, which
// lanes 0-5 store in dx_out.
// d (the PCG's solution) is only read, so no lane reads a value another
// lane writes.
struct RefineArgs {
  const int32_t *row_ptr, *col_idx;
  const float *values, *b, *Minv, *d;
  float *dx_out;
};
__global__ __launch_bounds__(256) void gn_refine_update_kernel(GnArgs a, RefineArgs r) {
  if (*a.stop) return;
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= a.N) return;  // the whole wave
  // the loads that do not depend on the residual go out first, so that they
  // travel with the row's block loads: lane q < 36 holds M^-1 entry
  // (q / 6, q % 6), lanes 0-5 the PCG's solution, lane 0 the transform and
  // the node position of the update
  const float mv = lane < 36 ? r.Minv[(size_t)i * 36 + lane] : 0.f;
  const float dl = lane < 6 ? r.d[6 * i + lane] : 0.f;
  DQ T0{};
  float3 xp = f3(0.f, 0.f, 0.f);
  if (lane == 0) {
    T0 = a.T[i];
    xp = f3(a.nodes[i].pos);
  }
  const double res = refine_row_residual(r.row_ptr, r.col_idx, r.values, r.b, r.d, i, lane);
  // x + M^-1 res for lanes 0-5 (refine_row_step's products, in its order)
  double z = 0.0;
  for (int c = 0; c < 6; ++c) {
    const double rc = __shfl(res, c, 64);
    const double m = (double)__shfl(mv, (lane * 6 + c) & 63, 64);
    if (lane < 6) z += m * rc;
  }
  const float x = lane < 6 ? (float)((double)dl + z) : 0.f;
  if (lane < 6) r.dx_out[6 * i + lane] = x;
  float xs[6];
  for (int c = 0; c < 6; ++c) xs[c] = __shfl(x, c, 64);
  if (lane == 0) {
    const DQ Tn = scaled_twist_update(xs, a.scale, a.max_rot, a.max_trans, T0);
    a.T[i] = Tn;
    if (a.centered) a.centered[i] = dq_centered(Tn, xp);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) a.res->gn_done += 1;
}

// ------------------------------------------------------ plain form
// The reference's GPU structure of the solve (kernels/solver_kernels.cu:
// 543-741), kept as the baseline that the optimised forms are measured
// against (solver option plain_form): one kernel per operation, float
// atomics into the BSR blocks (one lane per correspondence, one per node for
// the ARAP edges j > i), and a host loop around the PCG that reads each dot
// product back to the host, as thrust::inner_product does.  The dot
// products are summed in FP64 (the configuration's FP32 step with FP64 PCG
// dots); everything else is FP32.  Not graph-capturable (host reads).
constexpr int kPlainThreads = 256;

// find_block (kernels/solver_kernels.cu:28-44): binary search of column j
// in the sorted row i.
__device__ inline int plain_find_block(const int32_t *row_ptr, const int32_t *col_idx, int i, int j) {
  int lo = row_ptr[i], hi = row_ptr[i + 1] - 1;
  while (lo <= hi) {
    const int mid = (lo + hi) >> 1;
    const int c = col_idx[mid];
    if (c == j) return mid;
    if (c < j) lo = mid + 1;
    else hi = mid - 1;
  }
  return -1;
}

__global__ __launch_bounds__(kPlainThreads) void plain_clear_kernel(float *__restrict__ values, int nv,
                                                                   float *__restrict__ b, int n6) {
  for (int k = blockIdx.x * kPlainThreads + threadIdx.x; k < nv; k += gridDim.x * kPlainThreads) values[k] = 0.f;
  for (int k = blockIdx.x * kPlainThreads + threadIdx.x; k < n6; k += gridDim.x * kPlainThreads) b[k] = 0.f;
}

// assemble_data_term_kernel (kernels/solver_kernels.cu:51-138): the source
// re-warped from the canonical point when one is given (the multi-GN entry).
__global__ __launch_bounds__(kPlainThreads) void plain_data_kernel(AsmArgs a, int n) {
  const int c = blockIdx.x * kPlainThreads + threadIdx.x;
  if (c >= n) return;
  const nrf_corr &cr = a.corrs[c];
  if (!cr.valid) return;
  const float3 dst = f3(cr.dst), nrm = f3(cr.normal);
  const int ids[4] = {cr.node_ids[0], cr.node_ids[1], cr.node_ids[2], cr.node_ids[3]};
  const float ws[4] = {cr.node_ws[0], cr.node_ws[1], cr.node_ws[2], cr.node_ws[3]};
  float3 src;
  if (a.canonical) {
    src = f3(a.canonical[3 * c], a.canonical[3 * c + 1], a.canonical[3 * c + 2]);
    DQ bl;
    if (dq_blend4(a.centered, ids, ws, bl)) src = dq_transform(bl, src);
  } else {
    src = f3(cr.src);
  }
  const float r = dot(nrm, sub(src, dst));
  const float fw = fmaxf(cr.weight, 0.0f) * huber_w(r, 0.2f);
  if (fw <= 1e-8f) return;
  float J[4][6];
  bool ok[4];
  for (int k = 0; k < 4; ++k) {
    const int id = ids[k];
    ok[k] = id >= 0 && id < a.N && ws[k] > 1e-8f;
    if (!ok[k]) continue;
    const float3 cp = cross(sub(src, f3(a.nodes[id].pos)), nrm);
    const float w = ws[k];
    J[k][0] = w * cp.x; J[k][1] = w * cp.y; J[k][2] = w * cp.z;
    J[k][3] = w * nrm.x; J[k][4] = w * nrm.y; J[k][5] = w * nrm.z;
    for (int q = 0; q < 6; ++q) atomicAdd(&a.b[id * 6 + q], -fw * J[k][q] * r);
  }
  for (int ki = 0; ki < 4; ++ki) {
    if (!ok[ki]) continue;
    for (int kj = 0; kj < 4; ++kj) {
      if (!ok[kj]) continue;
      const int blk = plain_find_block(a.row_ptr, a.col_idx, ids[ki], ids[kj]);
      if (blk < 0) continue;  // outside the pattern: dropped (Q2)
      float *B = a.values + (size_t)blk * 36;
      for (int p = 0; p < 6; ++p)
        for (int q = 0; q < 6; ++q) atomicAdd(&B[p * 6 + q], fw * J[ki][p] * J[kj][q]);
    }
  }
}

// assemble_smooth_term_kernel (kernels/solver_kernels.cu:202-352): one lane
// per node i, its edges to graph neighbours j > i.
__global__ __launch_bounds__(kPlainThreads) void plain_arap_kernel(AsmArgs a) {
  const int ni = blockIdx.x * kPlainThreads + threadIdx.x;
  if (ni >= a.N) return;
  const nrf_node &nd = a.nodes[ni];
  const int nn = max(0, min((int)nd.num_neighbors, (int)NRF_K_GRAPH));
  for (int k = 0; k < nn; ++k) {
    const int nj = nd.neighbors[k];
    if (nj < 0 || nj >= a.N || nj <= ni) continue;
    float3 r;
    float Ji[6][3], Jj[6][3];
    const float lam = arap_edge(a, ni, nj, fmaxf(nd.neighbor_w[k], 0.05f), r, Ji, Jj);
    if (lam == 0.f) continue;
    const int bi[4] = {plain_find_block(a.row_ptr, a.col_idx, ni, ni), plain_find_block(a.row_ptr, a.col_idx, nj, nj),
                       plain_find_block(a.row_ptr, a.col_idx, ni, nj), plain_find_block(a.row_ptr, a.col_idx, nj, ni)};
    for (int t = 0; t < 4; ++t) {
      if (bi[t] < 0) continue;
      const float(*L)[3] = (t == 0 || t == 2) ? Ji : Jj;
      const float(*R)[3] = (t == 0 || t == 3) ? Ji : Jj;
      const float sg = t < 2 ? lam : -lam;
      float *B = a.values + (size_t)bi[t] * 36;
      for (int p = 0; p < 6; ++p)
        for (int q = 0; q < 6; ++q) {
          float v = 0.f;
          for (int c = 0; c < 3; ++c) v += L[p][c] * R[q][c];
          atomicAdd(&B[p * 6 + q], sg * v);
        }
    }
    const float rv[3] = {r.x, r.y, r.z};
    for (int p = 0; p < 6; ++p) {
      float vi = 0.f, vj = 0.f;
      for (int c = 0; c < 3; ++c) {
        vi += Ji[p][c] * rv[c];
        vj += Jj[p][c] * rv[c];
      }
      atomicAdd(&a.b[ni * 6 + p], -lam * vi);
      atomicAdd(&a.b[nj * 6 + p], lam * vj);
    }
  }
}

// add_diagonal_damping_kernel (:167-194) and compute_preconditioner_kernel
// (:358-399, FP64 Cholesky with precond_fp64), one lane per node.
__global__ __launch_bounds__(kPlainThreads) void plain_damping_kernel(AsmArgs a) {
  const int i = blockIdx.x * kPlainThreads + threadIdx.x;
  if (i >= a.N) return;
  const int blk = plain_find_block(a.row_ptr, a.col_idx, i, i);
  if (blk < 0) return;
  float *D = a.values + (size_t)blk * 36;
  for (int k = 0; k < 6; ++k) {
    const float cur = D[k * 7];
    D[k * 7] = cur + a.lambda_damping * fmaxf(fabsf(cur), k < 3 ? 1e-3f : 1e-2f);
  }
}
__global__ __launch_bounds__(kPlainThreads) void plain_precond_kernel(AsmArgs a) {
  const int i = blockIdx.x * kPlainThreads + threadIdx.x;
  if (i >= a.N) return;
  const int blk = plain_find_block(a.row_ptr, a.col_idx, i, i);
  float D[36];
  for (int k = 0; k < 36; ++k) D[k] = blk >= 0 ? a.values[(size_t)blk * 36 + k] : 0.f;
  for (int col = 0; col < 6; ++col) {
    float outc[6];
    bool ok = true;
    if (a.precond_fp64) {
      double d[36], x[6];
      for (int k = 0; k < 36; ++k) d[k] = D[k];
      for (int k = 0; k < 6; ++k) d[k * 7] += (double)1e-4f;
      chol_inv6<double>(d, x, col, ok);
      for (int k = 0; k < 6; ++k) outc[k] = (float)x[k];
      if (!ok)
        for (int k = 0; k < 6; ++k) outc[k] = (k == col) ? (float)(1.0 / (d[k * 7] + (double)1e-6f)) : 0.f;
    } else {
      float d[36];
      for (int k = 0; k < 36; ++k) d[k] = D[k];
      for (int k = 0; k < 6; ++k) d[k * 7] += 1e-4f;
      chol_inv6<float>(d, outc, col, ok);
      if (!ok)
        for (int k = 0; k < 6; ++k) outc[k] = (k == col) ? 1.f / (d[k * 7] + 1e-6f) : 0.f;
    }
    for (int k = 0; k < 6; ++k) a.Minv[(size_t)i * 36 + k * 6 + col] = outc[k];
  }
}

// bsr_spmv_kernel (:406-432): one lane per scalar row.
__global__ __launch_bounds__(kPlainThreads) void plain_spmv_kernel(const int32_t *__restrict__ row_ptr,
                                                                  const int32_t *__restrict__ col_idx,
                                                                  const float *__restrict__ values,
                                                                  const float *__restrict__ x, float *__restrict__ y,
                                                                  int N) {
  const int g = blockIdx.x * kPlainThreads + threadIdx.x;
  if (g >= 6 * N) return;
  const int i = g / 6, row = g % 6;
  float sum = 0.f;
  for (int b = row_ptr[i]; b < row_ptr[i + 1]; ++b) {
    const float *B = values + (size_t)b * 36 + row * 6;
    const float *xj = x + 6 * col_idx[b];
    for (int c = 0; c < 6; ++c) sum += B[c] * xj[c];
  }
  y[g] = sum;
}
// apply_precond_kernel (:438-455): z_i = M_i^-1 r_i, one lane per node.
__global__ __launch_bounds__(kPlainThreads) void plain_precond_apply_kernel(const float *__restrict__ Minv,
                                                                           const float *__restrict__ r,
                                                                           float *__restrict__ z, int N) {
  const int i = blockIdx.x * kPlainThreads + threadIdx.x;
  if (i >= N) return;
  for (int p = 0; p < 6; ++p) {
    float v = 0.f;
    for (int q = 0; q < 6; ++q) v += Minv[(size_t)i * 36 + p * 6 + q] * r[6 * i + q];
    z[6 * i + p] = v;
  }
}
// y += alpha x (axpy_kernel) and y = x + alpha y (xpay_kernel).
__global__ __launch_bounds__(kPlainThreads) void plain_axpy_kernel(float alpha, const float *__restrict__ x,
                                                                  float *__restrict__ y, int n) {
  const int k = blockIdx.x * kPlainThreads + threadIdx.x;
  if (k < n) y[k] += alpha * x[k];
}
__global__ __launch_bounds__(kPlainThreads) void plain_xpay_kernel(float alpha, const float *__restrict__ x,
                                                                  float *__restrict__ y, int n) {
  const int k = blockIdx.x * kPlainThreads + threadIdx.x;
  if (k < n) y[k] = x[k] + alpha * y[k];
}
// FP64 dot products: per-block partial sums (tree in LDS), then one block
// adds them in a fixed order.
__global__ __launch_bounds__(kPlainThreads) void plain_dot_kernel(const float *__restrict__ x,
                                                                 const float *__restrict__ y, int n,
                                                                 double *__restrict__ partial) {
  __shared__ double sh[kPlainThreads];
  const int k = blockIdx.x * kPlainThreads + threadIdx.x;
  sh[threadIdx.x] = k < n ? (double)x[k] * (double)y[k] : 0.0;
  __syncthreads();
  for (int w = kPlainThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}
__global__ __launch_bounds__(kPlainThreads) void plain_sum_kernel(const double *__restrict__ partial, int nblk,
                                                                 double *__restrict__ out) {
  __shared__ double sh[kPlainThreads];
  double v = 0.0;
  for (int k = threadIdx.x; k < nblk; k += kPlainThreads) v += partial[k];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int w = kPlainThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}
// r = b - A x in FP64 (row sums of the FP32 blocks), stored in FP32, with
// per-block partial sums of r.r and b.b (partial[2 blk], partial[2 blk + 1]).
__global__ __launch_bounds__(kPlainThreads) void plain_residual_kernel(const int32_t *__restrict__ row_ptr,
                                                                      const int32_t *__restrict__ col_idx,
                                                                      const float *__restrict__ values,
                                                                      const float *__restrict__ b,
                                                                      const float *__restrict__ x, int N,
                                                                      float *__restrict__ r_out,
                                                                      double *__restrict__ partial) {
  __shared__ double s0[kPlainThreads], s1[kPlainThreads];
  const int g = blockIdx.x * kPlainThreads + threadIdx.x;
  double rr = 0.0, bb = 0.0;
  if (g < 6 * N) {
    const int i = g / 6, row = g % 6;
    double acc = b[g];
    for (int blk = row_ptr[i]; blk < row_ptr[i + 1]; ++blk) {
      const float *B = values + (size_t)blk * 36 + row * 6;
      const float *xj = x + 6 * col_idx[blk];
      for (int c = 0; c < 6; ++c) acc -= (double)B[c] * (double)xj[c];
    }
    r_out[g] = (float)acc;
    rr = acc * acc;
    bb = (double)b[g] * (double)b[g];
  }
  s0[threadIdx.x] = rr;
  s1[threadIdx.x] = bb;
  __syncthreads();
  for (int w = kPlainThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s0[threadIdx.x] += s0[threadIdx.x + w];
      s1[threadIdx.x] += s1[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = s0[0];
    partial[2 * blockIdx.x + 1] = s1[0];
  }
}
// The run's record (pcg_tally / tr_decide semantics) written by one lane.
__global__ void plain_tally_kernel(nrf_pcg_result *res, int iters, int flags, double rr, double bb) {
  if (threadIdx.x != 0) return;
  res->total_iterations += iters;
  res->iterations = iters;
  res->flags |= flags;
  res->final_residual = sqrt(rr);
  res->rhs_norm = sqrt(bb);
  res->converged += (flags & 2) ? 1 : 0;
  res->breakdowns += (flags & 1) ? 1 : 0;
  res->runs += 1;
}

// ------------------------------------------------------------- batches
// Several independent solves stepped by common launches (nrf_batch): every
// launch covers all instances of the batch.  An instance's device state is
// one InstDev record; per-frame inputs (BatchCorrs) travel as kernel
// arguments.  Elementwise work runs on 2-D grids (blockIdx.y = instance,
// blocks past the instance's size leave at once); per-row work (bucket
// sort, assembly) and the persistent PCG take row / workgroup tables that
// map a block to (instance, local index).
struct InstDev {
  AsmArgs as;
  PcgArgs pcg;
  GnArgs gn;
  int32_t *m_cnt, *m_off, *m_cur, *m_keys;
  nrf_solver_totals *tot;
};

__global__ __launch_bounds__(256) void mem_fill_batch_kernel(const InstDev *__restrict__ inst) {
  const InstDev &d = inst[blockIdx.y];
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < d.as.N) d.m_cnt[i] = 0;
}
__global__ __launch_bounds__(256) void mem_count_batch_kernel(const InstDev *__restrict__ inst, BatchCorrs fr) {
  const int k = blockIdx.y, n = fr.n[k];
  if ((int)blockIdx.x * 256 >= n) return;  // whole workgroup: every lane reaches the barriers
  mem_count_body(fr.corrs[k], n, inst[k].as.N, inst[k].m_cnt, blockIdx.x * 256 + threadIdx.x);
}
__global__ __launch_bounds__(1024) void member_scan_batch_kernel(const InstDev *__restrict__ inst) {
  const InstDev &d = inst[blockIdx.x];
  member_scan_body(d.m_cnt, d.m_off, d.m_cur, const_cast<int32_t *>(d.as.chunk_off), d.as.N);
}
__global__ __launch_bounds__(256) void mem_scatter_batch_kernel(const InstDev *__restrict__ inst, BatchCorrs fr) {
  const int k = blockIdx.y, n = fr.n[k];
  if ((int)blockIdx.x * 256 >= n) return;
  mem_scatter_body(fr.corrs[k], n, inst[k].as.N, inst[k].m_cur, inst[k].m_keys, blockIdx.x * 256 + threadIdx.x);
}
__global__ __launch_bounds__(kSortThreads) void mem_sort_batch_kernel(const InstDev *__restrict__ inst,
                                                                      const int32_t *__restrict__ row_inst,
                                                                      const int32_t *__restrict__ row_local) {
  const InstDev &d = inst[row_inst[blockIdx.x]];
  mem_sort_body(d.m_off, d.m_keys, const_cast<int32_t *>(d.as.m_keys), row_local[blockIdx.x], d.as.chunk_off,
                const_cast<int32_t *>(d.as.chunk_row));
}
__global__ void gn_begin_batch_kernel(const InstDev *__restrict__ inst, int count) {
  const int k = threadIdx.x;
  if (k < count) gn_begin_body(const_cast<int32_t *>(inst[k].as.stop), inst[k].pcg.res, inst[k].tot);
}
__global__ __launch_bounds__(256) void corr_prep_batch_kernel(const InstDev *__restrict__ inst, BatchCorrs fr) {
  const int k = blockIdx.y;
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= fr.n[k]) return;
  AsmArgs a = inst[k].as;
  a.corrs = fr.corrs[k];
  a.canonical = fr.canonical[k];
  corr_prep_one(a, c, const_cast<float *>(a.crec));
}
// Chunks of all instances: workgroup b is chunk slot b - cbase[k] of the
// instance k with cbase[k] <= b < cbase[k + 1].
template <int NS, int PREC>
__global__ __launch_bounds__(64) void assemble_chunks_batch_kernel(const InstDev *__restrict__ inst, BatchCorrs fr,
                                                                   int count) {
  extern __shared__ float lds[];
  __shared__ float s_rhs[6];
  const int b = blockIdx.x;
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (fr.cbase[mid] <= b) lo = mid; else hi = mid - 1;
  }
  AsmArgs a = inst[lo].as;
  a.corrs = fr.corrs[lo];
  a.canonical = fr.canonical[lo];
  assemble_chunk<NS, PREC>(a, a.N, b - fr.cbase[lo], lds, s_rhs);
}
template <int NS, int PREC>
__global__ __launch_bounds__(64) void assemble_tail_batch_kernel(const InstDev *__restrict__ inst, BatchCorrs fr,
                                                                 const int32_t *__restrict__ row_inst,
                                                                 const int32_t *__restrict__ row_local) {
  extern __shared__ float lds[];
  __shared__ float s_rhs[6];
  const int k = row_inst[blockIdx.x];
  AsmArgs a = inst[k].as;
  a.corrs = fr.corrs[k];
  a.canonical = fr.canonical[k];
  assemble_tail<NS, PREC>(a, row_local[blockIdx.x], lds, s_rhs);
}
__global__ void pcg_prep_batch_kernel(const InstDev *__restrict__ inst, int count) {
  const int k = threadIdx.x;
  if (k < count) {
    inst[k].pcg.ctl->epoch += 1;
    inst[k].pcg.ctl->abort = 0;
  }
}
// Several independent solves in one persistent launch: workgroup b works on
// instance wg_inst[b] as its workgroup wg_local[b].  The instances share no
// data; each synchronises its own workgroups through its own records, so
// only the workgroups of one instance need to be co-resident.
template <int SUB, int FORM>
__global__ __launch_bounds__(kPcgThreads) void pcg_pipe_batch_kernel(const InstDev *__restrict__ inst,
                                                                     const int32_t *__restrict__ wg_inst,
                                                                     const int32_t *__restrict__ wg_local,
                                                                     int round) {
  extern __shared__ float lds[];
  __shared__ double red[kRedSize];
  __shared__ int s_bad;
  __shared__ unsigned s_base;
  // the instance and workgroup indices are uniform: read them into scalar
  // registers so that the arguments (and every loop bound and branch on
  // them) are uniform to the compiler, as kernel arguments are
  const int k = __builtin_amdgcn_readfirstlane(wg_inst[blockIdx.x]);
  const int g = __builtin_amdgcn_readfirstlane(wg_local[blockIdx.x]);
  PcgArgs a = inst[k].pcg;
  if (round > 0 && a.tr) {  // a correction round of the true-residual rounds: A d = r
    a.round = round;
    a.b = a.tr->r;
    a.x = a.tr->d;
  }
  pcg_pipe_body<SUB, FORM>(a, g, lds, red, s_bad, s_base);
}
// True-residual rounds of all instances (TrState of each): residual,
// decision and correction add, one launch each over the batch.
__global__ __launch_bounds__(kResThreads) void tr_residual_batch_kernel(const InstDev *__restrict__ inst) {
  const InstDev &d = inst[blockIdx.y];
  if (!d.pcg.tr || (d.pcg.stop && *d.pcg.stop) || d.pcg.tr->done) return;
  if ((int)blockIdx.x * kResThreads >= 6 * d.as.N) return;  // whole workgroup
  residual_fp64_body(d.as.row_ptr, d.as.col_idx, d.as.values, d.as.b, d.pcg.x, d.as.N, d.pcg.tr->r,
                     d.pcg.tr->partial);
}
__global__ void tr_decide_batch_kernel(const InstDev *__restrict__ inst, int count, int final_check) {
  const int k = threadIdx.x;
  if (k >= count || !inst[k].pcg.tr) return;
  const InstDev &d = inst[k];
  tr_decide_body(d.pcg.tr, (6 * d.as.N + kResThreads - 1) / kResThreads, d.pcg.tol, final_check, d.pcg.res,
                 d.pcg.stop);
}
__global__ __launch_bounds__(256) void tr_axpy_batch_kernel(const InstDev *__restrict__ inst) {
  const InstDev &d = inst[blockIdx.y];
  if (!d.pcg.tr || (d.pcg.stop && *d.pcg.stop) || d.pcg.tr->done) return;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < 6 * d.as.N) d.pcg.x[i] += d.pcg.tr->d[i];
}

// The persistent PCG kernels by lanes per row (1, 2, 4) and form (0
// overlapped, 1 single reduction, 2 pipelined: pcg_pipelined - 1).
template <int FORM>
static const void *pcg_fn_of(int sub, bool batch) {
  if (batch)
    return sub == 1   ? reinterpret_cast<const void *>(pcg_pipe_batch_kernel<1, FORM>)
           : sub == 2 ? reinterpret_cast<const void *>(pcg_pipe_batch_kernel<2, FORM>)
                      : reinterpret_cast<const void *>(pcg_pipe_batch_kernel<4, FORM>);
  return sub == 1   ? reinterpret_cast<const void *>(pcg_pipe_kernel<1, FORM>)
         : sub == 2 ? reinterpret_cast<const void *>(pcg_pipe_kernel<2, FORM>)
                    : reinterpret_cast<const void *>(pcg_pipe_kernel<4, FORM>);
}
static const void *pcg_pipe_fn(int sub, int form) {
  return form == 2 ? pcg_fn_of<2>(sub, false) : form == 1 ? pcg_fn_of<1>(sub, false) : pcg_fn_of<0>(sub, false);
}
static const void *pcg_batch_fn(int sub, int form) {
  return form == 2 ? pcg_fn_of<2>(sub, true) : form == 1 ? pcg_fn_of<1>(sub, true) : pcg_fn_of<0>(sub, true);
}
__global__ __launch_bounds__(1024) void gn_validate_batch_kernel(const InstDev *__restrict__ inst) {
  const GnArgs &g = inst[blockIdx.x].gn;
  if (g.validate) gn_validate_body(g);
}
__global__ __launch_bounds__(256) void gn_update_batch_kernel(const InstDev *__restrict__ inst) {
  gn_update_body(inst[blockIdx.y].gn);
}

}  // namespace

// ------------------------------------------------------- host-side solver
struct Pattern {
  int N = -1, nb = 0, E = 0, Rmax = 0;
  std::vector<int32_t> row_ptr, col_idx, diag_rel, e_i, e_j, es_off, es_entry, es_offrel;
  std::vector<float> e_w;
};

// build_sparsity_pattern (kernels/solver_kernels.cu:498-541) plus the ARAP
// edge lists of each block row, in the order a sequential pass over nodes
// and their neighbour slots visits them (solver_kernels.cu:225-233).
static void build_edges(const nrf_node *h, int N, Pattern &P);
static void build_pattern(const nrf_node *h, int N, Pattern &P) {
  std::vector<std::vector<int32_t>> rc(N);
  for (int i = 0; i < N; ++i) rc[i].push_back(i);
  auto nnb = [](const nrf_node &n) { return std::max(0, std::min((int)n.num_neighbors, (int)NRF_K_GRAPH)); };
  for (int i = 0; i < N; ++i)
    for (int k = 0; k < nnb(h[i]); ++k) {
      int j = h[i].neighbors[k];
      if (j < 0 || j >= N) continue;
      rc[i].push_back(j);
      rc[j].push_back(i);
    }
  P.N = N;
  P.row_ptr.assign(N + 1, 0);
  P.col_idx.clear();
  P.diag_rel.assign(N, 0);
  P.Rmax = 0;
  for (int i = 0; i < N; ++i) {
    auto &c = rc[i];
    std::sort(c.begin(), c.end());
    c.erase(std::unique(c.begin(), c.end()), c.end());
    P.row_ptr[i] = (int32_t)P.col_idx.size();
    for (size_t q = 0; q < c.size(); ++q) {
      if (c[q] == i) P.diag_rel[i] = (int32_t)q;
      P.col_idx.push_back(c[q]);
    }
    P.Rmax = std::max(P.Rmax, (int)c.size());
  }
  P.row_ptr[N] = (int32_t)P.col_idx.size();
  P.nb = (int)P.col_idx.size();
  build_edges(h, N, P);
}

// The ARAP edge-sides of the nodes' current edges against the pattern P
// (rows' column lists unchanged): an edge's block offset is -1 when the
// pattern lacks it.
static void build_edges(const nrf_node *h, int N, Pattern &P) {
  auto nnb = [](const nrf_node &n) { return std::max(0, std::min((int)n.num_neighbors, (int)NRF_K_GRAPH)); };
  auto rel = [&](int row, int col) {
    auto b = P.col_idx.begin() + P.row_ptr[row], e = P.col_idx.begin() + P.row_ptr[row + 1];
    auto it = std::lower_bound(b, e, col);
    return (int32_t)(it != e && *it == col ? it - b : -1);
  };
  P.e_i.clear(); P.e_j.clear(); P.e_w.clear();
  std::vector<std::vector<std::pair<int32_t, int32_t>>> es(N);
  for (int i = 0; i < N; ++i)
    for (int k = 0; k < nnb(h[i]); ++k) {
      int j = h[i].neighbors[k];
      if (j < 0 || j >= N || j <= i) continue;
      int e = (int)P.e_i.size();
      P.e_i.push_back(i);
      P.e_j.push_back(j);
      P.e_w.push_back(h[i].neighbor_w[k]);
      es[i].push_back({e * 2 + 0, rel(i, j)});
      es[j].push_back({e * 2 + 1, rel(j, i)});
    }
  P.E = (int)P.e_i.size();
  P.es_off.assign(N + 1, 0);
  P.es_entry.clear();
  P.es_offrel.clear();
  for (int i = 0; i < N; ++i) {
    P.es_off[i] = (int32_t)P.es_entry.size();
    for (auto &pr : es[i]) {
      P.es_entry.push_back(pr.first);
      P.es_offrel.push_back(pr.second);
    }
  }
  P.es_off[N] = (int32_t)P.es_entry.size();
}

// Counts reallocations of solver buffers: a captured step that holds the old
// addresses must be captured again (solver_state_generation).
static std::atomic<unsigned> g_realloc{0};

template <typename T>
static int dev_alloc(T *&p, size_t n, size_t &cap) {
  if (n <= cap && p) return NRF_OK;
  g_realloc.fetch_add(1);
  if (p) hipFree(p);
  p = nullptr;
  cap = 0;
  size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
  NRF_HIP_CHECK(hipMalloc(reinterpret_cast<void **>(&p), bytes));
  cap = n;
  return NRF_OK;
}

template <typename T>
static int upload(T *&p, size_t &cap, const std::vector<T> &v) {
  int st = dev_alloc(p, v.size(), cap);
  if (st) return st;
  if (!v.empty()) NRF_HIP_CHECK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return NRF_OK;
}

}  // namespace nrf

using namespace nrf;

struct nrf_solver {
  nrf_solver_params params;
  int max_nodes = 0;
  Pattern P;
  // pattern buffers
  int32_t *row_ptr = nullptr, *col_idx = nullptr, *diag_rel = nullptr, *e_i = nullptr, *e_j = nullptr,
          *es_off = nullptr, *es_entry = nullptr, *es_offrel = nullptr, *rows = nullptr;
  float *e_w = nullptr;
  size_t c_row_ptr = 0, c_col_idx = 0, c_diag_rel = 0, c_e_i = 0, c_e_j = 0, c_es_off = 0, c_es_entry = 0,
         c_es_offrel = 0, c_rows = 0, c_e_w = 0;
  // system
  float *values = nullptr, *b = nullptr, *Minv = nullptr;
  size_t c_values = 0, c_b = 0, c_Minv = 0;
  // pcg: published z / p vectors, per-workgroup sync records, halo lists
  ll_word *zpub = nullptr, *P0 = nullptr, *P1 = nullptr;
  size_t c_zpub = 0, c_P0 = 0, c_P1 = 0;
  SyncRec *recs = nullptr;
  size_t c_recs = 0;
  PcgCtl *ctl = nullptr;
  size_t c_ctl = 0;
  int32_t *lcol = nullptr, *halo_off = nullptr, *halo = nullptr;
  size_t c_lcol = 0, c_halo_off = 0, c_halo = 0;
  int G = 1, lds_a = 0;
  size_t pcg_lds = 0, lds_attr = 65536;
  // memberships
  int32_t *m_cnt = nullptr, *m_off = nullptr, *m_cur = nullptr, *m_keys = nullptr, *m_keys2 = nullptr;
  size_t c_m_cnt = 0, c_m_off = 0, c_m_cur = 0, c_m_keys = 0, c_m_keys2 = 0;
  bool m_cnt_zero = false;  // m_cnt holds zeros for its first m_cnt_n entries (cleared by the last scan)
  int m_cnt_n = 0;
  // GN state
  int32_t *stop = nullptr;
  size_t c_stop = 0;
  float *crec = nullptr;  // per-corr records of the MFMA assembly
  float *arec = nullptr;  // per-edge-side ARAP records of the MFMA assembly
  size_t c_arec = 0;
  int32_t *chunk_off = nullptr, *chunk_row = nullptr;  // membership chunks of the MFMA assembly
  float *part = nullptr;                                // partial sums of split rows
  size_t c_chunk_off = 0, c_chunk_row = 0, c_part = 0;
  size_t c_crec = 0;
  float *rres = nullptr, *dref = nullptr;  // refinement: residual and correction
  size_t c_rres = 0, c_dref = 0;
  double *rpart = nullptr, *rnorm = nullptr;  // true residual partials and norms
  size_t c_rpart = 0, c_rnorm = 0;
  TrState *tr = nullptr;  // true-residual rounds (over rres, dref, rpart)
  size_t c_tr = 0;
  // plain form: PCG vectors, FP64 dot partials, the host-read result
  float *pl_r = nullptr, *pl_z = nullptr, *pl_p = nullptr, *pl_Ap = nullptr;
  double *pl_part = nullptr, *pl_out = nullptr, *pl_host = nullptr;
  size_t c_pl_r = 0, c_pl_z = 0, c_pl_p = 0, c_pl_Ap = 0, c_pl_part = 0, c_pl_out = 0;
  nrf_pcg_result *d_res = nullptr;
  size_t c_res = 0;
  nrf_solver_totals *d_tot = nullptr;
  size_t c_tot = 0;
  int cus = 0, max_lds = 0;
  int forced_G = 0;
  int pcg_sub = 4;        // lanes per scalar row of the PCG (1, 2, 4)
  int pcg_wg_nodes = 24;  // target nodes per PCG workgroup (measured best on the full-scale body: 20-24)
  int asm_mfma = 1;  // data term on the matrix cores (rows of up to 21 blocks)
  unsigned gen = 0;  // advanced by every partition rebuild
  // optional profiling events (see solver_set_event_sink)
  hipEvent_t *sink = nullptr;
  int sink_n = 0, sink_i = 0, sink_skip = 0;
  hipEvent_t pcg_wait = nullptr;  // see solver_set_pcg_wait
  solver_hook_fn hook = nullptr;   // see solver_set_pcg_hook
  void *hook_ctx = nullptr;
  hipEvent_t hook_ev = nullptr;
  unsigned ticket = 0;  // carried by the next PCG launch (solver_set_pcg_ticket)
  nrf_dualquat *snap_T = nullptr, *snap_C = nullptr;  // see solver_set_snapshot
  uint8_t *vmask = nullptr;  // valid slots of the last membership build (AsmArgs::vmask)
  size_t c_vmask = 0;
};

namespace nrf {

void solver_set_pcg_wait(nrf_solver *s, hipEvent_t ev) { s->pcg_wait = ev; }
int solver_pcg_start_wait(nrf_solver *s, unsigned ticket, hipStream_t st) {
  pcg_started_wait_kernel<<<1, 64, 0, st>>>(s->ctl, ticket);
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}
void solver_set_pcg_ticket(nrf_solver *s, unsigned ticket) { s->ticket = ticket; }
void solver_set_pcg_hook(nrf_solver *s, solver_hook_fn fn, void *ctx, hipEvent_t ev) {
  s->hook = fn;
  s->hook_ctx = ctx;
  s->hook_ev = ev;
}
void solver_set_snapshot(nrf_solver *s, nrf_dualquat *T, nrf_dualquat *centered) {
  s->snap_T = T;
  s->snap_C = centered;
}

void solver_set_event_sink(nrf_solver *s, hipEvent_t *events, int count, int skip) {
  s->sink = events;
  s->sink_n = events ? count : 0;
  s->sink_i = 0;
  s->sink_skip = skip;
}

// Sink points of one solve: after the memberships, then per GN iteration
// after the assembly, before the PCG, after the PCG and after the update.
// The first `sink_skip` points record nothing.
static void sink_record(nrf_solver *s, hipStream_t st) {
  if (!s->sink) return;
  const int j = s->sink_i++ - s->sink_skip;
  if (j >= 0 && j < s->sink_n) (void)hipEventRecord(s->sink[j], st);
}

static int solver_build_partition(nrf_solver *s);

static int solver_upload_edges(nrf_solver *s) {
  const Pattern &P = s->P;
  int st;
  if ((st = upload(s->e_i, s->c_e_i, P.e_i))) return st;
  if ((st = upload(s->e_j, s->c_e_j, P.e_j))) return st;
  if ((st = upload(s->e_w, s->c_e_w, P.e_w))) return st;
  if ((st = upload(s->es_off, s->c_es_off, P.es_off))) return st;
  if ((st = upload(s->es_entry, s->c_es_entry, P.es_entry))) return st;
  if ((st = upload(s->es_offrel, s->c_es_offrel, P.es_offrel))) return st;
  if (!P.es_entry.empty() &&
      (st = dev_alloc(s->arec, P.es_entry.size() * (size_t)kArapRec, s->c_arec)))
    return st;
  ++s->gen;
  return NRF_OK;
}

static int solver_set_pattern_host(nrf_solver *s, const nrf_node *h, int N) {
  build_pattern(h, N, s->P);
  const Pattern &P = s->P;
  int st;
  if ((st = upload(s->row_ptr, s->c_row_ptr, P.row_ptr))) return st;
  if ((st = upload(s->col_idx, s->c_col_idx, P.col_idx))) return st;
  if ((st = upload(s->diag_rel, s->c_diag_rel, P.diag_rel))) return st;
  if ((st = upload(s->e_i, s->c_e_i, P.e_i))) return st;
  if ((st = upload(s->e_j, s->c_e_j, P.e_j))) return st;
  if ((st = upload(s->e_w, s->c_e_w, P.e_w))) return st;
  if ((st = upload(s->es_off, s->c_es_off, P.es_off))) return st;
  if ((st = upload(s->es_entry, s->c_es_entry, P.es_entry))) return st;
  if ((st = upload(s->es_offrel, s->c_es_offrel, P.es_offrel))) return st;
  if (!P.es_entry.empty() &&
      (st = dev_alloc(s->arec, P.es_entry.size() * (size_t)kArapRec, s->c_arec)))
    return st;
  if ((st = dev_alloc(s->values, (size_t)std::max(P.nb, 1) * 36, s->c_values))) return st;
  return solver_build_partition(s);
}

// PCG partition: contiguous node ranges of at most 32, 80 or 170 nodes for
// pcg_sub = 4, 2, 1 lanes per scalar row (whole nodes per wave for 4 and 2),
// each with its halo (the distinct columns of
// its rows) and the blocks' columns renumbered into that halo.  The target
// size (pcg_wg_nodes) trades workgroups for rows per workgroup.
static int solver_build_partition(nrf_solver *s) {
  ++s->gen;
  const Pattern &P = s->P;
  const int N = P.N;
  int st;
  // 2 and 4 lanes per row: whole nodes per wave (64 / (6 pcg_sub)) times the
  // waves; 1 lane per row: rows packed over all lanes
  const int max_nodes = s->pcg_sub == 1 ? kPcgThreads / 6 : kPcgWaves * (64 / (6 * s->pcg_sub));
  const int target = std::max(1, std::min(max_nodes, s->pcg_wg_nodes));
  int G = s->forced_G;
  const int min_G = std::max(1, (N + max_nodes - 1) / max_nodes);
  if (G < min_G) G = std::max(min_G, std::min(s->cus, (N + target - 1) / target));
  if (G > kPcgMaxWg) return NRF_ERR_TOO_LARGE;  // one record per workgroup
  std::vector<int32_t> rows(G + 1);
  for (int g = 0; g <= G; ++g) rows[g] = (int32_t)((long)N * g / G);
  std::vector<int32_t> lcol(P.nb), hoff(G + 1, 0), halo;
  size_t need_max = 0;
  for (int g = 0; g < G; ++g) {
    const int b0 = P.row_ptr[rows[g]], b1 = P.row_ptr[rows[g + 1]];
    std::vector<int32_t> h(P.col_idx.begin() + b0, P.col_idx.begin() + b1);
    std::sort(h.begin(), h.end());
    h.erase(std::unique(h.begin(), h.end()), h.end());
    for (int q = b0; q < b1; ++q)
      lcol[q] = (int32_t)(std::lower_bound(h.begin(), h.end(), P.col_idx[q]) - h.begin());
    hoff[g] = (int32_t)halo.size();
    halo.insert(halo.end(), h.begin(), h.end());
    const size_t nblk = (size_t)(b1 - b0);
    need_max = std::max(need_max, nblk * (36 * 4 + 4) + (h.size() + 1) * 24 + (size_t)kPcgThreads * 8);
  }
  hoff[G] = (int32_t)halo.size();
  size_t halo_max = 0;
  for (int g = 0; g < G; ++g) halo_max = std::max(halo_max, (size_t)(hoff[g + 1] - hoff[g]));
  s->G = G;
  const size_t budget = (size_t)std::max(0, s->max_lds - 1024);
  s->lds_a = need_max <= budget ? 1 : 0;
  s->pcg_lds = align_up(s->lds_a ? need_max : (halo_max + 1) * 24 + (size_t)kPcgThreads * 8, 16);
  if ((st = upload(s->rows, s->c_rows, rows))) return st;
  if ((st = upload(s->lcol, s->c_lcol, lcol))) return st;
  if ((st = upload(s->halo_off, s->c_halo_off, hoff))) return st;
  if ((st = upload(s->halo, s->c_halo, halo))) return st;
  if (s->pcg_lds > 65536 && s->pcg_lds > s->lds_attr) {
    bool ok = hipFuncSetAttribute(reinterpret_cast<const void *>(pcg_kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)s->pcg_lds) == hipSuccess;
    for (int form = 0; form < 3; ++form)
      for (int sub : {1, 2, 4})
        for (const void *fn : {pcg_pipe_fn(sub, form), pcg_batch_fn(sub, form)})
          ok = ok && hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)s->pcg_lds) == hipSuccess;
    if (ok) {
      s->lds_attr = s->pcg_lds;
    } else {
      (void)hipGetLastError();
      s->lds_a = 0;  // keep the matrix in global memory
      s->pcg_lds = align_up((halo_max + 1) * 24 + (size_t)kPcgThreads * 8, 16);
    }
  }
  return NRF_OK;
}

// begin: the scan kernel also opens the solve (gn_begin_kernel's work).
static int solver_build_memberships(nrf_solver *s, const nrf_corr *corrs, int n, int N, hipStream_t st,
                                    bool begin = false, SnapArgs sn = SnapArgs()) {
  int rc;
  const size_t cap0 = s->c_m_cnt;
  if ((rc = dev_alloc(s->m_cnt, (size_t)N, s->c_m_cnt))) return rc;
  if (s->c_m_cnt != cap0) s->m_cnt_zero = false;  // reallocated: not cleared yet
  if ((rc = dev_alloc(s->m_off, (size_t)N + 1, s->c_m_off))) return rc;
  if ((rc = dev_alloc(s->m_cur, (size_t)N, s->c_m_cur))) return rc;
  if ((rc = dev_alloc(s->m_keys, (size_t)n * 4, s->c_m_keys))) return rc;
  if ((rc = dev_alloc(s->m_keys2, (size_t)n * 4, s->c_m_keys2))) return rc;
  if ((rc = dev_alloc(s->chunk_off, (size_t)N + 1, s->c_chunk_off))) return rc;
  if ((rc = dev_alloc(s->chunk_row, asm_max_chunks(N, n), s->c_chunk_row))) return rc;
  // the counts are zero from the previous frame's scan, which clears them
  // behind its reads (every count of the buffer is read and cleared, since
  // the scan covers [0, N) and N never exceeds the allocation)
  if (!s->m_cnt_zero || N > s->m_cnt_n) {
    fill_i32<<<((int)s->c_m_cnt + 255) / 256, 256, 0, st>>>(s->m_cnt, (int)s->c_m_cnt, 0);
    NRF_LAUNCH_CHECK();
  }
  s->m_cnt_zero = true;
  s->m_cnt_n = N;
  if ((rc = dev_alloc(s->vmask, (size_t)std::max(n, 1), s->c_vmask))) return rc;
  if (n > 0 || sn.dst_T) {
    mem_count_kernel<<<(std::max(n, sn.dst_T ? sn.N : 0) + 255) / 256, 256, 0, st>>>(corrs, n, N, s->m_cnt, sn,
                                                                                    s->vmask);
    NRF_LAUNCH_CHECK();
  }
  // offsets, cursors and the MFMA assembly's chunk offsets in one pass
  member_scan_kernel<<<1, 1024, 0, st>>>(s->m_cnt, s->m_off, s->m_cur, s->chunk_off, N, begin ? s->stop : nullptr,
                                         s->d_res, s->d_tot);
  NRF_LAUNCH_CHECK();
  if (n > 0) {
    mem_scatter_kernel<<<(n + 255) / 256, 256, 0, st>>>(corrs, n, N, s->m_cur, s->m_keys, s->vmask);
    NRF_LAUNCH_CHECK();
  }
  // each row's bucket ordered by key; the row's chunk entries filled alongside
  mem_sort_kernel<<<N, kSortThreads, 0, st>>>(s->m_off, s->m_keys, s->m_keys2, s->chunk_off, s->chunk_row);
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}

static AsmArgs solver_asm_args(const nrf_solver *s, const nrf_corr *corrs, const float *canonical,
                               const nrf_node *nodes, const DQ *T, const DQ *centered, int N) {
  AsmArgs a;
  a.corrs = corrs;
  a.canonical = canonical;
  a.nodes = nodes;
  a.T = T;
  a.centered = centered;
  a.N = N;
  a.row_ptr = s->row_ptr;
  a.col_idx = s->col_idx;
  a.diag_rel = s->diag_rel;
  a.m_off = s->m_off;
  a.m_keys = s->m_keys2;
  a.es_off = s->es_off;
  a.es_entry = s->es_entry;
  a.es_offrel = s->es_offrel;
  a.e_i = s->e_i;
  a.e_j = s->e_j;
  a.e_w = s->e_w;
  a.lambda_smooth = s->params.lambda_smooth;
  a.arap_sign = s->params.arap_consistent_sign ? -1.0f : 1.0f;
  a.psd_pairs = s->params.psd_dropped_pairs;
  a.lambda_damping = s->params.lambda_damping;
  a.precond_fp64 = s->params.precond_fp64;
  a.Rmax = s->P.Rmax;
  a.values = s->values;
  a.b = s->b;
  a.Minv = s->Minv;
  a.stop = s->stop;
  a.crec = nullptr;
  a.arec = nullptr;
  a.n_es = 0;
  a.pcg_open = nullptr;
  a.chunk_off = s->chunk_off;
  a.chunk_row = s->chunk_row;
  a.part = s->part;
  return a;
}

// LDS bytes of one assembly workgroup and the MFMA column strips of a row
// of Rmax blocks (block columns + the RHS column, 16 per strip).
static int asm_strips(int Rmax) { return (6 * Rmax + 1 + 15) / 16; }
// accumulators and columns of a row, then the staging area: the scalar
// form's [64][kStage] (also the row tail's scratch) or the MFMA form's dense
// operands, [kAsmBatch][8] + [16 slabs][kAsmBcStride]
static size_t asm_lds_bytes(int Rmax) {
  const size_t stage = std::max((size_t)64 * kStage, (size_t)kAsmBatch * 8 + (size_t)16 * asm_strips(Rmax) * kAsmBcStride);
  return (size_t)Rmax * 36 * 4 + (size_t)Rmax * 4 + stage * 4;
}

// open_pcg: the assembly's last launch also opens the PCG launch that follows
// it (solver_pcg with `opened`).
static int solver_assemble(nrf_solver *s, const nrf_corr *corrs, int n, const float *canonical,
                           const nrf_node *nodes, const DQ *T, const DQ *centered, int N, hipStream_t st,
                           bool open_pcg = false) {
  AsmArgs a = solver_asm_args(s, corrs, canonical, nodes, T, centered, N);
  a.pcg_open = open_pcg ? s->ctl : nullptr;
  const size_t lds = asm_lds_bytes(s->P.Rmax);
  if (lds > 64 * 1024) return NRF_ERR_TOO_LARGE;
  const int ns = asm_strips(s->P.Rmax);
  const bool mfma = s->asm_mfma && ns <= 8;
  if (mfma && n > 0) {
    int rc;
    if ((rc = dev_alloc(s->crec, (size_t)n * kRec, s->c_crec))) return rc;
    a.crec = s->crec;
    const int n_es = s->P.es_off.empty() ? 0 : s->P.es_off.back();
    if (n_es > 0 && s->arec) {  // the row tails read the edge-side records
      a.arec = s->arec;
      a.n_es = n_es;
    }
    a.vmask = s->vmask;  // made by this solve's membership build
    corr_prep_kernel<<<(n + a.n_es + 255) / 256, 256, 0, st>>>(a, n, s->crec);
    NRF_LAUNCH_CHECK();
  }
  const int prec = s->params.jtj_precision;
  if (!mfma && prec != 0) return NRF_ERR_INVALID_ARGUMENT;  // reduced precision exists in the MFMA form only
  const int nch = (int)asm_max_chunks(N, n);  // chunk slots: an upper bound, the others leave
  if (mfma) {
    int rc;
    if ((rc = dev_alloc(s->part, (size_t)nch * part_floats(8), s->c_part))) return rc;
    a.part = s->part;
  }
#define NRF_ASM_CASE(NSV)                                                                       \
  case NSV:                                                                                     \
    if (prec == 1) {                                                                            \
      assemble_rows_mfma_kernel<NSV, 1><<<nch, 64, lds, st>>>(a);                               \
      assemble_rows_tail_kernel<NSV, 1><<<N, 64, lds, st>>>(a);                                 \
    } else if (prec == 2) {                                                                     \
      assemble_rows_mfma_kernel<NSV, 2><<<nch, 64, lds, st>>>(a);                               \
      assemble_rows_tail_kernel<NSV, 2><<<N, 64, lds, st>>>(a);                                 \
    } else if (prec == 3) {                                                                     \
      assemble_rows_mfma_kernel<NSV, 3><<<nch, 64, lds, st>>>(a);                               \
      assemble_rows_tail_kernel<NSV, 3><<<N, 64, lds, st>>>(a);                                 \
    } else if (prec == 4) {                                                                     \
      assemble_rows_mfma_kernel<NSV, 4><<<nch, 64, lds, st>>>(a);                               \
      assemble_rows_tail_kernel<NSV, 4><<<N, 64, lds, st>>>(a);                                 \
    } else {                                                                                    \
      assemble_rows_mfma_kernel<NSV, 0><<<nch, 64, lds, st>>>(a);                               \
      assemble_rows_tail_kernel<NSV, 0><<<N, 64, lds, st>>>(a);                                 \
    }                                                                                           \
    break;
  switch (mfma ? ns : 0) {
    NRF_ASM_CASE(1)
    NRF_ASM_CASE(2)
    NRF_ASM_CASE(3)
    NRF_ASM_CASE(4)
    NRF_ASM_CASE(5)
    NRF_ASM_CASE(6)
    NRF_ASM_CASE(7)
    NRF_ASM_CASE(8)
    default: assemble_rows_kernel<<<N, 64, lds, st>>>(a); break;
  }
#undef NRF_ASM_CASE
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}

static int pcg_launch(nrf_solver *s, const float *rhs, float *x, int N, int maxit, bool record, hipStream_t st,
                      bool opened = false, int round = 0, bool refine = false);
static int solver_tr_rounds(nrf_solver *s, float *x, int N, hipStream_t st);

// opened: the launch was opened by the assembly (solver_assemble's open_pcg)
static int solver_pcg(nrf_solver *s, float *x, int N, hipStream_t st, bool opened = false) {
  int rc;
  if (s->params.refine_jacobi) {
    // the PCG solves into the scratch vector; one kernel then writes
    // x = d + M^-1 (b - A d)
    if ((rc = dev_alloc(s->dref, (size_t)6 * N, s->c_dref))) return rc;
    if ((rc = pcg_launch(s, s->b, s->dref, N, s->params.pcg_iterations, true, st, opened))) return rc;
    refine_jacobi_kernel<<<(N + 3) / 4, 256, 0, st>>>(s->row_ptr, s->col_idx, s->values, s->b,
                                                                           s->Minv, s->dref, N, s->stop, x);
    NRF_LAUNCH_CHECK();
    return NRF_OK;
  }
  rc = pcg_launch(s, s->b, x, N, s->params.pcg_iterations, true, st, opened);
  if (rc) return rc;
  if (s->params.true_residual_rounds > 0) return solver_tr_rounds(s, x, N, st);
  if (s->params.refine_iterations <= 0) return rc;
  // one iterative-refinement step: r = b - A x accumulated in FP64, A d = r
  // by refine_iterations PCG iterations, x += d
  const int n6 = 6 * N, nblk = (n6 + kResThreads - 1) / kResThreads;
  if ((rc = dev_alloc(s->rres, (size_t)n6, s->c_rres))) return rc;
  if ((rc = dev_alloc(s->dref, (size_t)n6, s->c_dref))) return rc;
  residual_fp64_kernel<<<nblk, kResThreads, 0, st>>>(s->row_ptr, s->col_idx, s->values, s->b, x, N, s->rres,
                                                     nullptr);
  NRF_LAUNCH_CHECK();
  if ((rc = pcg_launch(s, s->rres, s->dref, N, s->params.refine_iterations, false, st))) return rc;
  axpy_kernel<<<nblk, kResThreads, 0, st>>>(x, s->dref, n6);
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}

// `record`: the profiling sink's event that opens the PCG group (one per solve).
static PcgArgs solver_pcg_args(const nrf_solver *s, const float *rhs, float *x, int N, int maxit) {
  PcgArgs a;
  a.row_ptr = s->row_ptr;
  a.lcol = s->lcol;
  a.values = s->values;
  a.b = rhs;
  a.Minv = s->Minv;
  a.x = x;
  a.zpub = s->zpub;
  a.ppub0 = s->P0;
  a.ppub1 = s->P1;
  a.recs = s->recs;
  a.ctl = s->ctl;
  a.rows = s->rows;
  a.halo_off = s->halo_off;
  a.halo = s->halo;
  a.N = N;
  a.maxit = maxit;
  a.G = s->G;
  a.lds_a = s->lds_a;
  a.tol = s->params.pcg_tolerance;
  a.rel = s->params.relative_tolerance;
  a.res = s->d_res;
  a.stop = s->stop;
  a.tr = s->params.true_residual_rounds > 0 ? s->tr : nullptr;
  a.refine = 0;
  a.round = 0;
  return a;
}

// The persistent PCG takes an ordinary launch, eager or captured into a
// graph: its grid is sized to be co-resident on the device (at most the CU
// count, or the occupancy bound for batches), and every wait in it is bounded.
static int pcg_launch(nrf_solver *s, const float *rhs, float *x, int N, int maxit, bool record, hipStream_t st,
                      bool opened, int round, bool refine) {
  if (!opened) {
    pcg_prep_kernel<<<1, 64, 0, st>>>(s->ctl);
    NRF_LAUNCH_CHECK();
  }
  if (record) sink_record(s, st);
  PcgArgs a = solver_pcg_args(s, rhs, x, N, maxit);
  a.round = round;
  a.refine = refine ? 1 : 0;
  a.ticket = s->ticket;
  s->ticket = 0;
  void *args[] = {&a};
  const bool pipe = s->params.pcg_pipelined != 0;
  if (!pipe && s->pcg_sub != kPcgSub) return NRF_ERR_INVALID_ARGUMENT;  // textbook form: 4 lanes per row only
  const void *fn = pipe ? pcg_pipe_fn(s->pcg_sub, s->params.pcg_pipelined - 1)
                        : reinterpret_cast<const void *>(pcg_kernel);
  NRF_HIP_CHECK(hipLaunchKernel(fn, dim3(s->G), dim3(kPcgThreads), args, (unsigned)s->pcg_lds, st));
  return NRF_OK;
}

// True-residual rounds after a PCG run that left its solution in x: up to
// true_residual_rounds times, r = b - A x in FP64 and, while ||r|| >=
// tol ||b||, a correction PCG (A d = r to the absolute tolerance tol ||b||,
// the same iteration budget) and x += d; then the final check records the
// run's outcome from the true residual.  Every kernel after the decision
// that the run converged is a no-op (device flag), so the launch sequence
// is fixed (graph-safe).
static int solver_tr_rounds(nrf_solver *s, float *x, int N, hipStream_t st) {
  const int n6 = 6 * N, nblk = (n6 + kResThreads - 1) / kResThreads;
  const int R = s->params.true_residual_rounds;
  for (int k = 1; k <= R + 1; ++k) {
    tr_residual_kernel<<<nblk, kResThreads, 0, st>>>(s->row_ptr, s->col_idx, s->values, s->b, x, N, s->tr, s->stop);
    NRF_LAUNCH_CHECK();
    tr_decide_kernel<<<1, 64, 0, st>>>(s->tr, nblk, s->params.pcg_tolerance, k == R + 1 ? 1 : 0, s->d_res, s->stop);
    NRF_LAUNCH_CHECK();
    if (k == R + 1) break;
    int rc = pcg_launch(s, s->rres, s->dref, N, s->params.pcg_iterations, false, st, false, k);
    if (rc) return rc;
    tr_axpy_kernel<<<(n6 + 255) / 256, 256, 0, st>>>(x, n6, s->tr, s->stop);
    NRF_LAUNCH_CHECK();
  }
  return NRF_OK;
}

// ---- plain form (the reference's structure; see plain_data_kernel)
static int plain_buffers(nrf_solver *s, int N) {
  const size_t n6 = (size_t)6 * N;
  const size_t nblk = (n6 + kPlainThreads - 1) / kPlainThreads;
  int rc;
  if ((rc = dev_alloc(s->pl_r, n6, s->c_pl_r)) || (rc = dev_alloc(s->pl_z, n6, s->c_pl_z)) ||
      (rc = dev_alloc(s->pl_p, n6, s->c_pl_p)) || (rc = dev_alloc(s->pl_Ap, n6, s->c_pl_Ap)) ||
      (rc = dev_alloc(s->pl_part, 2 * nblk, s->c_pl_part)) || (rc = dev_alloc(s->pl_out, 2, s->c_pl_out)))
    return rc;
  if (!s->pl_host) NRF_HIP_CHECK(hipHostMalloc((void **)&s->pl_host, 2 * sizeof(double), hipHostMallocDefault));
  return NRF_OK;
}

// The data term, ARAP term, damping and block-Jacobi inverses as the
// reference launches them (kernels/solver_kernels.cu:543-614), atomics
// into the cleared system.
static int plain_assemble(nrf_solver *s, const nrf_corr *corrs, int n, const float *canonical,
                          const nrf_node *nodes, const DQ *T, const DQ *centered, int N, hipStream_t st) {
  const AsmArgs a = solver_asm_args(s, corrs, canonical, nodes, T, centered, N);
  const int nv = s->P.nb * 36, gn = (N + kPlainThreads - 1) / kPlainThreads;
  plain_clear_kernel<<<std::max(1, std::min(2048, (nv + kPlainThreads - 1) / kPlainThreads)), kPlainThreads, 0, st>>>(
      s->values, nv, s->b, 6 * N);
  NRF_LAUNCH_CHECK();
  if (n > 0) {
    plain_data_kernel<<<(n + kPlainThreads - 1) / kPlainThreads, kPlainThreads, 0, st>>>(a, n);
    NRF_LAUNCH_CHECK();
  }
  plain_arap_kernel<<<gn, kPlainThreads, 0, st>>>(a);
  NRF_LAUNCH_CHECK();
  plain_damping_kernel<<<gn, kPlainThreads, 0, st>>>(a);
  NRF_LAUNCH_CHECK();
  plain_precond_kernel<<<gn, kPlainThreads, 0, st>>>(a);
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}

// x . y in FP64, read back to the host (thrust::inner_product's round trip).
static int plain_dot(nrf_solver *s, const float *x, const float *y, int n, hipStream_t st, double &out) {
  const int nblk = (n + kPlainThreads - 1) / kPlainThreads;
  plain_dot_kernel<<<nblk, kPlainThreads, 0, st>>>(x, y, n, s->pl_part);
  NRF_LAUNCH_CHECK();
  plain_sum_kernel<<<1, kPlainThreads, 0, st>>>(s->pl_part, nblk, s->pl_out);
  NRF_LAUNCH_CHECK();
  NRF_HIP_CHECK(hipMemcpyAsync(s->pl_host, s->pl_out, sizeof(double), hipMemcpyDeviceToHost, st));
  NRF_HIP_CHECK(hipStreamSynchronize(st));
  out = s->pl_host[0];
  return NRF_OK;
}

// pcg_solve (kernels/solver_kernels.cu:616-687) from x0 = 0 to ||r|| < tol:
// SpMV, dot, two axpys, dot, preconditioner, dot, xpay per iteration, each
// a launch, each dot read by the host.  iters / flags (bit 0 breakdown, bit 1
// converged) / rr (the recurrence's r.r at the end) are returned.
static int plain_pcg(nrf_solver *s, const float *rhs, float *x, int N, int maxit, double tol, hipStream_t st,
                     int &iters, int &flags, double &rr) {
  const int n = 6 * N, gv = (n + kPlainThreads - 1) / kPlainThreads, gn = (N + kPlainThreads - 1) / kPlainThreads;
  int rc;
  NRF_HIP_CHECK(hipMemsetAsync(x, 0, sizeof(float) * n, st));
  NRF_HIP_CHECK(hipMemcpyAsync(s->pl_r, rhs, sizeof(float) * n, hipMemcpyDeviceToDevice, st));
  plain_precond_apply_kernel<<<gn, kPlainThreads, 0, st>>>(s->Minv, s->pl_r, s->pl_z, N);
  NRF_LAUNCH_CHECK();
  NRF_HIP_CHECK(hipMemcpyAsync(s->pl_p, s->pl_z, sizeof(float) * n, hipMemcpyDeviceToDevice, st));
  double rz;
  if ((rc = plain_dot(s, s->pl_r, s->pl_z, n, st, rz))) return rc;
  rr = -1.0;
  iters = 0;
  flags = 0;
  const bool rel = s->params.relative_tolerance != 0;
  for (int it = 0; it < maxit; ++it) {
    plain_spmv_kernel<<<gv, kPlainThreads, 0, st>>>(s->row_ptr, s->col_idx, s->values, s->pl_p, s->pl_Ap, N);
    NRF_LAUNCH_CHECK();
    double pAp;
    if ((rc = plain_dot(s, s->pl_p, s->pl_Ap, n, st, pAp))) return rc;
    if (rel ? !(pAp > 0.0) : fabs(pAp) < 1e-12) {  // the breakdown stop of the chosen tolerance form
      flags |= 1;
      break;
    }
    const double alpha = rz / pAp;
    plain_axpy_kernel<<<gv, kPlainThreads, 0, st>>>((float)alpha, s->pl_p, x, n);
    NRF_LAUNCH_CHECK();
    plain_axpy_kernel<<<gv, kPlainThreads, 0, st>>>((float)-alpha, s->pl_Ap, s->pl_r, n);
    NRF_LAUNCH_CHECK();
    if ((rc = plain_dot(s, s->pl_r, s->pl_r, n, st, rr))) return rc;
    iters = it + 1;
    if (sqrt(rr) < tol) {
      flags |= 2;
      break;
    }
    plain_precond_apply_kernel<<<gn, kPlainThreads, 0, st>>>(s->Minv, s->pl_r, s->pl_z, N);
    NRF_LAUNCH_CHECK();
    double rz_new;
    if ((rc = plain_dot(s, s->pl_r, s->pl_z, n, st, rz_new))) return rc;
    plain_xpay_kernel<<<gv, kPlainThreads, 0, st>>>((float)(rz_new / rz), s->pl_z, s->pl_p, n);
    NRF_LAUNCH_CHECK();
    rz = rz_new;
  }
  if (rr < 0.0) return plain_dot(s, s->pl_r, s->pl_r, n, st, rr);  // stopped before its first r.r
  return NRF_OK;
}

// One solve of the plain form: the PCG on the assembled system and, with
// true_residual_rounds, correction rounds judged by the FP64 true residual
// (as solver_tr_rounds); the outcome goes to the solver's result record.
static int plain_solve(nrf_solver *s, float *x, int N, hipStream_t st) {
  int rc;
  if ((rc = plain_buffers(s, N))) return rc;
  const int n = 6 * N, gv = (n + kPlainThreads - 1) / kPlainThreads;
  double bb;
  if ((rc = plain_dot(s, s->b, s->b, n, st, bb))) return rc;
  const double tol = s->params.relative_tolerance ? (double)s->params.pcg_tolerance * sqrt(bb)
                                                  : (double)s->params.pcg_tolerance;
  int iters, flags;
  double rr;
  if ((rc = plain_pcg(s, s->b, x, N, s->params.pcg_iterations, tol, st, iters, flags, rr))) return rc;
  const int R = s->params.true_residual_rounds;
  if (R > 0) {
    std::vector<double> part(2 * (size_t)gv);
    for (int k = 0;; ++k) {
      plain_residual_kernel<<<gv, kPlainThreads, 0, st>>>(s->row_ptr, s->col_idx, s->values, s->b, x, N, s->rres,
                                                          s->pl_part);
      NRF_LAUNCH_CHECK();
      NRF_HIP_CHECK(hipMemcpyAsync(part.data(), s->pl_part, sizeof(double) * part.size(), hipMemcpyDeviceToHost, st));
      NRF_HIP_CHECK(hipStreamSynchronize(st));
      double tr = 0.0;
      for (int q = 0; q < gv; ++q) tr += part[2 * q];
      rr = tr;
      const bool conv = sqrt(tr) < tol || bb == 0.0;
      flags = conv ? ((flags & ~1) | 2) : (flags & ~2);
      if (conv || k == R) break;
      int it2, fl2;
      double rr2;
      if ((rc = plain_pcg(s, s->rres, s->dref, N, s->params.pcg_iterations, tol, st, it2, fl2, rr2))) return rc;
      iters += it2;
      flags = (flags & ~1) | (fl2 & 1);
      plain_axpy_kernel<<<gv, kPlainThreads, 0, st>>>(1.0f, s->dref, x, n);
      NRF_LAUNCH_CHECK();
    }
  }
  plain_tally_kernel<<<1, 64, 0, st>>>(s->d_res, iters, flags, rr, bb);
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}

static int plain_check_stream(hipStream_t st) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  NRF_HIP_CHECK(hipStreamIsCapturing(st, &cs));
  return cs == hipStreamCaptureStatusNone ? NRF_OK : NRF_ERR_INVALID_ARGUMENT;  // host reads inside
}

static int solver_fetch_result(nrf_solver *s, nrf_pcg_result *out, hipStream_t st) {
  if (!out) return NRF_OK;
  NRF_HIP_CHECK(hipMemcpyAsync(out, s->d_res, sizeof(nrf_pcg_result), hipMemcpyDeviceToHost, st));
  NRF_HIP_CHECK(hipStreamSynchronize(st));
  if (out->flags & 8) return NRF_ERR_BARRIER_TIMEOUT;
  return NRF_OK;
}

static int solver_upload_edges(nrf_solver *s);
static int solver_ensure_pattern(nrf_solver *s, const nrf_node *d_nodes, int N, hipStream_t st) {
  // Q4 of the reference: the pattern is rebuilt only when the node count
  // changes (kernels/solver_kernels.cu:554-561); the ARAP term still reads
  // the current nodes' edges (kernels/solver_kernels.cu:202-233), so a graph
  // change at the same count refreshes the edge tables against the kept
  // pattern (blocks it lacks are dropped).
  std::vector<nrf_node> h(N);
  NRF_HIP_CHECK(hipMemcpyAsync(h.data(), d_nodes, sizeof(nrf_node) * N, hipMemcpyDeviceToHost, st));
  NRF_HIP_CHECK(hipStreamSynchronize(st));
  if (N != s->P.N) return solver_set_pattern_host(s, h.data(), N);
  Pattern Q = s->P;
  build_edges(h.data(), N, Q);
  if (Q.e_i == s->P.e_i && Q.e_j == s->P.e_j && Q.e_w == s->P.e_w && Q.es_offrel == s->P.es_offrel) return NRF_OK;
  s->P.e_i.swap(Q.e_i);
  s->P.e_j.swap(Q.e_j);
  s->P.e_w.swap(Q.e_w);
  s->P.E = Q.E;
  s->P.es_off.swap(Q.es_off);
  s->P.es_entry.swap(Q.es_entry);
  s->P.es_offrel.swap(Q.es_offrel);
  return solver_upload_edges(s);
}

}  // namespace nrf

// A PCG launch has 1 + 2 * pcg_iterations synchronisation points, which must
// fit the 16 point bits of a published word's tag.
static bool params_ok(const nrf_solver_params *p) {
  // Both PCG runs of a solve (the main one and the refinement) tag their
  // synchronisation points in 16 bits: 1 + 2 * iterations must fit.
  return p->pcg_iterations >= 0 && 1 + 2 * (long)p->pcg_iterations <= kPcgMaxPoints &&
         p->refine_iterations >= 0 && 1 + 2 * (long)p->refine_iterations <= kPcgMaxPoints && p->jtj_precision >= 0 &&
         p->jtj_precision <= 4 && p->pcg_pipelined >= 0 && p->pcg_pipelined <= 3 &&
         (p->refine_jacobi == 0 || (p->refine_jacobi == 1 && p->refine_iterations == 0)) &&
         p->true_residual_rounds >= 0 && p->true_residual_rounds <= 8 &&
         (p->true_residual_rounds == 0 ||
          (p->relative_tolerance == 1 && p->refine_iterations == 0 && p->refine_jacobi == 0)) &&
         (p->psd_dropped_pairs == 0 || p->psd_dropped_pairs == 1) &&
         (p->plain_form == 0 ||
          (p->plain_form == 1 && p->jtj_precision == 0 && p->refine_iterations == 0 && p->refine_jacobi == 0 &&
           p->psd_dropped_pairs == 0));
}

extern "C" nrf_solver *nrf_solver_create(const nrf_solver_params *params, int max_nodes) {
  if (!params || max_nodes <= 0 || !params_ok(params)) return nullptr;
  nrf_solver *s = new nrf_solver();
  s->params = *params;
  s->max_nodes = max_nodes;
  s->cus = device_cu_count();
  s->max_lds = device_max_lds_per_block();
  // The persistent PCG kernels take an ordinary launch: their grids are
  // sized to be co-resident (at most the CU count at one 1024-lane workgroup
  // per CU, or the occupancy bound for a batch group) and every wait in them
  // is bounded, so a grid that could not become resident ends with the
  // barrier-timeout flag instead of hanging.
  // NRF_ASM_MFMA=0 assembles the data term with the scalar LDS form
  const char *fm = getenv("NRF_ASM_MFMA");
  if (fm && fm[0] == '0') s->asm_mfma = 0;
  const char *fg = getenv("NRF_PCG_WORKGROUPS");
  if (fg) s->forced_G = atoi(fg);
  const size_t n6 = (size_t)max_nodes * 6;
  bool ok = dev_alloc(s->b, n6, s->c_b) == 0 && dev_alloc(s->Minv, (size_t)max_nodes * 36, s->c_Minv) == 0 &&
            dev_alloc(s->zpub, n6, s->c_zpub) == 0 && dev_alloc(s->P0, n6, s->c_P0) == 0 &&
            dev_alloc(s->P1, n6, s->c_P1) == 0 && dev_alloc(s->recs, (size_t)kPcgMaxWg, s->c_recs) == 0 &&
            dev_alloc(s->ctl, 1, s->c_ctl) == 0 && dev_alloc(s->stop, 1, s->c_stop) == 0 &&
            dev_alloc(s->d_res, 1, s->c_res) == 0 && dev_alloc(s->d_tot, 1, s->c_tot) == 0 &&
            dev_alloc(s->rres, n6, s->c_rres) == 0 && dev_alloc(s->dref, n6, s->c_dref) == 0 &&
            dev_alloc(s->rpart, 2 * ((n6 + kResThreads - 1) / kResThreads), s->c_rpart) == 0 &&
            dev_alloc(s->tr, 1, s->c_tr) == 0;
  if (ok) {  // the rounds' buffers are sized for max_nodes: their addresses never change
    TrState h{1, 0, 0.0, s->rres, s->dref, s->rpart};
    ok = hipMemcpy(s->tr, &h, sizeof(h), hipMemcpyHostToDevice) == hipSuccess;
  }
  // published words start with tag 0, which no synchronisation point uses
  if (ok) ok = hipMemset(s->zpub, 0, sizeof(ll_word) * n6) == hipSuccess &&
               hipMemset(s->P0, 0, sizeof(ll_word) * n6) == hipSuccess &&
               hipMemset(s->P1, 0, sizeof(ll_word) * n6) == hipSuccess &&
               hipMemset(s->recs, 0, sizeof(SyncRec) * kPcgMaxWg) == hipSuccess &&
               hipMemset(s->ctl, 0, sizeof(PcgCtl)) == hipSuccess &&
               hipMemset(s->stop, 0, sizeof(int32_t)) == hipSuccess &&
               hipMemset(s->d_res, 0, sizeof(nrf_pcg_result)) == hipSuccess &&
               hipMemset(s->d_tot, 0, sizeof(nrf_solver_totals)) == hipSuccess;
  if (!ok) {
    nrf_solver_destroy(s);
    return nullptr;
  }
  return s;
}

extern "C" void nrf_solver_destroy(nrf_solver *s) {
  if (!s) return;
  void *ptrs[] = {s->row_ptr, s->col_idx, s->diag_rel, s->e_i, s->e_j, s->es_off, s->es_entry, s->es_offrel,
                  s->rows, s->e_w, s->values, s->b, s->Minv, s->zpub, s->P0, s->P1, s->recs, s->ctl, s->lcol,
                  s->halo_off, s->halo,
                  s->m_cnt, s->m_off, s->m_cur, s->m_keys, s->m_keys2, s->stop, s->d_res, s->d_tot,
                  s->crec, s->arec, s->rres, s->dref, s->rpart, s->rnorm, s->chunk_off, s->chunk_row, s->part,
                  s->tr, s->pl_r, s->pl_z, s->pl_p, s->pl_Ap, s->pl_part, s->pl_out, s->vmask};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (s->pl_host) (void)hipHostFree(s->pl_host);
  delete s;
}

extern "C" int nrf_solver_set_params(nrf_solver *s, const nrf_solver_params *params) {
  if (!s || !params || !params_ok(params)) return NRF_ERR_INVALID_ARGUMENT;
  s->params = *params;
  return NRF_OK;
}

extern "C" int nrf_solver_set_pattern(nrf_solver *s, const nrf_node *host_nodes, int num_nodes) {
  if (!s || num_nodes < 0 || num_nodes > s->max_nodes || (num_nodes > 0 && !host_nodes))
    return NRF_ERR_INVALID_ARGUMENT;
  return solver_set_pattern_host(s, host_nodes, num_nodes);
}

extern "C" int nrf_solver_solve(nrf_solver *s, const nrf_corr *corrs, int num_corrs, const nrf_node *nodes,
                                const nrf_dualquat *transforms, int num_nodes, float *delta_x,
                                nrf_pcg_result *result, void *stream) {
  if (!s || num_corrs < 0 || num_nodes < 0 || num_nodes > s->max_nodes) return NRF_ERR_INVALID_ARGUMENT;
  if (num_nodes == 0) return NRF_OK;
  if (!nodes || !transforms || !delta_x || (num_corrs > 0 && !corrs)) return NRF_ERR_INVALID_ARGUMENT;
  hipStream_t st = as_stream(stream);
  int rc;
  if ((rc = solver_ensure_pattern(s, nodes, num_nodes, st))) return rc;
  if (s->params.plain_form) {
    if ((rc = plain_check_stream(st))) return rc;
    gn_begin_kernel<<<1, 64, 0, st>>>(s->stop, s->d_res, s->d_tot);
    NRF_LAUNCH_CHECK();
    if ((rc = plain_assemble(s, corrs, num_corrs, nullptr, nodes, reinterpret_cast<const DQ *>(transforms), nullptr,
                             num_nodes, st)))
      return rc;
    if ((rc = plain_solve(s, delta_x, num_nodes, st))) return rc;
    return solver_fetch_result(s, result, st);
  }
  if ((rc = solver_build_memberships(s, corrs, num_corrs, num_nodes, st, true))) return rc;
  if ((rc = solver_assemble(s, corrs, num_corrs, nullptr, nodes, reinterpret_cast<const DQ *>(transforms), nullptr,
                            num_nodes, st, true)))
    return rc;
  if ((rc = solver_pcg(s, delta_x, num_nodes, st, true))) return rc;
  return solver_fetch_result(s, result, st);
}

namespace nrf {
int solver_solve_gn(nrf_solver *s, const nrf_corr *corrs, const float *canonical, int num_corrs,
                    const nrf_node *nodes, nrf_dualquat *transforms, nrf_dualquat *centered, int num_nodes,
                    const nrf_gn_params *gn, float *delta_x, nrf_pcg_result *result, hipStream_t st, bool begin) {
  if (!s || !gn || num_corrs < 0 || num_nodes < 0 || num_nodes > s->max_nodes) return NRF_ERR_INVALID_ARGUMENT;
  if (num_nodes == 0) return NRF_OK;
  if (!nodes || !transforms || !centered || !delta_x || (num_corrs > 0 && !corrs)) return NRF_ERR_INVALID_ARGUMENT;
  if (s->P.N != num_nodes) return NRF_ERR_NOT_READY;
  int rc;
  const bool plain = s->params.plain_form != 0;
  if (plain && (rc = plain_check_stream(st))) return rc;
  // the snapshot of the transforms (solver_set_snapshot) in the first kernel
  SnapArgs sn;
  if (s->snap_T) {
    sn.T = reinterpret_cast<const DQ *>(transforms);
    sn.C = reinterpret_cast<const DQ *>(centered);
    sn.dst_T = reinterpret_cast<DQ *>(s->snap_T);
    sn.dst_C = reinterpret_cast<DQ *>(s->snap_C);
    sn.N = num_nodes;
    s->snap_T = s->snap_C = nullptr;
    if (plain) {
      const size_t bytes = sizeof(DQ) * (size_t)num_nodes;
      NRF_HIP_CHECK(hipMemcpyAsync(sn.dst_T, sn.T, bytes, hipMemcpyDeviceToDevice, st));
      NRF_HIP_CHECK(hipMemcpyAsync(sn.dst_C, sn.C, bytes, hipMemcpyDeviceToDevice, st));
    }
  }
  if (!plain && (rc = solver_build_memberships(s, corrs, num_corrs, num_nodes, st, begin, sn))) return rc;
  if (begin && plain) {
    gn_begin_kernel<<<1, 64, 0, st>>>(s->stop, s->d_res, s->d_tot);
    NRF_LAUNCH_CHECK();
  }
  sink_record(s, st);
  GnArgs g;
  g.dx = delta_x;
  g.T = reinterpret_cast<DQ *>(transforms);
  g.centered = reinterpret_cast<DQ *>(centered);
  g.nodes = nodes;
  g.N = num_nodes;
  g.scale = gn->update_scale;
  g.max_rot = gn->max_update_rot;
  g.max_trans = gn->max_update_trans;
  g.max_dx_mean = gn->max_dx_mean;
  g.validate = gn->validate;
  g.stop = s->stop;
  g.res = s->d_res;
  // the caller's pending work that reads the centred table (solver_set_pcg_wait)
  // ends before the first validation or twist update writes anything it reads
  auto wait_before_update = [&]() -> int {
    if (s->pcg_wait) {
      NRF_HIP_CHECK(hipStreamWaitEvent(st, s->pcg_wait, 0));
      s->pcg_wait = nullptr;
    }
    return NRF_OK;
  };
  // the caller's event before the first PCG launch and its callback after
  // that launch is enqueued (solver_set_pcg_hook)
  auto hook_mark = [&]() -> int {
    if (s->hook && s->hook_ev) NRF_HIP_CHECK(hipEventRecord(s->hook_ev, st));
    return NRF_OK;
  };
  auto hook_run = [&]() -> int {
    if (!s->hook) return NRF_OK;
    const solver_hook_fn fn = s->hook;
    s->hook = nullptr;
    return fn(s->hook_ctx, st);
  };
  for (int it = 0; it < std::max(1, gn->gn_iterations); ++it) {
    if (plain) {  // the reference's structure: atomics, host-driven PCG
      if ((rc = plain_assemble(s, corrs, num_corrs, canonical, nodes, reinterpret_cast<const DQ *>(transforms),
                               reinterpret_cast<const DQ *>(centered), num_nodes, st)))
        return rc;
      sink_record(s, st);
      if ((rc = hook_mark())) return rc;
      if ((rc = plain_solve(s, delta_x, num_nodes, st))) return rc;
      sink_record(s, st);
      if ((rc = hook_run())) return rc;
      if ((rc = wait_before_update())) return rc;
      if (g.validate) {
        gn_validate_kernel<<<1, 1024, 0, st>>>(g);
        NRF_LAUNCH_CHECK();
      }
      gn_update_kernel<<<(num_nodes + 255) / 256, 256, 0, st>>>(g);
      NRF_LAUNCH_CHECK();
      sink_record(s, st);
      if (g.validate) {  // the caller's loop ends on a rejected step (host read, as the reference's)
        int32_t stop = 0;
        NRF_HIP_CHECK(hipMemcpyAsync(&stop, s->stop, sizeof(stop), hipMemcpyDeviceToHost, st));
        NRF_HIP_CHECK(hipStreamSynchronize(st));
        if (stop) break;
      }
      continue;
    }
    if ((rc = solver_assemble(s, corrs, num_corrs, canonical, nodes, reinterpret_cast<const DQ *>(transforms),
                              reinterpret_cast<const DQ *>(centered), num_nodes, st, true)))
      return rc;
    sink_record(s, st);
    if ((rc = hook_mark())) return rc;
    if (s->params.refine_jacobi && s->params.pcg_pipelined == 3 && s->pcg_sub != 1 &&
        1 + 2 * (long)s->params.pcg_iterations < kPcgMaxPoints) {
      // the refinement step at the end of the pipelined PCG launch
      if ((rc = pcg_launch(s, s->b, delta_x, num_nodes, s->params.pcg_iterations, true, st, true, 0, true)))
        return rc;
      sink_record(s, st);
      if ((rc = hook_run())) return rc;
      if ((rc = wait_before_update())) return rc;
      if (g.validate) {
        gn_validate_kernel<<<1, 1024, 0, st>>>(g);
        NRF_LAUNCH_CHECK();
      }
      gn_update_kernel<<<(num_nodes + 255) / 256, 256, 0, st>>>(g);
      NRF_LAUNCH_CHECK();
      sink_record(s, st);
      continue;
    }
    if (s->params.refine_jacobi && !g.validate) {
      // the refinement step fused into the update (no validation between)
      if ((rc = dev_alloc(s->dref, (size_t)6 * num_nodes, s->c_dref))) return rc;
      if ((rc = pcg_launch(s, s->b, s->dref, num_nodes, s->params.pcg_iterations, true, st, true))) return rc;
      sink_record(s, st);
      if ((rc = hook_run())) return rc;
      const RefineArgs ra{s->row_ptr, s->col_idx, s->values, s->b, s->Minv, s->dref, delta_x};
      if ((rc = wait_before_update())) return rc;
      gn_refine_update_kernel<<<(num_nodes + 3) / 4, 256, 0, st>>>(g, ra);
      NRF_LAUNCH_CHECK();
      sink_record(s, st);
      continue;
    }
    if ((rc = solver_pcg(s, delta_x, num_nodes, st, true))) return rc;
    sink_record(s, st);
    if ((rc = hook_run())) return rc;
    if ((rc = wait_before_update())) return rc;
    if (g.validate) {
      gn_validate_kernel<<<1, 1024, 0, st>>>(g);
      NRF_LAUNCH_CHECK();
    }
    gn_update_kernel<<<(num_nodes + 255) / 256, 256, 0, st>>>(g);
    NRF_LAUNCH_CHECK();
    sink_record(s, st);
  }
  return solver_fetch_result(s, result, st);
}

int solver_gn_begin(nrf_solver *s, hipStream_t st) {
  if (!s) return NRF_ERR_INVALID_ARGUMENT;
  gn_begin_kernel<<<1, 64, 0, st>>>(s->stop, s->d_res, s->d_tot);
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}
int32_t *solver_stop_flag(nrf_solver *s) { return s ? s->stop : nullptr; }
nrf_pcg_result *solver_result_device(nrf_solver *s) { return s ? s->d_res : nullptr; }

int solver_gn_linearise(nrf_solver *s, const nrf_corr *corrs, int num_corrs, const nrf_node *nodes,
                        const nrf_dualquat *transforms, const nrf_dualquat *centered, int num_nodes,
                        const nrf_gn_params *gn, float *delta_x, hipStream_t st) {
  if (!s || !gn || num_corrs < 0 || num_nodes <= 0 || num_nodes > s->max_nodes) return NRF_ERR_INVALID_ARGUMENT;
  if (!nodes || !transforms || !centered || !delta_x || (num_corrs > 0 && !corrs)) return NRF_ERR_INVALID_ARGUMENT;
  if (s->P.N != num_nodes) return NRF_ERR_NOT_READY;
  int rc;
  if ((rc = solver_build_memberships(s, corrs, num_corrs, num_nodes, st))) return rc;
  if ((rc = solver_assemble(s, corrs, num_corrs, nullptr, nodes, reinterpret_cast<const DQ *>(transforms),
                            reinterpret_cast<const DQ *>(centered), num_nodes, st, true)))
    return rc;
  sink_record(s, st);
  if ((rc = solver_pcg(s, delta_x, num_nodes, st, true))) return rc;
  sink_record(s, st);
  if (gn->validate) {
    GnArgs g;
    g.dx = delta_x;
    g.T = nullptr;
    g.centered = nullptr;
    g.nodes = nodes;
    g.N = num_nodes;
    g.scale = gn->update_scale;
    g.max_rot = gn->max_update_rot;
    g.max_trans = gn->max_update_trans;
    g.max_dx_mean = gn->max_dx_mean;
    g.validate = 1;
    g.stop = s->stop;
    g.res = s->d_res;
    gn_validate_kernel<<<1, 1024, 0, st>>>(g);
    NRF_LAUNCH_CHECK();
  }
  return NRF_OK;
}
}  // namespace nrf

extern "C" int nrf_solver_solve_gn(nrf_solver *s, const nrf_corr *corrs, const float *canonical, int num_corrs,
                                   const nrf_node *nodes, nrf_dualquat *transforms, nrf_dualquat *centered,
                                   int num_nodes, const nrf_gn_params *gn, float *delta_x,
                                   nrf_pcg_result *result, void *stream) {
  return solver_solve_gn(s, corrs, canonical, num_corrs, nodes, transforms, centered, num_nodes, gn, delta_x, result,
                         as_stream(stream), true);
}

extern "C" int nrf_solver_system(const nrf_solver *s, const int32_t **row_ptr, const int32_t **col_idx,
                                 const float **values, int *num_blocks, const float **rhs, const float **precond) {
  if (!s) return NRF_ERR_INVALID_ARGUMENT;
  if (s->P.N < 0) return NRF_ERR_NOT_READY;
  if (row_ptr) *row_ptr = s->row_ptr;
  if (col_idx) *col_idx = s->col_idx;
  if (values) *values = s->values;
  if (num_blocks) *num_blocks = s->P.nb;
  if (rhs) *rhs = s->b;
  if (precond) *precond = s->Minv;
  return NRF_OK;
}

extern "C" const nrf_pcg_result *nrf_solver_device_result(const nrf_solver *s) { return s ? s->d_res : nullptr; }

extern "C" int nrf_solver_set_pcg_layout(nrf_solver *s, int lanes_per_row, int wg_nodes) {
  if (!s || (lanes_per_row != 1 && lanes_per_row != 2 && lanes_per_row != 4) || wg_nodes < 1)
    return NRF_ERR_INVALID_ARGUMENT;
  s->pcg_sub = lanes_per_row;
  s->pcg_wg_nodes = wg_nodes;
  if (s->P.N <= 0) return NRF_OK;
  return solver_build_partition(s);
}

extern "C" int nrf_solver_resolve(nrf_solver *s, float *delta_x, nrf_pcg_result *result, void *stream) {
  if (!s || !delta_x) return NRF_ERR_INVALID_ARGUMENT;
  const int N = s->P.N;
  if (N <= 0) return NRF_ERR_NOT_READY;
  hipStream_t st = as_stream(stream);
  gn_begin_kernel<<<1, 64, 0, st>>>(s->stop, s->d_res, s->d_tot);
  NRF_LAUNCH_CHECK();
  int rc = solver_pcg(s, delta_x, N, st);
  if (rc) return rc;
  return solver_fetch_result(s, result, st);
}

extern "C" int nrf_solver_true_residual(nrf_solver *s, const float *x, double *out, void *stream) {
  if (!s || !x || !out) return NRF_ERR_INVALID_ARGUMENT;
  const int N = s->P.N;
  if (N <= 0) return NRF_ERR_NOT_READY;
  hipStream_t st = as_stream(stream);
  const int n6 = 6 * N, nblk = (n6 + kResThreads - 1) / kResThreads;
  int rc;
  if ((rc = dev_alloc(s->rpart, (size_t)nblk * 2, s->c_rpart))) return rc;
  if ((rc = dev_alloc(s->rnorm, 2, s->c_rnorm))) return rc;
  residual_fp64_kernel<<<nblk, kResThreads, 0, st>>>(s->row_ptr, s->col_idx, s->values, s->b, x, N, nullptr,
                                                     s->rpart);
  NRF_LAUNCH_CHECK();
  residual_norms_kernel<<<1, 64, 0, st>>>(s->rpart, nblk, s->rnorm);
  NRF_LAUNCH_CHECK();
  NRF_HIP_CHECK(hipMemcpyAsync(out, s->rnorm, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
  NRF_HIP_CHECK(hipStreamSynchronize(st));
  return NRF_OK;
}

extern "C" int nrf_solver_get_totals(nrf_solver *s, nrf_solver_totals *out, void *stream) {
  if (!s || !out) return NRF_ERR_INVALID_ARGUMENT;
  hipStream_t st = as_stream(stream);
  nrf_pcg_result last;
  NRF_HIP_CHECK(hipMemcpyAsync(out, s->d_tot, sizeof(nrf_solver_totals), hipMemcpyDeviceToHost, st));
  NRF_HIP_CHECK(hipMemcpyAsync(&last, s->d_res, sizeof(nrf_pcg_result), hipMemcpyDeviceToHost, st));
  NRF_HIP_CHECK(hipStreamSynchronize(st));
  // the last solve has not been folded in yet
  out->gn_done += last.gn_done;
  out->pcg_iterations += last.total_iterations;
  out->rejected += (last.flags & 4) ? 1 : 0;
  out->pcg_runs += last.runs;
  out->pcg_converged += last.converged;
  out->pcg_breakdowns += last.breakdowns;
  return NRF_OK;
}

// ------------------------------------------------------------------ batches

struct nrf::SolverBatch {
  std::vector<SolverBatchInstance> in;
  int count = 0, sub = 4, prec = 0, ns = 1, gn_iters = 1, any_validate = 0, tr_rounds = 0;
  int form = 0;  // the PCG form of every instance (pcg_pipelined - 1)
  int max_N = 0, rows = 0, wgs = 0;
  size_t asm_lds = 0, pcg_lds = 0;
  InstDev *d_inst = nullptr;
  int32_t *row_inst = nullptr, *row_local = nullptr, *wg_inst = nullptr, *wg_local = nullptr;
  std::vector<std::pair<int, int>> groups;  // [first, last) workgroup of each persistent launch
  nrf_pcg_result *save_res = nullptr;
  nrf_solver_totals *save_tot = nullptr;
};

namespace nrf {

void solver_batch_destroy(SolverBatch *b) {
  if (!b) return;
  void *ptrs[] = {b->d_inst, b->row_inst, b->row_local, b->wg_inst, b->wg_local, b->save_res, b->save_tot};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  delete b;
}

int solver_rmax(const nrf_solver *s) { return s ? s->P.Rmax : 0; }

int solver_reserve(nrf_solver *s, int max_corrs) {
  if (!s || max_corrs < 0) return NRF_ERR_INVALID_ARGUMENT;
  const size_t N = (size_t)std::max(1, s->P.N), nc = (size_t)std::max(1, max_corrs);
  int rc;
  if ((rc = dev_alloc(s->crec, nc * kRec, s->c_crec))) return rc;
  if ((rc = dev_alloc(s->m_cnt, N, s->c_m_cnt))) return rc;
  if ((rc = dev_alloc(s->m_off, N + 1, s->c_m_off))) return rc;
  if ((rc = dev_alloc(s->m_cur, N, s->c_m_cur))) return rc;
  if ((rc = dev_alloc(s->m_keys, nc * 4, s->c_m_keys))) return rc;
  if ((rc = dev_alloc(s->m_keys2, nc * 4, s->c_m_keys2))) return rc;
  if ((rc = dev_alloc(s->vmask, nc, s->c_vmask))) return rc;
  if ((rc = dev_alloc(s->rres, 6 * N, s->c_rres))) return rc;
  if ((rc = dev_alloc(s->chunk_off, N + 1, s->c_chunk_off))) return rc;
  const size_t nch = asm_max_chunks((int)N, (int)nc);
  if ((rc = dev_alloc(s->chunk_row, nch, s->c_chunk_row))) return rc;
  if ((rc = dev_alloc(s->part, nch * part_floats(8), s->c_part))) return rc;
  return dev_alloc(s->dref, 6 * N, s->c_dref);
}

unsigned solver_state_generation(const nrf_solver *s) { return (s ? s->gen : 0u) + g_realloc.load(); }

template <typename T>
static int upload_new(T *&p, const std::vector<T> &v) {
  size_t cap = 0;
  return upload(p, cap, v);
}

int solver_batch_create(const SolverBatchInstance *inst, int count, int lanes_per_row, int wg_nodes,
                        SolverBatch **out) {
  if (!inst || !out || count < 1 || count > kMaxBatch) return NRF_ERR_INVALID_ARGUMENT;
  if ((lanes_per_row != 1 && lanes_per_row != 2 && lanes_per_row != 4) || wg_nodes < 1)
    return NRF_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  SolverBatch *b = new SolverBatch();
  b->in.assign(inst, inst + count);
  b->count = count;
  b->sub = lanes_per_row;
  int rc = NRF_OK;
  std::vector<InstDev> h(count);
  std::vector<int32_t> row_inst, row_local, wg_inst, wg_local;
  for (int k = 0; k < count && !rc; ++k) {
    const SolverBatchInstance &q = inst[k];
    nrf_solver *s = q.solver;
    if (!s || q.N <= 0 || !q.nodes || !q.T || !q.centered || !q.delta_x || q.max_corrs < 0) {
      rc = NRF_ERR_INVALID_ARGUMENT;
      break;
    }
    if (s->P.N != q.N) {
      rc = NRF_ERR_NOT_READY;
      break;
    }
    // one kernel form for all: the overlapped PCG, the MFMA assembly, shared
    // precision and GN iteration count, no refinement step
    const nrf_solver_params &sp = s->params;
    const int ns = asm_strips(s->P.Rmax);
    if (!sp.pcg_pipelined || sp.plain_form || sp.refine_iterations != 0 || sp.refine_jacobi != 0 || !s->asm_mfma ||
        ns > 8 ||
        asm_lds_bytes(s->P.Rmax) > 64 * 1024 ||
        (k > 0 && (sp.jtj_precision != b->prec || std::max(1, q.gn.gn_iterations) != b->gn_iters ||
                   sp.pcg_pipelined - 1 != b->form || sp.true_residual_rounds != b->tr_rounds))) {
      rc = NRF_ERR_INVALID_ARGUMENT;
      break;
    }
    if (k == 0) {
      b->form = sp.pcg_pipelined - 1;
      b->prec = sp.jtj_precision;
      b->tr_rounds = sp.true_residual_rounds;
      b->gn_iters = std::max(1, q.gn.gn_iterations);
    }
    b->ns = std::max(b->ns, ns);
    b->asm_lds = std::max(b->asm_lds, asm_lds_bytes(s->P.Rmax));
    b->any_validate |= q.gn.validate ? 1 : 0;
    b->max_N = std::max(b->max_N, q.N);
    s->pcg_sub = lanes_per_row;
    s->pcg_wg_nodes = wg_nodes;
    if ((rc = solver_build_partition(s))) break;
    b->pcg_lds = std::max(b->pcg_lds, s->pcg_lds);
    // buffers sized for the largest frame, so a step allocates nothing
    if ((rc = solver_reserve(s, q.max_corrs))) break;
    InstDev &d = h[k];
    d.as = solver_asm_args(s, nullptr, nullptr, q.nodes, reinterpret_cast<const DQ *>(q.T),
                           reinterpret_cast<const DQ *>(q.centered), q.N);
    d.as.crec = s->crec;
    d.pcg = solver_pcg_args(s, s->b, q.delta_x, q.N, sp.pcg_iterations);
    GnArgs &g = d.gn;
    g.dx = q.delta_x;
    g.T = reinterpret_cast<DQ *>(q.T);
    g.centered = reinterpret_cast<DQ *>(q.centered);
    g.nodes = q.nodes;
    g.N = q.N;
    g.scale = q.gn.update_scale;
    g.max_rot = q.gn.max_update_rot;
    g.max_trans = q.gn.max_update_trans;
    g.max_dx_mean = q.gn.max_dx_mean;
    g.validate = q.gn.validate;
    g.stop = s->stop;
    g.res = s->d_res;
    d.m_cnt = s->m_cnt;
    d.m_off = s->m_off;
    d.m_cur = s->m_cur;
    d.m_keys = s->m_keys;
    d.tot = s->d_tot;
    for (int i = 0; i < q.N; ++i) {
      row_inst.push_back(k);
      row_local.push_back(i);
    }
    for (int g2 = 0; g2 < s->G; ++g2) {
      wg_inst.push_back(k);
      wg_local.push_back(g2);
    }
  }
  if (!rc && b->pcg_lds > 65536 &&
      hipFuncSetAttribute(pcg_batch_fn(b->sub, b->form), hipFuncAttributeMaxDynamicSharedMemorySize, (int)b->pcg_lds) !=
          hipSuccess) {
    (void)hipGetLastError();
    rc = NRF_ERR_TOO_LARGE;
  }
  if (!rc) {
    // launch groups: consecutive instances whose workgroups are co-resident
    int per_cu = 0;
    NRF_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, pcg_batch_fn(b->sub, b->form), kPcgThreads,
                                                               b->pcg_lds));
    const int cap = per_cu * device_cu_count();
    int first = 0, used = 0, wg = 0;
    for (int k = 0; k < count && !rc; ++k) {
      const int G = b->in[k].solver->G;
      if (G > cap) {
        rc = NRF_ERR_TOO_LARGE;
        break;
      }
      if (used + G > cap) {
        b->groups.push_back({first, wg});
        first = wg;
        used = 0;
      }
      used += G;
      wg += G;
    }
    if (!rc) b->groups.push_back({first, wg});
    b->wgs = wg;
  }
  b->rows = (int)row_inst.size();
  if (!rc) rc = upload_new(b->d_inst, h);
  if (!rc) rc = upload_new(b->row_inst, row_inst);
  if (!rc) rc = upload_new(b->row_local, row_local);
  if (!rc) rc = upload_new(b->wg_inst, wg_inst);
  if (!rc) rc = upload_new(b->wg_local, wg_local);
  if (!rc) {
    size_t c1 = 0, c2 = 0;
    if ((rc = dev_alloc(b->save_res, (size_t)count, c1)) == 0) rc = dev_alloc(b->save_tot, (size_t)count, c2);
  }
  if (rc) {
    solver_batch_destroy(b);
    return rc;
  }
  *out = b;
  return NRF_OK;
}

int solver_batch_launch_groups(const SolverBatch *b) { return b ? (int)b->groups.size() : 0; }
int solver_batch_workgroups(const SolverBatch *b) { return b ? b->wgs : 0; }

static int batch_memberships(SolverBatch *b, const BatchCorrs &fr, hipStream_t st) {
  const dim3 gN((b->max_N + 255) / 256, b->count), gC((std::max(1, fr.max_n) + 255) / 256, b->count);
  mem_fill_batch_kernel<<<gN, 256, 0, st>>>(b->d_inst);
  NRF_LAUNCH_CHECK();
  if (fr.max_n > 0) {
    mem_count_batch_kernel<<<gC, 256, 0, st>>>(b->d_inst, fr);
    NRF_LAUNCH_CHECK();
  }
  // offsets, cursors and chunk offsets in one pass (the counts are cleared
  // behind the reads, as in the single-instance build)
  member_scan_batch_kernel<<<b->count, 1024, 0, st>>>(b->d_inst);
  NRF_LAUNCH_CHECK();
  if (fr.max_n > 0) {
    mem_scatter_batch_kernel<<<gC, 256, 0, st>>>(b->d_inst, fr);
    NRF_LAUNCH_CHECK();
  }
  mem_sort_batch_kernel<<<b->rows, kSortThreads, 0, st>>>(b->d_inst, b->row_inst, b->row_local);
  NRF_LAUNCH_CHECK();
  gn_begin_batch_kernel<<<1, 64, 0, st>>>(b->d_inst, b->count);
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}

static int batch_assemble(SolverBatch *b, const BatchCorrs &fr, hipStream_t st) {
  if (fr.max_n > 0) {
    corr_prep_batch_kernel<<<dim3((fr.max_n + 255) / 256, b->count), 256, 0, st>>>(b->d_inst, fr);
    NRF_LAUNCH_CHECK();
  }
  const size_t lds = b->asm_lds;
  const int nch = fr.cbase[b->count];
#define NRF_ASMB_CASE(NSV)                                                                                    \
  case NSV:                                                                                                   \
    if (b->prec == 1) {                                                                                       \
      assemble_chunks_batch_kernel<NSV, 1><<<nch, 64, lds, st>>>(b->d_inst, fr, b->count);                    \
      assemble_tail_batch_kernel<NSV, 1><<<b->rows, 64, lds, st>>>(b->d_inst, fr, b->row_inst, b->row_local); \
    } else if (b->prec == 2) {                                                                                \
      assemble_chunks_batch_kernel<NSV, 2><<<nch, 64, lds, st>>>(b->d_inst, fr, b->count);                    \
      assemble_tail_batch_kernel<NSV, 2><<<b->rows, 64, lds, st>>>(b->d_inst, fr, b->row_inst, b->row_local); \
    } else if (b->prec == 3) {                                                                                \
      assemble_chunks_batch_kernel<NSV, 3><<<nch, 64, lds, st>>>(b->d_inst, fr, b->count);                    \
      assemble_tail_batch_kernel<NSV, 3><<<b->rows, 64, lds, st>>>(b->d_inst, fr, b->row_inst, b->row_local); \
    } else if (b->prec == 4) {                                                                                \
      assemble_chunks_batch_kernel<NSV, 4><<<nch, 64, lds, st>>>(b->d_inst, fr, b->count);                    \
      assemble_tail_batch_kernel<NSV, 4><<<b->rows, 64, lds, st>>>(b->d_inst, fr, b->row_inst, b->row_local); \
    } else {                                                                                                  \
      assemble_chunks_batch_kernel<NSV, 0><<<nch, 64, lds, st>>>(b->d_inst, fr, b->count);                    \
      assemble_tail_batch_kernel<NSV, 0><<<b->rows, 64, lds, st>>>(b->d_inst, fr, b->row_inst, b->row_local); \
    }                                                                                                         \
    break;
  switch (b->ns) {
    NRF_ASMB_CASE(1)
    NRF_ASMB_CASE(2)
    NRF_ASMB_CASE(3)
    NRF_ASMB_CASE(4)
    NRF_ASMB_CASE(5)
    NRF_ASMB_CASE(6)
    NRF_ASMB_CASE(7)
    NRF_ASMB_CASE(8)
    default: return NRF_ERR_INVALID_ARGUMENT;
  }
#undef NRF_ASMB_CASE
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}

static int batch_pcg_launch(SolverBatch *b, int round, hipStream_t st) {
  pcg_prep_batch_kernel<<<1, 64, 0, st>>>(b->d_inst, b->count);
  NRF_LAUNCH_CHECK();
  const void *fn = pcg_batch_fn(b->sub, b->form);
  for (const auto &g : b->groups) {
    const InstDev *inst = b->d_inst;
    const int32_t *wi = b->wg_inst + g.first, *wl = b->wg_local + g.first;
    void *args[] = {&inst, &wi, &wl, &round};
    const dim3 grid(g.second - g.first);
    NRF_HIP_CHECK(hipLaunchKernel(fn, grid, dim3(kPcgThreads), args, (unsigned)b->pcg_lds, st));
  }
  return NRF_OK;
}

// The PCG of every instance, then (true_residual_rounds > 0) the rounds of
// solver_tr_rounds over all instances at once.
int solver_batch_pcg(SolverBatch *b, hipStream_t st) {
  int rc = batch_pcg_launch(b, 0, st);
  if (rc || b->tr_rounds <= 0) return rc;
  const int maxblk = (6 * b->max_N + kResThreads - 1) / kResThreads;
  const dim3 gR(maxblk, b->count), gA((6 * b->max_N + 255) / 256, b->count);
  for (int k = 1; k <= b->tr_rounds + 1; ++k) {
    tr_residual_batch_kernel<<<gR, kResThreads, 0, st>>>(b->d_inst);
    NRF_LAUNCH_CHECK();
    tr_decide_batch_kernel<<<1, 64, 0, st>>>(b->d_inst, b->count, k == b->tr_rounds + 1 ? 1 : 0);
    NRF_LAUNCH_CHECK();
    if (k == b->tr_rounds + 1) break;
    if ((rc = batch_pcg_launch(b, k, st))) return rc;
    tr_axpy_batch_kernel<<<gA, 256, 0, st>>>(b->d_inst);
    NRF_LAUNCH_CHECK();
  }
  return NRF_OK;
}

int solver_batch_solve_gn(SolverBatch *b, const BatchCorrs &fr, hipStream_t st) {
  if (!b) return NRF_ERR_INVALID_ARGUMENT;
  int rc;
  if ((rc = batch_memberships(b, fr, st))) return rc;
  const dim3 gN((b->max_N + 255) / 256, b->count);
  for (int it = 0; it < b->gn_iters; ++it) {
    if ((rc = batch_assemble(b, fr, st))) return rc;
    if ((rc = solver_batch_pcg(b, st))) return rc;
    if (b->any_validate) {
      gn_validate_batch_kernel<<<b->count, 1024, 0, st>>>(b->d_inst);
      NRF_LAUNCH_CHECK();
    }
    gn_update_batch_kernel<<<gN, 256, 0, st>>>(b->d_inst);
    NRF_LAUNCH_CHECK();
  }
  return NRF_OK;
}

int solver_batch_prepare(SolverBatch *b, const BatchCorrs &fr, hipStream_t st) {
  if (!b) return NRF_ERR_INVALID_ARGUMENT;
  for (int k = 0; k < b->count; ++k) {
    const nrf_solver *s = b->in[k].solver;
    NRF_HIP_CHECK(hipMemcpyAsync(b->save_res + k, s->d_res, sizeof(nrf_pcg_result), hipMemcpyDeviceToDevice, st));
    NRF_HIP_CHECK(
        hipMemcpyAsync(b->save_tot + k, s->d_tot, sizeof(nrf_solver_totals), hipMemcpyDeviceToDevice, st));
  }
  int rc;
  if ((rc = batch_memberships(b, fr, st))) return rc;
  return batch_assemble(b, fr, st);
}

int solver_batch_restore(SolverBatch *b, hipStream_t st) {
  if (!b) return NRF_ERR_INVALID_ARGUMENT;
  for (int k = 0; k < b->count; ++k) {
    nrf_solver *s = b->in[k].solver;
    NRF_HIP_CHECK(hipMemcpyAsync(s->d_res, b->save_res + k, sizeof(nrf_pcg_result), hipMemcpyDeviceToDevice, st));
    NRF_HIP_CHECK(
        hipMemcpyAsync(s->d_tot, b->save_tot + k, sizeof(nrf_solver_totals), hipMemcpyDeviceToDevice, st));
  }
  return NRF_OK;
}

}  // namespace nrf
