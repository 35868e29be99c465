// learned.hip - learned preconditioner for the GN solve (build-only component,
// no counterpart in the reference).
//
// A 3-layer message-passing network over the deformation graph (hidden width
// 64, BF16 weights, FP32 accumulation) maps each node's features (position,
// normal, radius, diagonal JtJ block) and each edge's features (offset, ARAP
// weight) to the blocks of a block lower-triangular L; the preconditioner is
// M^-1 = BJ^-1 + L L^T with L_ii = C_i lower(Td_i), L_ij = C_i Te_ij for
// neighbours j < i, C_i C_i^T = BJ_i^-1.  Variant 0 keeps the diagonal blocks
// only (then M^-1 is block diagonal and is folded into the solver's
// block-Jacobi inverses, so the persistent PCG runs unchanged); variant 1 is
// the graph pattern.  The definition (features, layer order, weight layout,
// loss) is oracle/learned.cpp's.
//
// Mapping: one wavefront per node, lane k owning hidden unit k; the inputs a
// layer reads (a neighbour's hidden state, the edge features) are staged in
// LDS once per wave and read by every lane; BF16 weights are stored input-
// major (W[c][k]), so the 64 lanes read 128 contiguous bytes per input.
// Training runs forward, loss, backward and Adam as device kernels; weight
// gradients are sums over nodes / edges of outer products, formed per row
// chunk and then summed over the chunks in a fixed order (no atomics:
// gradients are bitwise reproducible).
#include <cmath>
#include <cstring>
#include <vector>

#include "nrf_internal.h"

namespace nrf {
namespace {

constexpr int H = 64, F = 28, EF = 4, KN = 8, LAYERS = 3, NTD = 21, NTE = 36;
constexpr int MIN = H + EF, UIN = 2 * H, EIN = 2 * H + EF;
constexpr int kChunks = 64;  // row chunks of the weight-gradient reductions

struct LpOff {
  int W0, b0, Wm[LAYERS], bm[LAYERS], Wu[LAYERS], bu[LAYERS], Wo, bo, We, be, total;
};
constexpr LpOff make_off() {
  LpOff f{};
  int o = 0;
  f.W0 = o; o += F * H;
  f.b0 = o; o += H;
  for (int l = 0; l < LAYERS; ++l) {
    f.Wm[l] = o; o += MIN * H;
    f.bm[l] = o; o += H;
    f.Wu[l] = o; o += UIN * H;
    f.bu[l] = o; o += H;
  }
  f.Wo = o; o += H * NTD;
  f.bo = o; o += NTD;
  f.We = o; o += EIN * NTE;
  f.be = o; o += NTE;
  f.total = o;
  return f;
}
constexpr LpOff OFF = make_off();
static_assert(OFF.total == 46025, "parameter count of the network");

__device__ inline float wt(const uint16_t *w, int idx) { return __uint_as_float((unsigned)w[idx] << 16); }
__device__ inline float relu(float v) { return v > 0.f ? v : 0.f; }

__device__ inline int lp_nbr(const nrf_node *nodes, int N, int i, int s) {
  if (s >= nodes[i].num_neighbors) return -1;
  const int j = nodes[i].neighbors[s];
  return (j >= 0 && j < N && j != i) ? j : -1;
}
// component q (0..3) of the edge feature (p_j - p_i) / radius_i, w_ij
__device__ inline float lp_edge_feat(const nrf_node *nodes, int i, int j, int s, int q) {
  if (q == 3) return nodes[i].neighbor_w[s];
  const float r = fmaxf(nodes[i].radius, 1e-6f);
  return (nodes[j].pos[q] - nodes[i].pos[q]) / r;
}
__device__ inline bool lp_ledge(int variant, const nrf_node *nodes, int N, int i, int s, int &j) {
  j = lp_nbr(nodes, N, i, s);
  return variant != 0 && j >= 0 && j < i;
}
__device__ inline int tri(int r, int c) { return r * (r + 1) / 2 + c; }

// ---------------------------------------------------------------- weights
__global__ void lp_to_bf16_kernel(const float *__restrict__ p, uint16_t *__restrict__ w, int n) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  unsigned u = __float_as_uint(p[q]);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) {
    w[q] = (uint16_t)((u >> 16) | 0x40u);
    return;
  }
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  w[q] = (uint16_t)(u >> 16);
}

// --------------------------------------------------------------- features
__global__ void lp_centroid_kernel(const nrf_node *__restrict__ nodes, int N, float *__restrict__ cen) {
  __shared__ double s[3][256];
  const int t = threadIdx.x;
  double a[3] = {0, 0, 0};
  for (int i = t; i < N; i += 256)
    for (int d = 0; d < 3; ++d) a[d] += nodes[i].pos[d];
  for (int d = 0; d < 3; ++d) s[d][t] = a[d];
  __syncthreads();
  if (t == 0) {
    for (int d = 0; d < 3; ++d) {
      double v = 0;
      for (int k = 0; k < 256; ++k) v += s[d][k];
      cen[d] = (float)(v / (double)max(1, N));
    }
  }
}

// Node features and C_i = chol(BJ_i^-1) (zero when not positive definite).
__global__ void lp_features_kernel(const nrf_node *__restrict__ nodes, int N, const int32_t *__restrict__ row_ptr,
                                   const int32_t *__restrict__ col_idx, const float *__restrict__ values,
                                   const float *__restrict__ Minv, const float *__restrict__ cen,
                                   float *__restrict__ x, float *__restrict__ C) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  float A[36];
  for (int k = 0; k < 36; ++k) A[k] = 0.f;
  for (int q = row_ptr[i]; q < row_ptr[i + 1]; ++q)
    if (col_idx[q] == i)
      for (int k = 0; k < 36; ++k) A[k] = values[(size_t)q * 36 + k];
  float *f = x + (size_t)i * F;
  for (int d = 0; d < 3; ++d) f[d] = nodes[i].pos[d] - cen[d];
  for (int d = 0; d < 3; ++d) f[3 + d] = nodes[i].normal[d];
  f[6] = nodes[i].radius / 0.02f;
  for (int d = 0; d < 6; ++d) f[7 + d] = log10f(fmaxf(A[d * 6 + d], 1e-12f)) / 4.f;
  int q = 13;
  for (int a = 1; a < 6; ++a)
    for (int b = 0; b < a; ++b) f[q++] = A[a * 6 + b] / sqrtf(fmaxf(A[a * 6 + a] * A[b * 6 + b], 1e-30f));
  float L[36];
  for (int k = 0; k < 36; ++k) L[k] = 0.f;
  const float *M = Minv + (size_t)i * 36;
  bool ok = true;
  for (int r = 0; r < 6 && ok; ++r)
    for (int c = 0; c <= r; ++c) {
      float s = M[r * 6 + c];
      for (int k = 0; k < c; ++k) s -= L[r * 6 + k] * L[c * 6 + k];
      if (r == c) {
        if (!(s > 0.f)) {
          ok = false;
          break;
        }
        L[r * 6 + r] = sqrtf(s);
      } else {
        L[r * 6 + c] = s / L[c * 6 + c];
      }
    }
  for (int k = 0; k < 36; ++k) C[(size_t)i * 36 + k] = ok ? L[k] : 0.f;
}

// ---------------------------------------------------------------- forward
// All forward / backward layer kernels: 256 lanes = 4 waves = 4 nodes; no
// early return (every lane reaches every barrier).
__global__ __launch_bounds__(256) void lp_layer0_kernel(const uint16_t *__restrict__ w, const float *__restrict__ x,
                                                        int N, float *__restrict__ z0, float *__restrict__ h0) {
  const int lane = threadIdx.x & 63, i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= N) return;  // no barrier in this kernel
  float z = wt(w, OFF.b0 + lane);
  for (int c = 0; c < F; ++c) z += wt(w, OFF.W0 + c * H + lane) * x[(size_t)i * F + c];
  z0[(size_t)i * H + lane] = z;
  h0[(size_t)i * H + lane] = relu(z);
}

__global__ __launch_bounds__(256) void lp_layer_kernel(const uint16_t *__restrict__ w, int l,
                                                       const nrf_node *__restrict__ nodes, int N,
                                                       const float *__restrict__ hin, float *__restrict__ zm,
                                                       float *__restrict__ mout, float *__restrict__ zu,
                                                       float *__restrict__ hout) {
  __shared__ float st[4][UIN];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, i = blockIdx.x * 4 + wv;
  const bool valid = i < N;
  const int ic = valid ? i : 0;
  int deg = 0;
  float msum = 0.f;
  for (int s = 0; s < KN; ++s) {
    const int j = valid ? lp_nbr(nodes, N, ic, s) : -1;
    if (j >= 0) {
      st[wv][lane] = hin[(size_t)j * H + lane];
      if (lane < EF) st[wv][H + lane] = lp_edge_feat(nodes, ic, j, s, lane);
    }
    __syncthreads();
    if (j >= 0) {
      float z = wt(w, OFF.bm[l] + lane);
      for (int c = 0; c < MIN; ++c) z += wt(w, OFF.Wm[l] + c * H + lane) * st[wv][c];
      zm[((size_t)ic * KN + s) * H + lane] = z;
      msum += relu(z);
      ++deg;
    } else if (valid) {
      zm[((size_t)ic * KN + s) * H + lane] = 0.f;
    }
    __syncthreads();
  }
  const float m = msum / (float)max(1, deg);
  if (valid) {
    mout[(size_t)ic * H + lane] = m;
    st[wv][lane] = hin[(size_t)ic * H + lane];
    st[wv][H + lane] = m;
  }
  __syncthreads();
  if (valid) {
    float z = wt(w, OFF.bu[l] + lane);
    for (int c = 0; c < UIN; ++c) z += wt(w, OFF.Wu[l] + c * H + lane) * st[wv][c];
    zu[(size_t)ic * H + lane] = z;
    hout[(size_t)ic * H + lane] = st[wv][lane] + relu(z);
  }
}

__global__ __launch_bounds__(256) void lp_heads_kernel(const uint16_t *__restrict__ w, int variant,
                                                       const nrf_node *__restrict__ nodes, int N,
                                                       const float *__restrict__ h3, float *__restrict__ Td,
                                                       float *__restrict__ Te) {
  __shared__ float st[4][EIN + 4];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, i = blockIdx.x * 4 + wv;
  const bool valid = i < N;
  const int ic = valid ? i : 0;
  if (valid) st[wv][lane] = h3[(size_t)ic * H + lane];
  __syncthreads();
  if (valid && lane < NTD) {
    float t = wt(w, OFF.bo + lane);
    for (int k = 0; k < H; ++k) t += wt(w, OFF.Wo + k * NTD + lane) * st[wv][k];
    Td[(size_t)ic * NTD + lane] = t;
  }
  for (int s = 0; s < KN; ++s) {
    int j = -1;
    const bool le = valid && lp_ledge(variant, nodes, N, ic, s, j);
    __syncthreads();
    if (le) {
      st[wv][H + lane] = h3[(size_t)j * H + lane];
      if (lane < EF) st[wv][2 * H + lane] = lp_edge_feat(nodes, ic, j, s, lane);
    }
    __syncthreads();
    if (valid && lane < NTE) {
      float t = 0.f;
      if (le) {
        t = wt(w, OFF.be + lane);
        for (int c = 0; c < EIN; ++c) t += wt(w, OFF.We + c * NTE + lane) * st[wv][c];
      }
      Te[((size_t)ic * KN + s) * NTE + lane] = t;
    }
  }
}

// L blocks: block b = 0 is L_ii = C_i lower(Td_i), b = 1 + s is L_is = C_i Te_is.
__global__ void lp_lblocks_kernel(int N, const float *__restrict__ C, const float *__restrict__ Td,
                                  const float *__restrict__ Te, float *__restrict__ Ld, float *__restrict__ Le) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= N * (KN + 1)) return;
  const int i = q / (KN + 1), b = q % (KN + 1);
  const float *Ci = C + (size_t)i * 36;
  float T[36];
  if (b == 0) {
    for (int r = 0; r < 6; ++r)
      for (int c = 0; c < 6; ++c) T[r * 6 + c] = c <= r ? Td[(size_t)i * NTD + tri(r, c)] : 0.f;
  } else {
    for (int k = 0; k < 36; ++k) T[k] = Te[((size_t)i * KN + b - 1) * NTE + k];
  }
  float *out = b == 0 ? Ld + (size_t)i * 36 : Le + ((size_t)i * KN + b - 1) * 36;
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 6; ++c) {
      float a = 0.f;
      for (int k = 0; k < 6; ++k) a += Ci[r * 6 + k] * T[k * 6 + c];
      out[r * 6 + c] = a;
    }
}

// Minv_i += L_ii L_ii^T (variant 0: the whole correction is block diagonal)
// One node per lane: Minv_i += L_ii L_ii^T, unless a sum of the correction
// or of the folded block is not finite; then node i keeps its block-Jacobi
// inverse (the fallback, counted in *fallbacks).
__global__ void lp_fold_kernel(int N, const float *__restrict__ Ld, float *__restrict__ Minv,
                               int32_t *__restrict__ fallbacks) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const float *L = Ld + (size_t)i * 36;
  float *Mi = Minv + (size_t)i * 36;
  float out[36];
  bool finite = true;
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 6; ++c) {
      float a = 0.f;
      for (int k = 0; k < 6; ++k) a += L[r * 6 + k] * L[c * 6 + k];
      out[r * 6 + c] = Mi[r * 6 + c] + a;
      finite = finite && isfinite(out[r * 6 + c]);
    }
  if (!finite) {
    atomicAdd(fallbacks, 1);
    return;
  }
  for (int k = 0; k < 36; ++k) Mi[k] = out[k];
}
__global__ void lp_zero_kernel(int32_t *p) { *p = 0; }

// ------------------------------------------------------------------- loss
// vectors are stored [P][6N]
__global__ void lp_spmv_kernel(const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                               const float *__restrict__ A, const float *__restrict__ V, float *__restrict__ U, int N,
                               int P) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 6 * N * P) return;
  const int p = q / (6 * N), row = q % (6 * N), i = row / 6, r = row % 6;
  const float *v = V + (size_t)p * 6 * N;
  float a = 0.f;
  for (int b = rp[i]; b < rp[i + 1]; ++b) {
    const float *blk = A + (size_t)b * 36 + r * 6;
    const float *vv = v + (size_t)ci[b] * 6;
    for (int c = 0; c < 6; ++c) a += blk[c] * vv[c];
  }
  U[q] = a;
}

// out = L^T in, per column block k: L_kk^T in_k + sum over (i, s) with
// neighbours[s] of i == k and k < i of L_is^T in_i (reverse adjacency).
__global__ void lp_lt_kernel(int variant, const nrf_node *__restrict__ nodes, int N, const float *__restrict__ Ld,
                             const float *__restrict__ Le, const int32_t *__restrict__ radj_off,
                             const int32_t *__restrict__ radj, const float *__restrict__ in, float *__restrict__ out,
                             int P) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 6 * N * P) return;
  const int p = q / (6 * N), row = q % (6 * N), k = row / 6, c = row % 6;
  const float *v = in + (size_t)p * 6 * N;
  float a = 0.f;
  for (int r = 0; r < 6; ++r) a += Ld[(size_t)k * 36 + r * 6 + c] * v[(size_t)k * 6 + r];
  if (variant != 0)
    for (int e = radj_off[k]; e < radj_off[k + 1]; ++e) {
      const int i = radj[e] / KN, s = radj[e] % KN;
      if (k >= i) continue;
      const float *L = Le + ((size_t)i * KN + s) * 36;
      for (int r = 0; r < 6; ++r) a += L[r * 6 + c] * v[(size_t)i * 6 + r];
    }
  out[q] = a;
}

// e = BJ^-1 u + L t - v; per-block partial sums of e^2 (FP64).
__global__ __launch_bounds__(256) void lp_residual_kernel(int variant, const nrf_node *__restrict__ nodes, int N,
                                                          const float *__restrict__ Minv,
                                                          const float *__restrict__ Ld,
                                                          const float *__restrict__ Le, const float *__restrict__ U,
                                                          const float *__restrict__ T, const float *__restrict__ V,
                                                          float *__restrict__ E, double *__restrict__ part, int P) {
  __shared__ double red[256];
  const int q = blockIdx.x * 256 + threadIdx.x;
  double e2 = 0;
  if (q < 6 * N * P) {
    const int p = q / (6 * N), row = q % (6 * N), i = row / 6, r = row % 6;
    const float *u = U + (size_t)p * 6 * N, *t = T + (size_t)p * 6 * N;
    float a = 0.f;
    for (int c = 0; c < 6; ++c) a += Minv[(size_t)i * 36 + r * 6 + c] * u[(size_t)i * 6 + c];
    for (int c = 0; c < 6; ++c) a += Ld[(size_t)i * 36 + r * 6 + c] * t[(size_t)i * 6 + c];
    for (int s = 0; s < KN; ++s) {
      int j;
      if (!lp_ledge(variant, nodes, N, i, s, j)) continue;
      const float *L = Le + ((size_t)i * KN + s) * 36 + r * 6;
      for (int c = 0; c < 6; ++c) a += L[c] * t[(size_t)j * 6 + c];
    }
    const float e = a - V[q];
    E[q] = e;
    e2 = (double)e * (double)e;
  }
  red[threadIdx.x] = e2;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void lp_sum_kernel(const double *__restrict__ part, int n, double scale, double *__restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0;
    for (int k = 0; k < n; ++k) s += part[k];
    *out = s * scale;
  }
}

// dT from dL_ik = 2 scale sum_p (e_i t_k^T + u_i s_k^T): dT = C_i^T dL
// (the diagonal block keeps its lower triangle).
__global__ void lp_dl_kernel(int variant, const nrf_node *__restrict__ nodes, int N, int P, float scale2,
                             const float *__restrict__ C, const float *__restrict__ U, const float *__restrict__ T,
                             const float *__restrict__ E, const float *__restrict__ S, float *__restrict__ dTd,
                             float *__restrict__ dTe) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= N * (KN + 1)) return;
  const int i = q / (KN + 1), b = q % (KN + 1);
  int k = i;
  if (b > 0) {
    int j;
    if (!lp_ledge(variant, nodes, N, i, b - 1, j)) {
      for (int o = 0; o < NTE; ++o) dTe[((size_t)i * KN + b - 1) * NTE + o] = 0.f;
      return;
    }
    k = j;
  }
  float dL[36];
  for (int o = 0; o < 36; ++o) dL[o] = 0.f;
  for (int p = 0; p < P; ++p) {
    const size_t base = (size_t)p * 6 * N;
    for (int r = 0; r < 6; ++r) {
      const float e = E[base + (size_t)i * 6 + r], u = U[base + (size_t)i * 6 + r];
      for (int c = 0; c < 6; ++c) dL[r * 6 + c] += e * T[base + (size_t)k * 6 + c] + u * S[base + (size_t)k * 6 + c];
    }
  }
  const float *Ci = C + (size_t)i * 36;
  float g[36];
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 6; ++c) {
      float a = 0.f;
      for (int m = 0; m < 6; ++m) a += Ci[m * 6 + r] * dL[m * 6 + c];
      g[r * 6 + c] = scale2 * a;
    }
  if (b == 0) {
    for (int r = 0; r < 6; ++r)
      for (int c = 0; c <= r; ++c) dTd[(size_t)i * NTD + tri(r, c)] = g[r * 6 + c];
  } else {
    for (int o = 0; o < 36; ++o) dTe[((size_t)i * KN + b - 1) * NTE + o] = g[o];
  }
}

// --------------------------------------------------------------- backward
// heads: dself_i = Wo dTd_i + sum_s We[0:64] dTe_is; dE_is = We[64:128] dTe_is
// (the part that belongs to node j, gathered later).
__global__ __launch_bounds__(256) void lp_heads_bwd_kernel(const uint16_t *__restrict__ w, int variant,
                                                           const nrf_node *__restrict__ nodes, int N,
                                                           const float *__restrict__ dTd,
                                                           const float *__restrict__ dTe, float *__restrict__ dself,
                                                           float *__restrict__ dE) {
  const int lane = threadIdx.x & 63, i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= N) return;
  float a = 0.f;
  for (int o = 0; o < NTD; ++o) a += wt(w, OFF.Wo + lane * NTD + o) * dTd[(size_t)i * NTD + o];
  for (int s = 0; s < KN; ++s) {
    int j;
    float b = 0.f;
    if (lp_ledge(variant, nodes, N, i, s, j)) {
      const float *g = dTe + ((size_t)i * KN + s) * NTE;
      for (int o = 0; o < NTE; ++o) {
        a += wt(w, OFF.We + lane * NTE + o) * g[o];
        b += wt(w, OFF.We + (H + lane) * NTE + o) * g[o];
      }
    }
    dE[((size_t)i * KN + s) * H + lane] = b;
  }
  dself[(size_t)i * H + lane] = a;
}

// dout_j = dself_j + sum over (i, s) with neighbours[s] of i == j of dE_is
__global__ void lp_gather_kernel(int N, const int32_t *__restrict__ radj_off, const int32_t *__restrict__ radj,
                                 const float *__restrict__ dE, const float *__restrict__ dself,
                                 float *__restrict__ dout) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= N * H) return;
  const int j = q / H, k = q % H;
  float a = dself[q];
  for (int e = radj_off[j]; e < radj_off[j + 1]; ++e) a += dE[(size_t)radj[e] * H + k];
  dout[q] = a;
}

// layer l: given dh = dL/dh_l, forms dzu, dzm (stored for the weight
// gradients), dself = dh + Wu[0:64] dzu (the residual and update paths into
// h_{l-1} of the node) and dE_is = Wm[0:64] dz_is (into h_{l-1} of j).
__global__ __launch_bounds__(256) void lp_layer_bwd_kernel(const uint16_t *__restrict__ w, int l,
                                                           const nrf_node *__restrict__ nodes, int N,
                                                           const float *__restrict__ dh,
                                                           const float *__restrict__ zu,
                                                           const float *__restrict__ zm, float *__restrict__ dzu,
                                                           float *__restrict__ dzm, float *__restrict__ dself,
                                                           float *__restrict__ dE) {
  __shared__ float st[4][H];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, i = blockIdx.x * 4 + wv;
  const bool valid = i < N;
  const int ic = valid ? i : 0;
  const float du = (valid && zu[(size_t)ic * H + lane] > 0.f) ? dh[(size_t)ic * H + lane] : 0.f;
  if (valid) dzu[(size_t)ic * H + lane] = du;
  st[wv][lane] = du;
  __syncthreads();
  float dlo = 0.f, dhi = 0.f;
  for (int k = 0; k < H; ++k) {
    const float d = st[wv][k];
    dlo += wt(w, OFF.Wu[l] + lane * H + k) * d;
    dhi += wt(w, OFF.Wu[l] + (H + lane) * H + k) * d;
  }
  if (valid) dself[(size_t)ic * H + lane] = (valid ? dh[(size_t)ic * H + lane] : 0.f) + dlo;
  int deg = 0;
  for (int s = 0; s < KN; ++s) deg += (valid && lp_nbr(nodes, N, ic, s) >= 0) ? 1 : 0;
  const float dm = dhi / (float)max(1, deg);
  for (int s = 0; s < KN; ++s) {
    const int j = valid ? lp_nbr(nodes, N, ic, s) : -1;
    const float dz = (j >= 0 && zm[((size_t)ic * KN + s) * H + lane] > 0.f) ? dm : 0.f;
    if (valid) dzm[((size_t)ic * KN + s) * H + lane] = dz;
    __syncthreads();
    st[wv][lane] = dz;
    __syncthreads();
    float a = 0.f;
    if (j >= 0)
      for (int k = 0; k < H; ++k) a += wt(w, OFF.Wm[l] + lane * H + k) * st[wv][k];
    if (valid) dE[((size_t)ic * KN + s) * H + lane] = a;
  }
}

__global__ void lp_relu_bwd_kernel(int n, const float *__restrict__ z, const float *__restrict__ d,
                                   float *__restrict__ out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) out[q] = z[q] > 0.f ? d[q] : 0.f;
}

// ------------------------------------------------------- weight gradients
// Inputs of each weighted layer, row r (node or edge = node * 8 + slot);
// column c == cols is the bias input 1.
struct InRows {  // x (F), h (H), or [h; m] (2H) of node r
  const float *a, *b;
  int wa, wb;
  __device__ float operator()(int r, int c) const {
    return c < wa ? a[(size_t)r * wa + c] : b[(size_t)r * wb + c - wa];
  }
};
struct InMsg {  // [h_j; e_ij] of edge r
  const float *h;
  const nrf_node *nodes;
  int N;
  __device__ float operator()(int r, int c) const {
    const int i = r / KN, s = r % KN;
    const int j = lp_nbr(nodes, N, i, s);
    if (j < 0) return 0.f;
    return c < H ? h[(size_t)j * H + c] : lp_edge_feat(nodes, i, j, s, c - H);
  }
};
struct InEdgeHead {  // [h_i; h_j; e_ij] of edge r (graph-pattern blocks only)
  const float *h;
  const nrf_node *nodes;
  int N, variant;
  __device__ float operator()(int r, int c) const {
    const int i = r / KN, s = r % KN;
    int j;
    if (!lp_ledge(variant, nodes, N, i, s, j)) return 0.f;
    if (c < H) return h[(size_t)i * H + c];
    if (c < 2 * H) return h[(size_t)j * H + c - H];
    return lp_edge_feat(nodes, i, j, s, c - 2 * H);
  }
};

// part[chunk][c][k] = sum over the chunk's rows of in(r, c) dz[r][k]
template <typename In>
__global__ void lp_wgrad_kernel(In in, int R, int cols, int K, const float *__restrict__ dz,
                                float *__restrict__ part) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = (cols + 1) * K;
  if (q >= n) return;
  const int c = q / K, k = q % K;
  const int per = (R + kChunks - 1) / kChunks;
  const int r0 = blockIdx.y * per, r1 = min(R, r0 + per);
  float a = 0.f;
  for (int r = r0; r < r1; ++r) a += (c == cols ? 1.f : in(r, c)) * dz[(size_t)r * K + k];
  part[(size_t)blockIdx.y * n + q] = a;
}
__global__ void lp_wreduce_kernel(const float *__restrict__ part, int cols, int K, float *__restrict__ gW,
                                  float *__restrict__ gb) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = (cols + 1) * K;
  if (q >= n) return;
  float a = 0.f;
  for (int ch = 0; ch < kChunks; ++ch) a += part[(size_t)ch * n + q];
  const int c = q / K, k = q % K;
  if (c < cols)
    gW[c * K + k] = a;
  else
    gb[k] = a;
}

__global__ void lp_adam_kernel(int n, float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                               float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const float gq = g[q];
  const float mq = b1 * m[q] + (1.f - b1) * gq;
  const float vq = b2 * v[q] + (1.f - b2) * gq * gq;
  m[q] = mq;
  v[q] = vq;
  p[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
}

}  // namespace
}  // namespace nrf

using namespace nrf;

struct nrf_lp {
  int variant = 0, max_nodes = 0, max_probes = 0, N = 0;
  float *params = nullptr, *grad = nullptr, *am = nullptr, *av = nullptr;
  uint16_t *wb = nullptr;
  int32_t *radj_off = nullptr, *radj = nullptr;
  float *cen = nullptr, *x = nullptr, *C = nullptr, *h[LAYERS + 1] = {}, *z0 = nullptr, *zm[LAYERS] = {},
        *m[LAYERS] = {}, *zu[LAYERS] = {}, *Td = nullptr, *Te = nullptr, *Ld = nullptr, *Le = nullptr;
  float *U = nullptr, *T = nullptr, *E = nullptr, *S = nullptr;
  float *dTd = nullptr, *dTe = nullptr, *dh = nullptr, *dself = nullptr, *dE = nullptr, *dz0 = nullptr,
        *dzu[LAYERS] = {}, *dzm[LAYERS] = {};
  float *part = nullptr;
  double *lpart = nullptr, *dloss = nullptr;
  int32_t *fallbacks = nullptr;  // nodes that kept block-Jacobi in the last fold
  std::vector<void *> owned;
  const nrf_solver *fwd_solver = nullptr;  // the system the last forward read
};

template <typename T>
static bool lp_alloc(nrf_lp *p, T *&ptr, size_t n) {
  if (hipMalloc(reinterpret_cast<void **>(&ptr), std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return false;
  p->owned.push_back(ptr);
  return true;
}

extern "C" int nrf_lp_param_count(void) { return OFF.total; }

extern "C" void nrf_lp_destroy(nrf_lp *p) {
  if (!p) return;
  for (void *q : p->owned) (void)hipFree(q);
  delete p;
}

extern "C" nrf_lp *nrf_lp_create(int variant, int max_nodes, int max_probes) {
  if ((variant != 0 && variant != 1) || max_nodes <= 0 || max_probes <= 0) return nullptr;
  nrf_lp *p = new nrf_lp();
  p->variant = variant;
  p->max_nodes = max_nodes;
  p->max_probes = max_probes;
  const size_t Nn = (size_t)max_nodes, Ne = Nn * KN, nv = (size_t)max_probes * 6 * Nn;
  const size_t part_n = (size_t)kChunks * (UIN + 1) * H;  // largest (cols + 1) x K
  bool ok = lp_alloc(p, p->params, OFF.total) && lp_alloc(p, p->grad, OFF.total) && lp_alloc(p, p->am, OFF.total) &&
            lp_alloc(p, p->av, OFF.total) && lp_alloc(p, p->wb, OFF.total) && lp_alloc(p, p->radj_off, Nn + 1) &&
            lp_alloc(p, p->radj, Ne) && lp_alloc(p, p->cen, 4) && lp_alloc(p, p->x, Nn * F) &&
            lp_alloc(p, p->C, Nn * 36) && lp_alloc(p, p->z0, Nn * H) && lp_alloc(p, p->Td, Nn * NTD) &&
            lp_alloc(p, p->Te, Ne * NTE) && lp_alloc(p, p->Ld, Nn * 36) && lp_alloc(p, p->Le, Ne * 36) &&
            lp_alloc(p, p->U, nv) && lp_alloc(p, p->T, nv) && lp_alloc(p, p->E, nv) && lp_alloc(p, p->S, nv) &&
            lp_alloc(p, p->dTd, Nn * NTD) && lp_alloc(p, p->dTe, Ne * NTE) && lp_alloc(p, p->dh, Nn * H) &&
            lp_alloc(p, p->dself, Nn * H) && lp_alloc(p, p->dE, Ne * H) && lp_alloc(p, p->dz0, Nn * H) &&
            lp_alloc(p, p->part, part_n) && lp_alloc(p, p->lpart, (nv + 255) / 256 + 1) && lp_alloc(p, p->dloss, 1) &&
            lp_alloc(p, p->fallbacks, 1);
  for (int l = 0; l <= LAYERS && ok; ++l) ok = lp_alloc(p, p->h[l], Nn * H);
  for (int l = 0; l < LAYERS && ok; ++l)
    ok = lp_alloc(p, p->zm[l], Ne * H) && lp_alloc(p, p->m[l], Nn * H) && lp_alloc(p, p->zu[l], Nn * H) &&
         lp_alloc(p, p->dzu[l], Nn * H) && lp_alloc(p, p->dzm[l], Ne * H);
  ok = ok && hipMemset(p->params, 0, sizeof(float) * OFF.total) == hipSuccess &&
       hipMemset(p->fallbacks, 0, sizeof(int32_t)) == hipSuccess &&
       hipMemset(p->am, 0, sizeof(float) * OFF.total) == hipSuccess &&
       hipMemset(p->av, 0, sizeof(float) * OFF.total) == hipSuccess &&
       hipMemset(p->grad, 0, sizeof(float) * OFF.total) == hipSuccess && hipDeviceSynchronize() == hipSuccess;
  if (!ok) {
    set_last_error_msg("nrf_lp_create: device allocation failed");
    nrf_lp_destroy(p);
    return nullptr;
  }
  return p;
}

extern "C" float *nrf_lp_params(nrf_lp *p) { return p ? p->params : nullptr; }
extern "C" float *nrf_lp_grads(nrf_lp *p) { return p ? p->grad : nullptr; }

extern "C" int nrf_lp_reset_optimizer(nrf_lp *p) {
  if (!p) return NRF_ERR_INVALID_ARGUMENT;
  NRF_HIP_CHECK(hipMemset(p->am, 0, sizeof(float) * OFF.total));
  NRF_HIP_CHECK(hipMemset(p->av, 0, sizeof(float) * OFF.total));
  return NRF_OK;
}

extern "C" int nrf_lp_set_graph(nrf_lp *p, const nrf_node *host_nodes, int num_nodes) {
  if (!p || num_nodes <= 0 || num_nodes > p->max_nodes || !host_nodes) return NRF_ERR_INVALID_ARGUMENT;
  const int N = num_nodes;
  std::vector<int32_t> cnt(N + 1, 0), off(N + 1, 0), radj;
  auto nbr = [&](int i, int s) -> int {
    if (s >= host_nodes[i].num_neighbors) return -1;
    const int j = host_nodes[i].neighbors[s];
    return (j >= 0 && j < N && j != i) ? j : -1;
  };
  for (int i = 0; i < N; ++i)
    for (int s = 0; s < KN; ++s) {
      const int j = nbr(i, s);
      if (j >= 0) ++cnt[j];
    }
  for (int j = 0; j < N; ++j) off[j + 1] = off[j] + cnt[j];
  radj.assign(std::max(1, off[N]), 0);
  std::vector<int32_t> cur(off.begin(), off.end() - 1);
  for (int i = 0; i < N; ++i)  // increasing (i, s): a fixed summation order
    for (int s = 0; s < KN; ++s) {
      const int j = nbr(i, s);
      if (j >= 0) radj[cur[j]++] = i * KN + s;
    }
  NRF_HIP_CHECK(hipMemcpy(p->radj_off, off.data(), sizeof(int32_t) * (N + 1), hipMemcpyHostToDevice));
  NRF_HIP_CHECK(hipMemcpy(p->radj, radj.data(), sizeof(int32_t) * radj.size(), hipMemcpyHostToDevice));
  p->N = N;
  return NRF_OK;
}

extern "C" int nrf_lp_forward(nrf_lp *p, const nrf_solver *s, const nrf_node *nodes, int num_nodes, void *stream) {
  if (!p || !s || !nodes || num_nodes != p->N || p->N <= 0) return NRF_ERR_INVALID_ARGUMENT;
  const int32_t *rp, *ci;
  const float *vals, *Minv;
  int nb = 0;
  int rc = nrf_solver_system(s, &rp, &ci, &vals, &nb, nullptr, &Minv);
  if (rc) return rc;
  if (nb <= 0) return NRF_ERR_NOT_READY;
  const int N = p->N;
  hipStream_t st = as_stream(stream);
  lp_to_bf16_kernel<<<(OFF.total + 255) / 256, 256, 0, st>>>(p->params, p->wb, OFF.total);
  NRF_LAUNCH_CHECK();
  lp_centroid_kernel<<<1, 256, 0, st>>>(nodes, N, p->cen);
  NRF_LAUNCH_CHECK();
  lp_features_kernel<<<(N + 127) / 128, 128, 0, st>>>(nodes, N, rp, ci, vals, Minv, p->cen, p->x, p->C);
  NRF_LAUNCH_CHECK();
  const int nw = (N + 3) / 4;
  lp_layer0_kernel<<<nw, 256, 0, st>>>(p->wb, p->x, N, p->z0, p->h[0]);
  NRF_LAUNCH_CHECK();
  for (int l = 0; l < LAYERS; ++l) {
    lp_layer_kernel<<<nw, 256, 0, st>>>(p->wb, l, nodes, N, p->h[l], p->zm[l], p->m[l], p->zu[l], p->h[l + 1]);
    NRF_LAUNCH_CHECK();
  }
  lp_heads_kernel<<<nw, 256, 0, st>>>(p->wb, p->variant, nodes, N, p->h[LAYERS], p->Td, p->Te);
  NRF_LAUNCH_CHECK();
  lp_lblocks_kernel<<<(N * (KN + 1) + 127) / 128, 128, 0, st>>>(N, p->C, p->Td, p->Te, p->Ld, p->Le);
  NRF_LAUNCH_CHECK();
  p->fwd_solver = s;
  return NRF_OK;
}

extern "C" int nrf_lp_loss_grad(nrf_lp *p, const nrf_solver *s, const nrf_node *nodes, const float *probes,
                                int num_probes, double *loss, void *stream) {
  if (!p || !s || !nodes || !probes || num_probes <= 0 || num_probes > p->max_probes || p->N <= 0)
    return NRF_ERR_INVALID_ARGUMENT;
  if (p->fwd_solver != s) return NRF_ERR_NOT_READY;  // nrf_lp_forward on this system first
  const int32_t *rp, *ci;
  const float *vals, *Minv;
  int nb = 0;
  int rc = nrf_solver_system(s, &rp, &ci, &vals, &nb, nullptr, &Minv);
  if (rc) return rc;
  const int N = p->N, P = num_probes, nv = 6 * N * P;
  hipStream_t st = as_stream(stream);
  const int vb = (nv + 255) / 256;
  lp_spmv_kernel<<<vb, 256, 0, st>>>(rp, ci, vals, probes, p->U, N, P);
  NRF_LAUNCH_CHECK();
  lp_lt_kernel<<<vb, 256, 0, st>>>(p->variant, nodes, N, p->Ld, p->Le, p->radj_off, p->radj, p->U, p->T, P);
  NRF_LAUNCH_CHECK();
  lp_residual_kernel<<<vb, 256, 0, st>>>(p->variant, nodes, N, Minv, p->Ld, p->Le, p->U, p->T, probes, p->E,
                                         p->lpart, P);
  NRF_LAUNCH_CHECK();
  const double scale = 1.0 / ((double)P * 6.0 * (double)N);
  lp_sum_kernel<<<1, 64, 0, st>>>(p->lpart, vb, scale, p->dloss);
  NRF_LAUNCH_CHECK();
  lp_lt_kernel<<<vb, 256, 0, st>>>(p->variant, nodes, N, p->Ld, p->Le, p->radj_off, p->radj, p->E, p->S, P);
  NRF_LAUNCH_CHECK();
  lp_dl_kernel<<<(N * (KN + 1) + 127) / 128, 128, 0, st>>>(p->variant, nodes, N, P, (float)(2.0 * scale), p->C,
                                                           p->U, p->T, p->E, p->S, p->dTd, p->dTe);
  NRF_LAUNCH_CHECK();
  const int nw = (N + 3) / 4, R = N * KN;
  lp_heads_bwd_kernel<<<nw, 256, 0, st>>>(p->wb, p->variant, nodes, N, p->dTd, p->dTe, p->dself, p->dE);
  NRF_LAUNCH_CHECK();
  lp_gather_kernel<<<(N * H + 255) / 256, 256, 0, st>>>(N, p->radj_off, p->radj, p->dE, p->dself, p->dh);
  NRF_LAUNCH_CHECK();
  auto wgrad = [&](auto in, int rows, int cols, int K, const float *dz, int offW, int offb) -> int {
    const int n = (cols + 1) * K;
    lp_wgrad_kernel<<<dim3((n + 255) / 256, kChunks), 256, 0, st>>>(in, rows, cols, K, dz, p->part);
    NRF_LAUNCH_CHECK();
    lp_wreduce_kernel<<<(n + 255) / 256, 256, 0, st>>>(p->part, cols, K, p->grad + offW, p->grad + offb);
    NRF_LAUNCH_CHECK();
    return NRF_OK;
  };
  // head weights
  if ((rc = wgrad(InRows{p->h[LAYERS], nullptr, H, 0}, N, H, NTD, p->dTd, OFF.Wo, OFF.bo))) return rc;
  if ((rc = wgrad(InEdgeHead{p->h[LAYERS], nodes, N, p->variant}, R, EIN, NTE, p->dTe, OFF.We, OFF.be)))
    return rc;
  for (int l = LAYERS - 1; l >= 0; --l) {
    lp_layer_bwd_kernel<<<nw, 256, 0, st>>>(p->wb, l, nodes, N, p->dh, p->zu[l], p->zm[l], p->dzu[l], p->dzm[l],
                                            p->dself, p->dE);
    NRF_LAUNCH_CHECK();
    if ((rc = wgrad(InRows{p->h[l], p->m[l], H, H}, N, UIN, H, p->dzu[l], OFF.Wu[l], OFF.bu[l]))) return rc;
    if ((rc = wgrad(InMsg{p->h[l], nodes, N}, R, MIN, H, p->dzm[l], OFF.Wm[l], OFF.bm[l]))) return rc;
    lp_gather_kernel<<<(N * H + 255) / 256, 256, 0, st>>>(N, p->radj_off, p->radj, p->dE, p->dself, p->dh);
    NRF_LAUNCH_CHECK();
  }
  lp_relu_bwd_kernel<<<(N * H + 255) / 256, 256, 0, st>>>(N * H, p->z0, p->dh, p->dz0);
  NRF_LAUNCH_CHECK();
  if ((rc = wgrad(InRows{p->x, nullptr, F, 0}, N, F, H, p->dz0, OFF.W0, OFF.b0))) return rc;
  if (loss) {
    NRF_HIP_CHECK(hipMemcpyAsync(loss, p->dloss, sizeof(double), hipMemcpyDeviceToHost, st));
    NRF_HIP_CHECK(hipStreamSynchronize(st));
  }
  return NRF_OK;
}

extern "C" int nrf_lp_adam(nrf_lp *p, float lr, int step, void *stream) {
  if (!p || step < 1 || !(lr > 0.f)) return NRF_ERR_INVALID_ARGUMENT;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  lp_adam_kernel<<<(OFF.total + 255) / 256, 256, 0, as_stream(stream)>>>(OFF.total, p->params, p->grad, p->am,
                                                                          p->av, lr, b1, b2, 1e-8f, c1, c2);
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}

extern "C" int nrf_lp_fold(nrf_lp *p, nrf_solver *s, void *stream) {
  if (!p || !s || p->variant != 0 || p->N <= 0) return NRF_ERR_INVALID_ARGUMENT;
  if (p->fwd_solver != s) return NRF_ERR_NOT_READY;
  const float *Minv = nullptr;
  int rc = nrf_solver_system(s, nullptr, nullptr, nullptr, nullptr, nullptr, &Minv);
  if (rc) return rc;
  lp_zero_kernel<<<1, 1, 0, as_stream(stream)>>>(p->fallbacks);
  NRF_LAUNCH_CHECK();
  lp_fold_kernel<<<(p->N + 255) / 256, 256, 0, as_stream(stream)>>>(p->N, p->Ld, const_cast<float *>(Minv),
                                                                     p->fallbacks);
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}

extern "C" int nrf_lp_fallbacks(nrf_lp *p, int *count, void *stream) {
  if (!p || !count) return NRF_ERR_INVALID_ARGUMENT;
  int32_t v = 0;
  NRF_HIP_CHECK(hipMemcpyAsync(&v, p->fallbacks, sizeof(v), hipMemcpyDeviceToHost, as_stream(stream)));
  NRF_HIP_CHECK(hipStreamSynchronize(as_stream(stream)));
  *count = v;
  return NRF_OK;
}

extern "C" int nrf_lp_blocks(const nrf_lp *p, const float **C, const float **Td, const float **Te,
                             const float **features) {
  if (!p) return NRF_ERR_INVALID_ARGUMENT;
  if (C) *C = p->C;
  if (Td) *Td = p->Td;
  if (Te) *Te = p->Te;
  if (features) *features = p->x;
  return NRF_OK;
}
