// learned.cpp - FP64 definition of the learned preconditioner.
//
// TEST INFRASTRUCTURE ONLY (linked into liboracle.so with oracle.cpp; never
// into the product library).  The learned preconditioner is a build-only
// component with no counterpart in the reference: this file is its
// definition, and the GPU kernels (cudafusion_amd/csrc/learned.hip) are
// checked against it.  Its own analytic gradient is pinned by central finite
// differences (tests/test_oracle_learned.py).
//
// Network (hidden width H = 64, three message-passing layers), per node i of
// the deformation graph:
//   features x_i (28): position - centroid (3), normal (3), radius / 0.02,
//     log10 of the six diagonal entries of the JtJ diagonal block / 4 (6),
//     the 15 correlations A_ab / sqrt(A_aa A_bb), a > b, of that block;
//   edge features e_ij (4), j = neighbours[s]: (p_j - p_i) / radius_i, w_ij;
//   h0 = relu(W0 x + b0);
//   layer l: z_ij = Wm [h_j; e_ij] + bm, m_i = mean_j relu(z_ij),
//            h_i <- h_i + relu(Wu [h_i; m_i] + bu);
//   heads: Td_i = Wo h_i + bo (21 = lower triangle of a 6x6 block, row major)
//          Te_ij = We [h_i; h_j; e_ij] + be (36) for neighbours j < i
//          (graph-pattern variant only).
// Preconditioner: M^-1 = BJ^-1 + L L^T with L block lower triangular,
//   L_ii = C_i lower(Td_i), L_ij = C_i Te_ij (j < i, j a neighbour of i),
// where C_i C_i^T = BJ_i^-1 (Cholesky of the block-Jacobi inverse), so the
// correction is expressed in the scale of BJ^-1 and M^-1 is SPD by
// construction.  Variant 0 keeps only the diagonal blocks.
// Loss over probe vectors v_p (u_p = A v_p):
//   loss = sum_p |M^-1 u_p - v_p|^2 / (P * 6N).
// Weights are stored input-major: W[c][k] multiplies input c into output k.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../include/nrf.h"

namespace {

constexpr int H = 64, F = 28, EF = 4, K = 8, LAYERS = 3, NTD = 21, NTE = 36;
constexpr int MIN = H + EF, UIN = 2 * H, EIN = 2 * H + EF;

struct Off {
  int W0, b0, Wm[LAYERS], bm[LAYERS], Wu[LAYERS], bu[LAYERS], Wo, bo, We, be, total;
  Off() {
    int o = 0;
    W0 = o; o += F * H;
    b0 = o; o += H;
    for (int l = 0; l < LAYERS; ++l) {
      Wm[l] = o; o += MIN * H;
      bm[l] = o; o += H;
      Wu[l] = o; o += UIN * H;
      bu[l] = o; o += H;
    }
    Wo = o; o += H * NTD;
    bo = o; o += NTD;
    We = o; o += EIN * NTE;
    be = o; o += NTE;
    total = o;
  }
};
const Off OFF;

inline double relu(double v) { return v > 0 ? v : 0; }

// valid neighbour slot of node i (id in range)
inline int nbr(const nrf_node *nodes, int N, int i, int s) {
  if (s >= nodes[i].num_neighbors) return -1;
  const int j = nodes[i].neighbors[s];
  return (j >= 0 && j < N && j != i) ? j : -1;
}

void edge_feat(const nrf_node *nodes, int i, int j, int s, double *e) {
  const double r = std::max((double)nodes[i].radius, 1e-6);
  for (int d = 0; d < 3; ++d) e[d] = ((double)nodes[j].pos[d] - (double)nodes[i].pos[d]) / r;
  e[3] = nodes[i].neighbor_w[s];
}

// lower triangle, row major: (r, c), c <= r -> r (r + 1) / 2 + c
inline int tri(int r, int c) { return r * (r + 1) / 2 + c; }

struct Net {
  int N;
  std::vector<double> h[LAYERS + 1];  // h[0] .. h[3], N x H
  std::vector<double> z0;             // N x H pre-activation of h0
  std::vector<double> zm[LAYERS];     // N x K x H message pre-activations
  std::vector<double> m[LAYERS];      // N x H means
  std::vector<double> zu[LAYERS];     // N x H update pre-activations
  std::vector<int> deg;
};

void forward(const double *P, int variant, const nrf_node *nodes, int N, const double *x, Net &net, double *Td,
             double *Te) {
  net.N = N;
  net.deg.assign(N, 0);
  for (int i = 0; i < N; ++i)
    for (int s = 0; s < K; ++s) net.deg[i] += nbr(nodes, N, i, s) >= 0;
  net.z0.assign((size_t)N * H, 0);
  net.h[0].assign((size_t)N * H, 0);
  for (int i = 0; i < N; ++i)
    for (int k = 0; k < H; ++k) {
      double z = P[OFF.b0 + k];
      for (int c = 0; c < F; ++c) z += P[OFF.W0 + c * H + k] * x[(size_t)i * F + c];
      net.z0[(size_t)i * H + k] = z;
      net.h[0][(size_t)i * H + k] = relu(z);
    }
  for (int l = 0; l < LAYERS; ++l) {
    const std::vector<double> &hp = net.h[l];
    std::vector<double> &hn = net.h[l + 1];
    net.zm[l].assign((size_t)N * K * H, 0);
    net.m[l].assign((size_t)N * H, 0);
    net.zu[l].assign((size_t)N * H, 0);
    hn.assign((size_t)N * H, 0);
    for (int i = 0; i < N; ++i) {
      double in[MIN];
      for (int s = 0; s < K; ++s) {
        const int j = nbr(nodes, N, i, s);
        if (j < 0) continue;
        for (int c = 0; c < H; ++c) in[c] = hp[(size_t)j * H + c];
        edge_feat(nodes, i, j, s, in + H);
        for (int k = 0; k < H; ++k) {
          double z = P[OFF.bm[l] + k];
          for (int c = 0; c < MIN; ++c) z += P[OFF.Wm[l] + c * H + k] * in[c];
          net.zm[l][((size_t)i * K + s) * H + k] = z;
          net.m[l][(size_t)i * H + k] += relu(z);
        }
      }
      const double inv = 1.0 / std::max(1, net.deg[i]);
      for (int k = 0; k < H; ++k) net.m[l][(size_t)i * H + k] *= inv;
      for (int k = 0; k < H; ++k) {
        double z = P[OFF.bu[l] + k];
        for (int c = 0; c < H; ++c) z += P[OFF.Wu[l] + c * H + k] * hp[(size_t)i * H + c];
        for (int c = 0; c < H; ++c) z += P[OFF.Wu[l] + (H + c) * H + k] * net.m[l][(size_t)i * H + c];
        net.zu[l][(size_t)i * H + k] = z;
        hn[(size_t)i * H + k] = hp[(size_t)i * H + k] + relu(z);
      }
    }
  }
  const std::vector<double> &h3 = net.h[LAYERS];
  for (int i = 0; i < N; ++i) {
    for (int o = 0; o < NTD; ++o) {
      double t = P[OFF.bo + o];
      for (int k = 0; k < H; ++k) t += P[OFF.Wo + k * NTD + o] * h3[(size_t)i * H + k];
      Td[(size_t)i * NTD + o] = t;
    }
    for (int s = 0; s < K; ++s) {
      double *te = Te + ((size_t)i * K + s) * NTE;
      const int j = nbr(nodes, N, i, s);
      if (variant == 0 || j < 0 || j >= i) {
        for (int o = 0; o < NTE; ++o) te[o] = 0;
        continue;
      }
      double in[EIN];
      for (int c = 0; c < H; ++c) in[c] = h3[(size_t)i * H + c];
      for (int c = 0; c < H; ++c) in[H + c] = h3[(size_t)j * H + c];
      edge_feat(nodes, i, j, s, in + 2 * H);
      for (int o = 0; o < NTE; ++o) {
        double t = P[OFF.be + o];
        for (int c = 0; c < EIN; ++c) t += P[OFF.We + c * NTE + o] * in[c];
        te[o] = t;
      }
    }
  }
}

// L blocks: Ld_i = C_i lower(Td_i); Le_is = C_i Te_is
void l_blocks(int N, const double *C, const double *Td, const double *Te, double *Ld, double *Le) {
  for (int i = 0; i < N; ++i) {
    const double *Ci = C + (size_t)i * 36;
    double T[36] = {0};
    for (int r = 0; r < 6; ++r)
      for (int c = 0; c <= r; ++c) T[r * 6 + c] = Td[(size_t)i * NTD + tri(r, c)];
    for (int r = 0; r < 6; ++r)
      for (int c = 0; c < 6; ++c) {
        double a = 0;
        for (int q = 0; q < 6; ++q) a += Ci[r * 6 + q] * T[q * 6 + c];
        Ld[(size_t)i * 36 + r * 6 + c] = a;
      }
    for (int s = 0; s < K; ++s) {
      const double *te = Te + ((size_t)i * K + s) * NTE;
      double *le = Le + ((size_t)i * K + s) * 36;
      for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) {
          double a = 0;
          for (int q = 0; q < 6; ++q) a += Ci[r * 6 + q] * te[q * 6 + c];
          le[r * 6 + c] = a;
        }
    }
  }
}

inline bool l_edge(int variant, const nrf_node *nodes, int N, int i, int s, int &j) {
  j = nbr(nodes, N, i, s);
  return variant != 0 && j >= 0 && j < i;
}

// t = L^T u (per column block)
void lt_mul(int variant, const nrf_node *nodes, int N, const double *Ld, const double *Le, const double *u,
            double *t) {
  std::fill(t, t + (size_t)6 * N, 0.0);
  for (int i = 0; i < N; ++i) {
    for (int c = 0; c < 6; ++c)
      for (int r = 0; r < 6; ++r) t[(size_t)i * 6 + c] += Ld[(size_t)i * 36 + r * 6 + c] * u[(size_t)i * 6 + r];
    for (int s = 0; s < K; ++s) {
      int j;
      if (!l_edge(variant, nodes, N, i, s, j)) continue;
      const double *le = Le + ((size_t)i * K + s) * 36;
      for (int c = 0; c < 6; ++c)
        for (int r = 0; r < 6; ++r) t[(size_t)j * 6 + c] += le[r * 6 + c] * u[(size_t)i * 6 + r];
    }
  }
}

// w = L t
void l_mul(int variant, const nrf_node *nodes, int N, const double *Ld, const double *Le, const double *t,
           double *w) {
  for (int i = 0; i < N; ++i) {
    for (int r = 0; r < 6; ++r) {
      double a = 0;
      for (int c = 0; c < 6; ++c) a += Ld[(size_t)i * 36 + r * 6 + c] * t[(size_t)i * 6 + c];
      for (int s = 0; s < K; ++s) {
        int j;
        if (!l_edge(variant, nodes, N, i, s, j)) continue;
        const double *le = Le + ((size_t)i * K + s) * 36;
        for (int c = 0; c < 6; ++c) a += le[r * 6 + c] * t[(size_t)j * 6 + c];
      }
      w[(size_t)i * 6 + r] = a;
    }
  }
}

void apply(int variant, const nrf_node *nodes, int N, const double *Minv, const double *Ld, const double *Le,
           const double *u, double *w) {
  std::vector<double> t((size_t)6 * N);
  lt_mul(variant, nodes, N, Ld, Le, u, t.data());
  l_mul(variant, nodes, N, Ld, Le, t.data(), w);
  for (int i = 0; i < N; ++i)
    for (int r = 0; r < 6; ++r) {
      double a = 0;
      for (int c = 0; c < 6; ++c) a += Minv[(size_t)i * 36 + r * 6 + c] * u[(size_t)i * 6 + c];
      w[(size_t)i * 6 + r] += a;
    }
}

void spmv(const int32_t *rp, const int32_t *ci, const double *A, const double *x, double *y, int N) {
  for (int i = 0; i < N; ++i)
    for (int r = 0; r < 6; ++r) {
      double a = 0;
      for (int q = rp[i]; q < rp[i + 1]; ++q)
        for (int c = 0; c < 6; ++c) a += A[(size_t)q * 36 + r * 6 + c] * x[(size_t)ci[q] * 6 + c];
      y[(size_t)i * 6 + r] = a;
    }
}

}  // namespace

extern "C" {

int orc_lp_param_count() { return OFF.total; }

int orc_lp_offsets(int32_t *out, int n) {
  const int v[] = {OFF.W0, OFF.b0, OFF.Wm[0], OFF.bm[0], OFF.Wu[0], OFF.bu[0], OFF.Wm[1], OFF.bm[1], OFF.Wu[1],
                   OFF.bu[1], OFF.Wm[2], OFF.bm[2], OFF.Wu[2], OFF.bu[2], OFF.Wo, OFF.bo, OFF.We, OFF.be,
                   OFF.total};
  const int cnt = (int)(sizeof(v) / sizeof(v[0]));
  for (int k = 0; k < n && k < cnt; ++k) out[k] = v[k];
  return cnt;
}

// Node features and C_i = chol(Minv_i) (lower, row major; zero when Minv_i
// is not positive definite, which leaves block-Jacobi for that node).
void orc_lp_features_f64(const nrf_node *nodes, int N, const double *Adiag, const double *Minv, double *x,
                         double *C) {
  double cen[3] = {0, 0, 0};
  for (int i = 0; i < N; ++i)
    for (int d = 0; d < 3; ++d) cen[d] += nodes[i].pos[d];
  for (int d = 0; d < 3; ++d) cen[d] /= std::max(1, N);
  for (int i = 0; i < N; ++i) {
    double *f = x + (size_t)i * F;
    const double *A = Adiag + (size_t)i * 36;
    for (int d = 0; d < 3; ++d) f[d] = (double)nodes[i].pos[d] - cen[d];
    for (int d = 0; d < 3; ++d) f[3 + d] = nodes[i].normal[d];
    f[6] = (double)nodes[i].radius / 0.02;
    for (int d = 0; d < 6; ++d) f[7 + d] = std::log10(std::max(A[d * 6 + d], 1e-12)) / 4.0;
    int q = 13;
    for (int a = 1; a < 6; ++a)
      for (int b = 0; b < a; ++b)
        f[q++] = A[a * 6 + b] / std::sqrt(std::max(A[a * 6 + a] * A[b * 6 + b], 1e-30));
    double L[36] = {0};
    bool ok = true;
    const double *M = Minv + (size_t)i * 36;
    for (int r = 0; r < 6 && ok; ++r)
      for (int c = 0; c <= r; ++c) {
        double s = M[r * 6 + c];
        for (int k = 0; k < c; ++k) s -= L[r * 6 + k] * L[c * 6 + k];
        if (r == c) {
          if (!(s > 0)) {
            ok = false;
            break;
          }
          L[r * 6 + r] = std::sqrt(s);
        } else {
          L[r * 6 + c] = s / L[c * 6 + c];
        }
      }
    for (int k = 0; k < 36; ++k) C[(size_t)i * 36 + k] = ok ? L[k] : 0.0;
  }
}

void orc_lp_forward_f64(const double *params, int variant, const nrf_node *nodes, int N, const double *x,
                        double *Td, double *Te) {
  Net net;
  forward(params, variant, nodes, N, x, net, Td, Te);
}

// z = (BJ^-1 + L L^T) r for P vectors stored [P][6N].
void orc_lp_apply_f64(int variant, const nrf_node *nodes, int N, const double *C, const double *Minv,
                      const double *Td, const double *Te, const double *R, int P, double *Z) {
  std::vector<double> Ld((size_t)N * 36), Le((size_t)N * K * 36);
  l_blocks(N, C, Td, Te, Ld.data(), Le.data());
  for (int p = 0; p < P; ++p)
    apply(variant, nodes, N, Minv, Ld.data(), Le.data(), R + (size_t)p * 6 * N, Z + (size_t)p * 6 * N);
}

// Loss over the probes V ([P][6N]) and, when grad != NULL, its gradient with
// respect to every parameter.
double orc_lp_loss_grad_f64(const double *params, int variant, const nrf_node *nodes, int N, const double *x,
                            const double *C, const double *Minv, const int32_t *rp, const int32_t *ci,
                            const double *A, const double *V, int P, double *grad) {
  Net net;
  std::vector<double> Td((size_t)N * NTD), Te((size_t)N * K * NTE);
  forward(params, variant, nodes, N, x, net, Td.data(), Te.data());
  std::vector<double> Ld((size_t)N * 36), Le((size_t)N * K * 36);
  l_blocks(N, C, Td.data(), Te.data(), Ld.data(), Le.data());
  const size_t n = (size_t)6 * N;
  const double scale = 1.0 / ((double)P * (double)n);
  std::vector<double> u(n), t(n), w(n), e(n), s(n);
  std::vector<double> dLd((size_t)N * 36, 0.0), dLe((size_t)N * K * 36, 0.0);
  double loss = 0;
  for (int p = 0; p < P; ++p) {
    const double *v = V + (size_t)p * n;
    spmv(rp, ci, A, v, u.data(), N);
    lt_mul(variant, nodes, N, Ld.data(), Le.data(), u.data(), t.data());
    apply(variant, nodes, N, Minv, Ld.data(), Le.data(), u.data(), w.data());
    for (size_t q = 0; q < n; ++q) {
      e[q] = w[q] - v[q];
      loss += e[q] * e[q];
    }
    if (!grad) continue;
    lt_mul(variant, nodes, N, Ld.data(), Le.data(), e.data(), s.data());
    // d/dL_ik of |e|^2 = 2 (e_i t_k^T + u_i s_k^T)
    for (int i = 0; i < N; ++i) {
      for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c)
          dLd[(size_t)i * 36 + r * 6 + c] +=
              2 * scale * (e[(size_t)i * 6 + r] * t[(size_t)i * 6 + c] + u[(size_t)i * 6 + r] * s[(size_t)i * 6 + c]);
      for (int sl = 0; sl < K; ++sl) {
        int j;
        if (!l_edge(variant, nodes, N, i, sl, j)) continue;
        double *d = dLe.data() + ((size_t)i * K + sl) * 36;
        for (int r = 0; r < 6; ++r)
          for (int c = 0; c < 6; ++c)
            d[r * 6 + c] += 2 * scale *
                            (e[(size_t)i * 6 + r] * t[(size_t)j * 6 + c] + u[(size_t)i * 6 + r] * s[(size_t)j * 6 + c]);
      }
    }
  }
  loss *= scale;
  if (!grad) return loss;
  const double *Pm = params;
  std::memset(grad, 0, sizeof(double) * OFF.total);
  // heads: dT = C^T dL (diagonal: lower triangle only)
  std::vector<double> dTd((size_t)N * NTD), dTe((size_t)N * K * NTE, 0.0);
  for (int i = 0; i < N; ++i) {
    const double *Ci = C + (size_t)i * 36;
    double g[36];
    for (int r = 0; r < 6; ++r)
      for (int c = 0; c < 6; ++c) {
        double a = 0;
        for (int q = 0; q < 6; ++q) a += Ci[q * 6 + r] * dLd[(size_t)i * 36 + q * 6 + c];
        g[r * 6 + c] = a;
      }
    for (int r = 0; r < 6; ++r)
      for (int c = 0; c <= r; ++c) dTd[(size_t)i * NTD + tri(r, c)] = g[r * 6 + c];
    for (int sl = 0; sl < K; ++sl) {
      int j;
      if (!l_edge(variant, nodes, N, i, sl, j)) continue;
      const double *d = dLe.data() + ((size_t)i * K + sl) * 36;
      for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) {
          double a = 0;
          for (int q = 0; q < 6; ++q) a += Ci[q * 6 + r] * d[q * 6 + c];
          dTe[((size_t)i * K + sl) * NTE + r * 6 + c] = a;
        }
    }
  }
  const std::vector<double> &h3 = net.h[LAYERS];
  std::vector<double> dh((size_t)N * H, 0.0);
  for (int i = 0; i < N; ++i) {
    for (int o = 0; o < NTD; ++o) {
      const double g = dTd[(size_t)i * NTD + o];
      grad[OFF.bo + o] += g;
      for (int k = 0; k < H; ++k) {
        grad[OFF.Wo + k * NTD + o] += h3[(size_t)i * H + k] * g;
        dh[(size_t)i * H + k] += Pm[OFF.Wo + k * NTD + o] * g;
      }
    }
    for (int sl = 0; sl < K; ++sl) {
      int j;
      if (!l_edge(variant, nodes, N, i, sl, j)) continue;
      double in[EIN];
      for (int c = 0; c < H; ++c) in[c] = h3[(size_t)i * H + c];
      for (int c = 0; c < H; ++c) in[H + c] = h3[(size_t)j * H + c];
      edge_feat(nodes, i, j, sl, in + 2 * H);
      for (int o = 0; o < NTE; ++o) {
        const double g = dTe[((size_t)i * K + sl) * NTE + o];
        grad[OFF.be + o] += g;
        for (int c = 0; c < EIN; ++c) grad[OFF.We + c * NTE + o] += in[c] * g;
        for (int c = 0; c < H; ++c) {
          dh[(size_t)i * H + c] += Pm[OFF.We + c * NTE + o] * g;
          dh[(size_t)j * H + c] += Pm[OFF.We + (H + c) * NTE + o] * g;
        }
      }
    }
  }
  for (int l = LAYERS - 1; l >= 0; --l) {
    const std::vector<double> &hp = net.h[l];
    std::vector<double> dhp(dh);  // residual path
    for (int i = 0; i < N; ++i) {
      double dzu[H], dm[H];
      for (int k = 0; k < H; ++k) dzu[k] = net.zu[l][(size_t)i * H + k] > 0 ? dh[(size_t)i * H + k] : 0.0;
      for (int c = 0; c < H; ++c) dm[c] = 0;
      for (int k = 0; k < H; ++k) {
        grad[OFF.bu[l] + k] += dzu[k];
        for (int c = 0; c < H; ++c) {
          grad[OFF.Wu[l] + c * H + k] += hp[(size_t)i * H + c] * dzu[k];
          grad[OFF.Wu[l] + (H + c) * H + k] += net.m[l][(size_t)i * H + c] * dzu[k];
          dhp[(size_t)i * H + c] += Pm[OFF.Wu[l] + c * H + k] * dzu[k];
          dm[c] += Pm[OFF.Wu[l] + (H + c) * H + k] * dzu[k];
        }
      }
      const double inv = 1.0 / std::max(1, net.deg[i]);
      for (int sl = 0; sl < K; ++sl) {
        const int j = nbr(nodes, N, i, sl);
        if (j < 0) continue;
        double in[MIN];
        for (int c = 0; c < H; ++c) in[c] = hp[(size_t)j * H + c];
        edge_feat(nodes, i, j, sl, in + H);
        for (int k = 0; k < H; ++k) {
          const double dz = net.zm[l][((size_t)i * K + sl) * H + k] > 0 ? dm[k] * inv : 0.0;
          if (dz == 0.0) continue;
          grad[OFF.bm[l] + k] += dz;
          for (int c = 0; c < MIN; ++c) grad[OFF.Wm[l] + c * H + k] += in[c] * dz;
          for (int c = 0; c < H; ++c) dhp[(size_t)j * H + c] += Pm[OFF.Wm[l] + c * H + k] * dz;
        }
      }
    }
    dh.swap(dhp);
  }
  for (int i = 0; i < N; ++i)
    for (int k = 0; k < H; ++k) {
      const double dz = net.z0[(size_t)i * H + k] > 0 ? dh[(size_t)i * H + k] : 0.0;
      grad[OFF.b0 + k] += dz;
      for (int c = 0; c < F; ++c) grad[OFF.W0 + c * H + k] += x[(size_t)i * F + c] * dz;
    }
  return loss;
}

// PCG (FP64) with M^-1 = BJ^-1 + L L^T, x0 = 0, stopping at
// |r| < tol |b| (p.Ap <= 0 is a breakdown).  Returns the iterations run.
int orc_lp_pcg_f64(int variant, const nrf_node *nodes, int N, const int32_t *rp, const int32_t *ci, const double *A,
                   const double *b, const double *C, const double *Minv, const double *Td, const double *Te,
                   int maxit, double tol, double *x, double *true_rel) {
  const size_t n = (size_t)6 * N;
  std::vector<double> Ld((size_t)N * 36), Le((size_t)N * K * 36);
  l_blocks(N, C, Td, Te, Ld.data(), Le.data());
  std::vector<double> r(b, b + n), z(n), p(n), Ap(n);
  std::fill(x, x + n, 0.0);
  apply(variant, nodes, N, Minv, Ld.data(), Le.data(), r.data(), z.data());
  p = z;
  double rz = 0, bb = 0;
  for (size_t q = 0; q < n; ++q) {
    rz += r[q] * z[q];
    bb += b[q] * b[q];
  }
  const double tl = tol * std::sqrt(bb);
  int it = 0;
  for (; it < maxit; ++it) {
    spmv(rp, ci, A, p.data(), Ap.data(), N);
    double pAp = 0;
    for (size_t q = 0; q < n; ++q) pAp += p[q] * Ap[q];
    if (!(pAp > 0)) break;
    const double al = rz / pAp;
    double rr = 0;
    for (size_t q = 0; q < n; ++q) {
      x[q] += al * p[q];
      r[q] -= al * Ap[q];
      rr += r[q] * r[q];
    }
    if (std::sqrt(rr) < tl) {
      ++it;
      break;
    }
    apply(variant, nodes, N, Minv, Ld.data(), Le.data(), r.data(), z.data());
    double rzn = 0;
    for (size_t q = 0; q < n; ++q) rzn += r[q] * z[q];
    const double be = rzn / rz;
    rz = rzn;
    for (size_t q = 0; q < n; ++q) p[q] = z[q] + be * p[q];
  }
  if (true_rel) {
    std::vector<double> ax(n);
    spmv(rp, ci, A, x, ax.data(), N);
    double rr = 0;
    for (size_t q = 0; q < n; ++q) rr += (b[q] - ax[q]) * (b[q] - ax[q]);
    *true_rel = std::sqrt(rr) / std::max(std::sqrt(bb), 1e-300);
  }
  return it;
}

}  // extern "C"
