// graph.hip - deformation-graph maintenance: node sampling on the surface,
// geodesic graph edges and pruning (the graph-update frames of the
// reference, src/warp_field.cu:407-472, 488-694, 709-954, 986-1104).
//
// Host parts (sequential by definition in the reference): the greedy node
// sampling over a spatial hash, the welded surface graph of the marching
// cubes soup and each node's nearest graph vertex, the neighbour filters and
// weights, and the connected components of the pruning.  Device parts:
//
// * the bounded geodesic distances: one workgroup per source node with the
//   source's local vertex set in LDS (an open-addressing table from graph
//   vertex to distance, 12288 slots), relaxed Bellman-Ford style from the
//   vertices whose distance fell in the previous round until no distance
//   changes.  The result is the fixed point of d(v) = min(d(v), d(u) + w_uv)
//   under the max-distance cut, which does not depend on the relaxation
//   order (FP32 addition is monotone), so it is deterministic and equals a
//   bounded Dijkstra's; the reference runs global-memory relaxations of 32
//   sources at a time with a host round trip per round and stops after 96
//   rounds, which equals it whenever 96 rounds converge.  The K_GRAPH
//   nearest other nodes by geodesic distance (ties to the lower node index)
//   are picked in the same workgroup.
// * the TSDF support score of every node (one wave per node over its voxel
//   ball, integer counts, so order independent).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "nrf_internal.h"
#include "nrf_math.h"

namespace nrf {
namespace {

constexpr float kMinGraphNormalDot = 0.70f;  // cos(45 deg), src/warp_field.cu:736-737

// ------------------------------------------------------------ host helpers
struct CellKey {
  int x, y, z;
  bool operator==(const CellKey &o) const { return x == o.x && y == o.y && z == o.z; }
};
struct CellKeyHash {
  size_t operator()(const CellKey &k) const {
    uint64_t h = (uint64_t)(uint32_t)k.x * 0x9E3779B97F4A7C15ull;
    h ^= (uint64_t)(uint32_t)k.y * 0xC2B2AE3D27D4EB4Full + (h >> 29);
    h ^= (uint64_t)(uint32_t)k.z * 0x165667B19E3779F9ull + (h >> 31);
    return (size_t)h;
  }
};
using Grid = std::unordered_map<CellKey, std::vector<int>, CellKeyHash>;

inline float h_dot(const float *a, const float *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline float h_norm(const float *a) { return sqrtf(h_dot(a, a)); }
inline CellKey cell_of(const float *p, float inv) {
  return {(int)floorf(p[0] * inv), (int)floorf(p[1] * inv), (int)floorf(p[2] * inv)};
}
inline float dist2(const float *a, const float *b) {
  const float d[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
  return d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
}

DQ h_identity() {
  DQ q;
  q.r = make_float4(0.f, 0.f, 0.f, 1.f);
  q.d = make_float4(0.f, 0.f, 0.f, 0.f);
  return q;
}
// weighted blend with the running-sign fix of the reference's transform
// initialisation (src/warp_field.cu:770-808, 895-946)
struct DqAccum {
  DQ acc;
  float wsum = 0.f;
  DqAccum() {
    acc.r = make_float4(0.f, 0.f, 0.f, 0.f);
    acc.d = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  void add(DQ t, float w) {
    if (qdot(t.r, acc.r) < 0.0f) {
      t.r = make_float4(-t.r.x, -t.r.y, -t.r.z, -t.r.w);
      t.d = make_float4(-t.d.x, -t.d.y, -t.d.z, -t.d.w);
    }
    acc.r.x += w * t.r.x; acc.r.y += w * t.r.y; acc.r.z += w * t.r.z; acc.r.w += w * t.r.w;
    acc.d.x += w * t.d.x; acc.d.y += w * t.d.y; acc.d.z += w * t.d.z; acc.d.w += w * t.d.w;
    wsum += w;
  }
};

// add_nodes_from_surface (src/warp_field.cu:709-841) on host arrays
int sample_nodes(nrf_node *nodes, nrf_dualquat *T, int &n, int max_nodes, const float *verts, const float *normals,
                 int nv, float min_dist, float node_radius) {
  const float inv_cell = 1.f / min_dist;
  const float min_dist2 = min_dist * min_dist;
  const float graph_dist = fmaxf(2.5f * min_dist, node_radius);
  const float graph_dist2 = graph_dist * graph_dist;
  Grid grid;
  grid.reserve((size_t)n * 2 + 16);
  for (int i = 0; i < n; ++i) grid[cell_of(nodes[i].pos, inv_cell)].push_back(i);
  const int first_new = n;
  int added = 0;
  std::vector<std::pair<float, int>> near;
  for (int vi = 0; vi < nv; ++vi) {
    const float *v = verts + 3 * (size_t)vi;
    if (n >= max_nodes) break;
    const CellKey c = cell_of(v, inv_cell);
    bool close = false;
    for (int dx = -1; dx <= 1 && !close; ++dx)
      for (int dy = -1; dy <= 1 && !close; ++dy)
        for (int dz = -1; dz <= 1 && !close; ++dz) {
          auto it = grid.find({c.x + dx, c.y + dy, c.z + dz});
          if (it == grid.end()) continue;
          for (int idx : it->second)
            if (dist2(v, nodes[idx].pos) < min_dist2) {
              close = true;
              break;
            }
        }
    if (close) continue;
    nrf_node node;
    std::memset(&node, 0, sizeof(node));
    for (int k = 0; k < 3; ++k) node.pos[k] = v[k];
    if (normals) {
      const float *nn = normals + 3 * (size_t)vi;
      const float ln = h_norm(nn);
      if (ln > 1e-6f)
        for (int k = 0; k < 3; ++k) node.normal[k] = nn[k] / ln;
    }
    node.radius = node_radius;
    for (int k = 0; k < NRF_K_GRAPH; ++k) node.neighbors[k] = -1;
    near.clear();
    for (int dx = -3; dx <= 3; ++dx)
      for (int dy = -3; dy <= 3; ++dy)
        for (int dz = -3; dz <= 3; ++dz) {
          auto it = grid.find({c.x + dx, c.y + dy, c.z + dz});
          if (it == grid.end()) continue;
          for (int idx : it->second) {
            const float d2 = dist2(v, nodes[idx].pos);
            if (d2 < graph_dist2) near.push_back({d2, idx});
          }
        }
    if (h_norm(node.normal) > 1e-6f)
      near.erase(std::remove_if(near.begin(), near.end(),
                                [&](const std::pair<float, int> &e) {
                                  const float *nb = nodes[e.second].normal;
                                  return h_norm(nb) > 1e-6f && h_dot(node.normal, nb) < kMinGraphNormalDot;
                                }),
                 near.end());
    std::sort(near.begin(), near.end());
    const int k = std::min((int)near.size(), (int)NRF_K_GRAPH);
    for (int i = 0; i < k; ++i) {
      node.neighbors[i] = near[i].second;
      node.neighbor_w[i] = 1.f;
    }
    node.num_neighbors = k;
    DQ t_init = h_identity();
    if (k > 0) {
      DqAccum a;
      for (int i = 0; i < k; ++i) {
        const int nb = near[i].second;
        if (nb < 0 || nb >= first_new) continue;
        a.add(load_dq(T[nb]), 1.f / (sqrtf(near[i].first) + 1e-6f));
      }
      if (a.wsum > 0) t_init = dq_normalize(a.acc);
    }
    nodes[n] = node;
    store_dq(T[n], t_init);
    grid[c].push_back(n);
    ++n;
    ++added;
  }
  return added;
}

// build_surface_graph (src/warp_field.cu:145-198): the soup's vertices
// welded on a 0.05 min_dist lattice (round to nearest cell), edges of every
// triangle (both directions, weight = length, zero-length skipped)
struct SurfaceGraph {
  std::vector<float> pos;  // 3 per graph vertex
  std::vector<int32_t> row_ptr, col;
  std::vector<float> w;
  Grid spatial;
  float inv_cell = 1.f;
  int count() const { return (int)(pos.size() / 3); }
};

SurfaceGraph build_graph(const float *verts, int nv, const int32_t *tris, int nt, float min_dist) {
  SurfaceGraph g;
  std::vector<int> rep(nv, -1);
  const float weld = fmaxf(1e-5f, 0.05f * min_dist);
  const float inv_weld = 1.0f / weld;
  std::unordered_map<CellKey, int, CellKeyHash> reps;
  reps.reserve((size_t)nv);
  for (int i = 0; i < nv; ++i) {
    const float *p = verts + 3 * (size_t)i;
    const CellKey key{(int)floorf(p[0] * inv_weld + 0.5f), (int)floorf(p[1] * inv_weld + 0.5f),
                      (int)floorf(p[2] * inv_weld + 0.5f)};
    auto it = reps.find(key);
    if (it == reps.end()) {
      const int r = g.count();
      reps.emplace(key, r);
      g.pos.insert(g.pos.end(), p, p + 3);
      rep[i] = r;
    } else {
      rep[i] = it->second;
    }
  }
  std::vector<std::vector<std::pair<int, float>>> adj(g.count());
  auto edge = [&](int a, int b) {
    if (a == b) return;
    const float d = sqrtf(dist2(&g.pos[3 * (size_t)a], &g.pos[3 * (size_t)b]));
    if (d <= 1e-8f) return;
    adj[a].push_back({b, d});
    adj[b].push_back({a, d});
  };
  for (int t = 0; t < nt; ++t) {
    const int32_t *tr = tris + 3 * (size_t)t;
    if (tr[0] < 0 || tr[1] < 0 || tr[2] < 0 || tr[0] >= nv || tr[1] >= nv || tr[2] >= nv) continue;
    const int a = rep[tr[0]], b = rep[tr[1]], c = rep[tr[2]];
    edge(a, b);
    edge(b, c);
    edge(c, a);
  }
  g.row_ptr.resize(g.count() + 1);
  int e = 0;
  for (int v = 0; v < g.count(); ++v) {
    g.row_ptr[v] = e;
    e += (int)adj[v].size();
  }
  g.row_ptr[g.count()] = e;
  g.col.reserve(e);
  g.w.reserve(e);
  for (const auto &a : adj)
    for (const auto &x : a) {
      g.col.push_back(x.first);
      g.w.push_back(x.second);
    }
  g.inv_cell = 1.0f / min_dist;
  for (int i = 0; i < g.count(); ++i) g.spatial[cell_of(&g.pos[3 * (size_t)i], g.inv_cell)].push_back(i);
  return g;
}

// nearest_graph_vertex (src/warp_field.cu:200-244): the nearest vertex of
// the first cube of cells (radius 1..4) holding any, else of all
int nearest_vertex(const SurfaceGraph &g, const float *p) {
  if (g.count() == 0) return -1;
  const CellKey c = cell_of(p, g.inv_cell);
  float best_d2 = INFINITY;
  int best = -1;
  for (int r = 1; r <= 4 && best < 0; ++r)
    for (int dx = -r; dx <= r; ++dx)
      for (int dy = -r; dy <= r; ++dy)
        for (int dz = -r; dz <= r; ++dz) {
          auto it = g.spatial.find({c.x + dx, c.y + dy, c.z + dz});
          if (it == g.spatial.end()) continue;
          for (int idx : it->second) {
            const float d2 = dist2(p, &g.pos[3 * (size_t)idx]);
            if (d2 < best_d2) {
              best_d2 = d2;
              best = idx;
            }
          }
        }
  if (best >= 0) return best;
  for (int i = 0; i < g.count(); ++i) {
    const float d2 = dist2(p, &g.pos[3 * (size_t)i]);
    if (d2 < best_d2) {
      best_d2 = d2;
      best = i;
    }
  }
  return best;
}

// ------------------------------------------------------------ device part
constexpr int kGeoCap = 12288;   // LDS distance-table slots per source
constexpr int kGeoCand = 1024;   // candidate neighbour nodes per source
constexpr int kGeoThreads = 256;
constexpr int kGeoMaxRounds = 1 << 20;

struct GeoArgs {
  const int32_t *row_ptr, *col;
  const float *w;
  const int32_t *reps;
  int N;
  float max_dist;
  int32_t *out_ids;
  float *out_d;
  int32_t *status;  // bit 0: a distance table overflowed, bit 1: candidates overflowed, bit 2: round cap
};

__device__ inline uint32_t slot_hash(int key) { return ((uint32_t)key * 2654435761u) % (uint32_t)kGeoCap; }
// the table slot of key (inserted when absent and insert), -1 when full / absent
__device__ inline int geo_slot(int32_t *keys, int key, bool insert) {
  uint32_t h = slot_hash(key);
  for (int probe = 0; probe < kGeoCap; ++probe) {
    const int k = keys[h];
    if (k == key) return (int)h;
    if (k == -1) {
      if (!insert) return -1;
      const int old = atomicCAS(&keys[h], -1, key);
      if (old == -1 || old == key) return (int)h;
    }
    h = h + 1 == (uint32_t)kGeoCap ? 0u : h + 1;
  }
  return -1;
}

__global__ __launch_bounds__(kGeoThreads) void geodesic_kernel(GeoArgs a) {
  extern __shared__ __align__(16) unsigned char lds[];
  int32_t *keys = reinterpret_cast<int32_t *>(lds);
  uint32_t *dist = reinterpret_cast<uint32_t *>(keys + kGeoCap);
  uint8_t *flag0 = reinterpret_cast<uint8_t *>(dist + kGeoCap);
  uint8_t *flag1 = flag0 + kGeoCap;
  float *cand_d = reinterpret_cast<float *>(flag1 + kGeoCap);
  int32_t *cand_id = reinterpret_cast<int32_t *>(cand_d + kGeoCand);
  __shared__ int s_changed, s_status, s_ncand;
  const int s = blockIdx.x, t = threadIdx.x;
  const int src = a.reps[s];
  for (int i = t; i < kGeoCap; i += kGeoThreads) {
    keys[i] = -1;
    dist[i] = 0x7f800000u;  // +inf
    flag0[i] = flag1[i] = 0;
  }
  if (t == 0) {
    s_status = 0;
    s_ncand = 0;
  }
  __syncthreads();
  if (src >= 0 && t == 0) {
    const int sl = geo_slot(keys, src, true);
    dist[sl] = 0u;
    flag0[sl] = 1;
  }
  uint8_t *cur = flag0, *nxt = flag1;
  int rounds = 0;
  for (;;) {
    if (t == 0) s_changed = 0;
    __syncthreads();
    for (int i = t; i < kGeoCap; i += kGeoThreads) {
      if (!cur[i]) continue;
      cur[i] = 0;
      const int u = keys[i];
      const float du = __uint_as_float(dist[i]);
      if (!(du <= a.max_dist)) continue;
      for (int e = a.row_ptr[u]; e < a.row_ptr[u + 1]; ++e) {
        const float nd = du + a.w[e];
        if (!(nd <= a.max_dist)) continue;
        const int sv = geo_slot(keys, a.col[e], true);
        if (sv < 0) {
          atomicOr(&s_status, 1);
          continue;
        }
        const uint32_t old = atomicMin(&dist[sv], __float_as_uint(nd));
        if (__float_as_uint(nd) < old) {
          nxt[sv] = 1;
          s_changed = 1;
        }
      }
    }
    __syncthreads();
    const bool more = s_changed != 0;
    __syncthreads();
    if (!more) break;
    if (++rounds >= kGeoMaxRounds) {
      if (t == 0) atomicOr(&s_status, 4);
      break;
    }
    uint8_t *tmp = cur;
    cur = nxt;
    nxt = tmp;
  }
  // select_geodesic_neighbors_kernel (src/warp_field.cu:540-590): the other
  // nodes whose representative vertex lies within (0, max_dist]
  for (int id = t; id < a.N; id += kGeoThreads) {
    if (id == s) continue;
    const int rep = a.reps[id];
    if (rep < 0) continue;
    const int sl = geo_slot(keys, rep, false);
    if (sl < 0) continue;
    const float gd = __uint_as_float(dist[sl]);
    if (!isfinite(gd) || gd <= 0.0f || gd > a.max_dist) continue;
    const int c = atomicAdd(&s_ncand, 1);
    if (c < kGeoCand) {
      cand_d[c] = gd;
      cand_id[c] = id;
    } else {
      atomicOr(&s_status, 2);
    }
  }
  __syncthreads();
  if (t == 0) {
    float bd[NRF_K_GRAPH];
    int bi[NRF_K_GRAPH];
    for (int k = 0; k < NRF_K_GRAPH; ++k) {
      bd[k] = 1e30f;
      bi[k] = -1;
    }
    const int nc = min(s_ncand, kGeoCand);
    for (int c = 0; c < nc; ++c) {  // smallest (distance, node id) first
      const float gd = cand_d[c];
      const int id = cand_id[c];
      int k = NRF_K_GRAPH - 1;
      if (gd > bd[k] || (gd == bd[k] && (bi[k] >= 0 && id > bi[k]))) continue;
      bd[k] = gd;
      bi[k] = id;
      for (; k > 0; --k) {
        if (bd[k - 1] < bd[k] || (bd[k - 1] == bd[k] && bi[k - 1] >= 0 && bi[k - 1] < bi[k])) break;
        const float td = bd[k - 1];
        bd[k - 1] = bd[k];
        bd[k] = td;
        const int ti = bi[k - 1];
        bi[k - 1] = bi[k];
        bi[k] = ti;
      }
    }
    for (int k = 0; k < NRF_K_GRAPH; ++k) {
      a.out_ids[(size_t)s * NRF_K_GRAPH + k] = bi[k];
      a.out_d[(size_t)s * NRF_K_GRAPH + k] = bd[k];
    }
    if (s_status) atomicOr(a.status, s_status);
  }
}

size_t geo_lds_bytes() {
  return (size_t)kGeoCap * (4 + 4 + 1 + 1) + (size_t)kGeoCand * (4 + 4);
}

// score_node_tsdf_support_kernel (src/warp_field.cu:407-472): per node, the
// observed / near-surface / empty voxels of its ball; one wave per node.
struct PruneArgs {
  const nrf_voxel *vox;
  int X, Y, Z;
  float ox, oy, oz, s;
  const nrf_node *nodes;
  int N;
  nrf_prune_params pp;
  uint8_t *keep;
};

__global__ __launch_bounds__(256) void prune_score_kernel(PruneArgs a) {
  const int lane = threadIdx.x & 63;
  const int ni = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (ni >= a.N) return;
  const nrf_node &node = a.nodes[ni];
  const float radius = fmaxf(node.radius * a.pp.support_radius_factor, a.s);
  const float radius2 = radius * radius;
  const int cx = (int)floorf((node.pos[0] - a.ox) / a.s), cy = (int)floorf((node.pos[1] - a.oy) / a.s),
            cz = (int)floorf((node.pos[2] - a.oz) / a.s);
  const int r = (int)ceilf(radius / a.s);
  const int x0 = max(cx - r, 0), y0 = max(cy - r, 0), z0 = max(cz - r, 0);
  const int x1 = min(cx + r, a.X - 1), y1 = min(cy + r, a.Y - 1), z1 = min(cz + r, a.Z - 1);
  int observed = 0, surface = 0, empty = 0;
  if (x0 <= x1 && y0 <= y1 && z0 <= z1) {
    const int nx = x1 - x0 + 1, ny = y1 - y0 + 1, nz = z1 - z0 + 1;
    const int total = nx * ny * nz;
    for (int q = lane; q < total; q += 64) {
      const int x = x0 + q % nx, y = y0 + (q / nx) % ny, z = z0 + q / (nx * ny);
      const float px = a.ox + (x + 0.5f) * a.s, py = a.oy + (y + 0.5f) * a.s, pz = a.oz + (z + 0.5f) * a.s;
      const float dx = px - node.pos[0], dy = py - node.pos[1], dz = pz - node.pos[2];
      if (dx * dx + dy * dy + dz * dz > radius2) continue;
      const nrf_voxel &v = a.vox[((size_t)z * a.Y + y) * a.X + x];
      if (v.weight <= 0.0f) continue;
      ++observed;
      if (fabsf(v.tsdf) <= a.pp.surface_tsdf_abs) ++surface;
      if (v.tsdf >= a.pp.empty_tsdf) ++empty;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    observed += __shfl_xor(observed, off, 64);
    surface += __shfl_xor(surface, off, 64);
    empty += __shfl_xor(empty, off, 64);
  }
  if (lane == 0) {
    const bool unknown = observed < a.pp.min_observed_voxels;
    const bool supported = surface >= a.pp.min_surface_voxels;
    const bool confidently_empty = empty >= a.pp.min_observed_voxels && surface == 0;
    a.keep[ni] = (unknown || supported || !confidently_empty) ? 1 : 0;
  }
}

template <typename T>
struct DevBuf {
  T *p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t n) {
    NRF_HIP_CHECK(hipMalloc(reinterpret_cast<void **>(&p), std::max<size_t>(n, 1) * sizeof(T)));
    return NRF_OK;
  }
};

}  // namespace
}  // namespace nrf

using namespace nrf;

extern "C" int nrf_graph_add_nodes(nrf_node *nodes, nrf_dualquat *transforms, int *num_nodes, int max_nodes,
                                   const float *vertices, const float *normals, int num_vertices, float min_dist,
                                   float node_radius, int *added) {
  if (!nodes || !transforms || !num_nodes || *num_nodes < 0 || *num_nodes > max_nodes || num_vertices < 0 ||
      (num_vertices > 0 && !vertices) || !(min_dist > 0.f))
    return NRF_ERR_INVALID_ARGUMENT;
  int n = *num_nodes;
  const int a = sample_nodes(nodes, transforms, n, max_nodes, vertices, normals, num_vertices, min_dist, node_radius);
  *num_nodes = n;
  if (added) *added = a;
  return NRF_OK;
}

extern "C" int nrf_graph_add_nodes_geodesic(nrf_node *nodes, nrf_dualquat *transforms, int *num_nodes,
                                            int max_nodes, const float *vertices, const float *normals,
                                            int num_vertices, const int32_t *triangles, int num_triangles,
                                            float min_dist, float node_radius, int *added, void *stream) {
  if (!nodes || !transforms || !num_nodes || *num_nodes < 0 || *num_nodes > max_nodes || num_vertices < 0 ||
      num_triangles < 0 || (num_vertices > 0 && !vertices) || (num_triangles > 0 && !triangles) ||
      !(min_dist > 0.f))
    return NRF_ERR_INVALID_ARGUMENT;
  int n = *num_nodes;
  const int a = sample_nodes(nodes, transforms, n, max_nodes, vertices, normals, num_vertices, min_dist, node_radius);
  *num_nodes = n;
  if (added) *added = a;
  // src/warp_field.cu:849-851
  if (a <= 0 || n == 0 || num_vertices == 0 || num_triangles == 0) return NRF_OK;
  SurfaceGraph g = build_graph(vertices, num_vertices, triangles, num_triangles, min_dist);
  if (g.count() == 0) return NRF_OK;
  std::vector<int32_t> reps(n);
  for (int i = 0; i < n; ++i) reps[i] = nearest_vertex(g, nodes[i].pos);
  const float max_dist = fmaxf(3.5f * min_dist, 1.5f * node_radius);
  hipStream_t st = as_stream(stream);
  DevBuf<int32_t> d_rp, d_col, d_reps, d_ids, d_status;
  DevBuf<float> d_w, d_d;
  int rc;
  if ((rc = d_rp.alloc(g.row_ptr.size())) || (rc = d_col.alloc(g.col.size())) || (rc = d_w.alloc(g.w.size())) ||
      (rc = d_reps.alloc(n)) || (rc = d_ids.alloc((size_t)n * NRF_K_GRAPH)) ||
      (rc = d_d.alloc((size_t)n * NRF_K_GRAPH)) || (rc = d_status.alloc(1)))
    return rc;
  NRF_HIP_CHECK(hipMemcpyAsync(d_rp.p, g.row_ptr.data(), g.row_ptr.size() * 4, hipMemcpyHostToDevice, st));
  if (!g.col.empty()) {
    NRF_HIP_CHECK(hipMemcpyAsync(d_col.p, g.col.data(), g.col.size() * 4, hipMemcpyHostToDevice, st));
    NRF_HIP_CHECK(hipMemcpyAsync(d_w.p, g.w.data(), g.w.size() * 4, hipMemcpyHostToDevice, st));
  }
  NRF_HIP_CHECK(hipMemcpyAsync(d_reps.p, reps.data(), (size_t)n * 4, hipMemcpyHostToDevice, st));
  NRF_HIP_CHECK(hipMemsetAsync(d_status.p, 0, 4, st));
  GeoArgs ga{d_rp.p, d_col.p, d_w.p, d_reps.p, n, max_dist, d_ids.p, d_d.p, d_status.p};
  const size_t lds = geo_lds_bytes();
  NRF_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(geodesic_kernel),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  geodesic_kernel<<<n, kGeoThreads, lds, st>>>(ga);
  NRF_LAUNCH_CHECK();
  std::vector<int32_t> ids((size_t)n * NRF_K_GRAPH);
  std::vector<float> gds((size_t)n * NRF_K_GRAPH);
  int32_t status = 0;
  NRF_HIP_CHECK(hipMemcpyAsync(ids.data(), d_ids.p, ids.size() * 4, hipMemcpyDeviceToHost, st));
  NRF_HIP_CHECK(hipMemcpyAsync(gds.data(), d_d.p, gds.size() * 4, hipMemcpyDeviceToHost, st));
  NRF_HIP_CHECK(hipMemcpyAsync(&status, d_status.p, 4, hipMemcpyDeviceToHost, st));
  NRF_HIP_CHECK(hipStreamSynchronize(st));
  if (status) {
    set_last_error_msg(std::string("geodesic graph: ") + ((status & 1) ? "a local distance table overflowed; " : "") +
                       ((status & 2) ? "too many candidate neighbours; " : "") +
                       ((status & 4) ? "relaxation round cap reached" : ""));
    return NRF_ERR_TOO_LARGE;
  }
  // edges from the geodesic neighbours (src/warp_field.cu:866-892)
  for (int ni = 0; ni < n; ++ni) {
    nrf_node &node = nodes[ni];
    node.num_neighbors = 0;
    for (int k = 0; k < NRF_K_GRAPH; ++k) {
      node.neighbors[k] = -1;
      node.neighbor_w[k] = 0.f;
    }
    for (int k = 0; k < NRF_K_GRAPH; ++k) {
      const int nb = ids[(size_t)ni * NRF_K_GRAPH + k];
      const float gd = gds[(size_t)ni * NRF_K_GRAPH + k];
      if (nb < 0 || !std::isfinite(gd)) continue;
      if (nb >= n || nb == ni) continue;
      if (h_norm(node.normal) > 1e-6f && h_norm(nodes[nb].normal) > 1e-6f &&
          h_dot(node.normal, nodes[nb].normal) < kMinGraphNormalDot)
        continue;
      const int kk = node.num_neighbors;
      if (kk >= NRF_K_GRAPH) break;
      node.neighbors[kk] = nb;
      const float r = fmaxf(node.radius, 1e-6f);
      node.neighbor_w[kk] = expf(-(gd * gd) / (2.0f * r * r));
      node.num_neighbors++;
    }
  }
  // transforms of the new nodes from their pre-existing geodesic
  // neighbours (src/warp_field.cu:893-946)
  const int first_new = n - a;
  for (int ni = first_new; ni < n; ++ni) {
    DqAccum acc;
    int used = 0;
    for (int k = 0; k < NRF_K_GRAPH; ++k) {
      const int nb = ids[(size_t)ni * NRF_K_GRAPH + k];
      const float gd = gds[(size_t)ni * NRF_K_GRAPH + k];
      if (nb < 0 || !std::isfinite(gd)) continue;
      if (nb >= first_new) continue;
      if (used >= NRF_K_GRAPH) break;
      acc.add(load_dq(transforms[nb]), 1.f / (sqrtf(gd) + 1e-6f));
      ++used;
    }
    if (acc.wsum > 0.0f) store_dq(transforms[ni], dq_normalize(acc.acc));
  }
  return NRF_OK;
}

extern "C" int nrf_graph_prune(const nrf_volume_params *vp, const nrf_voxel *voxels, nrf_node *nodes,
                               nrf_dualquat *transforms, int *num_nodes, const nrf_prune_params *pp, int *removed,
                               void *stream) {
  if (!volume_ok(vp) || !voxels || !nodes || !transforms || !num_nodes || *num_nodes < 0 || !pp)
    return NRF_ERR_INVALID_ARGUMENT;
  if (removed) *removed = 0;
  const int n = *num_nodes;
  if (!pp->enabled || n == 0) return NRF_OK;
  hipStream_t st = as_stream(stream);
  DevBuf<nrf_node> d_nodes;
  DevBuf<uint8_t> d_keep;
  int rc;
  if ((rc = d_nodes.alloc(n)) || (rc = d_keep.alloc(n))) return rc;
  NRF_HIP_CHECK(hipMemcpyAsync(d_nodes.p, nodes, sizeof(nrf_node) * n, hipMemcpyHostToDevice, st));
  PruneArgs a;
  a.vox = voxels;
  a.X = vp->dims[0]; a.Y = vp->dims[1]; a.Z = vp->dims[2];
  a.ox = vp->origin[0]; a.oy = vp->origin[1]; a.oz = vp->origin[2];
  a.s = vp->voxel_size;
  a.nodes = d_nodes.p;
  a.N = n;
  a.pp = *pp;
  a.keep = d_keep.p;
  prune_score_kernel<<<(n + 3) / 4, 256, 0, st>>>(a);
  NRF_LAUNCH_CHECK();
  std::vector<uint8_t> keep(n);
  NRF_HIP_CHECK(hipMemcpyAsync(keep.data(), d_keep.p, n, hipMemcpyDeviceToHost, st));
  NRF_HIP_CHECK(hipStreamSynchronize(st));
  if (pp->remove_disconnected) {  // src/warp_field.cu:1005-1050
    std::vector<std::vector<int>> adj(n);
    for (int i = 0; i < n; ++i) {
      if (!keep[i]) continue;
      for (int k = 0; k < nodes[i].num_neighbors; ++k) {
        const int nb = nodes[i].neighbors[k];
        if (nb < 0 || nb >= n || !keep[nb]) continue;
        adj[i].push_back(nb);
        adj[nb].push_back(i);
      }
    }
    std::vector<uint8_t> seen(n, 0);
    std::vector<int> stack, comp;
    for (int s0 = 0; s0 < n; ++s0) {
      if (!keep[s0] || seen[s0]) continue;
      stack.assign(1, s0);
      comp.clear();
      seen[s0] = 1;
      while (!stack.empty()) {
        const int u = stack.back();
        stack.pop_back();
        comp.push_back(u);
        for (int v : adj[u])
          if (!seen[v]) {
            seen[v] = 1;
            stack.push_back(v);
          }
      }
      if ((int)comp.size() < pp->min_component_size)
        for (int idx : comp) keep[idx] = 0;
    }
  }
  std::vector<int> remap(n, -1);
  int m = 0;
  for (int i = 0; i < n; ++i)
    if (keep[i]) remap[i] = m++;
  if (m == n) return NRF_OK;
  int out = 0;
  for (int i = 0; i < n; ++i) {
    if (!keep[i]) continue;
    nrf_node node = nodes[i];
    const nrf_dualquat t = transforms[i];
    int old_nb[NRF_K_GRAPH];
    float old_w[NRF_K_GRAPH];
    std::memcpy(old_nb, node.neighbors, sizeof(old_nb));
    std::memcpy(old_w, node.neighbor_w, sizeof(old_w));
    for (int k = 0; k < NRF_K_GRAPH; ++k) {
      node.neighbors[k] = -1;
      node.neighbor_w[k] = 0.f;
    }
    int o = 0;
    for (int k = 0; k < node.num_neighbors && o < NRF_K_GRAPH; ++k) {
      const int nb = old_nb[k];
      if (nb < 0 || nb >= n || remap[nb] < 0) continue;
      node.neighbors[o] = remap[nb];
      node.neighbor_w[o] = old_w[k];
      ++o;
    }
    node.num_neighbors = o;
    nodes[out] = node;
    transforms[out] = t;
    ++out;
  }
  *num_nodes = m;
  if (removed) *removed = n - m;
  return NRF_OK;
}
