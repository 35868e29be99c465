// graph.cpp - CPU restatement of the reference's deformation-graph
// maintenance (TEST INFRASTRUCTURE ONLY: loaded by tests/ as the checker of
// the device library; nothing in the product calls it).
//
//   add_nodes_from_surface          src/warp_field.cu:709-841
//   add_nodes_from_mesh_geodesic    src/warp_field.cu:843-954
//     build_surface_graph / nearest_graph_vertex   :145-244
//     geodesic distances + K_GRAPH selection        :488-694
//   prune_nodes + score kernel      src/warp_field.cu:407-472, 986-1104
//
// The geodesic distances are computed here by a bounded Dijkstra in FP32
// (d(v) = d(u) + w with the max-distance cut), not by relaxation rounds: both
// give the least fixed point of the FP32 recurrence, so they agree bit for
// bit whenever the relaxation converges.  Spatial hashing uses an ordered
// map (lookups only; results do not depend on the container).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <map>
#include <queue>
#include <tuple>
#include <vector>

#include "../include/nrf.h"

namespace {

using Key = std::tuple<int, int, int>;
using Cells = std::map<Key, std::vector<int>>;

float g_dot(const float *a, const float *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
float g_len(const float *a) { return std::sqrt(g_dot(a, a)); }
float g_d2(const float *a, const float *b) {
  float x = a[0] - b[0], y = a[1] - b[1], z = a[2] - b[2];
  return x * x + y * y + z * z;
}
Key key_of(const float *p, float inv) {
  return Key((int)std::floor(p[0] * inv), (int)std::floor(p[1] * inv), (int)std::floor(p[2] * inv));
}

// DualQuat helpers of the transform initialisation (se3_math.cuh:284-330)
struct Q8 {
  float r[4], d[4];
};
Q8 q_load(const nrf_dualquat &q) {
  Q8 o;
  for (int k = 0; k < 4; ++k) {
    o.r[k] = q.real[k];
    o.d[k] = q.dual[k];
  }
  return o;
}
void q_store(nrf_dualquat &q, const Q8 &o) {
  for (int k = 0; k < 4; ++k) {
    q.real[k] = o.r[k];
    q.dual[k] = o.d[k];
  }
}
// dq_normalize: both parts times 1/|real|, identity below 1e-12
// (se3_math.cuh:302-328)
Q8 q_normalize(Q8 q) {
  float n = std::sqrt(q.r[0] * q.r[0] + q.r[1] * q.r[1] + q.r[2] * q.r[2] + q.r[3] * q.r[3]);
  if (n < 1e-12f) {
    Q8 id{{0, 0, 0, 1}, {0, 0, 0, 0}};
    return id;
  }
  float inv = 1.0f / n;
  for (int k = 0; k < 4; ++k) {
    q.r[k] *= inv;
    q.d[k] *= inv;
  }
  return q;
}
struct Blend {
  Q8 acc{{0, 0, 0, 0}, {0, 0, 0, 0}};
  float wsum = 0;
  void add(Q8 t, float w) {
    float dot = t.r[0] * acc.r[0] + t.r[1] * acc.r[1] + t.r[2] * acc.r[2] + t.r[3] * acc.r[3];
    if (dot < 0.0f)
      for (int k = 0; k < 4; ++k) {
        t.r[k] = -t.r[k];
        t.d[k] = -t.d[k];
      }
    // component order of the reference (x, y, z, w of real, then dual)
    for (int k = 0; k < 4; ++k) acc.r[k] += w * t.r[k];
    for (int k = 0; k < 4; ++k) acc.d[k] += w * t.d[k];
    wsum += w;
  }
};

int surface_nodes(nrf_node *nodes, nrf_dualquat *T, int &n, int max_nodes, const float *V, const float *Nrm, int nv,
                  float min_dist, float radius) {
  const float inv = 1.f / min_dist, md2 = min_dist * min_dist;
  const float gdist = std::fmax(2.5f * min_dist, radius), gd2 = gdist * gdist;
  Cells cells;
  for (int i = 0; i < n; ++i) cells[key_of(nodes[i].pos, inv)].push_back(i);
  const int first_new = n;
  int added = 0;
  for (int vi = 0; vi < nv && n < max_nodes; ++vi) {
    const float *v = V + 3 * (long)vi;
    const Key c = key_of(v, inv);
    auto visit = [&](int R, const std::function<bool(int)> &f) {
      for (int dx = -R; dx <= R; ++dx)
        for (int dy = -R; dy <= R; ++dy)
          for (int dz = -R; dz <= R; ++dz) {
            auto it = cells.find(Key(std::get<0>(c) + dx, std::get<1>(c) + dy, std::get<2>(c) + dz));
            if (it == cells.end()) continue;
            for (int idx : it->second)
              if (f(idx)) return;
          }
    };
    bool close = false;
    visit(1, [&](int idx) { return close = g_d2(v, nodes[idx].pos) < md2; });
    if (close) continue;
    nrf_node nd;
    std::memset(&nd, 0, sizeof(nd));
    std::memcpy(nd.pos, v, sizeof(nd.pos));
    if (Nrm && g_len(Nrm + 3 * (long)vi) > 1e-6f) {
      const float l = g_len(Nrm + 3 * (long)vi);
      for (int k = 0; k < 3; ++k) nd.normal[k] = Nrm[3 * (long)vi + k] / l;
    }
    nd.radius = radius;
    for (int k = 0; k < NRF_K_GRAPH; ++k) nd.neighbors[k] = -1;
    std::vector<std::pair<float, int>> near;
    visit(3, [&](int idx) {
      float d2 = g_d2(v, nodes[idx].pos);
      if (d2 < gd2) near.push_back({d2, idx});
      return false;
    });
    if (g_len(nd.normal) > 1e-6f) {
      std::vector<std::pair<float, int>> kept;
      for (auto &e : near) {
        const float *nb = nodes[e.second].normal;
        if (!(g_len(nb) > 1e-6f && g_dot(nd.normal, nb) < 0.70f)) kept.push_back(e);
      }
      near.swap(kept);
    }
    std::sort(near.begin(), near.end());
    const int k = std::min<int>((int)near.size(), NRF_K_GRAPH);
    for (int i = 0; i < k; ++i) {
      nd.neighbors[i] = near[i].second;
      nd.neighbor_w[i] = 1.f;
    }
    nd.num_neighbors = k;
    Q8 t{{0, 0, 0, 1}, {0, 0, 0, 0}};
    Blend b;
    for (int i = 0; i < k; ++i)
      if (near[i].second < first_new) b.add(q_load(T[near[i].second]), 1.f / (std::sqrt(near[i].first) + 1e-6f));
    if (k > 0 && b.wsum > 0) t = q_normalize(b.acc);
    nodes[n] = nd;
    q_store(T[n], t);
    cells[c].push_back(n);
    ++n;
    ++added;
  }
  return added;
}

struct Mesh {
  std::vector<float> P;
  std::vector<std::vector<std::pair<int, float>>> adj;
  Cells cells;
  float inv = 1;
};

Mesh weld(const float *V, int nv, const int32_t *tri, int nt, float min_dist) {
  Mesh m;
  const float cell = std::fmax(1e-5f, 0.05f * min_dist), inv = 1.0f / cell;
  std::map<Key, int> id;
  std::vector<int> rep(nv);
  for (int i = 0; i < nv; ++i) {
    const float *p = V + 3 * (long)i;
    Key k((int)std::floor(p[0] * inv + 0.5f), (int)std::floor(p[1] * inv + 0.5f), (int)std::floor(p[2] * inv + 0.5f));
    auto it = id.find(k);
    if (it != id.end()) {
      rep[i] = it->second;
      continue;
    }
    rep[i] = (int)(m.P.size() / 3);
    id[k] = rep[i];
    m.P.insert(m.P.end(), p, p + 3);
  }
  const int G = (int)(m.P.size() / 3);
  m.adj.resize(G);
  for (int t = 0; t < nt; ++t) {
    const int32_t *x = tri + 3 * (long)t;
    bool bad = false;
    for (int k = 0; k < 3; ++k) bad = bad || x[k] < 0 || x[k] >= nv;
    if (bad) continue;
    const int r[3] = {rep[x[0]], rep[x[1]], rep[x[2]]};
    for (int k = 0; k < 3; ++k) {
      const int a = r[k], b = r[(k + 1) % 3];
      if (a == b) continue;
      const float w = std::sqrt(g_d2(&m.P[3 * (long)a], &m.P[3 * (long)b]));
      if (w <= 1e-8f) continue;
      m.adj[a].push_back({b, w});
      m.adj[b].push_back({a, w});
    }
  }
  m.inv = 1.0f / min_dist;
  for (int i = 0; i < G; ++i) m.cells[key_of(&m.P[3 * (long)i], m.inv)].push_back(i);
  return m;
}

int nearest(const Mesh &m, const float *p) {
  const int G = (int)(m.P.size() / 3);
  if (G == 0) return -1;
  const Key c = key_of(p, m.inv);
  float best = INFINITY;
  int bi = -1;
  for (int R = 1; R <= 4 && bi < 0; ++R)
    for (int dx = -R; dx <= R; ++dx)
      for (int dy = -R; dy <= R; ++dy)
        for (int dz = -R; dz <= R; ++dz) {
          auto it = m.cells.find(Key(std::get<0>(c) + dx, std::get<1>(c) + dy, std::get<2>(c) + dz));
          if (it == m.cells.end()) continue;
          for (int i : it->second) {
            float d2 = g_d2(p, &m.P[3 * (long)i]);
            if (d2 < best) {
              best = d2;
              bi = i;
            }
          }
        }
  if (bi >= 0) return bi;
  for (int i = 0; i < G; ++i) {
    float d2 = g_d2(p, &m.P[3 * (long)i]);
    if (d2 < best) {
      best = d2;
      bi = i;
    }
  }
  return bi;
}

// bounded FP32 Dijkstra from src; unreached vertices stay +inf
std::vector<float> geodesic(const Mesh &m, int src, float max_dist) {
  const int G = (int)(m.P.size() / 3);
  std::vector<float> d(G, INFINITY);
  if (src < 0) return d;
  using E = std::pair<float, int>;
  std::priority_queue<E, std::vector<E>, std::greater<E>> q;
  d[src] = 0.f;
  q.push({0.f, src});
  while (!q.empty()) {
    auto [du, u] = q.top();
    q.pop();
    if (du != d[u]) continue;
    for (auto &e : m.adj[u]) {
      const float nd = du + e.second;
      if (nd <= max_dist && nd < d[e.first]) {
        d[e.first] = nd;
        q.push({nd, e.first});
      }
    }
  }
  return d;
}

}  // namespace

extern "C" {

int orc_graph_add_nodes(nrf_node *nodes, nrf_dualquat *T, int *num_nodes, int max_nodes, const float *V,
                        const float *Nrm, int nv, float min_dist, float radius) {
  int n = *num_nodes;
  int a = surface_nodes(nodes, T, n, max_nodes, V, Nrm, nv, min_dist, radius);
  *num_nodes = n;
  return a;
}

// Returns the added count; geo_ids / geo_d (optional, 8 per node) receive
// each node's selected geodesic neighbours before the normal filter.
int orc_graph_add_nodes_geodesic(nrf_node *nodes, nrf_dualquat *T, int *num_nodes, int max_nodes, const float *V,
                                 const float *Nrm, int nv, const int32_t *tri, int nt, float min_dist, float radius,
                                 int32_t *geo_ids, float *geo_d) {
  int n = *num_nodes;
  const int added = surface_nodes(nodes, T, n, max_nodes, V, Nrm, nv, min_dist, radius);
  *num_nodes = n;
  if (added <= 0 || n == 0 || nv == 0 || nt == 0) return added;
  Mesh m = weld(V, nv, tri, nt, min_dist);
  if (m.P.empty()) return added;
  std::vector<int> rep(n);
  for (int i = 0; i < n; ++i) rep[i] = nearest(m, nodes[i].pos);
  const float max_dist = std::fmax(3.5f * min_dist, 1.5f * radius);
  std::vector<std::vector<std::pair<float, int>>> nb(n);
  for (int s = 0; s < n; ++s) {
    std::vector<float> d = geodesic(m, rep[s], max_dist);
    std::vector<std::pair<float, int>> c;
    for (int j = 0; j < n; ++j) {
      if (j == s || rep[j] < 0) continue;
      const float gd = d[rep[j]];
      if (!std::isfinite(gd) || gd <= 0.0f || gd > max_dist) continue;
      c.push_back({gd, j});
    }
    std::sort(c.begin(), c.end());  // nearest first, ties to the lower node index
    if (c.size() > (size_t)NRF_K_GRAPH) c.resize(NRF_K_GRAPH);
    nb[s] = c;
    if (geo_ids)
      for (int k = 0; k < NRF_K_GRAPH; ++k) {
        geo_ids[(long)s * NRF_K_GRAPH + k] = k < (int)c.size() ? c[k].second : -1;
        geo_d[(long)s * NRF_K_GRAPH + k] = k < (int)c.size() ? c[k].first : 1e30f;
      }
  }
  for (int i = 0; i < n; ++i) {
    nrf_node &nd = nodes[i];
    nd.num_neighbors = 0;
    for (int k = 0; k < NRF_K_GRAPH; ++k) {
      nd.neighbors[k] = -1;
      nd.neighbor_w[k] = 0.f;
    }
    for (auto &e : nb[i]) {
      const int j = e.second;
      if (g_len(nd.normal) > 1e-6f && g_len(nodes[j].normal) > 1e-6f && g_dot(nd.normal, nodes[j].normal) < 0.70f)
        continue;
      const float r = std::fmax(nd.radius, 1e-6f);
      nd.neighbors[nd.num_neighbors] = j;
      nd.neighbor_w[nd.num_neighbors] = std::exp(-(e.first * e.first) / (2.0f * r * r));
      nd.num_neighbors++;
    }
  }
  const int first_new = n - added;
  for (int i = first_new; i < n; ++i) {
    Blend b;
    for (auto &e : nb[i])
      if (e.second < first_new) b.add(q_load(T[e.second]), 1.f / (std::sqrt(e.first) + 1e-6f));
    if (b.wsum > 0.0f) q_store(T[i], q_normalize(b.acc));
  }
  return added;
}

// prune_nodes with the score of score_node_tsdf_support_kernel; returns the
// removed count, keep (optional, the input count) the flags before compaction
int orc_graph_prune(const nrf_volume_params *vp, const nrf_voxel *vox, nrf_node *nodes, nrf_dualquat *T,
                    int *num_nodes, const nrf_prune_params *pp, uint8_t *keep_out) {
  const int n = *num_nodes;
  if (!pp->enabled || n == 0) return 0;
  std::vector<uint8_t> keep(n);
  const float s = vp->voxel_size;
  for (int i = 0; i < n; ++i) {
    const nrf_node &nd = nodes[i];
    const float rad = std::fmax(nd.radius * pp->support_radius_factor, s), rad2 = rad * rad;
    int c[3], lo[3], hi[3];
    for (int k = 0; k < 3; ++k) c[k] = (int)std::floor((nd.pos[k] - vp->origin[k]) / s);
    const int r = (int)std::ceil(rad / s);
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::max(c[k] - r, 0);
      hi[k] = std::min(c[k] + r, vp->dims[k] - 1);
    }
    int obs = 0, surf = 0, empty = 0;
    for (int z = lo[2]; z <= hi[2]; ++z)
      for (int y = lo[1]; y <= hi[1]; ++y)
        for (int x = lo[0]; x <= hi[0]; ++x) {
          const float q[3] = {vp->origin[0] + (x + 0.5f) * s, vp->origin[1] + (y + 0.5f) * s,
                              vp->origin[2] + (z + 0.5f) * s};
          if (g_d2(q, nd.pos) > rad2) continue;
          const nrf_voxel &v = vox[((long)z * vp->dims[1] + y) * vp->dims[0] + x];
          if (v.weight <= 0.0f) continue;
          ++obs;
          if (std::fabs(v.tsdf) <= pp->surface_tsdf_abs) ++surf;
          if (v.tsdf >= pp->empty_tsdf) ++empty;
        }
    const bool unknown = obs < pp->min_observed_voxels, supported = surf >= pp->min_surface_voxels;
    const bool empty_conf = empty >= pp->min_observed_voxels && surf == 0;
    keep[i] = (unknown || supported || !empty_conf) ? 1 : 0;
  }
  if (keep_out) std::copy(keep.begin(), keep.end(), keep_out);
  if (pp->remove_disconnected) {
    std::vector<std::vector<int>> adj(n);
    for (int i = 0; i < n; ++i)
      if (keep[i])
        for (int k = 0; k < nodes[i].num_neighbors; ++k) {
          const int j = nodes[i].neighbors[k];
          if (j < 0 || j >= n || !keep[j]) continue;
          adj[i].push_back(j);
          adj[j].push_back(i);
        }
    std::vector<int> comp_of(n, -1);
    for (int st = 0; st < n; ++st) {
      if (!keep[st] || comp_of[st] >= 0) continue;
      std::vector<int> todo{st}, comp;
      comp_of[st] = st;
      while (!todo.empty()) {
        int u = todo.back();
        todo.pop_back();
        comp.push_back(u);
        for (int v : adj[u])
          if (comp_of[v] < 0) {
            comp_of[v] = st;
            todo.push_back(v);
          }
      }
      if ((int)comp.size() < pp->min_component_size)
        for (int u : comp) keep[u] = 0;
    }
  }
  std::vector<int> map(n, -1);
  int m = 0;
  for (int i = 0; i < n; ++i)
    if (keep[i]) map[i] = m++;
  if (m == n) return 0;
  std::vector<nrf_node> nn;
  std::vector<nrf_dualquat> nt;
  for (int i = 0; i < n; ++i) {
    if (!keep[i]) continue;
    nrf_node nd = nodes[i];
    int o = 0;
    nrf_node src = nd;
    for (int k = 0; k < NRF_K_GRAPH; ++k) {
      nd.neighbors[k] = -1;
      nd.neighbor_w[k] = 0.f;
    }
    for (int k = 0; k < src.num_neighbors && o < NRF_K_GRAPH; ++k) {
      const int j = src.neighbors[k];
      if (j < 0 || j >= n || map[j] < 0) continue;
      nd.neighbors[o] = map[j];
      nd.neighbor_w[o] = src.neighbor_w[k];
      ++o;
    }
    nd.num_neighbors = o;
    nn.push_back(nd);
    nt.push_back(T[i]);
  }
  std::copy(nn.begin(), nn.end(), nodes);
  std::copy(nt.begin(), nt.end(), T);
  *num_nodes = m;
  return n - m;
}

}  // extern "C"
