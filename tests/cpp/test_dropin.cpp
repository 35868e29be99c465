// test_dropin.cpp - the reference's class API (TSDFVolume, WarpField,
// GaussNewtonSolver) driven the way its own tests drive it:
//   * tests/test_mesh_extraction.cu:180-219,556-558: analytic sphere TSDF
//     written into device_data(), extract_surface -> verts > 100, tris > 50;
//   * tests/test_tsdf_integration.cu:34-62,341-345: a 640x480 sphere depth
//     frame integrated into a 128^3 / 1 cm volume, camera at z = -1.5;
//   * tests/test_warp_field.cu:429-455: fresh transforms are the identity;
//   * GaussNewtonSolver::solve on correspondences of a small node graph:
//     finite update, symmetric system.
// Prints one line per check and exits non-zero on the first failure.
#include <cmath>
#include <cstdio>
#include <vector>

#include "solver.h"
#include "tsdf_volume.h"
#include "warp_field.h"

static int fails = 0;
#define CHECK(cond, ...)                    \
  do {                                      \
    if (cond) {                             \
      printf("PASS ");                      \
    } else {                                \
      printf("FAIL ");                      \
      ++fails;                              \
    }                                       \
    printf(__VA_ARGS__);                    \
    printf("\n");                           \
    fflush(stdout);                         \
  } while (0)

static void mesh_extraction() {
  TSDFVolume::Params p;
  p.dims = make_int3(128, 128, 128);
  p.voxel_size = 0.01f;
  p.truncation = 0.03f;
  p.origin = make_float3(-0.64f, -0.64f, 0.36f);
  TSDFVolume vol(p);
  std::vector<TSDFVoxel> v(vol.total_voxels());
  for (int idx = 0; idx < vol.total_voxels(); ++idx) {
    float3 w = vol.voxel_to_world(vol.idx_to_xyz(idx));
    float d = std::sqrt(w.x * w.x + w.y * w.y + (w.z - 1.f) * (w.z - 1.f)) - 0.3f;
    if (std::fabs(d) > p.truncation) d = d > 0 ? p.truncation : -p.truncation;
    v[idx].tsdf = d;
    v[idx].weight = 1.f;
  }
  NRF_DEVICE_CHECK(hipMemcpy(vol.device_data(), v.data(), v.size() * sizeof(TSDFVoxel), hipMemcpyHostToDevice));
  std::vector<float3> verts, normals;
  std::vector<int3> tris;
  vol.extract_surface(verts, normals, tris);
  CHECK(verts.size() > 100 && tris.size() > 50 && verts.size() == 3 * tris.size() && normals.size() == verts.size(),
        "mesh_extraction: %zu vertices, %zu triangles", verts.size(), tris.size());
  double area = 0;
  for (const int3 &t : tris) {
    float3 a = verts[t.x], b = verts[t.y], c = verts[t.z];
    float3 u = make_float3(b.x - a.x, b.y - a.y, b.z - a.z), w = make_float3(c.x - a.x, c.y - a.y, c.z - a.z);
    float3 n = make_float3(u.y * w.z - u.z * w.y, u.z * w.x - u.x * w.z, u.x * w.y - u.y * w.x);
    area += 0.5 * std::sqrt((double)n.x * n.x + (double)n.y * n.y + (double)n.z * n.z);
  }
  CHECK(std::fabs(area - 4 * M_PI * 0.09) < 0.01 * 4 * M_PI * 0.09, "mesh_extraction: area %.4f m^2 (4 pi r^2 = %.4f)",
        area, 4 * M_PI * 0.09);
  DeviceArray<float3> dv;
  DeviceArray<int3> dt;
  vol.extract_surface_device(dv, dt);
  std::vector<float3> hv;
  dv.download(hv);
  bool same = hv.size() == verts.size();
  for (size_t i = 0; same && i < hv.size(); ++i) same = hv[i].x == verts[i].x && hv[i].y == verts[i].y && hv[i].z == verts[i].z;
  CHECK(same, "extract_surface_device matches extract_surface (%zu vertices)", hv.size());
}

static void tsdf_integration() {
  TSDFVolume::Params p;
  p.dims = make_int3(128, 128, 128);
  p.voxel_size = 0.01f;
  p.truncation = 0.03f;
  p.origin = make_float3(-0.64f, -0.64f, -0.64f);
  TSDFVolume vol(p);
  CameraIntrinsics cam{525.f, 525.f, 319.5f, 239.5f, 640, 480};
  std::vector<float> depth(640 * 480, 0.f);
  const float r = 0.3f, cz = 1.5f;  // sphere of radius 0.3 m, 1.5 m in front of the camera
  for (int v = 0; v < 480; ++v)
    for (int u = 0; u < 640; ++u) {
      float dx = (u - cam.cx) / cam.fx, dy = (v - cam.cy) / cam.fy;
      float a = dx * dx + dy * dy + 1, b = -2 * cz, c = cz * cz - r * r;
      float disc = b * b - 4 * a * c;
      if (disc > 0) depth[v * 640 + u] = (-b - std::sqrt(disc)) / (2 * a);
    }
  DeviceArray<float> d;
  d.upload(depth);
  DeviceArray<float3> normals;
  Mat4 pose;
  pose.m[2][3] = -1.5f;  // camera 1.5 m behind the world origin (test_tsdf_integration.cu:426-427)
  vol.integrate(d, normals, cam, pose);
  std::vector<TSDFVoxel> h(vol.total_voxels());
  NRF_DEVICE_CHECK(hipMemcpy(h.data(), vol.device_data(), h.size() * sizeof(TSDFVoxel), hipMemcpyDeviceToHost));
  long obs = 0, near = 0;
  for (const TSDFVoxel &x : h) {
    obs += x.weight > 0;
    near += x.weight > 0 && std::fabs(x.tsdf) < 1.f;
  }
  CHECK(obs > 50000 && near > 5000, "tsdf_integration: %ld observed voxels, %ld inside the truncation band", obs, near);
  // raycast of the integrated sphere (the valid-ray check of
  // tests/test_tsdf_integration.cu:600-610): rays hit the sphere's surface
  DeviceArray<float3> rv, rn;
  vol.raycast(rv, rn, cam, pose);
  std::vector<float3> hv;
  rv.download(hv);
  long hits = 0, on_sphere = 0, disk = 0;
  for (int i = 0; i < 640 * 480; ++i) {
    disk += depth[i] > 0;
    const float3 q = hv[i];
    if (q.x == 0.f && q.y == 0.f && q.z == 0.f) continue;
    ++hits;
    on_sphere += std::fabs(std::sqrt(q.x * q.x + q.y * q.y + q.z * q.z) - r) < 2 * p.voxel_size;
  }
  CHECK(hits > disk / 2 && on_sphere == hits, "raycast: %ld of %ld sphere pixels hit, %ld within 2 voxels of r",
        hits, disk, on_sphere);
  // graph maintenance on the extracted surface (initialize_node_graph,
  // include/fusion_app.h:994-1003): geodesic node graph, then pruning
  std::vector<float3> sv, sn;
  std::vector<int3> st;
  vol.extract_surface(sv, sn, st);
  WarpField wf(0.04f);
  const int added = wf.add_nodes_from_mesh_geodesic(sv, st, 0.025f, &sn);
  std::vector<DeformNode> gn = wf.download_nodes();
  long edges = 0;
  for (const DeformNode &x : gn) edges += x.num_neighbors;
  CHECK(added > 50 && wf.num_nodes() == added && edges > 3L * added,
        "graph: %d nodes sampled on the surface, %ld geodesic edges", added, edges);
  const int removed = wf.prune_nodes(vol, WarpField::PruneParams());
  CHECK(removed >= 0 && wf.num_nodes() == added - removed && (int)wf.download_transforms().size() == wf.num_nodes(),
        "graph: pruning removed %d of %d nodes", removed, added);
}

static std::vector<DeformNode> ring_nodes(int n) {
  std::vector<DeformNode> nodes(n);
  for (int i = 0; i < n; ++i) {
    float a = 2 * M_PI * i / n;
    nodes[i].pos = make_float3(0.1f * std::cos(a), 0.1f * std::sin(a), 1.f);
    nodes[i].normal = make_float3(std::cos(a), std::sin(a), 0.f);
    nodes[i].radius = 0.05f;
    for (int k = 0; k < K_GRAPH; ++k) {
      nodes[i].neighbors[k] = -1;
      nodes[i].neighbor_w[k] = 0.f;
    }
    nodes[i].neighbors[0] = (i + 1) % n;
    nodes[i].neighbors[1] = (i + n - 1) % n;
    nodes[i].neighbor_w[0] = nodes[i].neighbor_w[1] = 1.f;
    nodes[i].num_neighbors = 2;
  }
  return nodes;
}

static void warp_field_and_solver() {
  const int n = 16;
  WarpField wf(0.05f);
  wf.upload_nodes(ring_nodes(n));
  std::vector<DualQuat> T = wf.download_transforms();
  bool ident = (int)T.size() == n;
  for (const DualQuat &q : T)
    ident = ident && q.real.x == 0 && q.real.y == 0 && q.real.z == 0 && q.real.w == 1 && q.dual.x == 0 &&
            q.dual.y == 0 && q.dual.z == 0 && q.dual.w == 0;
  CHECK(ident, "warp_field: %d fresh transforms are the identity", n);
  // correspondences on the ring, pulled 5 mm outwards along the normal
  std::vector<Correspondence> c(n * 8);
  for (int k = 0; k < n * 8; ++k) {
    float a = 2 * M_PI * k / (n * 8);
    int i0 = (int)std::floor(a / (2 * M_PI) * n) % n;
    Correspondence &x = c[k];
    x.src = make_float3(0.1f * std::cos(a), 0.1f * std::sin(a), 1.f);
    x.normal = make_float3(std::cos(a), std::sin(a), 0.f);
    x.dst = make_float3(x.src.x + 0.005f * x.normal.x, x.src.y + 0.005f * x.normal.y, 1.f);
    x.node_ids[0] = i0;
    x.node_ids[1] = (i0 + 1) % n;
    x.node_ids[2] = -1;
    x.node_ids[3] = -1;
    x.node_ws[0] = x.node_ws[1] = 0.5f;
    x.node_ws[2] = x.node_ws[3] = 0.f;
    x.weight = 1.f;
    x.valid = true;
  }
  DeviceArray<Correspondence> dc;
  dc.upload(c);
  GaussNewtonSolver::Params sp;
  sp.pcg_iterations = 30;
  sp.pcg_tolerance = 0.f;
  GaussNewtonSolver solver(sp, WarpField::MAX_NODES);
  DeviceArray<float> dx;
  solver.solve(dc, (int)c.size(), wf.device_nodes(), wf.device_transforms(), wf.num_nodes(), dx);
  std::vector<float> h;
  dx.download(h);
  bool finite = (int)h.size() == 6 * n;
  double radial = 0;
  for (int i = 0; i < n; ++i) {
    for (int k = 0; k < 6; ++k) finite = finite && std::isfinite(h[6 * i + k]);
    float a = 2 * M_PI * i / n;
    radial += h[6 * i + 3] * std::cos(a) + h[6 * i + 4] * std::sin(a);
  }
  radial /= n;
  CHECK(finite && radial > 0.002 && radial < 0.008, "solver: %zu finite updates, mean radial step %.5f m (target 0.005)",
        h.size(), radial);
  const BSRMatrix &A = solver.system_matrix();
  std::vector<int> rp, ci;
  std::vector<float> va;
  A.row_ptr.download(rp);
  A.col_idx.download(ci);
  A.values.download(va);
  double asym = 0, amax = 0;
  for (int i = 0; i < n; ++i)
    for (int q = rp[i]; q < rp[i + 1]; ++q) {
      int j = ci[q];
      int qt = -1;
      for (int s = rp[j]; s < rp[j + 1]; ++s)
        if (ci[s] == i) qt = s;
      for (int a = 0; a < 6; ++a)
        for (int b = 0; b < 6; ++b) {
          asym = std::fmax(asym, std::fabs(va[q * 36 + a * 6 + b] - va[qt * 36 + b * 6 + a]));
          amax = std::fmax(amax, std::fabs(va[q * 36 + a * 6 + b]));
        }
    }
  CHECK(A.num_blocks == 3 * n && asym <= 1e-6 * amax, "solver: %d blocks, symmetric to %.2e (max %.2e)", A.num_blocks,
        asym, amax);
  wf.apply_twist_increment(dx, 0.1f, 0.05f, 1.f);
  std::vector<DualQuat> T2 = wf.download_transforms();
  double moved = 0;
  for (int i = 0; i < n; ++i) {
    // translation of a small-rotation DQ ~ 2 * dual.xyz
    float a = 2 * M_PI * i / n;
    moved += 2 * (T2[i].dual.x * std::cos(a) + T2[i].dual.y * std::sin(a));
  }
  CHECK(moved / n > 0.002, "warp_field: twist update moves the nodes outwards by %.5f m", moved / n);
  wf.save_transforms();
  wf.reset_transforms();
  wf.restore_transforms();
  std::vector<DualQuat> T3 = wf.download_transforms();
  CHECK(T3[3].dual.x == T2[3].dual.x && T3[3].real.w == T2[3].real.w, "warp_field: save / restore round trip");
}

int main() {
  int dev = 0;
  if (hipGetDeviceCount(&dev) != hipSuccess || dev == 0) {
    printf("SKIP no HIP device\n");
    return 0;
  }
  mesh_extraction();
  tsdf_integration();
  warp_field_and_solver();
  printf("%s: %d failure(s)\n", fails ? "FAILED" : "OK", fails);
  return fails ? 1 : 0;
}
