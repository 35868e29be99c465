// oracle.cpp - CPU restatement of the reference's non-rigid fusion step.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity oracle and the timed
// CPU baseline; it is never linked into, loaded by or called from the
// product library.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg use it.
//
// Every function is a scalar restatement of one reference routine (cited by
// file:line, paths relative to the Hydran00/CUDAFusion tree), written as a
// template over the arithmetic type R:
//   R = float  follows the reference's FP32 operation order (one CUDA thread
//              run sequentially, atomics replaced by in-order sums);
//   R = double is the FP64 oracle the GPU results are measured against.
// Loops over independent voxels / rows are split with OpenMP; sums whose
// order matters are kept sequential, so results do not depend on the thread
// count.
//
// Build: oracle/Makefile (g++ -O2 -fopenmp -ffp-contract=off, no fast-math).
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <string>
#include <type_traits>
#include <vector>

#include "../include/nrf.h"

namespace orc {

template <typename R>
struct V3 {
  R x, y, z;
};
template <typename R>
struct Q4 {
  R x, y, z, w;
};
template <typename R>
struct DQ {
  Q4<R> r, d;
};

template <typename R> V3<R> v3(R x, R y, R z) { return {x, y, z}; }
template <typename R> V3<R> v3f(const float *p) { return {(R)p[0], (R)p[1], (R)p[2]}; }
template <typename R> V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename R> R dot3(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename R> V3<R> cross3(V3<R> a, V3<R> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// quat_mul / quat_conj / quat_normalize / quat_dot: include/se3_math.cuh:205-235
template <typename R> Q4<R> qmul(Q4<R> a, Q4<R> b) {
  return {a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
          a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
template <typename R> Q4<R> qconj(Q4<R> q) { return {-q.x, -q.y, -q.z, q.w}; }
template <typename R> R qdot(Q4<R> a, Q4<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
template <typename R> Q4<R> qnorm(Q4<R> q) {
  R n = std::sqrt(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
  if (n < (R)1e-12f) return {0, 0, 0, 1};
  return {q.x / n, q.y / n, q.z / n, q.w / n};
}
template <typename R> DQ<R> dq_id() { return {{0, 0, 0, 1}, {0, 0, 0, 0}}; }

// dq_normalize: se3_math.cuh:302-328
template <typename R> DQ<R> dq_normalize(DQ<R> q) {
  R n = std::sqrt(q.r.x * q.r.x + q.r.y * q.r.y + q.r.z * q.r.z + q.r.w * q.r.w);
  if (n < (R)1e-12f) return dq_id<R>();
  R inv = (R)1 / n;
  q.r.x *= inv; q.r.y *= inv; q.r.z *= inv; q.r.w *= inv;
  q.d.x *= inv; q.d.y *= inv; q.d.z *= inv; q.d.w *= inv;
  return q;
}
// dq_mul: se3_math.cuh:330-340
template <typename R> DQ<R> dq_mul(DQ<R> a, DQ<R> b) {
  DQ<R> o;
  o.r = qmul(a.r, b.r);
  Q4<R> ad = qmul(a.r, b.d), bd = qmul(a.d, b.r);
  o.d = {ad.x + bd.x, ad.y + bd.y, ad.z + bd.z, ad.w + bd.w};
  return dq_normalize(o);
}
// dq_rotate_vector / dq_translation / dq_transform_point: se3_math.cuh:349-377
template <typename R> V3<R> dq_rot(const DQ<R> &q, V3<R> p) {
  Q4<R> r = qmul(qmul(q.r, Q4<R>{p.x, p.y, p.z, 0}), qconj(q.r));
  return {r.x, r.y, r.z};
}
template <typename R> V3<R> dq_trans(const DQ<R> &q) {
  Q4<R> t = qmul(q.d, qconj(q.r));
  return {(R)2 * t.x, (R)2 * t.y, (R)2 * t.z};
}
template <typename R> V3<R> dq_xf(const DQ<R> &q, V3<R> p) {
  V3<R> r = dq_rot(q, p), t = dq_trans(q);
  return {r.x + t.x, r.y + t.y, r.z + t.z};
}
// dq_transform_point_centered: se3_math.cuh:379-388
template <typename R> V3<R> dq_xf_c(const DQ<R> &q, V3<R> p, V3<R> c) {
  V3<R> rp = dq_xf(q, p), rc = dq_rot(q, c);
  return {rp.x + c.x - rc.x, rp.y + c.y - rc.y, rp.z + c.z - rc.z};
}
// dq_from_rotation_translation: se3_math.cuh:390-403
template <typename R> DQ<R> dq_from_rt(Q4<R> qr, V3<R> t) {
  DQ<R> o;
  o.r = qnorm(qr);
  o.d = qmul(Q4<R>{t.x, t.y, t.z, 0}, o.r);
  o.d.x *= (R)0.5f; o.d.y *= (R)0.5f; o.d.z *= (R)0.5f; o.d.w *= (R)0.5f;
  return dq_normalize(o);
}
// dq_centered: se3_math.cuh:405-415
template <typename R> DQ<R> dq_centered(const DQ<R> &q, V3<R> c) {
  V3<R> t = dq_trans(q), rc = dq_rot(q, c);
  return dq_from_rt(q.r, V3<R>{t.x + c.x - rc.x, t.y + c.y - rc.y, t.z + c.z - rc.z});
}

// exp_se3: se3_math.cuh:40-98 (rows 0-2 of the 4x4)
template <typename R> void exp_se3(const R tw[6], R T[3][4]) {
  R ox = tw[0], oy = tw[1], oz = tw[2], vx = tw[3], vy = tw[4], vz = tw[5];
  R th = std::sqrt(ox * ox + oy * oy + oz * oz);
  R A, B, C;
  if (th < (R)1e-6f) {
    A = 1; B = (R)0.5f; C = (R)1 / (R)6;
  } else {
    R it = (R)1 / th, it2 = it * it;
    A = std::sin(th) * it;
    B = ((R)1 - std::cos(th)) * it2;
    C = (th - std::sin(th)) * it2 * it;
  }
  T[0][0] = (R)1 - B * (oy * oy + oz * oz); T[0][1] = B * ox * oy - A * oz; T[0][2] = B * ox * oz + A * oy;
  T[1][0] = B * ox * oy + A * oz; T[1][1] = (R)1 - B * (ox * ox + oz * oz); T[1][2] = B * oy * oz - A * ox;
  T[2][0] = B * ox * oz - A * oy; T[2][1] = B * oy * oz + A * ox; T[2][2] = (R)1 - B * (ox * ox + oy * oy);
  R Vxx = (R)1 - C * (oy * oy + oz * oz), Vxy = C * ox * oy - B * oz, Vxz = C * ox * oz + B * oy;
  R Vyx = C * ox * oy + B * oz, Vyy = (R)1 - C * (ox * ox + oz * oz), Vyz = C * oy * oz - B * ox;
  R Vzx = C * ox * oz - B * oy, Vzy = C * oy * oz + B * ox, Vzz = (R)1 - C * (ox * ox + oy * oy);
  T[0][3] = Vxx * vx + Vxy * vy + Vxz * vz;
  T[1][3] = Vyx * vx + Vyy * vy + Vyz * vz;
  T[2][3] = Vzx * vx + Vzy * vy + Vzz * vz;
}

// mat4_to_dq: se3_math.cuh:244-300
template <typename R> DQ<R> mat_to_dq(const R T[3][4]) {
  R tr = T[0][0] + T[1][1] + T[2][2];
  Q4<R> q;
  if (tr > 0) {
    R s = std::sqrt(tr + (R)1) * (R)2;
    q.w = (R)0.25f * s; q.x = (T[2][1] - T[1][2]) / s; q.y = (T[0][2] - T[2][0]) / s; q.z = (T[1][0] - T[0][1]) / s;
  } else if (T[0][0] > T[1][1] && T[0][0] > T[2][2]) {
    R s = std::sqrt((R)1 + T[0][0] - T[1][1] - T[2][2]) * (R)2;
    q.w = (T[2][1] - T[1][2]) / s; q.x = (R)0.25f * s; q.y = (T[0][1] + T[1][0]) / s; q.z = (T[0][2] + T[2][0]) / s;
  } else if (T[1][1] > T[2][2]) {
    R s = std::sqrt((R)1 + T[1][1] - T[0][0] - T[2][2]) * (R)2;
    q.w = (T[0][2] - T[2][0]) / s; q.x = (T[0][1] + T[1][0]) / s; q.y = (R)0.25f * s; q.z = (T[1][2] + T[2][1]) / s;
  } else {
    R s = std::sqrt((R)1 + T[2][2] - T[0][0] - T[1][1]) * (R)2;
    q.w = (T[1][0] - T[0][1]) / s; q.x = (T[0][2] + T[2][0]) / s; q.y = (T[1][2] + T[2][1]) / s; q.z = (R)0.25f * s;
  }
  q = qnorm(q);
  Q4<R> qd = qmul(Q4<R>{T[0][3], T[1][3], T[2][3], 0}, q);
  return {q, {qd.x * (R)0.5f, qd.y * (R)0.5f, qd.z * (R)0.5f, qd.w * (R)0.5f}};
}

// dq_to_mat4: se3_math.cuh:417-443
template <typename R> void dq_to_mat(DQ<R> q, R T[3][4]) {
  q = dq_normalize(q);
  R x = q.r.x, y = q.r.y, z = q.r.z, w = q.r.w;
  V3<R> t = dq_trans(q);
  T[0][0] = (R)1 - (R)2 * (y * y + z * z); T[0][1] = (R)2 * (x * y - z * w); T[0][2] = (R)2 * (x * z + y * w);
  T[0][3] = t.x;
  T[1][0] = (R)2 * (x * y + z * w); T[1][1] = (R)1 - (R)2 * (x * x + z * z); T[1][2] = (R)2 * (y * z - x * w);
  T[1][3] = t.y;
  T[2][0] = (R)2 * (x * z - y * w); T[2][1] = (R)2 * (y * z + x * w); T[2][2] = (R)1 - (R)2 * (x * x + y * y);
  T[2][3] = t.z;
}

template <typename R> DQ<R> load_dq(const nrf_dualquat &q) {
  return {{(R)q.real[0], (R)q.real[1], (R)q.real[2], (R)q.real[3]},
          {(R)q.dual[0], (R)q.dual[1], (R)q.dual[2], (R)q.dual[3]}};
}
template <typename R> DQ<R> load_dq8(const R *p) { return {{p[0], p[1], p[2], p[3]}, {p[4], p[5], p[6], p[7]}}; }
template <typename R> void store_dq8(R *p, const DQ<R> &q) {
  p[0] = q.r.x; p[1] = q.r.y; p[2] = q.r.z; p[3] = q.r.w;
  p[4] = q.d.x; p[5] = q.d.y; p[6] = q.d.z; p[7] = q.d.w;
}

// warp_point: se3_math.cuh:450-510 (host twin include/fusion_math.h:121-167),
// with the centred node transforms precomputed (same values).
template <typename R>
V3<R> warp_point(V3<R> p, const DQ<R> *centered, int N, const int32_t ids[4], const float ws[4]) {
  DQ<R> b{{0, 0, 0, 0}, {0, 0, 0, 0}};
  R wsum = 0;
  for (int k = 0; k < 4; ++k) {
    int id = ids[k];
    R w = (R)ws[k];
    if (id < 0 || id >= N || ws[k] < 1e-8f) continue;
    DQ<R> q = centered[id];
    if (qdot(q.r, b.r) < 0) {
      q.r = {-q.r.x, -q.r.y, -q.r.z, -q.r.w};
      q.d = {-q.d.x, -q.d.y, -q.d.z, -q.d.w};
    }
    b.r.x += w * q.r.x; b.r.y += w * q.r.y; b.r.z += w * q.r.z; b.r.w += w * q.r.w;
    b.d.x += w * q.d.x; b.d.y += w * q.d.y; b.d.z += w * q.d.z; b.d.w += w * q.d.w;
    wsum += w;
  }
  if (wsum < (R)1e-8f) return p;
  return dq_xf(dq_normalize(b), p);
}

// warp_normal: se3_math.cuh:512-558 (uncentred real parts, no weight-sum test)
template <typename R>
V3<R> warp_normal(V3<R> n, const DQ<R> *T, int N, const int32_t ids[4], const float ws[4]) {
  DQ<R> b{{0, 0, 0, 0}, {0, 0, 0, 0}};
  for (int k = 0; k < 4; ++k) {
    int id = ids[k];
    if (id < 0 || id >= N || ws[k] < 1e-8f) continue;
    R w = (R)ws[k];
    Q4<R> q = T[id].r;
    if (qdot(q, b.r) < 0) q = {-q.x, -q.y, -q.z, -q.w};
    b.r.x += w * q.x; b.r.y += w * q.y; b.r.z += w * q.z; b.r.w += w * q.w;
  }
  b = dq_normalize(b);
  Q4<R> r = qmul(qmul(b.r, Q4<R>{n.x, n.y, n.z, 0}), qconj(b.r));
  R len = std::sqrt(r.x * r.x + r.y * r.y + r.z * r.z);
  if (len > (R)1e-9f) return {r.x / len, r.y / len, r.z / len};
  return {r.x, r.y, r.z};
}

template <typename R>
std::vector<DQ<R>> centered_table(const nrf_node *nodes, const DQ<R> *T, int N) {
  std::vector<DQ<R>> c(N);
  for (int i = 0; i < N; ++i) c[i] = dq_centered(T[i], v3f<R>(nodes[i].pos));
  return c;
}

// Rigid inverse of the camera pose: src/tsdf_volume.cu:36-49 (FP32 on host).
static void invert_pose(const float pose[16], float T[3][4]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) T[i][j] = pose[j * 4 + i];
  for (int i = 0; i < 3; ++i) T[i][3] = -(T[i][0] * pose[3] + T[i][1] * pose[7] + T[i][2] * pose[11]);
}

// tsdf_integrate_kernel: kernels/tsdf_kernels.cu:85-174, one voxel at a time.
template <typename R>
long integrate(const nrf_volume_params &vp, R *tsdf, R *weight, const float *depth, const nrf_camera &cam,
               const float pose[16], const nrf_node *nodes, const DQ<R> *T, int N, const int32_t *knn,
               const float *knn_w, const int32_t *opt, int min_opt) {
  const int X = vp.dims[0], Y = vp.dims[1], Z = vp.dims[2];
  const bool warp = nodes != nullptr && N > 0;
  float Tc[3][4];
  invert_pose(pose, Tc);
  std::vector<DQ<R>> cen;
  if (warp) cen = centered_table(nodes, T, N);
  const R s = (R)vp.voxel_size, tr = (R)vp.truncation, decay = (R)vp.decay_alpha, wmax = (R)vp.max_weight;
  long updated = 0;
#pragma omp parallel for schedule(static) reduction(+ : updated)
  for (int z = 0; z < Z; ++z)
    for (int y = 0; y < Y; ++y)
      for (int x = 0; x < X; ++x) {
        const long idx = ((long)z * Y + y) * X + x;
        if (warp && opt && opt[idx] < min_opt) continue;
        V3<R> p = {(R)vp.origin[0] + ((R)x + (R)0.5f) * s, (R)vp.origin[1] + ((R)y + (R)0.5f) * s,
                   (R)vp.origin[2] + ((R)z + (R)0.5f) * s};
        if (warp) p = warp_point(p, cen.data(), N, knn + idx * 4, knn_w + idx * 4);
        R cx = (R)Tc[0][0] * p.x + (R)Tc[0][1] * p.y + (R)Tc[0][2] * p.z + (R)Tc[0][3];
        R cy = (R)Tc[1][0] * p.x + (R)Tc[1][1] * p.y + (R)Tc[1][2] * p.z + (R)Tc[1][3];
        R cz = (R)Tc[2][0] * p.x + (R)Tc[2][1] * p.y + (R)Tc[2][2] * p.z + (R)Tc[2][3];
        if (cz <= 0) continue;
        R fu = (R)cam.fx * cx / cz + (R)cam.cx;
        R fv = (R)cam.fy * cy / cz + (R)cam.cy;
        if (!(fu > (R)-1e9f && fu < (R)1e9f && fv > (R)-1e9f && fv < (R)1e9f)) continue;
        int u = (int)std::nearbyint(fu), v = (int)std::nearbyint(fv);  // round half to even
        if (u < 0 || u >= cam.width || v < 0 || v >= cam.height) continue;
        R d = (R)depth[(long)v * cam.width + u];
        if (d <= 0) continue;
        R sdf = d - cz;
        if (sdf < -tr) continue;
        R t = std::min((R)1, sdf / tr);
        R dw = weight[idx] * decay;
        R nw = dw + (R)1;
        tsdf[idx] = (tsdf[idx] * dw + t * (R)1) / nw;
        weight[idx] = std::min(nw, wmax);
        ++updated;
      }
  return updated;
}

// ------------------------------------------------------- marching cubes
static int g_tri[256][16];
static int g_ntri[256];
static bool g_tables = false;
static const int kCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                  {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
static const int kEdge[12][2] = {{0, 1}, {1, 2}, {3, 2}, {0, 3}, {4, 5}, {5, 6},
                                 {7, 6}, {4, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

static int load_tables(const char *path) {
  std::ifstream f(path);
  if (!f) return -1;
  std::string line;
  int c = 0;
  while (std::getline(f, line) && c < 256) {
    if (line.empty() || line[0] == '#') continue;
    for (int k = 0; k < 16; ++k) g_tri[c][k] = -1;
    int n = 0;
    if (line != ".")
      for (char ch : line) g_tri[c][n++] = (ch >= 'a') ? ch - 'a' + 10 : ch - '0';
    g_ntri[c] = n / 3;
    ++c;
  }
  g_tables = c == 256;
  return g_tables ? 0 : -2;
}

// load_cube / count_surface_kernel / generate_surface_kernel:
// src/marching_cubes.cu:315-409, emitted in increasing cube index order.
template <typename R>
long marching_cubes(const nrf_volume_params &vp, const R *tsdf, const R *weight, uint8_t *cases, R *verts,
                    R *normals, long cap) {
  const int X = vp.dims[0], Y = vp.dims[1], Z = vp.dims[2];
  const long V = (long)X * Y * Z;
  const R s = (R)vp.voxel_size;
  std::vector<long> rowcnt((long)Y * Z, 0);
#pragma omp parallel for schedule(static)
  for (long row = 0; row < (long)Y * Z; ++row) {
    const int y = (int)(row % Y), z = (int)(row / Y);
    long cnt = 0;
    for (int x = 0; x < X; ++x) {
      const long idx = row * X + x;
      int cube = 0;
      if (x < X - 1 && y < Y - 1 && z < Z - 1) {
        for (int c = 0; c < 8; ++c) {
          long vi = idx + kCorner[c][0] + (long)kCorner[c][1] * X + (long)kCorner[c][2] * X * Y;
          if (weight[vi] <= 0) { cube = -1; break; }
          if (tsdf[vi] > 0) cube |= 1 << c;
        }
        if (cube <= 0 || cube >= 255) cube = 0;
      }
      if (cases) cases[idx] = (uint8_t)cube;
      cnt += g_ntri[cube];
    }
    rowcnt[row] = cnt;
  }
  (void)V;
  std::vector<long> rowoff((long)Y * Z + 1, 0);
  for (long r = 0; r < (long)Y * Z; ++r) rowoff[r + 1] = rowoff[r] + rowcnt[r];
  const long total = rowoff[(long)Y * Z];
  if (!verts) return total;
#pragma omp parallel for schedule(dynamic, 64)
  for (long row = 0; row < (long)Y * Z; ++row) {
    if (rowcnt[row] == 0) continue;
    const int y = (int)(row % Y), z = (int)(row / Y);
    long t = rowoff[row];
    for (int x = 0; x < X - 1; ++x) {
      const long idx = row * X + x;
      int cube = 0;
      R val[8];
      V3<R> pts[8];
      if (y < Y - 1 && z < Z - 1) {
        for (int c = 0; c < 8; ++c) {
          long vi = idx + kCorner[c][0] + (long)kCorner[c][1] * X + (long)kCorner[c][2] * X * Y;
          if (weight[vi] <= 0) { cube = -1; break; }
          val[c] = tsdf[vi];
          pts[c] = {(R)(float)(x + kCorner[c][0]) + (R)0.5f, (R)(float)(y + kCorner[c][1]) + (R)0.5f,
                    (R)(float)(z + kCorner[c][2]) + (R)0.5f};
          if (val[c] > 0) cube |= 1 << c;
        }
      }
      if (cube <= 0 || cube >= 255) continue;
      V3<R> ev[12];
      for (int e = 0; e < 12; ++e) {  // interpolate_edge, marching_cubes.cu:315-321
        int c0 = kEdge[e][0], c1 = kEdge[e][1];
        R v0 = val[c0], v1 = val[c1], den = v1 - v0;
        R tt = std::fabs(den) > (R)1e-12f ? -v0 / den : (R)0.5f;
        tt = std::min((R)1, std::max((R)0, tt));
        V3<R> p0 = pts[c0], p1 = pts[c1];
        V3<R> q = {p0.x + (p1.x - p0.x) * tt, p0.y + (p1.y - p0.y) * tt, p0.z + (p1.z - p0.z) * tt};
        ev[e] = {(R)vp.origin[0] + q.x * s, (R)vp.origin[1] + q.y * s, (R)vp.origin[2] + q.z * s};
      }
      for (int k = 0; k < g_ntri[cube]; ++k, ++t) {
        if (t >= cap) continue;
        V3<R> a = ev[g_tri[cube][3 * k]], b = ev[g_tri[cube][3 * k + 1]], c = ev[g_tri[cube][3 * k + 2]];
        R *vo = verts + t * 9;
        vo[0] = a.x; vo[1] = a.y; vo[2] = a.z; vo[3] = b.x; vo[4] = b.y; vo[5] = b.z;
        vo[6] = c.x; vo[7] = c.y; vo[8] = c.z;
        if (normals) {
          V3<R> n = cross3(b - a, c - a);
          R len = std::sqrt(n.x * n.x + n.y * n.y + n.z * n.z);
          V3<R> nn = {0, 0, 0};
          if (!(len <= (R)1e-12f)) {
            R inv = (R)1 / len;
            nn = {n.x * inv, n.y * inv, n.z * inv};
          }
          R *no = normals + t * 9;
          for (int j = 0; j < 3; ++j) { no[3 * j] = nn.x; no[3 * j + 1] = nn.y; no[3 * j + 2] = nn.z; }
        }
      }
    }
  }
  return total;
}

// compute_voxel_knn_kernel: src/warp_field.cu:299-405.
template <typename R>
void voxel_knn(const nrf_volume_params &vp, const float *weight, const nrf_node *nodes, int N, int32_t *knn,
               R *knn_w) {
  const int X = vp.dims[0], Y = vp.dims[1], Z = vp.dims[2];
  const R s = (R)vp.voxel_size;
#pragma omp parallel for schedule(dynamic, 1)
  for (int z = 0; z < Z; ++z)
    for (int y = 0; y < Y; ++y)
      for (int x = 0; x < X; ++x) {
        const long idx = ((long)z * Y + y) * X + x;
        if (weight[idx] <= 0 || N == 0) {
          for (int k = 0; k < 4; ++k) { knn[idx * 4 + k] = -1; knn_w[idx * 4 + k] = 0; }
          continue;
        }
        V3<R> p = {(R)vp.origin[0] + ((R)x + (R)0.5f) * s, (R)vp.origin[1] + ((R)y + (R)0.5f) * s,
                   (R)vp.origin[2] + ((R)z + (R)0.5f) * s};
        R bd[4] = {(R)1e30f, (R)1e30f, (R)1e30f, (R)1e30f};
        int bi[4] = {-1, -1, -1, -1};
        for (int n = 0; n < N; ++n) {
          R dx = p.x - (R)nodes[n].pos[0], dy = p.y - (R)nodes[n].pos[1], dz = p.z - (R)nodes[n].pos[2];
          R d2 = dx * dx + dy * dy + dz * dz;
          if (d2 < bd[3]) {
            bd[3] = d2; bi[3] = n;
            for (int k = 2; k >= 0; --k)
              if (bd[k + 1] < bd[k]) { std::swap(bd[k], bd[k + 1]); std::swap(bi[k], bi[k + 1]); }
          }
        }
        R w[4], wsum = 0;
        for (int k = 0; k < 4; ++k) {
          if (bi[k] < 0) { w[k] = 0; continue; }
          R r = std::max((R)nodes[bi[k]].radius, (R)1e-6f);
          w[k] = std::exp(-bd[k] / ((R)2 * r * r));
          wsum += w[k];
        }
        R inv = wsum > (R)1e-8f ? (R)1 / wsum : (R)0;
        for (int k = 0; k < 4; ++k) { knn[idx * 4 + k] = bi[k]; knn_w[idx * 4 + k] = w[k] * inv; }
      }
}

// ------------------------------------------------------------- solver
// build_sparsity_pattern: kernels/solver_kernels.cu:498-541.
static int nnb(const nrf_node &n) { return std::max(0, std::min((int)n.num_neighbors, (int)NRF_K_GRAPH)); }
static void pattern(const nrf_node *h, int N, std::vector<int32_t> &rp, std::vector<int32_t> &ci) {
  std::vector<std::vector<int32_t>> rc(N);
  for (int i = 0; i < N; ++i) rc[i].push_back(i);
  for (int i = 0; i < N; ++i)
    for (int k = 0; k < nnb(h[i]); ++k) {
      int j = h[i].neighbors[k];
      if (j < 0 || j >= N) continue;
      rc[i].push_back(j);
      rc[j].push_back(i);
    }
  rp.assign(N + 1, 0);
  ci.clear();
  for (int i = 0; i < N; ++i) {
    std::sort(rc[i].begin(), rc[i].end());
    rc[i].erase(std::unique(rc[i].begin(), rc[i].end()), rc[i].end());
    rp[i] = (int32_t)ci.size();
    for (int c : rc[i]) ci.push_back(c);
  }
  rp[N] = (int32_t)ci.size();
}

// find_block: kernels/solver_kernels.cu:25-44
static int find_block(int row, int col, const int32_t *rp, const int32_t *ci) {
  int lo = rp[row], hi = rp[row + 1] - 1;
  while (lo <= hi) {
    int mid = (lo + hi) >> 1;
    if (ci[mid] == col) return mid;
    if (ci[mid] < col) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

template <typename R> R huber(R r, R delta) {
  R a = std::fabs(r);
  return a <= delta ? (R)1 : delta / a;
}

// assemble_system (kernels/solver_kernels.cu:572-604): data term (:52-138),
// smoothness (:202-352), damping (:167-194), then the block-Jacobi inverse
// (:358-399, se3_math.cuh:152-202).  The atomics become in-order sums:
// corrections in slot order, then edges in (node, slot) order.
template <typename R>
void assemble(const nrf_corr *corrs, int n, const float *canonical, const nrf_node *nodes, const DQ<R> *T, int N,
              R lam_s, R lam_d, const int32_t *rp, const int32_t *ci, R *A, R *b, R *Minv, int precond_fp64,
              int consistent_sign, int psd_pairs = 0) {
  const int nb = rp[N];
  std::fill(A, A + (long)nb * 36, (R)0);
  std::fill(b, b + (long)N * 6, (R)0);
  std::vector<DQ<R>> cen;
  if (canonical) cen = centered_table(nodes, T, N);
  for (int c = 0; c < n; ++c) {
    const nrf_corr &cr = corrs[c];
    if (!cr.valid) continue;
    V3<R> src = v3f<R>(cr.src);
    if (canonical) src = warp_point(v3f<R>(canonical + 3 * c), cen.data(), N, cr.node_ids, cr.node_ws);
    V3<R> dst = v3f<R>(cr.dst), nrm = v3f<R>(cr.normal);
    R r = dot3(nrm, src - dst);
    R fw = std::max((R)cr.weight, (R)0) * huber(r, (R)0.2f);
    if (fw <= (R)1e-8f) continue;
    R J[4][6];
    bool ok[4] = {false, false, false, false};
    for (int k = 0; k < 4; ++k) {
      int id = cr.node_ids[k];
      if (id >= 0 && id < N && cr.node_ws[k] > 1e-8f) {
        R w = (R)cr.node_ws[k];
        V3<R> cp = cross3(src - v3f<R>(nodes[id].pos), nrm);
        J[k][0] = w * cp.x; J[k][1] = w * cp.y; J[k][2] = w * cp.z;
        J[k][3] = w * nrm.x; J[k][4] = w * nrm.y; J[k][5] = w * nrm.z;
        ok[k] = true;
        for (int a = 0; a < 6; ++a) b[id * 6 + a] += -fw * J[k][a] * r;
      }
    }
    for (int ki = 0; ki < 4; ++ki) {
      if (!ok[ki]) continue;
      for (int kj = 0; kj < 4; ++kj) {
        if (!ok[kj]) continue;
        int bp = find_block(cr.node_ids[ki], cr.node_ids[kj], rp, ci);
        if (bp < 0) {
          // Q2: the pair is dropped; psd_pairs (solver option, no reference
          // counterpart) adds f J_ki^T J_ki to the diagonal block instead
          if (!psd_pairs) continue;
          bp = find_block(cr.node_ids[ki], cr.node_ids[ki], rp, ci);
          R *blk = A + (long)bp * 36;
          for (int a = 0; a < 6; ++a)
            for (int bb = 0; bb < 6; ++bb) blk[a * 6 + bb] += fw * J[ki][a] * J[ki][bb];
          continue;
        }
        R *blk = A + (long)bp * 36;
        for (int a = 0; a < 6; ++a)
          for (int bb = 0; bb < 6; ++bb) blk[a * 6 + bb] += fw * J[ki][a] * J[kj][bb];
      }
    }
  }
  for (int ni = 0; ni < N; ++ni) {
    const nrf_node &nd = nodes[ni];
    for (int k = 0; k < nnb(nd); ++k) {
      int nj = nd.neighbors[k];
      if (nj < 0 || nj >= N) continue;
      if (nj <= ni) continue;
      R nw = std::max((R)nd.neighbor_w[k], (R)0.05f);
      V3<R> xj = v3f<R>(nodes[nj].pos), xi = v3f<R>(nd.pos);
      V3<R> Ti_xj = dq_xf_c(T[ni], xj, xi), Ti_xi = dq_xf_c(T[ni], xi, xi), Tj_xj = dq_xf_c(T[nj], xj, xj);
      V3<R> rest = xi - xj, curr = Ti_xi - Tj_xj;
      R dr = std::sqrt(rest.x * rest.x + rest.y * rest.y + rest.z * rest.z);
      R dc = std::sqrt(curr.x * curr.x + curr.y * curr.y + curr.z * curr.z);
      R strain = std::fabs(dc - dr) / std::max(dr, (R)1e-6f);
      R ratio = strain / (R)0.25f;
      R sw = std::exp(-(ratio * ratio));
      V3<R> r = Ti_xj - Tj_xj;
      R rn = std::sqrt(r.x * r.x + r.y * r.y + r.z * r.z);
      R lam = lam_s * nw * huber(rn, (R)0.05f) * sw;
      if (lam <= (R)1e-8f) continue;
      // compute_arap_jacobian :140-165; consistent_sign negates the rotation
      // rows so that they equal d(omega x q)/d(omega)
      auto jac = [consistent_sign](V3<R> q, R J[6][3]) {
        if (consistent_sign) q = V3<R>{-q.x, -q.y, -q.z};
        R t[6][3] = {{0, q.z, -q.y}, {-q.z, 0, q.x}, {q.y, -q.x, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        std::memcpy(J, t, sizeof(t));
      };
      R Ji[6][3], Jj[6][3];
      jac(Ti_xj - xi, Ji);
      jac(Tj_xj - xj, Jj);
      int ii = find_block(ni, ni, rp, ci), jj = find_block(nj, nj, rp, ci);
      int ij = find_block(ni, nj, rp, ci), ji = find_block(nj, ni, rp, ci);
      auto add = [&](int bp, R Ja[6][3], R Jb[6][3], R sc) {
        if (bp < 0) return;
        for (int a = 0; a < 6; ++a)
          for (int bb = 0; bb < 6; ++bb) {
            R v = 0;
            for (int c = 0; c < 3; ++c) v += Ja[a][c] * Jb[bb][c];
            A[(long)bp * 36 + a * 6 + bb] += sc * v;
          }
      };
      add(ii, Ji, Ji, lam);
      add(jj, Jj, Jj, lam);
      add(ij, Ji, Jj, -lam);
      add(ji, Jj, Ji, -lam);
      R ra[3] = {r.x, r.y, r.z};
      for (int a = 0; a < 6; ++a) {
        R vi = 0, vj = 0;
        for (int c = 0; c < 3; ++c) { vi += Ji[a][c] * ra[c]; vj += Jj[a][c] * ra[c]; }
        b[ni * 6 + a] += -lam * vi;
        b[nj * 6 + a] += lam * vj;
      }
    }
  }
  for (int ni = 0; ni < N; ++ni) {
    int bp = find_block(ni, ni, rp, ci);
    if (bp < 0) continue;
    R *d = A + (long)bp * 36;
    for (int k = 0; k < 6; ++k) {
      R cur = d[k * 7];
      R fl = k < 3 ? (R)1e-3f : (R)1e-2f;
      d[k * 7] = cur + lam_d * std::max(std::fabs(cur), fl);
    }
  }
  // compute_preconditioner_kernel
  for (int ni = 0; ni < N; ++ni) {
    R dg[36] = {};
    for (int q = rp[ni]; q < rp[ni + 1]; ++q)
      if (ci[q] == ni) {
        for (int k = 0; k < 36; ++k) dg[k] = A[(long)q * 36 + k];
        break;
      }
    auto inv6 = [](auto *D, auto *Out) -> bool {
      using S = std::remove_reference_t<decltype(D[0])>;
      S L[6][6] = {};
      for (int i = 0; i < 6; ++i)
        for (int j = 0; j <= i; ++j) {
          S s = D[i * 6 + j];
          for (int k = 0; k < j; ++k) s -= L[i][k] * L[j][k];
          if (i == j) {
            if (s <= 0) return false;
            L[i][j] = std::sqrt(s);
          } else {
            L[i][j] = s / L[j][j];
          }
        }
      for (int col = 0; col < 6; ++col) {
        S y[6] = {}, x[6] = {};
        for (int i = 0; i < 6; ++i) {
          S s = (i == col) ? (S)1 : (S)0;
          for (int k = 0; k < i; ++k) s -= L[i][k] * y[k];
          y[i] = s / L[i][i];
        }
        for (int i = 5; i >= 0; --i) {
          S s = y[i];
          for (int k = i + 1; k < 6; ++k) s -= L[k][i] * x[k];
          x[i] = s / L[i][i];
        }
        for (int i = 0; i < 6; ++i) Out[i * 6 + col] = x[i];
      }
      return true;
    };
    R *out = Minv + (long)ni * 36;
    if (precond_fp64 && sizeof(R) == 4) {
      double d64[36], o64[36];
      for (int k = 0; k < 36; ++k) d64[k] = dg[k];
      for (int k = 0; k < 6; ++k) d64[k * 7] += (double)1e-4f;
      if (inv6(d64, o64)) {
        for (int k = 0; k < 36; ++k) out[k] = (R)o64[k];
      } else {
        for (int k = 0; k < 36; ++k) out[k] = 0;
        for (int k = 0; k < 6; ++k) out[k * 7] = (R)(1.0 / (d64[k * 7] + (double)1e-6f));
      }
      continue;
    }
    for (int k = 0; k < 6; ++k) dg[k * 7] += (R)1e-4f;
    if (!inv6(dg, out)) {
      for (int k = 0; k < 36; ++k) out[k] = 0;
      for (int k = 0; k < 6; ++k) out[k * 7] = (R)1 / (dg[k * 7] + (R)1e-6f);
    }
  }
}

// bsr_spmv_kernel (:406-432), apply_precond_kernel (:438-461).
template <typename R> void spmv(const int32_t *rp, const int32_t *ci, const R *A, const R *x, R *y, int N) {
#pragma omp parallel for schedule(static)
  for (int g = 0; g < 6 * N; ++g) {
    int ni = g / 6, row = g % 6;
    R sum = 0;
    for (int bk = rp[ni]; bk < rp[ni + 1]; ++bk) {
      const R *blk = A + (long)bk * 36;
      const R *xj = x + ci[bk] * 6;
      for (int c = 0; c < 6; ++c) sum += blk[row * 6 + c] * xj[c];
    }
    y[g] = sum;
  }
}
template <typename R> void precond(const R *M, const R *r, R *z, int N) {
  for (int ni = 0; ni < N; ++ni)
    for (int row = 0; row < 6; ++row) {
      R acc = 0;
      for (int c = 0; c < 6; ++c) acc += M[(long)ni * 36 + row * 6 + c] * r[ni * 6 + c];
      z[ni * 6 + row] = acc;
    }
}
template <typename R> R dotv(const R *a, const R *b, int n) {
  R s = 0;
  for (int i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

// pcg_solve: kernels/solver_kernels.cu:616-687 with x0 = 0.  tol is absolute
// unless rel (then tol * ||b||, and the breakdown test is p.Ap <= 0).
template <typename R>
R pcg(const int32_t *rp, const int32_t *ci, const R *A, const R *b, const R *M, int N, int maxit, R tol, int rel,
      R *x, int *iters, int *flags, std::vector<double> *trace) {
  const int n = 6 * N;
  std::vector<R> r(b, b + n), z(n), p(n), Ap(n);
  std::fill(x, x + n, (R)0);
  precond(M, r.data(), z.data(), N);
  p = z;
  R rz = dotv(r.data(), z.data(), n);
  R bn = std::sqrt(dotv(b, b, n));
  R tl = rel ? tol * bn : tol;
  int used = 0, fl = 0;
  for (int it = 0; it < maxit; ++it) {
    spmv(rp, ci, A, p.data(), Ap.data(), N);
    R pAp = dotv(p.data(), Ap.data(), n);
    // reference guard |p.Ap| < 1e-12; with the relative rule only p.Ap <= 0
    if (rel ? !(pAp > (R)0) : std::fabs(pAp) < (R)1e-12f) { fl |= 1; used = it; break; }
    R alpha = rz / pAp;
    for (int i = 0; i < n; ++i) x[i] += alpha * p[i];
    for (int i = 0; i < n; ++i) r[i] += -alpha * Ap[i];
    R rn = std::sqrt(dotv(r.data(), r.data(), n));
    if (trace) trace->push_back((double)rn);
    used = it + 1;
    if (rn < tl) { fl |= 2; break; }
    precond(M, r.data(), z.data(), N);
    R rzn = dotv(r.data(), z.data(), n);
    R beta = rzn / rz;
    for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    rz = rzn;
  }
  if (iters) *iters = used;
  if (flags) *flags = fl;
  return std::sqrt(dotv(r.data(), r.data(), n));
}

// Caller GN step: validation (include/fusion_app.h:409-491), scaled + clamped
// update (fusion_app.h:1353-1375), apply_twist_kernel (warp_field.cu:250-291).
template <typename R>
bool validate_step(const R *dx, int N, const nrf_gn_params &g) {
  R rs = 0, ts = 0, ers = 0, ets = 0, rm = 0, tm = 0, erm = 0, etm = 0;
  bool fin = true;
  for (int i = 0; i < N; ++i) {
    const R *d = dx + i * 6;
    R rn = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    R tn = std::sqrt(d[3] * d[3] + d[4] * d[4] + d[5] * d[5]);
    fin = fin && std::isfinite(rn) && std::isfinite(tn);
    rs += rn; ts += tn; rm = std::max(rm, rn); tm = std::max(tm, tn);
    R er = (R)g.update_scale * rn, et = (R)g.update_scale * tn;
    if (g.max_update_rot > 0.0f) er = std::min(er, (R)g.max_update_rot);
    if (g.max_update_trans > 0.0f) et = std::min(et, (R)g.max_update_trans);
    ers += er; ets += et; erm = std::max(erm, er); etm = std::max(etm, et);
  }
  R n = (R)std::max(1, N);
  R erm_mean = ers / n, etm_mean = ets / n;
  R rej_r = g.max_update_rot > 0.0f ? (R)5 * (R)g.max_update_rot : std::numeric_limits<R>::infinity();
  R rej_t = g.max_update_trans > 0.0f ? (R)5 * (R)g.max_update_trans : std::numeric_limits<R>::infinity();
  bool spike = ((R)g.update_scale * rm > rej_r) || ((R)g.update_scale * tm > rej_t);
  bool mean_ok = (g.max_dx_mean <= 0.0f) || (erm_mean <= (R)g.max_dx_mean && etm_mean <= (R)g.max_dx_mean);
  return fin && std::isfinite(erm_mean) && std::isfinite(etm_mean) && std::isfinite(erm) && std::isfinite(etm) &&
         erm < rej_r && etm < rej_t && mean_ok && !spike;
}

template <typename R> DQ<R> twist_update(const R dx[6], const DQ<R> &T, R max_rot, R max_trans, R scale) {
  R tw[6];
  for (int k = 0; k < 6; ++k) tw[k] = scale * dx[k];
  R rn = std::sqrt(tw[0] * tw[0] + tw[1] * tw[1] + tw[2] * tw[2]);
  if (max_rot > 0 && rn > max_rot) { R f = max_rot / rn; tw[0] *= f; tw[1] *= f; tw[2] *= f; }
  R tn = std::sqrt(tw[3] * tw[3] + tw[4] * tw[4] + tw[5] * tw[5]);
  if (max_trans > 0 && tn > max_trans) { R f = max_trans / tn; tw[3] *= f; tw[4] *= f; tw[5] *= f; }
  R M[3][4];
  exp_se3(tw, M);
  return dq_mul(mat_to_dq(M), T);
}

template <typename R>
void apply_step(const R *dx, DQ<R> *T, int N, const nrf_gn_params &g) {
  for (int i = 0; i < N; ++i) {
    R d[6];
    for (int k = 0; k < 6; ++k) d[k] = dx[i * 6 + k];
    R rn = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    R tn = std::sqrt(d[3] * d[3] + d[4] * d[4] + d[5] * d[5]);
    R rsc = 1, tsc = 1;
    if (g.max_update_rot > 0.0f && rn > (R)g.max_update_rot) rsc = (R)g.max_update_rot / rn;
    if (g.max_update_trans > 0.0f && tn > (R)g.max_update_trans) tsc = (R)g.max_update_trans / tn;
    for (int k = 0; k < 3; ++k) d[k] *= rsc * (R)g.update_scale;
    for (int k = 3; k < 6; ++k) d[k] *= tsc * (R)g.update_scale;
    T[i] = twist_update(d, T[i], (R)g.max_update_rot, (R)g.max_update_trans, (R)1);
  }
}

template <typename R>
int solve_gn(const nrf_corr *corrs, const float *canonical, int n, const nrf_node *nodes, int N, DQ<R> *T,
             const nrf_gn_params &g, const nrf_solver_params &sp, R *dx_out, int *total_iters, int *gn_done,
             int *last_iters) {
  std::vector<int32_t> rp, ci;
  pattern(nodes, N, rp, ci);
  const int nb = rp[N];
  std::vector<R> A((long)nb * 36), b(6 * N), M((long)N * 36), x(6 * N);
  int tot = 0, done = 0, li = 0;
  for (int it = 0; it < std::max(1, g.gn_iterations); ++it) {
    assemble(corrs, n, canonical, nodes, T, N, (R)sp.lambda_smooth, (R)sp.lambda_damping, rp.data(), ci.data(),
             A.data(), b.data(), M.data(), sp.precond_fp64, sp.arap_consistent_sign, sp.psd_dropped_pairs);
    int iters = 0, flags = 0;
    pcg(rp.data(), ci.data(), A.data(), b.data(), M.data(), N, sp.pcg_iterations, (R)sp.pcg_tolerance,
        sp.relative_tolerance, x.data(), &iters, &flags, nullptr);
    tot += iters;
    li = iters;
    if (sp.refine_iterations > 0 || sp.refine_jacobi) {
      // the solver's optional iterative-refinement step (no reference
      // counterpart): r = b - A x, then the correction d from A d = r, by
      // refine_iterations PCG iterations from d = 0 with the same stopping
      // rules, or d = M^-1 r (refine_jacobi); x += d
      std::vector<R> r(6 * N), d(6 * N);
      spmv(rp.data(), ci.data(), A.data(), x.data(), r.data(), N);
      for (int k = 0; k < 6 * N; ++k) r[k] = b[k] - r[k];
      if (sp.refine_jacobi) {
        precond(M.data(), r.data(), d.data(), N);
      } else {
        int it2 = 0, fl2 = 0;
        pcg(rp.data(), ci.data(), A.data(), r.data(), M.data(), N, sp.refine_iterations, (R)sp.pcg_tolerance,
            sp.relative_tolerance, d.data(), &it2, &fl2, nullptr);
      }
      for (int k = 0; k < 6 * N; ++k) x[k] += d[k];
    }
    if (g.validate && !validate_step(x.data(), N, g)) break;
    apply_step(x.data(), T, N, g);
    ++done;
  }
  if (dx_out) std::copy(x.begin(), x.end(), dx_out);
  if (total_iters) *total_iters = tot;
  if (gn_done) *gn_done = done;
  if (last_iters) *last_iters = li;
  return 0;
}

template <typename R> std::vector<DQ<R>> load_T(const nrf_dualquat *T, int N) {
  std::vector<DQ<R>> o(N);
  for (int i = 0; i < N; ++i) o[i] = load_dq<R>(T[i]);
  return o;
}

}  // namespace orc

using namespace orc;

// ======================================================================
// C entry points (ctypes).  f32 = reference FP32 order, f64 = FP64 oracle.
// Transforms in FP64 are arrays of 8 doubles (real xyzw, dual xyzw).
extern "C" {

int orc_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
}

int orc_load_mc_table(const char *path) { return load_tables(path); }

void orc_exp_se3_f64(const double tw[6], double out[12]) {
  double T[3][4];
  exp_se3(tw, T);
  std::memcpy(out, T, sizeof(T));
}
void orc_exp_se3_f32(const float tw[6], float out[12]) {
  float T[3][4];
  exp_se3(tw, T);
  std::memcpy(out, T, sizeof(T));
}
void orc_mat_to_dq_f64(const double m[12], double out[8]) {
  double T[3][4];
  std::memcpy(T, m, sizeof(T));
  store_dq8(out, mat_to_dq(T));
}
void orc_dq_to_mat_f64(const double q[8], double out[12]) {
  double T[3][4];
  dq_to_mat(load_dq8(q), T);
  std::memcpy(out, T, sizeof(T));
}
void orc_dq_mul_f64(const double a[8], const double b[8], double out[8]) {
  store_dq8(out, dq_mul(load_dq8(a), load_dq8(b)));
}
void orc_dq_centered_f64(const double q[8], const double c[3], double out[8]) {
  store_dq8(out, dq_centered(load_dq8(q), V3<double>{c[0], c[1], c[2]}));
}
void orc_dq_transform_f64(const double q[8], const double p[3], double out[3]) {
  V3<double> r = dq_xf(load_dq8(q), V3<double>{p[0], p[1], p[2]});
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void orc_twist_update_f64(const double dx[6], const double T[8], double max_rot, double max_trans, double scale,
                          double out[8]) {
  store_dq8(out, twist_update(dx, load_dq8(T), max_rot, max_trans, scale));
}
void orc_twist_update_f32(const float dx[6], const nrf_dualquat *T, float max_rot, float max_trans, float scale,
                          nrf_dualquat *out) {
  DQ<float> r = twist_update(dx, load_dq<float>(*T), max_rot, max_trans, scale);
  out->real[0] = r.r.x; out->real[1] = r.r.y; out->real[2] = r.r.z; out->real[3] = r.r.w;
  out->dual[0] = r.d.x; out->dual[1] = r.d.y; out->dual[2] = r.d.z; out->dual[3] = r.d.w;
}

void orc_warp_points_f64(const float *pts, const int32_t *ids, const float *ws, int count, const nrf_node *nodes,
                         const double *T8, int N, double *out) {
  std::vector<DQ<double>> T(N);
  for (int i = 0; i < N; ++i) T[i] = load_dq8(T8 + 8 * i);
  auto cen = centered_table(nodes, T.data(), N);
  for (int i = 0; i < count; ++i) {
    V3<double> p = warp_point(v3f<double>(pts + 3 * i), cen.data(), N, ids + 4 * i, ws + 4 * i);
    out[3 * i] = p.x; out[3 * i + 1] = p.y; out[3 * i + 2] = p.z;
  }
}
void orc_warp_points_f32(const float *pts, const int32_t *ids, const float *ws, int count, const nrf_node *nodes,
                         const nrf_dualquat *Tq, int N, float *out) {
  auto T = load_T<float>(Tq, N);
  auto cen = centered_table(nodes, T.data(), N);
  for (int i = 0; i < count; ++i) {
    V3<float> p = warp_point(v3f<float>(pts + 3 * i), cen.data(), N, ids + 4 * i, ws + 4 * i);
    out[3 * i] = p.x; out[3 * i + 1] = p.y; out[3 * i + 2] = p.z;
  }
}

void orc_warp_normals_f32(const float *nin, const int32_t *ids, const float *ws, int count,
                          const nrf_dualquat *Tq, int N, float *out) {
  auto T = load_T<float>(Tq, N);
  for (int i = 0; i < count; ++i) {
    V3<float> r = warp_normal(v3f<float>(nin + 3 * i), T.data(), N, ids + 4 * i, ws + 4 * i);
    out[3 * i] = r.x; out[3 * i + 1] = r.y; out[3 * i + 2] = r.z;
  }
}
void orc_warp_normals_f64(const float *nin, const int32_t *ids, const float *ws, int count, const double *T8,
                          int N, double *out) {
  std::vector<DQ<double>> T(N);
  for (int i = 0; i < N; ++i) T[i] = load_dq8(T8 + 8 * i);
  for (int i = 0; i < count; ++i) {
    V3<double> r = warp_normal(v3f<double>(nin + 3 * i), T.data(), N, ids + 4 * i, ws + 4 * i);
    out[3 * i] = r.x; out[3 * i + 1] = r.y; out[3 * i + 2] = r.z;
  }
}

// Distance of each voxel's FP64 projection from the rounding boundaries of
// the integration test (for parity checks of FP32 kernels): px = distance of
// (u, v) from the nearest half-integer (where round-half-even flips), for
// voxels projecting into the image or its 1-pixel border; sdf =
// |sdf + truncation| / truncation at the sampled pixel; +inf elsewhere.
void orc_integrate_margin_f64(const nrf_volume_params *vp, const float *depth, const nrf_camera *cam,
                              const float pose[16], const nrf_node *nodes, const double *T8, int N,
                              const int32_t *knn, const float *knn_w, double *m_px, double *m_sdf) {
  const int X = vp->dims[0], Y = vp->dims[1], Z = vp->dims[2];
  const bool warp = nodes != nullptr && N > 0;
  float Tc[3][4];
  invert_pose(pose, Tc);
  std::vector<DQ<double>> T(N), cen;
  for (int i = 0; i < N; ++i) T[i] = load_dq8(T8 + 8 * i);
  if (warp) cen = centered_table(nodes, T.data(), N);
  const double s = vp->voxel_size, tr = vp->truncation;
#pragma omp parallel for schedule(static)
  for (int z = 0; z < Z; ++z)
    for (int y = 0; y < Y; ++y)
      for (int x = 0; x < X; ++x) {
        const long idx = ((long)z * Y + y) * X + x;
        m_px[idx] = m_sdf[idx] = INFINITY;
        V3<double> p = {(double)vp->origin[0] + (x + 0.5) * s, (double)vp->origin[1] + (y + 0.5) * s,
                        (double)vp->origin[2] + (z + 0.5) * s};
        if (warp) p = warp_point(p, cen.data(), N, knn + idx * 4, knn_w + idx * 4);
        double cx = Tc[0][0] * p.x + Tc[0][1] * p.y + Tc[0][2] * p.z + Tc[0][3];
        double cy = Tc[1][0] * p.x + Tc[1][1] * p.y + Tc[1][2] * p.z + Tc[1][3];
        double cz = Tc[2][0] * p.x + Tc[2][1] * p.y + Tc[2][2] * p.z + Tc[2][3];
        if (cz <= 0) continue;
        double fu = cam->fx * cx / cz + cam->cx, fv = cam->fy * cy / cz + cam->cy;
        if (!(fu > -1.5 && fu < cam->width + 0.5 && fv > -1.5 && fv < cam->height + 0.5)) continue;
        m_px[idx] = std::min(std::fabs(fu - std::floor(fu) - 0.5), std::fabs(fv - std::floor(fv) - 0.5));
        int u = (int)std::nearbyint(fu), v = (int)std::nearbyint(fv);
        if (u < 0 || u >= cam->width || v < 0 || v >= cam->height) continue;
        double d = depth[(long)v * cam->width + u];
        if (d <= 0) continue;
        m_sdf[idx] = std::fabs(d - cz + tr) / tr;
      }
}

long orc_integrate_f32(const nrf_volume_params *vp, nrf_voxel *vox, const float *depth, const nrf_camera *cam,
                       const float pose[16], const nrf_node *nodes, const nrf_dualquat *Tq, int N,
                       const int32_t *knn, const float *knn_w, const int32_t *opt, int min_opt) {
  const long V = (long)vp->dims[0] * vp->dims[1] * vp->dims[2];
  std::vector<float> t(V), w(V);
  for (long i = 0; i < V; ++i) { t[i] = vox[i].tsdf; w[i] = vox[i].weight; }
  auto T = load_T<float>(Tq, N);
  long u = integrate(*vp, t.data(), w.data(), depth, *cam, pose, nodes, T.data(), N, knn, knn_w, opt, min_opt);
  for (long i = 0; i < V; ++i) { vox[i].tsdf = t[i]; vox[i].weight = w[i]; }
  return u;
}
long orc_integrate_f64(const nrf_volume_params *vp, double *tsdf, double *weight, const float *depth,
                       const nrf_camera *cam, const float pose[16], const nrf_node *nodes, const double *T8, int N,
                       const int32_t *knn, const float *knn_w, const int32_t *opt, int min_opt) {
  std::vector<DQ<double>> T(N);
  for (int i = 0; i < N; ++i) T[i] = load_dq8(T8 + 8 * i);
  return integrate(*vp, tsdf, weight, depth, *cam, pose, nodes, T.data(), N, knn, knn_w, opt, min_opt);
}

long orc_mc_f32(const nrf_volume_params *vp, const nrf_voxel *vox, uint8_t *cases, float *verts, float *normals,
                long cap) {
  if (!g_tables) return -1;
  const long V = (long)vp->dims[0] * vp->dims[1] * vp->dims[2];
  std::vector<float> t(V), w(V);
  for (long i = 0; i < V; ++i) { t[i] = vox[i].tsdf; w[i] = vox[i].weight; }
  return marching_cubes(*vp, t.data(), w.data(), cases, verts, normals, cap);
}
long orc_mc_f64(const nrf_volume_params *vp, const double *tsdf, const double *weight, uint8_t *cases,
                double *verts, double *normals, long cap) {
  if (!g_tables) return -1;
  return marching_cubes(*vp, tsdf, weight, cases, verts, normals, cap);
}

void orc_knn_f32(const nrf_volume_params *vp, const nrf_voxel *vox, const nrf_node *nodes, int N, int32_t *knn,
                 float *knn_w) {
  const long V = (long)vp->dims[0] * vp->dims[1] * vp->dims[2];
  std::vector<float> w(V);
  for (long i = 0; i < V; ++i) w[i] = vox[i].weight;
  voxel_knn(*vp, w.data(), nodes, N, knn, knn_w);
}
void orc_knn_f64(const nrf_volume_params *vp, const nrf_voxel *vox, const nrf_node *nodes, int N, int32_t *knn,
                 double *knn_w) {
  const long V = (long)vp->dims[0] * vp->dims[1] * vp->dims[2];
  std::vector<float> w(V);
  for (long i = 0; i < V; ++i) w[i] = vox[i].weight;
  voxel_knn(*vp, w.data(), nodes, N, knn, knn_w);
}

int orc_pattern(const nrf_node *nodes, int N, int32_t *row_ptr, int32_t *col_idx, int cap) {
  std::vector<int32_t> rp, ci;
  pattern(nodes, N, rp, ci);
  if (row_ptr) std::copy(rp.begin(), rp.end(), row_ptr);
  if (col_idx && (int)ci.size() <= cap) std::copy(ci.begin(), ci.end(), col_idx);
  return (int)ci.size();
}

int orc_assemble_f32(const nrf_corr *corrs, int n, const float *canonical, const nrf_node *nodes,
                     const nrf_dualquat *Tq, int N, float lam_s, float lam_d, int precond_fp64, float *A, float *b,
                     float *M, int consistent_sign, int psd_pairs) {
  std::vector<int32_t> rp, ci;
  pattern(nodes, N, rp, ci);
  auto T = load_T<float>(Tq, N);
  assemble(corrs, n, canonical, nodes, T.data(), N, lam_s, lam_d, rp.data(), ci.data(), A, b, M, precond_fp64,
           consistent_sign, psd_pairs);
  return rp[N];
}
// assemble_system with a given pattern (rp, ci), as after a graph change
// that kept the node count (Q4: the pattern is not rebuilt)
void orc_assemble_pattern_f32(const nrf_corr *corrs, int n, const float *canonical, const nrf_node *nodes,
                              const nrf_dualquat *Tq, int N, float lam_s, float lam_d, const int32_t *rp,
                              const int32_t *ci, float *A, float *b, float *M, int consistent_sign) {
  auto T = load_T<float>(Tq, N);
  assemble(corrs, n, canonical, nodes, T.data(), N, lam_s, lam_d, rp, ci, A, b, M, 0, consistent_sign, 0);
}
int orc_assemble_f64(const nrf_corr *corrs, int n, const float *canonical, const nrf_node *nodes,
                     const double *T8, int N, double lam_s, double lam_d, double *A, double *b, double *M,
                     int consistent_sign, int psd_pairs) {
  std::vector<int32_t> rp, ci;
  pattern(nodes, N, rp, ci);
  std::vector<DQ<double>> T(N);
  for (int i = 0; i < N; ++i) T[i] = load_dq8(T8 + 8 * i);
  assemble(corrs, n, canonical, nodes, T.data(), N, lam_s, lam_d, rp.data(), ci.data(), A, b, M, 1,
           consistent_sign, psd_pairs);
  return rp[N];
}

double orc_pcg_f32(const int32_t *rp, const int32_t *ci, const float *A, const float *b, const float *M, int N,
                   int maxit, float tol, int rel, float *x, int *iters, int *flags, double *trace) {
  std::vector<double> tr;
  float r = pcg(rp, ci, A, b, M, N, maxit, tol, rel, x, iters, flags, trace ? &tr : nullptr);
  if (trace) std::copy(tr.begin(), tr.end(), trace);
  return r;
}
double orc_pcg_f64(const int32_t *rp, const int32_t *ci, const double *A, const double *b, const double *M, int N,
                   int maxit, double tol, int rel, double *x, int *iters, int *flags, double *trace) {
  std::vector<double> tr;
  double r = pcg(rp, ci, A, b, M, N, maxit, tol, rel, x, iters, flags, trace ? &tr : nullptr);
  if (trace) std::copy(tr.begin(), tr.end(), trace);
  return r;
}
void orc_spmv_f64(const int32_t *rp, const int32_t *ci, const double *A, const double *x, double *y, int N) {
  spmv(rp, ci, A, x, y, N);
}
// skin_vertices_from_voxel_knn_kernel (src/skinning_kernels.cu:29-63): each
// vertex warped (warp_point) through the 4-NN of the voxel containing it,
// unchanged outside the volume or without nodes.
void orc_skin_vertices_f32(const float *canon, int n, const nrf_volume_params *vp, const nrf_node *nodes,
                           const nrf_dualquat *Tq, int N, const int32_t *knn, const float *knn_w, float *out) {
  std::vector<DQ<float>> cen;
  if (N > 0) {
    auto T = load_T<float>(Tq, N);
    cen = centered_table(nodes, T.data(), N);
  }
  for (int i = 0; i < n; ++i) {
    V3<float> p = v3f<float>(canon + 3 * (long)i);
    if (N > 0) {
      const int vx = (int)std::floor((p.x - vp->origin[0]) / vp->voxel_size);
      const int vy = (int)std::floor((p.y - vp->origin[1]) / vp->voxel_size);
      const int vz = (int)std::floor((p.z - vp->origin[2]) / vp->voxel_size);
      if (vx >= 0 && vx < vp->dims[0] && vy >= 0 && vy < vp->dims[1] && vz >= 0 && vz < vp->dims[2]) {
        const long v = ((long)vz * vp->dims[1] + vy) * vp->dims[0] + vx;
        p = warp_point(p, cen.data(), N, knn + 4 * v, knn_w + 4 * v);
      }
    }
    out[3 * i] = p.x; out[3 * i + 1] = p.y; out[3 * i + 2] = p.z;
  }
}
// raycast_kernel (kernels/tsdf_kernels.cu:236-361) with
// sample_tsdf_trilinear (:14-64), FP32 in the reference's order: the ray of
// integer pixel (u, v) marched from t = 0.05 to 3.5 m (0.8 voxel steps, 0.5
// once |tsdf| < 0.5), the first + to - crossing (previous sample in (0, 1))
// interpolated linearly, the normal from six trilinear samples one voxel
// apart, warped through the crossing's base-voxel 4-NN when N > 0.
static bool ray_sample(const nrf_voxel *vox, const nrf_volume_params &vp, float px, float py, float pz, float &tsdf,
                       int &base) {
  const float gx = (px - vp.origin[0]) / vp.voxel_size, gy = (py - vp.origin[1]) / vp.voxel_size,
              gz = (pz - vp.origin[2]) / vp.voxel_size;
  auto fl = [](float v) { return (int)std::fmin(std::fmax(std::floor(v), -1073741824.f), 1073741824.f); };
  const int x0 = fl(gx), y0 = fl(gy), z0 = fl(gz);
  const int X = vp.dims[0], Y = vp.dims[1], Z = vp.dims[2];
  if (x0 < 0 || x0 >= X - 1 || y0 < 0 || y0 >= Y - 1 || z0 < 0 || z0 >= Z - 1) return false;
  const float fx = gx - (float)x0, fy = gy - (float)y0, fz = gz - (float)z0;
  float acc = 0.0f;
  for (int dz = 0; dz <= 1; ++dz) {
    const float wz = dz ? fz : (1.0f - fz);
    for (int dy = 0; dy <= 1; ++dy) {
      const float wy = dy ? fy : (1.0f - fy);
      for (int dx = 0; dx <= 1; ++dx) {
        const float wx = dx ? fx : (1.0f - fx);
        const nrf_voxel &v = vox[((long)(z0 + dz) * Y + (y0 + dy)) * X + (x0 + dx)];
        if (v.weight <= 0.0f) return false;
        acc += wx * wy * wz * v.tsdf;
      }
    }
  }
  tsdf = acc;
  base = (z0 * Y + y0) * X + x0;
  return true;
}

void orc_raycast_f32(const nrf_volume_params *vp, const nrf_voxel *vox, const nrf_camera *cam, const float pose[16],
                     const nrf_node *nodes, const nrf_dualquat *Tq, int N, const int32_t *knn, const float *knn_w,
                     float *verts, float *normals, int32_t *vidx_out) {
  std::vector<DQ<float>> T, cen;
  if (N > 0) {
    T = load_T<float>(Tq, N);
    cen = centered_table(nodes, T.data(), N);
  }
  const float s = vp->voxel_size;
  auto nrm = [](float x, float y, float z) {
    const float n = std::sqrt(x * x + y * y + z * z);
    return n > 1e-9f ? V3<float>{x / n, y / n, z / n} : V3<float>{x, y, z};
  };
#pragma omp parallel for schedule(dynamic, 4)
  for (int v = 0; v < cam->height; ++v)
    for (int u = 0; u < cam->width; ++u) {
      const long px = (long)v * cam->width + u;
      V3<float> ov{0, 0, 0}, on{0, 0, 0};
      int oi = -1;
      const float dcx = 1.0f * ((float)u - cam->cx) / cam->fx, dcy = 1.0f * ((float)v - cam->cy) / cam->fy,
                  dcz = 1.0f;
      const float ox = pose[0] * 0.f + pose[1] * 0.f + pose[2] * 0.f + pose[3];
      const float oy = pose[4] * 0.f + pose[5] * 0.f + pose[6] * 0.f + pose[7];
      const float oz = pose[8] * 0.f + pose[9] * 0.f + pose[10] * 0.f + pose[11];
      const V3<float> d = nrm(pose[0] * dcx + pose[1] * dcy + pose[2] * dcz, pose[4] * dcx + pose[5] * dcy + pose[6] * dcz,
                              pose[8] * dcx + pose[9] * dcy + pose[10] * dcz);
      float t = 0.05f, step = s * 0.8f, prev = 0.f, t_prev = t;
      bool have_prev = false;
      for (; t < 3.5f; t += step) {
        float cur = 1.0f;
        int vi = -1;
        if (!ray_sample(vox, *vp, ox + t * d.x, oy + t * d.y, oz + t * d.z, cur, vi)) {
          have_prev = false;
          continue;
        }
        if (have_prev && prev > 0 && cur <= 0 && prev < 1.0f) {
          const float ts = t - (t - t_prev) * cur / (cur - prev);
          const float sx = ox + ts * d.x, sy = oy + ts * d.y, sz = oz + ts * d.z;
          float g[6];
          int un = -1;
          const bool ok = ray_sample(vox, *vp, sx + s, sy, sz, g[0], un) && ray_sample(vox, *vp, sx - s, sy, sz, g[1], un) &&
                          ray_sample(vox, *vp, sx, sy + s, sz, g[2], un) && ray_sample(vox, *vp, sx, sy - s, sz, g[3], un) &&
                          ray_sample(vox, *vp, sx, sy, sz + s, g[4], un) && ray_sample(vox, *vp, sx, sy, sz - s, g[5], un);
          if (!ok) break;
          V3<float> n = nrm(g[0] - g[1], g[2] - g[3], g[4] - g[5]);
          oi = vi;
          V3<float> p{sx, sy, sz};
          if (N > 0 && nodes) {
            p = warp_point(p, cen.data(), N, knn + 4 * (long)vi, knn_w + 4 * (long)vi);
            n = warp_normal(n, T.data(), N, knn + 4 * (long)vi, knn_w + 4 * (long)vi);
          }
          ov = p;
          on = n;
          break;
        }
        prev = cur;
        have_prev = true;
        t_prev = t;
        step = std::fabs(cur) < 0.5f ? s * 0.5f : s * 0.8f;
      }
      verts[3 * px] = ov.x; verts[3 * px + 1] = ov.y; verts[3 * px + 2] = ov.z;
      normals[3 * px] = on.x; normals[3 * px + 1] = on.y; normals[3 * px + 2] = on.z;
      if (vidx_out) vidx_out[px] = oi;
    }
}

// upload_scaled_delta_x(scale) (include/fusion_app.h:1353-1375), then
// apply_twist_increment(trial, max_rot, max_trans, 1) (src/warp_field.cu:
// 1121-1141) on every node, in place.
void orc_apply_step_f32(const float *dx, nrf_dualquat *Tq, int N, const nrf_gn_params *g, float scale) {
  nrf_gn_params gs = *g;
  gs.update_scale = scale;
  auto T = load_T<float>(Tq, N);
  apply_step(dx, T.data(), N, gs);
  for (int i = 0; i < N; ++i) {
    Tq[i].real[0] = T[i].r.x; Tq[i].real[1] = T[i].r.y; Tq[i].real[2] = T[i].r.z; Tq[i].real[3] = T[i].r.w;
    Tq[i].dual[0] = T[i].d.x; Tq[i].dual[1] = T[i].d.y; Tq[i].dual[2] = T[i].d.z; Tq[i].dual[3] = T[i].d.w;
  }
}
// Step validation of the caller (include/fusion_app.h:409-491): 1 = accept.
int orc_validate_f32(const float *dx, int N, const nrf_gn_params *g) { return validate_step(dx, N, *g) ? 1 : 0; }
int orc_validate_f64(const double *dx, int N, const nrf_gn_params *g) { return validate_step(dx, N, *g) ? 1 : 0; }

int orc_solve_gn_f32(const nrf_corr *corrs, const float *canonical, int n, const nrf_node *nodes, int N,
                     nrf_dualquat *Tq, const nrf_gn_params *g, const nrf_solver_params *sp, float *dx,
                     int *total_iters, int *gn_done, int *last_iters) {
  auto T = load_T<float>(Tq, N);
  int rc = solve_gn(corrs, canonical, n, nodes, N, T.data(), *g, *sp, dx, total_iters, gn_done, last_iters);
  for (int i = 0; i < N; ++i) {
    Tq[i].real[0] = T[i].r.x; Tq[i].real[1] = T[i].r.y; Tq[i].real[2] = T[i].r.z; Tq[i].real[3] = T[i].r.w;
    Tq[i].dual[0] = T[i].d.x; Tq[i].dual[1] = T[i].d.y; Tq[i].dual[2] = T[i].d.z; Tq[i].dual[3] = T[i].d.w;
  }
  return rc;
}
int orc_solve_gn_f64(const nrf_corr *corrs, const float *canonical, int n, const nrf_node *nodes, int N,
                     double *T8, const nrf_gn_params *g, const nrf_solver_params *sp, double *dx, int *total_iters,
                     int *gn_done, int *last_iters) {
  std::vector<DQ<double>> T(N);
  for (int i = 0; i < N; ++i) T[i] = load_dq8(T8 + 8 * i);
  int rc = solve_gn(corrs, canonical, n, nodes, N, T.data(), *g, *sp, dx, total_iters, gn_done, last_iters);
  for (int i = 0; i < N; ++i) store_dq8(T8 + 8 * i, T[i]);
  return rc;
}

}  // extern "C"
