// types.h - the reference's POD types and device buffer, for HIP / gfx950.
//
// Same names, members and byte layouts as include/types.h of the reference
// (:26-246), so code written against the reference's TSDFVolume / WarpField /
// GaussNewtonSolver compiles against this library with hipcc.  Device errors
// abort the process (the reference's CUDA_CHECK behaviour, types.h:14-20)
// through NRF_DEVICE_CHECK.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "nrf.h"

#define NRF_DEVICE_CHECK(call)                                                              \
  do {                                                                                      \
    hipError_t err_ = (call);                                                               \
    if (err_ != hipSuccess) {                                                               \
      fprintf(stderr, "HIP error at %s:%d: %s\n", __FILE__, __LINE__, hipGetErrorString(err_)); \
      exit(1);                                                                              \
    }                                                                                       \
  } while (0)

// Status of a C ABI call: abort like the reference on failure.
#define NRF_STATUS_CHECK(call)                                                              \
  do {                                                                                      \
    int st_ = (call);                                                                       \
    if (st_ != NRF_OK) {                                                                    \
      fprintf(stderr, "nrf error at %s:%d: %s (%s)\n", __FILE__, __LINE__,                  \
              nrf_status_string(st_), nrf_last_error());                                    \
      exit(1);                                                                              \
    }                                                                                       \
  } while (0)

struct Mat4 {
  float m[4][4];
  __host__ __device__ Mat4() {
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) m[i][j] = (i == j) ? 1.f : 0.f;
  }
  __host__ __device__ static Mat4 identity() { return Mat4(); }
  __host__ __device__ float3 transform_point(float3 p) const {
    return make_float3(m[0][0] * p.x + m[0][1] * p.y + m[0][2] * p.z + m[0][3],
                       m[1][0] * p.x + m[1][1] * p.y + m[1][2] * p.z + m[1][3],
                       m[2][0] * p.x + m[2][1] * p.y + m[2][2] * p.z + m[2][3]);
  }
  __host__ __device__ float3 transform_normal(float3 n) const {
    return make_float3(m[0][0] * n.x + m[0][1] * n.y + m[0][2] * n.z, m[1][0] * n.x + m[1][1] * n.y + m[1][2] * n.z,
                       m[2][0] * n.x + m[2][1] * n.y + m[2][2] * n.z);
  }
  __host__ __device__ Mat4 operator*(const Mat4 &o) const {
    Mat4 r;
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) {
        r.m[i][j] = 0;
        for (int k = 0; k < 4; k++) r.m[i][j] += m[i][k] * o.m[k][j];
      }
    return r;
  }
};

struct DualQuat {
  float4 real;  // rotation quaternion (x, y, z, w)
  float4 dual;  // dual part, translation t = 2 dual real*
  __host__ __device__ DualQuat() : real{0.f, 0.f, 0.f, 1.f}, dual{0.f, 0.f, 0.f, 0.f} {}
  __host__ __device__ static DualQuat identity() { return DualQuat(); }
};

// Move-only RAII device buffer (include/types.h:100-142 of the reference).
template <typename T>
struct DeviceArray {
  T *data = nullptr;
  size_t count = 0;
  DeviceArray() = default;
  explicit DeviceArray(size_t n) { allocate(n); }
  DeviceArray(const DeviceArray &) = delete;
  DeviceArray &operator=(const DeviceArray &) = delete;
  DeviceArray(DeviceArray &&o) noexcept : data(o.data), count(o.count) {
    o.data = nullptr;
    o.count = 0;
  }
  ~DeviceArray() { free(); }
  void allocate(size_t n) {
    free();
    count = n;
    if (n > 0) NRF_DEVICE_CHECK(hipMalloc(reinterpret_cast<void **>(&data), n * sizeof(T)));
  }
  void free() {
    if (data) {
      NRF_DEVICE_CHECK(hipFree(data));
      data = nullptr;
      count = 0;
    }
  }
  void zero() { NRF_DEVICE_CHECK(hipMemset(data, 0, count * sizeof(T))); }
  void upload(const std::vector<T> &v) {
    if (count != v.size()) allocate(v.size());
    NRF_DEVICE_CHECK(hipMemcpy(data, v.data(), count * sizeof(T), hipMemcpyHostToDevice));
  }
  void download(std::vector<T> &v) const {
    v.resize(count);
    NRF_DEVICE_CHECK(hipMemcpy(v.data(), data, count * sizeof(T), hipMemcpyDeviceToHost));
  }
  size_t size() const { return count; }
  size_t bytes() const { return count * sizeof(T); }
};

struct CameraIntrinsics {
  float fx, fy, cx, cy;
  int width, height;
  __host__ __device__ float2 project(float3 p) const {
    return make_float2(fx * p.x / p.z + cx, fy * p.y / p.z + cy);
  }
  __host__ __device__ float3 unproject(float u, float v, float z) const {
    return make_float3(z * (u - cx) / fx, z * (v - cy) / fy, z);
  }
};

struct TSDFVoxel {
  float tsdf;
  float weight;
  uchar3 color;
  uint8_t pad;
  __host__ __device__ TSDFVoxel() : tsdf(1.f), weight(0.f), color{128, 128, 128}, pad(0) {}
};

static constexpr int K_NEIGHBORS = 4;
static constexpr int K_GRAPH = 8;

struct DeformNode {
  float3 pos;
  float3 normal;
  float radius;
  int neighbors[K_GRAPH];
  float neighbor_w[K_GRAPH];
  int num_neighbors;
};

struct Correspondence {
  float3 src;
  float3 dst;
  float3 normal;
  int node_ids[K_NEIGHBORS];
  float node_ws[K_NEIGHBORS];
  float weight;
  bool valid;
};

static constexpr int BLOCK_DIM = 6;
static constexpr int BLOCK_SIZE = BLOCK_DIM * BLOCK_DIM;

struct BSRMatrix {
  int num_block_rows = 0;
  int num_blocks = 0;
  DeviceArray<int> row_ptr;
  DeviceArray<int> col_idx;
  DeviceArray<float> values;
  void allocate(int n_rows, int n_blocks) {
    num_block_rows = n_rows;
    num_blocks = n_blocks;
    row_ptr.allocate(n_rows + 1);
    col_idx.allocate(n_blocks);
    values.allocate((size_t)n_blocks * BLOCK_SIZE);
  }
  void zero_values() { values.zero(); }
};

struct BBoxFilter {
  float3 min_pt;
  float3 max_pt;
  bool enabled = false;
};

inline dim3 grid1d(int n, int block = 256) { return dim3((n + block - 1) / block); }

// The reference's layouts are the C ABI's.
static_assert(sizeof(TSDFVoxel) == sizeof(nrf_voxel), "TSDFVoxel layout");
static_assert(sizeof(DeformNode) == sizeof(nrf_node), "DeformNode layout");
static_assert(sizeof(DualQuat) == sizeof(nrf_dualquat), "DualQuat layout");
static_assert(sizeof(Correspondence) == sizeof(nrf_corr), "Correspondence layout");
static_assert(sizeof(CameraIntrinsics) == sizeof(nrf_camera), "CameraIntrinsics layout");
static_assert(offsetof(Correspondence, valid) == offsetof(nrf_corr, valid), "Correspondence.valid offset");
static_assert(offsetof(DeformNode, num_neighbors) == offsetof(nrf_node, num_neighbors), "DeformNode offsets");
