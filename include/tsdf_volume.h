// tsdf_volume.h - TSDFVolume with the reference's interface
// (include/tsdf_volume.h:10-81), backed by the gfx950 kernels of libnrf.
#pragma once
#include <string>

#include "types.h"

class TSDFVolume {
 public:
  struct Params {
    int3 dims = {256, 256, 256};
    float voxel_size = 0.006f;
    float truncation = 0.03f;
    float3 origin = {-0.768f, -0.768f, 0.5f};
    float decay_alpha = 0.90f;
    float max_weight = 50.0f;
  };

  explicit TSDFVolume(const Params &p);
  ~TSDFVolume();

  // kernels/tsdf_kernels.cu:85-174 via nrf_tsdf_integrate; rigid when nodes
  // is null or num_nodes == 0.  `normals` is unused, as in the reference.
  void integrate(const DeviceArray<float> &depth, const DeviceArray<float3> &normals, const CameraIntrinsics &cam,
                 const Mat4 &camera_pose, const DeformNode *nodes = nullptr, const DualQuat *transforms = nullptr,
                 int num_nodes = 0, const int *voxel_knn = nullptr, const float *voxel_knn_w = nullptr,
                 const int *voxel_opt_counts = nullptr, int min_opt_count = 0);

  // kernels/tsdf_kernels.cu:236-361 via nrf_raycast (src/tsdf_volume.cu:
  // 65-92): warped when nodes is set and num_nodes > 0.
  void raycast(DeviceArray<float3> &vertices, DeviceArray<float3> &normals, const CameraIntrinsics &cam,
               const Mat4 &camera_pose, const DeformNode *nodes = nullptr, const DualQuat *transforms = nullptr,
               int num_nodes = 0, const int *voxel_knn = nullptr, const float *voxel_knn_w = nullptr,
               int *out_canonical_vidx = nullptr);

  // Triangle soup: 3 vertices and 3 copies of the face normal per triangle.
  void extract_surface(std::vector<float3> &vertices, std::vector<float3> &normals,
                       std::vector<int3> &triangles) const;
  void extract_surface_device(DeviceArray<float3> &vertices, DeviceArray<int3> &triangles) const;

  const Params &params() const { return params_; }
  TSDFVoxel *device_data() { return d_voxels_.data; }
  const TSDFVoxel *device_data() const { return d_voxels_.data; }
  int total_voxels() const { return params_.dims.x * params_.dims.y * params_.dims.z; }

  __host__ int3 idx_to_xyz(int idx) const {
    int x = idx % params_.dims.x;
    int y = (idx / params_.dims.x) % params_.dims.y;
    int z = idx / (params_.dims.x * params_.dims.y);
    return make_int3(x, y, z);
  }
  // no +0.5 here (include/tsdf_volume.h:72-76 of the reference)
  __host__ float3 voxel_to_world(int3 xyz) const {
    return make_float3(params_.origin.x + xyz.x * params_.voxel_size, params_.origin.y + xyz.y * params_.voxel_size,
                       params_.origin.z + xyz.z * params_.voxel_size);
  }

 private:
  nrf_volume_params vp() const;
  Params params_;
  DeviceArray<TSDFVoxel> d_voxels_;
  mutable DeviceArray<uint8_t> mc_workspace_;
  mutable DeviceArray<DualQuat> centered_;
};
