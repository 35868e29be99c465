// warp_field.h - WarpField with the reference's interface
// (include/warp_field.h:12-126): node / transform storage, graph
// maintenance (node sampling, geodesic edges, pruning), voxel 4-NN
// snapshot, twist update, save / restore.
#pragma once
#include <vector>

#include "tsdf_volume.h"
#include "types.h"

class WarpField {
 public:
  static constexpr int MAX_NODES = 8192;
  static constexpr float NODE_RADIUS_FACTOR = 2.5f;

  WarpField(float node_radius, int max_nodes = MAX_NODES);
  ~WarpField() = default;

  // src/warp_field.cu:709-841 via nrf_graph_add_nodes
  int add_nodes_from_surface(const std::vector<float3> &surface_vertices, float min_dist_between_nodes,
                             const std::vector<float3> *surface_normals = nullptr);
  // src/warp_field.cu:843-954 via nrf_graph_add_nodes_geodesic
  int add_nodes_from_mesh_geodesic(const std::vector<float3> &surface_vertices,
                                   const std::vector<int3> &surface_triangles, float min_dist_between_nodes,
                                   const std::vector<float3> *surface_normals = nullptr);

  struct PruneParams {
    bool enabled = true;
    bool remove_disconnected = true;
    int min_observed_voxels = 8;
    int min_surface_voxels = 2;
    int min_component_size = 4;
    float support_radius_factor = 1.0f;
    float surface_tsdf_abs = 0.45f;
    float empty_tsdf = 0.75f;
  };
  // src/warp_field.cu:986-1104 via nrf_graph_prune
  int prune_nodes(const TSDFVolume &volume, const PruneParams &params);

  int num_nodes() const { return num_nodes_; }

  // src/warp_field.cu:957-984 via nrf_voxel_knn
  void compute_voxel_knn(const TSDFVolume &volume, DeviceArray<int> &d_voxel_knn, DeviceArray<float> &d_voxel_knn_w);

  DeformNode *device_nodes() { return d_nodes_.data; }
  DualQuat *device_transforms() { return d_transforms_.data; }
  DualQuat *device_transforms_prev() { return d_transforms_prev_.data; }
  const DeformNode *device_nodes() const { return d_nodes_.data; }
  const DualQuat *device_transforms() const { return d_transforms_.data; }

  void save_transforms();
  void restore_transforms();
  // T_i <- exp(clamp(update_scale * dx_i)) * T_i (src/warp_field.cu:1121-1141)
  void apply_twist_increment(const DeviceArray<float> &delta_x, float max_rot, float max_trans,
                             float update_scale = 1.0f);
  void reset_transforms();

  std::vector<DeformNode> download_nodes() const;
  std::vector<DualQuat> download_transforms() const;
  void upload_nodes(const std::vector<DeformNode> &nodes);
  void upload_transforms(const std::vector<DualQuat> &transforms);

 private:
  void sync_graph(int old_count);
  float node_radius_;
  int num_nodes_;
  int max_nodes_;
  DeviceArray<DeformNode> d_nodes_;
  DeviceArray<DualQuat> d_transforms_;
  DeviceArray<DualQuat> d_transforms_prev_;
  std::vector<DeformNode> h_nodes_;
  std::vector<DualQuat> h_transforms_;
};
