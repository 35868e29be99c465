// warp_field.h - WarpField with the reference's interface
// (include/warp_field.h:12-126) for the per-frame step: node / transform
// storage, voxel 4-NN snapshot, twist update, save / restore.  Node sampling
// and graph maintenance (add_nodes_from_surface, prune_nodes) run on graph
// updates, not per frame, and are not part of this library: nodes are
// supplied with upload_nodes.
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
  float node_radius_;
  int num_nodes_;
  int max_nodes_;
  DeviceArray<DeformNode> d_nodes_;
  DeviceArray<DualQuat> d_transforms_;
  DeviceArray<DualQuat> d_transforms_prev_;
  std::vector<DeformNode> h_nodes_;
  std::vector<DualQuat> h_transforms_;
};
