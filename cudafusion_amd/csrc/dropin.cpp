// dropin.cpp - the reference's TSDFVolume / WarpField / GaussNewtonSolver
// classes implemented over the C ABI (include/nrf.h).  Synchronisation
// points follow the reference: integrate / solve / twist update complete
// before returning, downloads synchronise.
#include <algorithm>
#include <cstring>

#include "solver.h"
#include "tsdf_volume.h"
#include "warp_field.h"

namespace {

void pose16(const Mat4 &m, float out[16]) {
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) out[i * 4 + j] = m.m[i][j];
}

nrf_camera to_cam(const CameraIntrinsics &c) {
  nrf_camera o;
  std::memcpy(&o, &c, sizeof(o));
  return o;
}

}  // namespace

// ------------------------------------------------------------ TSDFVolume
nrf_volume_params TSDFVolume::vp() const {
  nrf_volume_params v;
  v.dims[0] = params_.dims.x;
  v.dims[1] = params_.dims.y;
  v.dims[2] = params_.dims.z;
  v.voxel_size = params_.voxel_size;
  v.truncation = params_.truncation;
  v.origin[0] = params_.origin.x;
  v.origin[1] = params_.origin.y;
  v.origin[2] = params_.origin.z;
  v.decay_alpha = params_.decay_alpha;
  v.max_weight = params_.max_weight;
  return v;
}

TSDFVolume::TSDFVolume(const Params &p) : params_(p) {
  d_voxels_.allocate((size_t)total_voxels());
  nrf_volume_params v = vp();
  NRF_STATUS_CHECK(nrf_tsdf_reset(&v, reinterpret_cast<nrf_voxel *>(d_voxels_.data), nullptr));
  NRF_DEVICE_CHECK(hipDeviceSynchronize());
}

TSDFVolume::~TSDFVolume() {}

void TSDFVolume::integrate(const DeviceArray<float> &depth, const DeviceArray<float3> &, const CameraIntrinsics &cam,
                           const Mat4 &camera_pose, const DeformNode *nodes, const DualQuat *transforms,
                           int num_nodes, const int *voxel_knn, const float *voxel_knn_w,
                           const int *voxel_opt_counts, int min_opt_count) {
  nrf_volume_params v = vp();
  nrf_camera c = to_cam(cam);
  float pose[16];
  pose16(camera_pose, pose);
  const bool warp = nodes != nullptr && num_nodes > 0;
  const nrf_dualquat *cen = nullptr;
  if (warp) {
    if (centered_.size() < (size_t)num_nodes) centered_.allocate((size_t)num_nodes);
    NRF_STATUS_CHECK(nrf_warp_centered_table(reinterpret_cast<const nrf_node *>(nodes),
                                             reinterpret_cast<const nrf_dualquat *>(transforms), num_nodes,
                                             reinterpret_cast<nrf_dualquat *>(centered_.data), nullptr));
    cen = reinterpret_cast<const nrf_dualquat *>(centered_.data);
  }
  NRF_STATUS_CHECK(nrf_tsdf_integrate(&v, reinterpret_cast<nrf_voxel *>(d_voxels_.data), depth.data, &c, pose,
                                      reinterpret_cast<const nrf_node *>(nodes),
                                      reinterpret_cast<const nrf_dualquat *>(transforms), num_nodes, cen, voxel_knn,
                                      voxel_knn_w, voxel_opt_counts, min_opt_count, nullptr, nullptr));
  NRF_DEVICE_CHECK(hipDeviceSynchronize());
}

void TSDFVolume::raycast(DeviceArray<float3> &vertices, DeviceArray<float3> &normals, const CameraIntrinsics &cam,
                         const Mat4 &camera_pose, const DeformNode *nodes, const DualQuat *transforms, int num_nodes,
                         const int *voxel_knn, const float *voxel_knn_w, int *out_canonical_vidx) {
  const size_t npx = (size_t)cam.width * cam.height;
  vertices.allocate(npx);
  normals.allocate(npx);
  nrf_volume_params v = vp();
  nrf_camera c = to_cam(cam);
  float pose[16];
  pose16(camera_pose, pose);
  NRF_STATUS_CHECK(nrf_raycast(&v, reinterpret_cast<const nrf_voxel *>(d_voxels_.data), &c, pose,
                               reinterpret_cast<const nrf_node *>(nodes),
                               reinterpret_cast<const nrf_dualquat *>(transforms), num_nodes, voxel_knn, voxel_knn_w,
                               reinterpret_cast<float *>(vertices.data), reinterpret_cast<float *>(normals.data),
                               out_canonical_vidx, nullptr));
  NRF_DEVICE_CHECK(hipDeviceSynchronize());
}

void TSDFVolume::extract_surface(std::vector<float3> &vertices, std::vector<float3> &normals,
                                 std::vector<int3> &triangles) const {
  vertices.clear();
  normals.clear();
  triangles.clear();
  if (params_.dims.x < 2 || params_.dims.y < 2 || params_.dims.z < 2) return;
  nrf_volume_params v = vp();
  const size_t ws = nrf_mc_workspace_size(&v);
  if (mc_workspace_.size() < ws) mc_workspace_.allocate(ws);
  DeviceArray<int32_t> counts(4);
  NRF_STATUS_CHECK(nrf_mc_extract(&v, reinterpret_cast<const nrf_voxel *>(d_voxels_.data), mc_workspace_.data,
                                  mc_workspace_.size(), nullptr, nullptr, nullptr, 0, counts.data, nullptr));
  int32_t h[4];
  NRF_DEVICE_CHECK(hipMemcpy(h, counts.data, sizeof(h), hipMemcpyDeviceToHost));
  const int T = h[1];
  if (T <= 0) return;
  DeviceArray<float3> dv((size_t)3 * T), dn((size_t)3 * T);
  DeviceArray<int3> dt((size_t)T);
  NRF_STATUS_CHECK(nrf_mc_extract(&v, reinterpret_cast<const nrf_voxel *>(d_voxels_.data), mc_workspace_.data,
                                  mc_workspace_.size(), reinterpret_cast<float *>(dv.data),
                                  reinterpret_cast<float *>(dn.data), reinterpret_cast<int32_t *>(dt.data), T,
                                  counts.data, nullptr));
  dv.download(vertices);
  dn.download(normals);
  dt.download(triangles);
}

void TSDFVolume::extract_surface_device(DeviceArray<float3> &vertices, DeviceArray<int3> &triangles) const {
  vertices.free();
  triangles.free();
  if (params_.dims.x < 2 || params_.dims.y < 2 || params_.dims.z < 2) return;
  nrf_volume_params v = vp();
  const size_t ws = nrf_mc_workspace_size(&v);
  if (mc_workspace_.size() < ws) mc_workspace_.allocate(ws);
  DeviceArray<int32_t> counts(4);
  NRF_STATUS_CHECK(nrf_mc_extract(&v, reinterpret_cast<const nrf_voxel *>(d_voxels_.data), mc_workspace_.data,
                                  mc_workspace_.size(), nullptr, nullptr, nullptr, 0, counts.data, nullptr));
  int32_t h[4];
  NRF_DEVICE_CHECK(hipMemcpy(h, counts.data, sizeof(h), hipMemcpyDeviceToHost));
  const int T = h[1];
  if (T <= 0) return;
  vertices.allocate((size_t)3 * T);
  triangles.allocate((size_t)T);
  NRF_STATUS_CHECK(nrf_mc_extract(&v, reinterpret_cast<const nrf_voxel *>(d_voxels_.data), mc_workspace_.data,
                                  mc_workspace_.size(), reinterpret_cast<float *>(vertices.data), nullptr,
                                  reinterpret_cast<int32_t *>(triangles.data), T, counts.data, nullptr));
  NRF_DEVICE_CHECK(hipDeviceSynchronize());
}

// ------------------------------------------------------------- WarpField
WarpField::WarpField(float node_radius, int max_nodes)
    : node_radius_(node_radius), num_nodes_(0), max_nodes_(max_nodes) {
  d_nodes_.allocate(max_nodes);
  d_transforms_.allocate(max_nodes);
  d_transforms_prev_.allocate(max_nodes);
  h_nodes_.reserve(max_nodes);
  h_transforms_.reserve(max_nodes);
}

// the host mirror after a graph change, uploaded as the reference does
// (nodes always; transforms and the previous transforms)
void WarpField::sync_graph(int) {
  NRF_DEVICE_CHECK(hipMemcpy(d_nodes_.data, h_nodes_.data(), num_nodes_ * sizeof(DeformNode), hipMemcpyHostToDevice));
  NRF_DEVICE_CHECK(hipMemcpy(d_transforms_.data, h_transforms_.data(), num_nodes_ * sizeof(DualQuat),
                             hipMemcpyHostToDevice));
  NRF_DEVICE_CHECK(hipMemcpy(d_transforms_prev_.data, h_transforms_.data(), num_nodes_ * sizeof(DualQuat),
                             hipMemcpyHostToDevice));
}

int WarpField::add_nodes_from_surface(const std::vector<float3> &surface_vertices, float min_dist,
                                      const std::vector<float3> *surface_normals) {
  const int old = num_nodes_;
  h_nodes_.resize(max_nodes_);
  h_transforms_.resize(max_nodes_);
  int n = num_nodes_, added = 0;
  const float *nrm = surface_normals && surface_normals->size() >= surface_vertices.size()
                         ? reinterpret_cast<const float *>(surface_normals->data())
                         : nullptr;
  NRF_STATUS_CHECK(nrf_graph_add_nodes(reinterpret_cast<nrf_node *>(h_nodes_.data()),
                                       reinterpret_cast<nrf_dualquat *>(h_transforms_.data()), &n, max_nodes_,
                                       reinterpret_cast<const float *>(surface_vertices.data()), nrm,
                                       (int)surface_vertices.size(), min_dist, node_radius_, &added));
  num_nodes_ = n;
  h_nodes_.resize(n);
  h_transforms_.resize(n);
  if (added > 0) sync_graph(old);
  return added;
}

int WarpField::add_nodes_from_mesh_geodesic(const std::vector<float3> &surface_vertices,
                                            const std::vector<int3> &surface_triangles, float min_dist,
                                            const std::vector<float3> *surface_normals) {
  const int old = num_nodes_;
  h_nodes_.resize(max_nodes_);
  h_transforms_.resize(max_nodes_);
  int n = num_nodes_, added = 0;
  const float *nrm = surface_normals && surface_normals->size() >= surface_vertices.size()
                         ? reinterpret_cast<const float *>(surface_normals->data())
                         : nullptr;
  NRF_STATUS_CHECK(nrf_graph_add_nodes_geodesic(
      reinterpret_cast<nrf_node *>(h_nodes_.data()), reinterpret_cast<nrf_dualquat *>(h_transforms_.data()), &n,
      max_nodes_, reinterpret_cast<const float *>(surface_vertices.data()), nrm, (int)surface_vertices.size(),
      reinterpret_cast<const int32_t *>(surface_triangles.data()), (int)surface_triangles.size(), min_dist,
      node_radius_, &added, nullptr));
  num_nodes_ = n;
  h_nodes_.resize(n);
  h_transforms_.resize(n);
  if (added > 0) sync_graph(old);
  return added;
}

int WarpField::prune_nodes(const TSDFVolume &volume, const PruneParams &params) {
  if (!params.enabled || num_nodes_ == 0) return 0;
  const auto &p = volume.params();
  nrf_volume_params v;
  v.dims[0] = p.dims.x; v.dims[1] = p.dims.y; v.dims[2] = p.dims.z;
  v.voxel_size = p.voxel_size;
  v.truncation = p.truncation;
  v.origin[0] = p.origin.x; v.origin[1] = p.origin.y; v.origin[2] = p.origin.z;
  v.decay_alpha = p.decay_alpha;
  v.max_weight = p.max_weight;
  nrf_prune_params pp;
  pp.enabled = params.enabled ? 1 : 0;
  pp.remove_disconnected = params.remove_disconnected ? 1 : 0;
  pp.min_observed_voxels = params.min_observed_voxels;
  pp.min_surface_voxels = params.min_surface_voxels;
  pp.min_component_size = params.min_component_size;
  pp.support_radius_factor = params.support_radius_factor;
  pp.surface_tsdf_abs = params.surface_tsdf_abs;
  pp.empty_tsdf = params.empty_tsdf;
  const int old = num_nodes_;
  int n = num_nodes_, removed = 0;
  h_transforms_.resize(num_nodes_);
  NRF_STATUS_CHECK(nrf_graph_prune(&v, reinterpret_cast<const nrf_voxel *>(volume.device_data()),
                                   reinterpret_cast<nrf_node *>(h_nodes_.data()),
                                   reinterpret_cast<nrf_dualquat *>(h_transforms_.data()), &n, &pp, &removed,
                                   nullptr));
  num_nodes_ = n;
  h_nodes_.resize(n);
  h_transforms_.resize(n);
  if (removed > 0) sync_graph(old);
  return removed;
}

void WarpField::compute_voxel_knn(const TSDFVolume &volume, DeviceArray<int> &d_voxel_knn,
                                  DeviceArray<float> &d_voxel_knn_w) {
  if (num_nodes_ == 0) return;
  const auto &p = volume.params();
  const size_t total = (size_t)p.dims.x * p.dims.y * p.dims.z;
  d_voxel_knn.allocate(total * K_NEIGHBORS);
  d_voxel_knn_w.allocate(total * K_NEIGHBORS);
  nrf_volume_params v;
  v.dims[0] = p.dims.x; v.dims[1] = p.dims.y; v.dims[2] = p.dims.z;
  v.voxel_size = p.voxel_size;
  v.truncation = p.truncation;
  v.origin[0] = p.origin.x; v.origin[1] = p.origin.y; v.origin[2] = p.origin.z;
  v.decay_alpha = p.decay_alpha;
  v.max_weight = p.max_weight;
  NRF_STATUS_CHECK(nrf_voxel_knn(&v, reinterpret_cast<const nrf_voxel *>(volume.device_data()),
                                 reinterpret_cast<const nrf_node *>(d_nodes_.data), num_nodes_, d_voxel_knn.data,
                                 d_voxel_knn_w.data, nullptr));
  NRF_DEVICE_CHECK(hipDeviceSynchronize());
}

void WarpField::save_transforms() {
  NRF_DEVICE_CHECK(hipMemcpy(d_transforms_prev_.data, d_transforms_.data, num_nodes_ * sizeof(DualQuat),
                             hipMemcpyDeviceToDevice));
}

void WarpField::restore_transforms() {
  NRF_DEVICE_CHECK(hipMemcpy(d_transforms_.data, d_transforms_prev_.data, num_nodes_ * sizeof(DualQuat),
                             hipMemcpyDeviceToDevice));
  h_transforms_.resize(num_nodes_);
  NRF_DEVICE_CHECK(hipMemcpy(h_transforms_.data(), d_transforms_.data, num_nodes_ * sizeof(DualQuat),
                             hipMemcpyDeviceToHost));
}

void WarpField::apply_twist_increment(const DeviceArray<float> &delta_x, float max_rot, float max_trans,
                                      float update_scale) {
  if (num_nodes_ == 0) return;
  NRF_STATUS_CHECK(nrf_warp_apply_twist(delta_x.data, reinterpret_cast<nrf_dualquat *>(d_transforms_.data),
                                        num_nodes_, max_rot, max_trans, update_scale, nullptr, nullptr, nullptr));
  NRF_DEVICE_CHECK(hipDeviceSynchronize());
  h_transforms_.resize(num_nodes_);
  NRF_DEVICE_CHECK(hipMemcpy(h_transforms_.data(), d_transforms_.data, num_nodes_ * sizeof(DualQuat),
                             hipMemcpyDeviceToHost));
}

void WarpField::reset_transforms() {
  h_transforms_.assign(num_nodes_, DualQuat::identity());
  NRF_DEVICE_CHECK(hipMemcpy(d_transforms_.data, h_transforms_.data(), num_nodes_ * sizeof(DualQuat),
                             hipMemcpyHostToDevice));
}

std::vector<DeformNode> WarpField::download_nodes() const { return h_nodes_; }
std::vector<DualQuat> WarpField::download_transforms() const { return h_transforms_; }

void WarpField::upload_nodes(const std::vector<DeformNode> &nodes) {
  num_nodes_ = std::min((int)nodes.size(), max_nodes_);
  h_nodes_.assign(nodes.begin(), nodes.begin() + num_nodes_);
  NRF_DEVICE_CHECK(hipMemcpy(d_nodes_.data, h_nodes_.data(), num_nodes_ * sizeof(DeformNode), hipMemcpyHostToDevice));
  if ((int)h_transforms_.size() != num_nodes_) {
    h_transforms_.assign(num_nodes_, DualQuat::identity());
    NRF_DEVICE_CHECK(hipMemcpy(d_transforms_.data, h_transforms_.data(), num_nodes_ * sizeof(DualQuat),
                               hipMemcpyHostToDevice));
    NRF_DEVICE_CHECK(hipMemcpy(d_transforms_prev_.data, h_transforms_.data(), num_nodes_ * sizeof(DualQuat),
                               hipMemcpyHostToDevice));
  }
}

void WarpField::upload_transforms(const std::vector<DualQuat> &transforms) {
  int n = std::min((int)transforms.size(), max_nodes_);
  if (num_nodes_ == 0) num_nodes_ = n;
  n = std::min(n, num_nodes_);
  h_transforms_.assign(transforms.begin(), transforms.begin() + n);
  NRF_DEVICE_CHECK(hipMemcpy(d_transforms_.data, h_transforms_.data(), n * sizeof(DualQuat), hipMemcpyHostToDevice));
  NRF_DEVICE_CHECK(
      hipMemcpy(d_transforms_prev_.data, h_transforms_.data(), n * sizeof(DualQuat), hipMemcpyHostToDevice));
}

// ----------------------------------------------------- GaussNewtonSolver
GaussNewtonSolver::GaussNewtonSolver(const Params &p, int max_nodes) : params_(p), max_nodes_(max_nodes) {
  nrf_solver_params sp;
  std::memset(&sp, 0, sizeof(sp));
  sp.gn_iterations = p.gn_iterations;
  sp.pcg_iterations = p.pcg_iterations;
  sp.pcg_tolerance = p.pcg_tolerance;
  sp.lambda_smooth = p.lambda_smooth;
  sp.lambda_rot = p.lambda_rot;
  sp.lambda_damping = p.lambda_damping;
  sp.debug = p.debug ? 1 : 0;
  solver_ = nrf_solver_create(&sp, max_nodes);
  if (!solver_) {
    fprintf(stderr, "nrf_solver_create failed: %s\n", nrf_last_error());
    exit(1);
  }
}

GaussNewtonSolver::~GaussNewtonSolver() { nrf_solver_destroy(solver_); }

void GaussNewtonSolver::solve(const DeviceArray<Correspondence> &corrs, int num_corrs, const DeformNode *d_nodes,
                              const DualQuat *d_transforms, int num_nodes, DeviceArray<float> &delta_x) {
  delta_x.allocate((size_t)num_nodes * BLOCK_DIM);
  NRF_STATUS_CHECK(nrf_solver_solve(solver_, reinterpret_cast<const nrf_corr *>(corrs.data), num_corrs,
                                    reinterpret_cast<const nrf_node *>(d_nodes),
                                    reinterpret_cast<const nrf_dualquat *>(d_transforms), num_nodes, delta_x.data,
                                    &result_, nullptr));
  last_num_nodes_ = num_nodes;
  if (params_.debug)
    printf("[PCG] nodes=%d iters=%d final_r=%e ||b||=%e flags=%d\n", num_nodes, result_.iterations,
           result_.final_residual, result_.rhs_norm, result_.flags);
}

const BSRMatrix &GaussNewtonSolver::system_matrix() const {
  const int32_t *rp, *ci;
  const float *va, *rhs, *pc;
  int nb = 0;
  if (nrf_solver_system(solver_, &rp, &ci, &va, &nb, &rhs, &pc) != NRF_OK) return A_;
  const int N = last_num_nodes_;
  if (A_.num_block_rows != N || A_.num_blocks != nb) A_.allocate(N, nb);
  NRF_DEVICE_CHECK(hipMemcpy(A_.row_ptr.data, rp, (N + 1) * sizeof(int), hipMemcpyDeviceToDevice));
  NRF_DEVICE_CHECK(hipMemcpy(A_.col_idx.data, ci, nb * sizeof(int), hipMemcpyDeviceToDevice));
  NRF_DEVICE_CHECK(hipMemcpy(A_.values.data, va, (size_t)nb * BLOCK_SIZE * sizeof(float), hipMemcpyDeviceToDevice));
  return A_;
}

const DeviceArray<float> &GaussNewtonSolver::rhs() const {
  const int32_t *rp, *ci;
  const float *va, *rhs, *pc;
  int nb = 0;
  if (nrf_solver_system(solver_, &rp, &ci, &va, &nb, &rhs, &pc) != NRF_OK) return b_;
  const size_t n = (size_t)last_num_nodes_ * BLOCK_DIM;
  if (b_.size() != n) b_.allocate(n);
  NRF_DEVICE_CHECK(hipMemcpy(b_.data, rhs, n * sizeof(float), hipMemcpyDeviceToDevice));
  return b_;
}
