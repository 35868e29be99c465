// nrf_internal.h - shared host-side helpers of the C ABI implementation.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "nrf.h"

namespace nrf {

void set_last_error(const char *where, hipError_t e);
void set_last_error_msg(const std::string &msg);

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

inline int64_t volume_count(const nrf_volume_params *vp) {
  return (int64_t)vp->dims[0] * vp->dims[1] * vp->dims[2];
}

inline bool volume_ok(const nrf_volume_params *vp) {
  if (!vp) return false;
  if (vp->dims[0] <= 0 || vp->dims[1] <= 0 || vp->dims[2] <= 0) return false;
  // voxel indices are 32-bit in the kernels (as in the reference).
  return volume_count(vp) <= (int64_t)INT32_MAX / 4;
}

// 3x4 rows of the world->camera transform, inverted on the host exactly as
// TSDFVolume::integrate does (src/tsdf_volume.cu:36-49): R^T, -R^T t.
struct CamXf {
  float m[3][4];
};
CamXf invert_pose(const float pose[16]);

int device_cu_count();

// Optional event sink of nrf_solver_solve_gn (used by the pipeline's
// per-kernel profiling): when set, events are recorded on the stream after
// the membership build, then per GN iteration after the assembly, just
// before the PCG kernel, after the PCG kernel and after the update, i.e.
// 1 + 4 * gn_iterations events.
void solver_set_event_sink(nrf_solver *s, hipEvent_t *events, int count, int skip = 0);
// The next nrf_solver_solve_gn makes its stream wait for `ev` before its
// first validation or twist update, the first of its kernels that writes the
// node transforms or their centred table (null: no wait).  One solve
// consumes it.
void solver_set_pcg_wait(nrf_solver *s, hipEvent_t ev);
// The next nrf_solver_solve_gn records `ev` on its stream right before its
// first PCG launch (after the memberships and the first assembly) and, once
// that launch is enqueued, calls fn(ctx, stream) on the host, which may
// order work of its own after `ev`.  A non-zero return ends the solve with
// that status.  One solve consumes it; it is dropped when the solve ends
// before a PCG launch.
typedef int (*solver_hook_fn)(void *ctx, hipStream_t stream);
void solver_set_pcg_hook(nrf_solver *s, solver_hook_fn fn, void *ctx, hipEvent_t ev);
// The solver's next PCG launch carries `ticket` (non-zero): once all its
// workgroups have started, it records the ticket.  solver_pcg_start_wait
// enqueues on `stream` a one-wave kernel that ends when that ticket is
// recorded (or after a bounded wait): work behind it starts after the stream
// work that preceded the launch, and cannot take the device before the
// launch's workgroups.
void solver_set_pcg_ticket(nrf_solver *s, unsigned ticket);
int solver_pcg_start_wait(nrf_solver *s, unsigned ticket, hipStream_t stream);
// The next nrf_solver_solve_gn copies the node transforms and their centred
// table, as they are when it starts, into T / centered (num_nodes entries
// each), in its first kernel.  One solve consumes it.
void solver_set_snapshot(nrf_solver *s, nrf_dualquat *T, nrf_dualquat *centered);
int device_max_lds_per_block();

// ---------------------------------------------------------------- batches
// Independent instances (one pipeline each) stepped by common launches.
// Per-frame inputs travel as kernel arguments, so a captured batch step
// depends on nothing but them and the instances' device state.
// Marching-cubes bricks: 64 x 4 x 4 voxels (cubes by their corner 0),
// numbered x fastest.  The integration marks the bricks it updates in an
// occupancy map of one byte per brick, which marching cubes reads.
constexpr int kMcBrickX = 64, kMcBrickY = 4, kMcBrickZ = 4;
__host__ __device__ inline int mc_bricks_x(int X) { return (X + kMcBrickX - 1) / kMcBrickX; }
__host__ __device__ inline int mc_bricks_y(int Y) { return (Y + kMcBrickY - 1) / kMcBrickY; }
__host__ __device__ inline int mc_brick_of(int x, int y, int z, int X, int Y) {
  return ((z / kMcBrickZ) * mc_bricks_y(Y) + y / kMcBrickY) * mc_bricks_x(X) + x / kMcBrickX;
}
inline size_t mc_brick_count(int X, int Y, int Z) {
  return (size_t)mc_bricks_x(X) * mc_bricks_y(Y) * ((Z + kMcBrickZ - 1) / kMcBrickZ);
}

constexpr int kMaxBatch = 32;
// memberships per chunk of the MFMA assembly (rows with more are split)
constexpr int kAsmChunk = 128;
struct BatchCorrs {
  const nrf_corr *corrs[kMaxBatch];
  const float *canonical[kMaxBatch];
  int32_t n[kMaxBatch];
  int32_t max_n;
  // first assembly chunk slot of each instance in the batched launch:
  // cbase[k + 1] - cbase[k] = N_k + ceil(4 n_k / kAsmChunk)
  int32_t cbase[kMaxBatch + 1];
};
struct BatchDepth {
  const float *depth[kMaxBatch];
  CamXf xf[kMaxBatch];
};

struct SolverBatch;
struct SolverBatchInstance {
  nrf_solver *solver;
  const nrf_node *nodes;  // device
  nrf_dualquat *T, *centered;
  int N;
  nrf_gn_params gn;
  float *delta_x;
  int max_corrs;
};
// Sets every solver's PCG layout (lanes per scalar row, target nodes per
// workgroup), groups the instances into persistent launches whose
// workgroups fit on the device together, and uploads the device records.
int solver_batch_create(const SolverBatchInstance *inst, int count, int lanes_per_row, int wg_nodes,
                        SolverBatch **out);
void solver_batch_destroy(SolverBatch *b);
// The GN loop of every instance (memberships, gn_iterations x (records,
// assembly, PCG, validation, update)), all stream-ordered.
int solver_batch_solve_gn(SolverBatch *b, const BatchCorrs &fr, hipStream_t st);
// Pieces used by the tuner: memberships + assembly at the current transforms
// (results and totals are saved first), PCG alone, and the restore.
int solver_batch_prepare(SolverBatch *b, const BatchCorrs &fr, hipStream_t st);
int solver_batch_pcg(SolverBatch *b, hipStream_t st);
int solver_batch_restore(SolverBatch *b, hipStream_t st);
int solver_batch_launch_groups(const SolverBatch *b);
int solver_batch_workgroups(const SolverBatch *b);
int solver_rmax(const nrf_solver *s);
// Sizes the per-frame buffers of the solver (correspondence records,
// memberships, refinement vectors) for frames of up to max_corrs records at
// the current node count, so that a step (possibly under capture) allocates
// nothing.
int solver_reserve(nrf_solver *s, int max_corrs);
// Changes whenever a solver's partition is rebuilt or any solver buffer is
// reallocated: captured steps compare it before a replay.
unsigned solver_state_generation(const nrf_solver *s);

struct IntegrateBatch;
struct IntegrateBatchInstance {
  nrf_volume_params vp;
  nrf_camera cam;
  nrf_voxel *vox;
  const nrf_node *nodes;
  const nrf_dualquat *T, *centered;
  int N;
  const int32_t *knn;
  const float *knn_w;
  const uint32_t *knn_rows;  // row occupancy of knn (knn_row_flags), or null
  int32_t *updated;
  uint8_t *occ;  // marching-cubes brick occupancy to mark (mc_occupancy_bytes), or null
};
int integrate_batch_create(const IntegrateBatchInstance *inst, int count, int z_chunk, IntegrateBatch **out);
void integrate_batch_destroy(IntegrateBatch *b);
int integrate_batch_run(IntegrateBatch *b, const BatchDepth &fr, hipStream_t st);

// Row occupancy of a voxel kNN: one word per run of 64 voxels along x (the
// runs the integration's waves cover; run bx of row (y, z) at word
// (z Y + y) ceil(X / 64) + bx), nonzero when some voxel of the run has an
// id >= 0.  The integration skips the kNN reads of the other runs, whose
// voxels all integrate unwarped.  Valid until the kNN changes.
int64_t knn_row_words(const nrf_volume_params *vp);
int knn_row_flags(const nrf_volume_params *vp, const int32_t *knn, uint32_t *rows, hipStream_t st);
// nrf_solver_solve_gn; begin = false continues the solve of the previous
// call (no reset of its stop flag and record: a rejected step stops the
// later calls too, and the record sums them)
int solver_solve_gn(nrf_solver *s, const nrf_corr *corrs, const float *canonical, int num_corrs,
                    const nrf_node *nodes, nrf_dualquat *transforms, nrf_dualquat *centered, int num_nodes,
                    const nrf_gn_params *gn, float *delta_x, nrf_pcg_result *result, hipStream_t st, bool begin);
// The start of a solve's record (clears the stop flag and the device
// result, folding the previous solve into the totals).
int solver_gn_begin(nrf_solver *s, hipStream_t st);
// The device flag that turns every later kernel of the solve into a no-op
// (raised by a rejected step; the tracking loop raises it too).  Written
// only by single-workgroup kernels, so that no multi-workgroup kernel sees it
// change while it runs.
int32_t *solver_stop_flag(nrf_solver *s);
nrf_pcg_result *solver_result_device(nrf_solver *s);
// One GN linearisation without the update: memberships, assembly of the
// stored corrs[i].src (no canonical points), PCG with the configured
// refinement or true-residual rounds (the fused refinement-update form runs
// unfused), and with gn->validate the step validation (a rejected step
// raises the stop flag).  delta_x receives the step.
int solver_gn_linearise(nrf_solver *s, const nrf_corr *corrs, int num_corrs, const nrf_node *nodes,
                        const nrf_dualquat *transforms, const nrf_dualquat *centered, int num_nodes,
                        const nrf_gn_params *gn, float *delta_x, hipStream_t st);
// live surface prediction (raster.hip): nrf_skin_vertices and
// nrf_rasterize_mesh with the depth/triangle key buffer (8 B per pixel).
// d_ntri (optional): the triangle count of a triangle-soup mesh on the
// device (vertices 3 per triangle), capped by num_vertices / 3 resp.
// num_triangles, so that no host read of the count is needed.
int skin_vertices(const float *canonical, int n, const nrf_volume_params *vp, const nrf_dualquat *centered,
                  int num_nodes, const int32_t *voxel_knn, const float *voxel_knn_w, float *out, hipStream_t st,
                  const int32_t *d_ntri = nullptr, const int32_t *skip = nullptr);
int rasterize_mesh(const float *canonical, const float *warped, const int32_t *triangles, int num_vertices,
                   int num_triangles, const nrf_camera *cam, const float camera_pose[16], const nrf_volume_params *vp,
                   unsigned long long *ztri, float *live_vertices, float *live_normals, int32_t *hit_voxel,
                   float *canonical_points, hipStream_t st, const int32_t *d_ntri = nullptr,
                   const int32_t *skip = nullptr);
// nrf_find_correspondences; with pixel_knn NULL and hit_voxel set, the
// pixel 4-NN is read through the hit voxel (the pixel kNN step fused).
// skip (optional, device): nonzero leaves every output as it is.
int find_correspondences(const float *live_vertices, const float *live_normals, const float *depth,
                         const float *depth_normals, const nrf_camera *cam, const float camera_pose[16],
                         const int32_t *pixel_knn, const float *pixel_knn_w, const int32_t *hit_voxel,
                         const int32_t *voxel_knn, const float *voxel_knn_w, const nrf_assoc_params *params,
                         nrf_corr *corrs, int32_t *num_valid, int32_t *stats, const int32_t *skip, hipStream_t st);
// tracking-loop control (track.hip): the device record of a tracking frame
// and its decisions, each gated by the solver's stop flag
struct TrackDev {
  nrf_track_stats stats;
  int32_t counter;  // valid correspondences of the running association
  float best_scale;
  double best_rmse;
};
int track_begin(TrackDev *t, hipStream_t st);
int track_count_reset(TrackDev *t, const int32_t *stop, hipStream_t st);
int track_gate(TrackDev *t, int g, int min_valid, int32_t *stop, hipStream_t st);
int track_rmse_blocks(int n);
// trial = false: the base RMSE of iteration g; true: the trial test of scale
int track_rmse(TrackDev *t, const nrf_corr *corrs, int n, double *part, int g, bool trial, float scale,
               int min_valid, const int32_t *stop, hipStream_t st);
int track_ls_begin(TrackDev *t, int g, float update_scale, const nrf_dualquat *T, nrf_dualquat *Tsave, int N,
                   int32_t *stop, hipStream_t st);
int track_ls_apply(TrackDev *t, const float *dx, const nrf_dualquat *Tsave, nrf_dualquat *T, nrf_dualquat *centered,
                   const nrf_node *nodes, int N, float scale, bool use_best, float max_rot, float max_trans,
                   const int32_t *stop, hipStream_t st);
int track_ls_commit(TrackDev *t, int g, int32_t *stop, hipStream_t st);
int track_end(TrackDev *t, const nrf_pcg_result *res, bool line_search, const int32_t *stop, hipStream_t st);
// nrf_tsdf16_integrate with the kNN's row occupancy (null: none) and the
// optimised-voxel gate of nrf_tsdf_integrate (warped only)
int tsdf16_integrate(const nrf_volume_params *vp, uint16_t *tsdf16, uint8_t *weight8, const float *depth,
                     const nrf_camera *cam, const float camera_pose[16], const nrf_dualquat *centered, int num_nodes,
                     const uint16_t *knn16, const uint16_t *knn_w16, const uint32_t *knn_rows,
                     int32_t *updated_count, hipStream_t st, const int32_t *voxel_opt_counts = nullptr,
                     int min_opt_count = 0, const void *dev_pose = nullptr, uint8_t *occ = nullptr,
                     float *depth_blocks = nullptr);
// dev_pose (tsdf_integrate / tsdf16_integrate): a device CamXf (the
// world-to-camera transform, as invert_pose gives it) that the kernel reads
// instead of camera_pose's inverse, so a captured launch can take each
// replay's pose from memory (set_device_pose writes it stream-ordered).
int set_device_pose(CamXf *dev, const float camera_pose[16], hipStream_t st);
// Marching cubes (nrf_mc_extract / nrf_mc16_extract) with an optional brick
// occupancy map: one byte per 64 x 4 x 4 voxel brick (mc_occupancy_bytes),
// nonzero when a voxel of the brick may have weight > 0; bricks with a zero
// byte are skipped (none of their cubes can emit).  The map is kept by the
// integration calls that take it (they set the byte of every brick holding
// an updated voxel and never clear one).
size_t mc_occupancy_bytes(const nrf_volume_params *vp);
int mc_extract(const nrf_volume_params *vp, const nrf_voxel *voxels, const uint8_t *occ, void *workspace,
               size_t workspace_bytes, float *vertices, float *normals, int32_t *triangles, int max_triangles,
               int32_t *counts, hipStream_t st);
int mc16_extract(const nrf_volume_params *vp, const uint16_t *tsdf16, const uint8_t *weight8, const uint8_t *occ,
                 void *workspace, size_t workspace_bytes, float *vertices, int32_t *triangles, int max_triangles,
                 int32_t *counts, hipStream_t st);
// depth_blocks (tsdf_integrate / tsdf16_integrate): scratch of
// depth_block_count floats; when given, the call first writes the depth's
// block maxima there and every wave of the integration whose voxels all
// integrate unwarped and cannot reach a usable depth within the truncation
// skips its loads (the same result).
size_t depth_block_count(const nrf_camera *cam);
// nrf_tsdf_integrate with the kNN's row occupancy (null: none)
int tsdf_integrate(const nrf_volume_params *vp, nrf_voxel *voxels, const float *depth, const nrf_camera *cam,
                   const float camera_pose[16], const nrf_node *nodes, const nrf_dualquat *transforms, int num_nodes,
                   const nrf_dualquat *centered, const int32_t *voxel_knn, const float *voxel_knn_w,
                   const uint32_t *knn_rows, const int32_t *voxel_opt_counts, int min_opt_count,
                   int32_t *updated_count, hipStream_t st, const void *dev_pose = nullptr, uint8_t *occ = nullptr,
                   float *depth_blocks = nullptr);

}  // namespace nrf

#define NRF_HIP_CHECK(call)                        \
  do {                                             \
    hipError_t e_ = (call);                        \
    if (e_ != hipSuccess) {                        \
      nrf::set_last_error(#call, e_);              \
      return NRF_ERR_HIP;                          \
    }                                              \
  } while (0)

#define NRF_LAUNCH_CHECK()                          \
  do {                                              \
    hipError_t e_ = hipGetLastError();              \
    if (e_ != hipSuccess) {                         \
      nrf::set_last_error("kernel launch", e_);     \
      return NRF_ERR_HIP;                           \
    }                                               \
  } while (0)
