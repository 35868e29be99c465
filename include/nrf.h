/*
 * nrf.h - C ABI of the MI355X (gfx950) non-rigid fusion step.
 *
 * Plain C: POD descriptors, raw device pointers, sizes and a HIP stream
 * passed as `void*` (a hipStream_t; NULL = the legacy default stream).
 * Every entry point returns an nrf_status; all device work is enqueued on
 * the given stream and nothing synchronises unless the comment says so.
 * No entry point allocates memory the caller can see, except the handle
 * constructors (nrf_solver_create, nrf_pipeline_create), which own their
 * scratch space.  Functions that need scratch take a caller workspace whose
 * size is returned by the matching *_workspace_size query.
 *
 * The structs below reproduce the byte layout of the reference's POD types
 * (include/types.h:26-206) so that buffers written by either side can be
 * exchanged without repacking.
 */
#ifndef NRF_H
#define NRF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
typedef enum nrf_status {
  NRF_OK = 0,
  NRF_ERR_INVALID_ARGUMENT = 1, /* null pointer, negative size, bad dims   */
  NRF_ERR_TOO_LARGE = 2,        /* exceeds a documented capacity          */
  NRF_ERR_HIP = 3,              /* a HIP runtime call failed              */
  NRF_ERR_WORKSPACE = 4,        /* caller workspace smaller than required */
  NRF_ERR_NOT_READY = 5,        /* pattern / graph not set                */
  NRF_ERR_BARRIER_TIMEOUT = 6,  /* a persistent kernel's bounded wait ran out */
  NRF_ERR_NO_KERNEL = 7         /* the HIP code object has no gfx950 kernels */
} nrf_status;

const char *nrf_status_string(int status);
/* Last HIP error text seen by the library on this thread ("" if none). */
const char *nrf_last_error(void);
/* Build string: target, compile flags that affect numerics. */
const char *nrf_build_info(void);

/* ------------------------------------------------------------- POD types */
enum { NRF_K_NEIGHBORS = 4, NRF_K_GRAPH = 8, NRF_BLOCK_DIM = 6, NRF_BLOCK_SIZE = 36 };

/* = TSDFVoxel (include/types.h:168-176), 12 bytes. Initial state tsdf=1, w=0. */
typedef struct nrf_voxel {
  float tsdf;
  float weight;
  uint8_t color[3];
  uint8_t pad;
} nrf_voxel;

/* = DeformNode (include/types.h:185-192), 96 bytes. */
typedef struct nrf_node {
  float pos[3];
  float normal[3];
  float radius;
  int32_t neighbors[NRF_K_GRAPH];
  float neighbor_w[NRF_K_GRAPH];
  int32_t num_neighbors;
} nrf_node;

/* = DualQuat (include/types.h:86-94), 32 bytes, quaternions (x, y, z, w). */
typedef struct nrf_dualquat {
  float real[4];
  float dual[4];
} nrf_dualquat;

/* = Correspondence (include/types.h:198-206), 76 bytes. */
typedef struct nrf_corr {
  float src[3];
  float dst[3];
  float normal[3];
  int32_t node_ids[NRF_K_NEIGHBORS];
  float node_ws[NRF_K_NEIGHBORS];
  float weight;
  uint8_t valid;
  uint8_t pad[3];
} nrf_corr;

/* = CameraIntrinsics (include/types.h:148-162), 24 bytes. */
typedef struct nrf_camera {
  float fx, fy, cx, cy;
  int32_t width, height;
} nrf_camera;

/* = TSDFVolume::Params (include/tsdf_volume.h:12-20). */
typedef struct nrf_volume_params {
  int32_t dims[3];
  float voxel_size;
  float truncation;
  float origin[3];
  float decay_alpha;
  float max_weight;
} nrf_volume_params;

/* = GaussNewtonSolver::Params (include/solver.h:13-22) + build options. */
typedef struct nrf_solver_params {
  int32_t gn_iterations;  /* 3   */
  int32_t pcg_iterations; /* 10  */
  float pcg_tolerance;    /* 1e-4, absolute ||r||_2 unless relative_tolerance */
  float lambda_smooth;    /* 1   */
  float lambda_rot;       /* 1, unused (as in the reference) */
  float lambda_damping;   /* 1e-2 */
  int32_t debug;          /* 0   */
  /* options with no reference counterpart (0 = reference behaviour) */
  int32_t relative_tolerance; /* 1: stop when ||r|| < tol * ||b||; the breakdown
                                 stop is then p.Ap <= 0 instead of the
                                 reference's |p.Ap| < 1e-12                   */
  int32_t precond_fp64;       /* 1: 6x6 Cholesky inverses computed in FP64 */
  int32_t pcg_pipelined;       /* 1: A p by the recurrence A z + beta A p, so that the
                                  product overlaps the (r.r, r.z) all-reduce; 2: in
                                  addition p.Ap from z.Az (Chronopoulos-Gear), reduced
                                  with r.r and r.z: one all-reduce per iteration; 3:
                                  pipelined (Ghysels-Vanroose): A M^-1 A p and the like
                                  as recurrences, so that the one all-reduce overlaps
                                  the halo exchange and the product (FP64 recurrences).
                                  The same iterates in exact arithmetic.  Form 3 suits
                                  fixed iteration counts: its recurrence residual drifts
                                  from the true one near 1e-6, so a tolerance stop
                                  should use form 2 */
  int32_t arap_consistent_sign; /* 1: ARAP rotation Jacobian rows with the sign of
                                   d(omega x q)/d(omega), as the data term and the
                                   exp map use; 0 keeps the reference's negated rows
                                   (kernels/solver_kernels.cu:145-153) */
  int32_t jtj_precision;       /* data-term JtJ blocks on the matrix cores in FP32 (0),
                                  BF16 (1) or FP16 (2), or with each operand split
                                  into two BF16 (3) / FP16 (4) parts and three
                                  products per pair (about 16 / 22 significant bits);
                                  FP32 accumulation; the RHS stays FP32 (precision
                                  form).  The FP16 forms need |weight * J| < 1000 */
  int32_t refine_iterations;   /* > 0: after each solve, one iterative-refinement
                                  step: r = b - A x accumulated in FP64, A d = r by
                                  this many PCG iterations, x += d */
  int32_t refine_jacobi;       /* 1: one iterative-refinement step after each solve
                                  whose correction solve is the block-Jacobi inverse:
                                  x += M^-1 (b - A x), residual and correction in
                                  FP64 (one kernel); refine_iterations must be 0 */
  int32_t true_residual_rounds; /* > 0 (at most 8, needs relative_tolerance = 1 and
                                  no refinement step): convergence judged by the
                                  true residual.  After the PCG, r = b - A x in FP64;
                                  while ||r|| >= tol ||b||, up to this many times, a
                                  correction PCG solves A d = r to tol ||b|| (same
                                  iteration budget) and x += d.  The result's
                                  converged / breakdowns / final_residual are then
                                  those of the true residual */
  int32_t psd_dropped_pairs;   /* 1: a data-term pair of skinning slots whose nodes
                                  are not adjacent in the pattern adds f J_k^T J_k to
                                  the diagonal block of slot k instead of being
                                  dropped (kernels/solver_kernels.cu:123-125 drops it,
                                  which can leave the system indefinite); the system
                                  is then the full J^T J plus the sum over dropped
                                  pairs of (J_k - J_l)^T (J_k - J_l): positive
                                  semidefinite before the damping */
  int32_t plain_form;          /* 1: the reference's own GPU structure of the solve
                                  (kernels/solver_kernels.cu:543-741): float atomics
                                  into the BSR blocks, one kernel per operation and a
                                  host loop around the PCG that reads every dot
                                  product back (FP64 sums); the baseline the other
                                  forms are measured against.  pcg_pipelined is
                                  ignored; needs jtj_precision, refine_* and
                                  psd_dropped_pairs 0; not graph-capturable
                                  (NRF_ERR_INVALID_ARGUMENT under capture) */
} nrf_solver_params;

/* Caller-side GN loop settings (include/fusion_app.h:72-119 names). */
typedef struct nrf_gn_params {
  int32_t gn_iterations;
  float update_scale;      /* multiplies delta_x before the clamps */
  float max_update_rot;    /* per-node ||omega|| clamp, <=0 disables */
  float max_update_trans;  /* per-node ||v|| clamp, <=0 disables */
  int32_t validate;        /* 1: reject non-finite / too-large steps (fusion_app.h:409-491) */
  float max_dx_mean;       /* mean-step limit used by the validation */
} nrf_gn_params;

/* PCG outcome of the last solve (no reference counterpart: the reference
   discards the final residual, kernels/solver_kernels.cu:569). */
typedef struct nrf_pcg_result {
  int32_t iterations;      /* PCG iterations of the last GN iteration */
  int32_t total_iterations;/* summed over GN iterations */
  int32_t flags;           /* bit0 pAp breakdown, bit1 converged, bit2 step rejected, bit3 barrier timeout */
  int32_t gn_done;         /* GN iterations whose step was applied */
  double final_residual;   /* ||r||_2 of the PCG recurrence */
  double rhs_norm;         /* ||b||_2 */
  int32_t converged;       /* PCG runs of this solve that met the tolerance */
  int32_t breakdowns;      /* PCG runs of this solve that stopped on the breakdown test */
  int32_t runs;            /* PCG runs of this solve */
  int32_t pad;
} nrf_pcg_result;

/* ----------------------------------------------------- TSDF integration */
/* Replaces TSDFVolume::integrate (src/tsdf_volume.cu:23-63) and
   tsdf_integrate_kernel (kernels/tsdf_kernels.cu:85-174).  camera_pose is
   the 4x4 row-major camera-to-world pose; it is inverted on the host exactly
   as the reference does.  Warping is on iff nodes != NULL and num_nodes > 0.
   centered (optional, device, num_nodes entries) is the per-node table of
   dq_centered(transforms[k], nodes[k].pos) from nrf_warp_centered_table; when
   NULL the kernel recomputes it per voxel.  updated_count (optional, device
   int) receives the number of voxels written (added to its current value). */
int nrf_tsdf_integrate(const nrf_volume_params *vp, nrf_voxel *voxels,
                       const float *depth, const nrf_camera *cam,
                       const float camera_pose[16], const nrf_node *nodes,
                       const nrf_dualquat *transforms, int num_nodes,
                       const nrf_dualquat *centered, const int32_t *voxel_knn,
                       const float *voxel_knn_w, const int32_t *voxel_opt_counts,
                       int min_opt_count, int32_t *updated_count, void *stream);

/* Replaces TSDFVolume::raycast (src/tsdf_volume.cu:65-92) and raycast_kernel
   (kernels/tsdf_kernels.cu:236-361, with sample_tsdf_trilinear :14-64): per
   pixel (integer coordinates) the ray from the camera centre marched from
   0.05 to 3.5 m in 0.8-voxel steps (0.5 once |tsdf| < 0.5), the first
   crossing from tsdf in (0, 1) to tsdf <= 0 interpolated linearly, the
   normal from central differences of trilinear samples one voxel apart.
   Trilinear samples read voxel k at origin + k s (no half-voxel offset) and
   fail when any of the 8 voxels has weight <= 0.  Warped (warp_point /
   warp_normal of the crossing's base-voxel 4-NN) when nodes != NULL and
   num_nodes > 0.  Outputs (device, 3 floats per pixel; canonical_vidx one
   int, optional): zero vectors and -1 where no surface is found.
   camera_pose: camera-to-world, row-major. */
int nrf_raycast(const nrf_volume_params *vp, const nrf_voxel *voxels, const nrf_camera *cam,
                const float camera_pose[16], const nrf_node *nodes, const nrf_dualquat *transforms, int num_nodes,
                const int32_t *voxel_knn, const float *voxel_knn_w, float *vertices, float *normals,
                int32_t *canonical_vidx, void *stream);

/* Fills voxels with the reference's initial state (tsdf=1, w=0, colour 128). */
int nrf_tsdf_reset(const nrf_volume_params *vp, nrf_voxel *voxels, void *stream);

/* -------------------------------------------------------- marching cubes */
/* Replaces TSDFVolume::extract_surface / extract_surface_device
   (src/marching_cubes.cu:455-541, kernels K2-K4 :349-451).  Output is the
   reference's triangle soup (3 vertices per triangle, per-face normals,
   triangles[t] = {3t, 3t+1, 3t+2}) in increasing cube-index order, so it is
   deterministic.  counts (device, 2 ints) receives {vertices, triangles}
   (the full count even if it exceeds max_triangles; writes are clamped).
   normals may be NULL (extract_surface_device form). */
size_t nrf_mc_workspace_size(const nrf_volume_params *vp);
int nrf_mc_extract(const nrf_volume_params *vp, const nrf_voxel *voxels,
                   void *workspace, size_t workspace_bytes, float *vertices,
                   float *normals, int32_t *triangles, int max_triangles,
                   int32_t *counts, void *stream);
/* Optional: case index per cube (device, dims product bytes; 0 for cubes the
   reference skips), from the workspace of the last nrf_mc_extract call. */
const uint8_t *nrf_mc_cases(const void *workspace);
/* No counterpart in the reference.  For callers that cannot allocate device
   memory (code under HIP graph capture, serving loops, harnesses that forbid
   hipMalloc): nrf_mc_extract with the same outputs, bit for bit, and no
   workspace argument.  Its scratch (row counts, scan status words, row
   offsets; no case array, the emission recomputes each cube's case) is a
   library-owned device array of 2.1 MB, one per device, held until the
   process ends: dims[1] * dims[2] must not exceed 262144 rows
   (NRF_ERR_TOO_LARGE otherwise, with the reason in nrf_last_error).  The
   call allocates nothing and clears the scratch with a kernel (no memset
   node).  Every call on a device shares that device's scratch: calls that
   overlap on one GPU (other streams, other threads) race. */
int nrf_mc_extract_noalloc(const nrf_volume_params *vp, const nrf_voxel *voxels, float *vertices,
                           float *normals, int32_t *triangles, int max_triangles, int32_t *counts,
                           void *stream);

/* ---------------------------------------------- compact (precision) form */
/* The TSDF stored as FP16 value + uint8 weight (3 B/voxel) and the voxel kNN
   as uint16 ids (0xffff = none) + FP16 weights (16 B/voxel).  Integration and
   marching cubes run the FP32 arithmetic of nrf_tsdf_integrate /
   nrf_mc_extract on the decoded state and round once when storing (FP16
   round-to-nearest-even; weight rounded to the nearest integer, exact when
   decay_alpha = 1).  tsdf16 holds IEEE binary16 bit patterns. */
int nrf_tsdf16_reset(const nrf_volume_params *vp, uint16_t *tsdf16, uint8_t *weight8, void *stream);
int nrf_tsdf16_from_voxels(const nrf_volume_params *vp, const nrf_voxel *voxels, uint16_t *tsdf16,
                           uint8_t *weight8, void *stream);
int nrf_knn_compact(const nrf_volume_params *vp, const int32_t *voxel_knn, const float *voxel_knn_w,
                    uint16_t *knn16, uint16_t *knn_w16, void *stream);
/* Warped when num_nodes > 0 (centered = nrf_warp_centered_table output). */
int nrf_tsdf16_integrate(const nrf_volume_params *vp, uint16_t *tsdf16, uint8_t *weight8, const float *depth,
                         const nrf_camera *cam, const float camera_pose[16], const nrf_dualquat *centered,
                         int num_nodes, const uint16_t *knn16, const uint16_t *knn_w16,
                         int32_t *updated_count, void *stream);
/* Same workspace and output conventions as nrf_mc_extract (no normals). */
int nrf_mc16_extract(const nrf_volume_params *vp, const uint16_t *tsdf16, const uint8_t *weight8,
                     void *workspace, size_t workspace_bytes, float *vertices, int32_t *triangles,
                     int max_triangles, int32_t *counts, void *stream);

/* ------------------------------------------------------------ warp field */
/* Replaces WarpField::compute_voxel_knn (src/warp_field.cu:957-984) and
   compute_voxel_knn_kernel (:299-405): brute-force 4-NN over all nodes for
   voxels with weight > 0, ids -1 / weights 0 elsewhere. */
int nrf_voxel_knn(const nrf_volume_params *vp, const nrf_voxel *voxels,
                  const nrf_node *nodes, int num_nodes, int32_t *voxel_knn,
                  float *voxel_knn_w, void *stream);

/* Replaces WarpField::apply_twist_increment (src/warp_field.cu:1121-1141) and
   apply_twist_kernel (:250-291): T_i <- dq(exp(clamp(s * dx_i))) * T_i.
   centered (optional) is refreshed with dq_centered(T_i, nodes[i].pos). */
int nrf_warp_apply_twist(const float *delta_x, nrf_dualquat *transforms,
                         int num_nodes, float max_rot, float max_trans,
                         float update_scale, const nrf_node *nodes,
                         nrf_dualquat *centered, void *stream);

/* Per-node table dq_centered(transforms[k], nodes[k].pos)
   (include/se3_math.cuh:405-415), the operand of every DQ blend. */
int nrf_warp_centered_table(const nrf_node *nodes, const nrf_dualquat *transforms,
                            int num_nodes, nrf_dualquat *centered, void *stream);

/* Warps points through their 4-NN: out[i] = warp_point(points[i], ...)
   (include/se3_math.cuh:450-510). */
int nrf_warp_points(const float *points, const int32_t *knn_ids,
                    const float *knn_ws, int count, const nrf_node *nodes,
                    const nrf_dualquat *transforms, int num_nodes, float *out,
                    void *stream);

/* Rotates normals through their 4-NN as warp_normal does
   (include/se3_math.cuh:512-558): blend of the uncentred transforms' real
   parts, no weight-sum test. */
int nrf_warp_normals(const float *normals, const int32_t *knn_ids,
                     const float *knn_ws, int count,
                     const nrf_dualquat *transforms, int num_nodes, float *out,
                     void *stream);

/* Association input: packs one correspondence per point (src = canonical
   point, 4-NN copied from the voxel containing it, weight 1 or weights[i])
   as compute_pixel_knn_kernel (kernels/tsdf_kernels.cu:573-604) and
   find_correspondences_kernel (:386-563) fill Correspondence.  A point is
   valid iff valid[i] != 0, it lies in the volume and its voxel has nodes. */
int nrf_corr_pack(const nrf_volume_params *vp, const int32_t *voxel_knn,
                  const float *voxel_knn_w, const float *canonical,
                  const float *dst, const float *normals, const float *weights,
                  const uint8_t *valid, int count, nrf_corr *out, void *stream);

/* ------------------------------------------------------ graph maintenance */
/* The graph-update frames of the reference.  nodes / transforms are HOST
   arrays of capacity max_nodes (the reference keeps this host mirror and
   uploads it afterwards); *num_nodes is updated.  A pipeline picks up a
   changed graph with nrf_pipeline_set_nodes. */

/* Replaces WarpField::add_nodes_from_surface (src/warp_field.cu:709-841):
   greedy, in vertex order, a vertex becomes a node unless an existing node
   lies closer than min_dist; its graph neighbours are the (at most 8)
   nearest nodes within max(2.5 min_dist, node_radius) whose normals agree
   (dot >= 0.7), weight 1; its transform the 1/(d + 1e-6)-weighted blend of
   those that existed before the call (identity without any).  normals may
   be NULL. */
int nrf_graph_add_nodes(nrf_node *nodes, nrf_dualquat *transforms, int *num_nodes, int max_nodes,
                        const float *vertices, const float *normals, int num_vertices, float min_dist,
                        float node_radius, int *added);

/* Replaces WarpField::add_nodes_from_mesh_geodesic (src/warp_field.cu:
   843-954, with geodesic_neighbors_for_nodes_gpu :596-694 and kernels K25-
   K27 :488-590): nrf_graph_add_nodes, then, when nodes were added, every
   node's edges rebuilt from geodesic distances over the mesh (vertices
   welded on a 0.05 min_dist lattice, triangle edges weighted by length,
   each node at its nearest graph vertex): the (at most 8) nearest other
   nodes within max(3.5 min_dist, 1.5 node_radius) whose normals agree,
   weight exp(-d^2 / 2 r^2); new nodes' transforms blended from their
   pre-existing geodesic neighbours (weights 1 / (sqrt d + 1e-6)).  The
   distances are the converged bounded shortest paths (the reference stops
   its relaxation after 96 rounds).  triangles: 3 indices into vertices
   each (the marching-cubes soup: {3t, 3t+1, 3t+2}).  NRF_ERR_TOO_LARGE when
   a node's neighbourhood exceeds the device table (12288 graph vertices
   within the geodesic radius). */
int nrf_graph_add_nodes_geodesic(nrf_node *nodes, nrf_dualquat *transforms, int *num_nodes, int max_nodes,
                                 const float *vertices, const float *normals, int num_vertices,
                                 const int32_t *triangles, int num_triangles, float min_dist, float node_radius,
                                 int *added, void *stream);

/* = WarpField::PruneParams (include/warp_field.h:44-53 and defaults). */
typedef struct nrf_prune_params {
  int32_t enabled;               /* 1 */
  int32_t remove_disconnected;   /* 1 */
  int32_t min_observed_voxels;   /* 8 */
  int32_t min_surface_voxels;    /* 2 */
  int32_t min_component_size;    /* 4 */
  float support_radius_factor;   /* 1.0 */
  float surface_tsdf_abs;        /* 0.45 */
  float empty_tsdf;              /* 0.75 */
} nrf_prune_params;

/* Replaces WarpField::prune_nodes (src/warp_field.cu:986-1104) and
   score_node_tsdf_support_kernel (:407-472): a node is dropped when the
   observed voxels of its ball (radius max(radius * factor, voxel)) are
   confidently empty without surface support, then (remove_disconnected)
   every graph component smaller than min_component_size; the survivors are
   compacted in order with their neighbour lists remapped.  voxels: device;
   nodes / transforms: host. */
int nrf_graph_prune(const nrf_volume_params *vp, const nrf_voxel *voxels, nrf_node *nodes,
                    nrf_dualquat *transforms, int *num_nodes, const nrf_prune_params *params, int *removed,
                    void *stream);

/* ------------------------------------------------ projective association */
/* The step that produces the solver's correspondences and the integration's
   optimised-voxel counts from the live surface prediction (per-pixel live
   vertices, normals and canonical hit voxel, world frame) and the depth
   frame.  Vectors are 3 floats per pixel, x fastest; camera_pose is the
   row-major 4x4 camera-to-world pose (inverted on the host as
   include/fusion_math.h:80-95 does). */

/* = the association fields of the pipeline parameters
   (include/fusion_app.h:79-82 and their defaults). */
typedef struct nrf_assoc_params {
  float dist_threshold;   /* 0.05: point-to-plane distance limit (m) */
  float angle_threshold;  /* 0.7: minimum |cos| between live and depth normals */
  float view_threshold;   /* 0.2: minimum |cos| between live normal and view ray */
  int32_t search_radius_px; /* 3: half-size of the search window around the projection */
} nrf_assoc_params;

/* Replaces compute_depth_normals_kernel (kernels/tsdf_kernels.cu:180-230):
   the normal of each pixel from its right and lower neighbours' back
   projections, pointing to the camera; zero where a depth is missing and in
   the last row and column. */
int nrf_depth_normals(const float *depth, const nrf_camera *cam, float *normals, void *stream);

/* Replaces compute_pixel_knn_kernel (kernels/tsdf_kernels.cu:573-604): the
   4-NN of each pixel's canonical hit voxel (hit_voxel[px] < 0: ids -1,
   weights 0). */
int nrf_pixel_knn(const int32_t *hit_voxel, int num_pixels, const int32_t *voxel_knn, const float *voxel_knn_w,
                  int32_t *pixel_knn, float *pixel_knn_w, void *stream);

/* Replaces find_correspondences_kernel (kernels/tsdf_kernels.cu:386-563):
   per pixel, the live vertex projected into the depth frame and the best
   depth pixel of the (2 r + 1)^2 window by point-to-plane distance (ties to
   the first in row-major window order), with the reference's distance,
   normal-angle and view-angle tests and confidence weight; corrs[px] gets
   src = live vertex, dst / normal in the world frame and the pixel's 4-NN.
   Pixels without one get valid = 0, weight 0, zero vectors and ids -1 (the
   reference leaves those fields unset).  num_valid (device int) and stats
   (optional, device, 6 ints: invalid live, out of projection, no depth, none
   close, angle or view, valid) are added to. */
int nrf_find_correspondences(const float *live_vertices, const float *live_normals, const float *depth,
                             const float *depth_normals, const nrf_camera *cam, const float camera_pose[16],
                             const int32_t *pixel_knn, const float *pixel_knn_w, const nrf_assoc_params *params,
                             nrf_corr *corrs, int32_t *num_valid, int32_t *stats, void *stream);

/* Replaces mark_optimized_voxels_kernel (kernels/tsdf_kernels.cu:606-651):
   every voxel of the 3x3x3 neighbourhood of a valid correspondence's hit
   voxel gets voxel_opt_counts += 1 once per frame (voxel_opt_frame holds the
   last frame that counted it).  The counts do not depend on the order of the
   updates. */
int nrf_mark_optimized_voxels(const nrf_corr *corrs, const int32_t *hit_voxel, int num_pixels,
                              const nrf_volume_params *vp, int frame_id, int32_t *voxel_opt_counts,
                              int32_t *voxel_opt_frame, void *stream);

/* --------------------------------------------- live surface prediction */
/* Replaces skin_vertices_from_voxel_knn_kernel (src/skinning_kernels.cu:
   29-63): each canonical vertex warped through the 4-NN of the voxel that
   contains it (floor of (p - origin) / voxel_size); unchanged outside the
   volume or when num_nodes = 0.  centered = nrf_warp_centered_table of the
   current transforms (the same blend as warp_point). */
int nrf_skin_vertices(const float *canonical_vertices, int num_vertices, const nrf_volume_params *vp,
                      const nrf_dualquat *centered, int num_nodes, const int32_t *voxel_knn,
                      const float *voxel_knn_w, float *warped_vertices, void *stream);

/* Replaces rasterize_triangles_z_kernel + resolve_rasterized_triangles_kernel
   (src/mesh_raster_kernels.cu:76-200) and the caller's resets
   (include/fusion_app.h:1226-1230): the warped mesh is rasterised into the
   camera (pixel centres at +0.5, barycentric test -1e-5, nearest depth then
   lowest triangle index wins); each covered pixel gets the interpolated live
   vertex, the face normal and the voxel of the interpolated canonical point
   (and, when canonical_points != NULL, that point); the others zero vectors
   and hit voxel -1.  triangles: 3 vertex indices each.  The workspace holds
   one 64-bit key per pixel (nrf_raster_workspace_size). */
size_t nrf_raster_workspace_size(const nrf_camera *cam);
int nrf_rasterize_mesh(const float *canonical_vertices, const float *warped_vertices, const int32_t *triangles,
                       int num_vertices, int num_triangles, const nrf_camera *cam, const float camera_pose[16],
                       const nrf_volume_params *vp, void *workspace, size_t workspace_bytes,
                       float *live_vertices, float *live_normals, int32_t *hit_voxel, float *canonical_points,
                       void *stream);

/* ---------------------------------------------------------------- solver */
typedef struct nrf_solver nrf_solver;

/* GaussNewtonSolver(const Params&, int max_nodes) (kernels/solver_kernels.cu:486-496).
   Returns NULL (and set_params NRF_ERR_INVALID_ARGUMENT) for pcg_iterations
   or refine_iterations outside [0, 32767], jtj_precision outside [0, 4] or
   pcg_pipelined outside [0, 3]. */
nrf_solver *nrf_solver_create(const nrf_solver_params *params, int max_nodes);
void nrf_solver_destroy(nrf_solver *s);
int nrf_solver_set_params(nrf_solver *s, const nrf_solver_params *params);

/* build_sparsity_pattern (kernels/solver_kernels.cu:498-541) from HOST nodes.
   Synchronous (uploads the pattern). */
int nrf_solver_set_pattern(nrf_solver *s, const nrf_node *host_nodes, int num_nodes);

/* GaussNewtonSolver::solve (kernels/solver_kernels.cu:543-570): one
   linearisation with the stored corrs[i].src, delta_x (6N floats, device) is
   zeroed and overwritten.  Rebuilds the pattern (D2H of nodes, sync) only when
   num_nodes differs from the previous call.  If result != NULL the call
   synchronises the stream and fills it. */
int nrf_solver_solve(nrf_solver *s, const nrf_corr *corrs, int num_corrs,
                     const nrf_node *nodes, const nrf_dualquat *transforms,
                     int num_nodes, float *delta_x, nrf_pcg_result *result,
                     void *stream);

/* The caller-side GN loop of include/fusion_app.h:256-646 with the
   association held fixed: for gn in [0, gn_iterations):
     src_i = warp_point(canonical_i) through corrs[i].node_ids/node_ws,
     assemble, precondition, PCG, validate, T <- exp(clamp(scale*dx)) * T.
   canonical (device, 3*num_corrs floats) replaces the stored src; if NULL
   the stored src is used and not re-warped.  transforms are updated in
   place.  centered (device, num_nodes) must hold the table of the incoming
   transforms and is kept current.  The pattern must have been set. */
int nrf_solver_solve_gn(nrf_solver *s, const nrf_corr *corrs,
                        const float *canonical, int num_corrs,
                        const nrf_node *nodes, nrf_dualquat *transforms,
                        nrf_dualquat *centered, int num_nodes,
                        const nrf_gn_params *gn, float *delta_x,
                        nrf_pcg_result *result, void *stream);

/* system_matrix() / rhs() (include/solver.h:38-39): device pointers owned by
   the solver, valid until the next solve or pattern change. */
int nrf_solver_system(const nrf_solver *s, const int32_t **row_ptr,
                      const int32_t **col_idx, const float **values,
                      int *num_blocks, const float **rhs, const float **precond);
/* Device pointer to the last nrf_pcg_result written by the solver (for
   stream-ordered readers that do not want a host sync). */
const nrf_pcg_result *nrf_solver_device_result(const nrf_solver *s);

/* Layout of the persistent PCG (no reference counterpart): lanes per scalar
   row (1, 2 or 4) and the target number of nodes per workgroup (capped at
   32, 80 and 170 for 4, 2 and 1 lanes per row); rebuilds the partition of
   the current pattern.
   Default: 4 lanes per row, 24 nodes per workgroup.  The textbook PCG form
   (pcg_pipelined = 0) needs 4 lanes per row.  Synchronous; no solve of this
   solver may be in flight, and a pipeline captured into a graph is captured
   again at its next step. */
int nrf_solver_set_pcg_layout(nrf_solver *s, int lanes_per_row, int wg_nodes);

/* PCG alone on the last assembled system (A, b, block-Jacobi inverse), with
   the solver's current parameters: for timing the solve and its stopping
   rules in isolation (no reference counterpart).  Stream-ordered; fills
   result (and synchronises) when result != NULL. */
int nrf_solver_resolve(nrf_solver *s, float *delta_x, nrf_pcg_result *result, void *stream);

/* True residual of the last assembled system for a device vector x:
   out[0] = ||b - A x||_2, out[1] = ||b||_2, products and sums in FP64
   (no reference counterpart: the reference discards the residual,
   kernels/solver_kernels.cu:569).  Synchronises the stream. */
int nrf_solver_true_residual(nrf_solver *s, const float *x, double out[2], void *stream);

/* Running totals over every solve since creation (no reference counterpart:
   the reference prints per-iteration diagnostics instead).  Accumulated on
   the device by the next solve's first kernel, so reading them costs no
   per-step synchronisation; this call synchronises the stream. */
typedef struct nrf_solver_totals {
  int64_t solves;          /* nrf_solver_solve + nrf_solver_solve_gn calls */
  int64_t gn_done;         /* GN steps applied */
  int64_t pcg_iterations;  /* PCG iterations run */
  int64_t rejected;        /* solves that ended on a rejected step */
  int64_t pcg_runs;        /* PCG runs (one per GN iteration that reached the PCG) */
  int64_t pcg_converged;   /* PCG runs that met the tolerance */
  int64_t pcg_breakdowns;  /* PCG runs that stopped on the breakdown test */
} nrf_solver_totals;
int nrf_solver_get_totals(nrf_solver *s, nrf_solver_totals *out, void *stream);

/* ------------------------------------------------------------- pipeline */
/* The per-frame non-rigid step (include/fusion_app.h:253-669 with the
   association as a caller input): GN solve, DQ-warped integration, marching
   cubes.  The handle owns the volume, node table, kNN field and mesh. */
typedef struct nrf_pipeline nrf_pipeline;

typedef struct nrf_pipeline_config {
  nrf_volume_params volume;
  nrf_camera camera;
  nrf_solver_params solver;
  nrf_gn_params gn;
  int32_t max_nodes;
  int32_t max_corrs;
  int32_t max_triangles;
  int32_t min_opt_count;   /* integrate_min_optimized_count */
  int32_t use_graph;       /* 1: replay the step as a captured hipGraph (nrf_pipeline_graph_captures) */
  int32_t storage16;       /* 1: precision form of the per-frame step: FP16 TSDF + u8
                              weight and u16 / FP16 voxel kNN (nrf_tsdf16_*, nrf_mc16_*);
                              frame 0 and the kNN snapshot use the FP32 volume, which
                              is converted once by nrf_pipeline_set_nodes */
} nrf_pipeline_config;

nrf_pipeline *nrf_pipeline_create(const nrf_pipeline_config *cfg);
void nrf_pipeline_destroy(nrf_pipeline *p);
/* Device pointers owned by the pipeline. */
nrf_voxel *nrf_pipeline_voxels(nrf_pipeline *p);
nrf_dualquat *nrf_pipeline_transforms(nrf_pipeline *p);
nrf_node *nrf_pipeline_nodes(nrf_pipeline *p);
/* The voxel kNN (computed by nrf_pipeline_set_nodes): read-only for the
   caller; the pipeline keeps a map of the 64-voxel runs along x that hold an
   id >= 0 (the warped integration reads the kNN of those runs only), made
   from the kNN at nrf_pipeline_set_nodes. */
int32_t *nrf_pipeline_knn(nrf_pipeline *p);
float *nrf_pipeline_knn_w(nrf_pipeline *p);
float *nrf_pipeline_mesh(nrf_pipeline *p, int32_t **counts);
nrf_solver *nrf_pipeline_solver(nrf_pipeline *p);
/* Compact volume and kNN of a storage16 pipeline (NULLs otherwise). */
int nrf_pipeline_volume16(nrf_pipeline *p, uint16_t **tsdf16, uint8_t **weight8, uint16_t **knn16,
                          uint16_t **knn_w16);
/* Unwarped integration of a frame (frame 0 of fusion_app.h:233). */
int nrf_pipeline_integrate_rigid(nrf_pipeline *p, const float *depth,
                                 const float camera_pose[16], void *stream);
/* Uploads host nodes (identity transforms), builds the pattern and the voxel
   kNN snapshot (fusion_app.h:994-1003 minus node sampling).  Synchronous. */
int nrf_pipeline_set_nodes(nrf_pipeline *p, const nrf_node *host_nodes, int num_nodes);
/* One non-rigid frame: solve_gn on (corrs, canonical), warped integrate of
   depth, marching cubes.  Inputs are device pointers. */
int nrf_pipeline_step(nrf_pipeline *p, const float *depth,
                      const float camera_pose[16], const nrf_corr *corrs,
                      const float *canonical, int num_corrs, void *stream);
/* Per-stage device times of the last step in ms (hipEvent based; valid after
   the stream has been synchronised): [solve, integrate, mc].  After a graph
   replay (use_graph) out_ms[0] is the whole step and the others are -1. */
int nrf_pipeline_stage_ms(nrf_pipeline *p, float out_ms[3]);
/* With use_graph: the number of times the step was captured.  A step copies
   its depth, correspondences and canonical points into pipeline-owned
   buffers and writes its pose to device memory ahead of the replay (on the
   caller's stream), so one capture serves every frame; it is captured again
   only when num_corrs, the presence of canonical points or the solver's
   buffers / pattern change. */
long long nrf_pipeline_graph_captures(nrf_pipeline *p);
/* No counterpart in the reference (which runs one frame after the other on
   the legacy stream).  enable = 1: an eager step's warped integration and
   marching cubes are deferred and run on a stream of the pipeline's own
   alongside the next step's first PCG (enqueued when that step reaches it),
   from a copy of the step's depth (made on that stream; the caller's stream
   waits for the copy, so the depth buffer may be overwritten once the step
   returns) and a snapshot of the node transforms as the step's solve left
   them.  The next step's first twist update waits only for the snapshot, so
   every frame computes what it computes unoverlapped (tested bit for bit).
   The persistent PCG then shares the device with them: its workgroups may
   start late, which every wait in it tolerates (they are bounded, not
   timed).  A step's volume and mesh are complete only after
   nrf_pipeline_join on the reader's stream or any other pipeline call,
   including nrf_pipeline_voxels / _mesh / _volume16 (which enqueue a
   deferred fusion and wait for it); a device synchronisation alone does not
   run a deferred fusion, and the step's stream must outlive it (a join or
   another pipeline call enqueues it behind that stream's work).
   Graph-replayed and per-kernel profiled steps do not overlap.
   Synchronises the host when switching. */
int nrf_pipeline_set_overlap(nrf_pipeline *p, int enable);
/* Enqueues a deferred fusion of the last overlapped step and makes `stream`
   wait for it (a no-op otherwise). */
int nrf_pipeline_join(nrf_pipeline *p, void *stream);
/* Per-kernel profiling: with ring_steps > 0, every step records hipEvents
   around each kernel group on the step's stream (eager launches, no graph)
   into a ring of ring_steps steps; 0 disables.  nrf_pipeline_profile_read
   returns, for the step `step_back` steps before the last one, the device
   time in ms of [membership build, assembly, PCG, twist update, integration,
   marching cubes], assembly / PCG / update summed over the GN iterations. */
int nrf_pipeline_profile(nrf_pipeline *p, int ring_steps);
/* No counterpart in the reference; for timing runs that must not be slowed by
   the events (each costs about 3 us of device time here): with ring_steps > 0
   every step records only two hipEvents, around the PCG launch of GN
   iteration (step mod gn_iterations), and nrf_pipeline_profile_read returns
   that one launch's ms in out_ms[2] (the other entries 0). */
int nrf_pipeline_profile_pcg(nrf_pipeline *p, int ring_steps);
/* No counterpart in the reference: the profiling events of either form are
   recorded on every k-th step only (k >= 1; 1 after each profile set-up);
   the ring then holds the sampled steps, and step_back of
   nrf_pipeline_profile_read counts sampled steps.  Restarts the count. */
int nrf_pipeline_profile_every(nrf_pipeline *p, int k);
int nrf_pipeline_profile_read(nrf_pipeline *p, int step_back, float out_ms[6]);
/* Voxels written by the warped integrations of all steps so far (syncs). */
long long nrf_pipeline_updated_voxels(nrf_pipeline *p);

/* ------------------------------------------------------- tracking frame */
#define NRF_TRACK_MAX_GN 16

/* The tracking fields of the pipeline parameters (include/fusion_app.h:
   79-83 and the line search of :494-641). */
typedef struct nrf_track_params {
  nrf_assoc_params assoc;
  int32_t min_valid_corrs; /* 2000: an association with fewer valid correspondences ends the GN loop */
  int32_t line_search;     /* 1: the reference's ten-scale search per GN iteration; 0: the step is
                              applied at gn.update_scale (the pipeline step's policy) */
} nrf_track_params;

/* What the last nrf_pipeline_track did (device-side record). */
typedef struct nrf_track_stats {
  int32_t gn_run;      /* GN iterations whose association ran */
  int32_t gn_applied;  /* GN iterations whose step was applied */
  int32_t stop_reason; /* 0 all gn_iterations ran, 1 too few correspondences, 2 step rejected by the
                          validation, 3 no trial scale lowered the RMSE, 4 base RMSE not finite or
                          update_scale <= 0 */
  int32_t last_valid;  /* valid correspondences of the association that marked the voxels */
  int32_t num_valid[NRF_TRACK_MAX_GN]; /* of GN iteration g's association */
  float best_scale[NRF_TRACK_MAX_GN];  /* line search: the scale applied at iteration g (0: none) */
  double base_rmse[NRF_TRACK_MAX_GN];  /* line search: correspondence RMSE before the step */
  double best_rmse[NRF_TRACK_MAX_GN];  /* line search: RMSE at the chosen scale */
} nrf_track_stats;

/* One non-rigid frame with its own association, the caller contract of
   DynamicFusionPipeline::process_frame (include/fusion_app.h:253-669) for
   frames after the first, without rigid tracking, SIFT and the bounding
   box: depth normals; per GN iteration (gn.gn_iterations, at most
   NRF_TRACK_MAX_GN) the live surface prediction of the canonical mesh
   (nrf_skin_vertices + nrf_rasterize_mesh of the last marching cubes output,
   extracted first when the volume changed since), the association
   (nrf_find_correspondences, all W x H slots passed to the solver as
   fusion_app.h:387 does), the min_valid_corrs test, one linearisation of
   the stored src and the step (validated when gn.validate; with
   line_search the ten trial scales, each with its own prediction and
   association, and the lowest RMSE kept); then nrf_mark_optimized_voxels
   with frame_id (the reference's frame_count_; the per-voxel frame record
   starts at 0), the warped integration gated by min_opt_count and marching
   cubes.  Stream-ordered: every decision is taken on the device, and no
   host synchronisation happens after the first call (which allocates the
   tracking buffers: association maps, W x H correspondence slots, 8 B per
   voxel of optimised-voxel counts).  depth: device, W x H. */
int nrf_pipeline_track(nrf_pipeline *p, const float *depth, const float camera_pose[16],
                       const nrf_track_params *params, int frame_id, void *stream);
/* The record of the last nrf_pipeline_track (synchronises the device). */
int nrf_pipeline_track_stats(nrf_pipeline *p, nrf_track_stats *out);
/* Device maps of the last nrf_pipeline_track, owned by the pipeline (NULL
   before the first): live vertices / normals (3 floats per pixel), canonical
   hit voxel per pixel, the W x H correspondence slots, the per-voxel
   optimised counts and the depth normals. */
int nrf_pipeline_track_maps(nrf_pipeline *p, float **live_vertices, float **live_normals, int32_t **hit_voxel,
                            nrf_corr **corrs, int32_t **voxel_opt_counts, float **depth_normals);

/* ---------------------------------------------------------------- batch
   Build-only form (no counterpart in the reference): up to NRF_MAX_BATCH
   independent pipelines (their own volume, camera, nodes and solver
   settings) stepped together, so that small instances share the device
   instead of each leaving most of it idle.  Per frame, the membership
   build, the correspondence records, the assembly, the PCG, the GN update
   and the warped integration each run as one launch over all instances
   (the persistent PCG as one launch per group of instances whose
   workgroups fit on the device together); marching cubes runs per
   instance.  A batch step of an instance gives the same bits as that
   pipeline's own nrf_pipeline_step with the same PCG layout.
   Requirements: every pipeline has its nodes set, the FP32 storage form,
   one of the persistent PCG forms (pcg_pipelined 1, 2 or 3, the same for
   all), no refinement step, and all share jtj_precision and gn_iterations.  Changing a member's nodes or PCG layout
   afterwards invalidates the batch (destroy and create it again). */
#define NRF_MAX_BATCH 32
typedef struct nrf_batch nrf_batch;
typedef struct nrf_batch_frame {
  const float *depth;       /* device, that instance's camera size */
  float camera_pose[16];    /* host values, row-major camera-to-world */
  const nrf_corr *corrs;    /* device, num_corrs records */
  const float *canonical;   /* device, 3 floats per corr, or NULL (stored src) */
  int32_t num_corrs;        /* <= that pipeline's max_corrs */
  int32_t pad;
} nrf_batch_frame;
typedef struct nrf_batch_tuning {
  int32_t lanes_per_row;    /* PCG lanes per scalar row: 1, 2 or 4 */
  int32_t wg_nodes;         /* PCG target nodes per workgroup */
  int32_t z_chunk;          /* integration voxels per lane along z: 1, 2, 4, 8, 16 */
  int32_t launch_groups;    /* out: persistent PCG launches per iteration */
  float pcg_ms;             /* out (nrf_batch_autotune): PCG stage time, all instances */
  float integrate_ms;       /* out (nrf_batch_autotune): integration time, all instances */
} nrf_batch_tuning;
/* tuning NULL: 4 lanes per row, 32 nodes per workgroup, z_chunk 4.
   use_graph: replay each step as a captured hipGraph (re-captured when the
   frame pointers, counts or poses change). */
nrf_batch *nrf_batch_create(nrf_pipeline *const *pipelines, int count, const nrf_batch_tuning *tuning,
                            int use_graph);
void nrf_batch_destroy(nrf_batch *b);
int nrf_batch_set_tuning(nrf_batch *b, const nrf_batch_tuning *tuning);
int nrf_batch_get_tuning(const nrf_batch *b, nrf_batch_tuning *tuning);
/* Times every candidate PCG layout on the systems of `frames` at the
   current transforms, and every integration z chunk on scratch copies of
   the volumes, then applies the fastest.  Leaves the instances' state
   (transforms, volumes, solver totals) as it was.  use_cache: reuse the
   choice made earlier in this process for the same shapes (node counts,
   row lengths, volume and image sizes).  Synchronous. */
int nrf_batch_autotune(nrf_batch *b, const nrf_batch_frame *frames, int use_cache, nrf_batch_tuning *out,
                       void *stream);
/* One frame of every instance (frames[k] for pipeline k). */
int nrf_batch_step(nrf_batch *b, const nrf_batch_frame *frames, void *stream);
/* Device times of the last step in ms (after a synchronisation):
   [solve, integrate, mc], all instances; after a graph replay out_ms[0] is
   the whole step and the others are -1. */
int nrf_batch_stage_ms(nrf_batch *b, float out_ms[3]);

/* ------------------------------------------------ learned preconditioner
   Build-only component (no counterpart in the reference).  A 3-layer
   message-passing network over the deformation graph (hidden width 64,
   BF16 weights, FP32 accumulation) gives the blocks of a block lower-
   triangular L; the preconditioner is M^-1 = BJ^-1 + L L^T with
   L_ii = C_i lower(Td_i), L_ij = C_i Te_ij (neighbours j < i) and
   C_i C_i^T = BJ_i^-1.  variant 0: diagonal blocks only (folded into the
   solver's block-Jacobi inverses by nrf_lp_fold, so the solver's PCG applies
   it); variant 1: the graph pattern.  Parameters are a flat FP32 vector of
   nrf_lp_param_count() floats (layout of oracle/learned.cpp), trained on the
   device: nrf_lp_forward + nrf_lp_loss_grad (probe loss
   sum_p |M^-1 A v_p - v_p|^2 / (P 6N) and its gradient) + nrf_lp_adam. */
typedef struct nrf_lp nrf_lp;
nrf_lp *nrf_lp_create(int variant, int max_nodes, int max_probes);
void nrf_lp_destroy(nrf_lp *lp);
int nrf_lp_param_count(void);
float *nrf_lp_params(nrf_lp *lp);   /* device, FP32 master parameters */
float *nrf_lp_grads(nrf_lp *lp);    /* device, gradient of the last nrf_lp_loss_grad */
int nrf_lp_reset_optimizer(nrf_lp *lp);
/* The graph (host nodes): reverse adjacency for the transposed products. */
int nrf_lp_set_graph(nrf_lp *lp, const nrf_node *host_nodes, int num_nodes);
/* Features, network and L blocks for the solver's assembled system (its
   diagonal blocks and block-Jacobi inverses); nodes on the device. */
int nrf_lp_forward(nrf_lp *lp, const nrf_solver *s, const nrf_node *nodes, int num_nodes, void *stream);
/* Loss over num_probes device vectors [num_probes][6N] and its gradient
   (into nrf_lp_grads), after nrf_lp_forward on the same solver; loss (host,
   may be NULL) synchronises the stream. */
int nrf_lp_loss_grad(nrf_lp *lp, const nrf_solver *s, const nrf_node *nodes, const float *probes,
                     int num_probes, double *loss, void *stream);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) on the FP32 parameters. */
int nrf_lp_adam(nrf_lp *lp, float lr, int step, void *stream);
/* variant 0: the solver's block-Jacobi inverses become BJ_i^-1 + L_ii L_ii^T
   (until its next assembly), after nrf_lp_forward on that solver.  A node
   whose correction or folded block is not finite keeps its block-Jacobi
   inverse (the fallback; counted for nrf_lp_fallbacks). */
int nrf_lp_fold(nrf_lp *lp, nrf_solver *s, void *stream);
/* Nodes that kept block-Jacobi in the last nrf_lp_fold (synchronises). */
int nrf_lp_fallbacks(nrf_lp *lp, int *count, void *stream);
/* Device pointers of the last forward: C (N x 36), Td (N x 21), Te
   (N x 8 x 36), features (N x 28). */
int nrf_lp_blocks(const nrf_lp *lp, const float **C, const float **Td, const float **Te, const float **features);

#ifdef __cplusplus
}
#endif
#endif /* NRF_H */
