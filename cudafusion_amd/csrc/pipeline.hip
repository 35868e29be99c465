// pipeline.hip - the per-frame non-rigid step as one stream-ordered call.
//
// Mirrors the caller contract of DynamicFusionPipeline::process_frame
// (include/fusion_app.h:253-669) with the association held as an input:
// solve (GN loop with validation and clamped twist updates) -> warped
// integration -> marching cubes (the extract_surface_device form).  Nothing
// synchronises the host inside a step, so a step can be captured into a
// hipGraph and replayed.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "nrf_internal.h"

using namespace nrf;

struct nrf_pipeline {
  nrf_pipeline_config cfg;
  nrf_voxel *vox = nullptr;
  int32_t *knn = nullptr;
  float *knn_w = nullptr;
  uint32_t *knn_rows = nullptr;  // row occupancy of knn (knn_row_flags), set with the kNN
  nrf_node *nodes = nullptr;
  nrf_dualquat *T = nullptr;
  nrf_dualquat *centered = nullptr;
  int N = 0;
  float *delta_x = nullptr;
  void *mc_ws = nullptr;
  size_t mc_bytes = 0;
  // marching-cubes brick occupancy: set by the pipeline's integrations for
  // every brick holding an updated voxel (a superset of the bricks holding a
  // voxel with weight > 0, since only an update makes a weight positive);
  // marching cubes skips the other bricks
  uint8_t *occ = nullptr;
  float *dblk = nullptr;  // depth block maxima of the frame being integrated (integration's wave cull)
  float *verts = nullptr;
  int32_t *tris = nullptr;
  int32_t *counts = nullptr;
  int32_t *updated = nullptr;
  // precision form (storage16)
  uint16_t *t16 = nullptr, *k16 = nullptr, *kw16 = nullptr;
  uint8_t *w8 = nullptr;
  nrf_solver *solver = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  // graph replay state
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  // per-kernel profiling ring (eager mode)
  std::vector<hipEvent_t> prof;
  int prof_ring = 0, prof_per = 0, prof_G = 0, prof_pcg = 0, prof_every = 1;
  long prof_steps = 0, prof_calls = 0;  // sampled steps, all steps since the set-up
  int g_n = -1;
  bool g_has_can = false;
  unsigned g_gen = 0;
  long long g_captures = 0;
  // frame inputs of the captured step: each step copies its depth,
  // correspondences and canonical points here and writes its pose to g_pose,
  // so one capture serves every frame
  float *g_depth = nullptr;
  nrf_corr *g_corrs = nullptr;
  float *g_canon = nullptr;
  CamXf *g_pose = nullptr;
  hipStream_t cap = nullptr;  // capture stream (the caller's may be the null stream)
  // overlapped fusion (nrf_pipeline_set_overlap).  Step F's integration and
  // marching cubes are deferred (fz_pending) and enqueued on `side` at step
  // F+1's first PCG launch (pipeline_fuse_hook), so that they run alongside
  // that step's PCG:
  //  - step F+1's first kernel copies the node transforms as step F left
  //    them into snapshot slot (F+1) & 1 (solver_set_snapshot), which F's
  //    fusion reads while F+1 updates the live tables;
  //  - `side` starts its part of step F once all workgroups of F's first
  //    PCG launch have started (that launch carries the ticket fz_ticket;
  //    solver_pcg_start_wait), so after F's work before that launch and
  //    without taking the device first; it then runs step F-1's fusion and
  //    copies F's depth into depth slot F & 1; ev_copy marks the copy, and
  //    F's stream waits for it before the step returns (the caller may then
  //    overwrite its depth buffer, and every later write of a slot on the
  //    step's stream follows the fusion that read it);
  //  - a fusion that no step takes (a join, a drain) is flushed after an
  //    event recorded then on its step's stream, with its snapshot copied on
  //    `side`; the next step then waits for it (fz_flushed) before its
  //    snapshot.
  // ev_fs / ev_fused mark the last enqueued fusion's start and end.
  bool overlap = false, side_pending = false, fz_pending = false, fz_events = false, fz_flushed = false;
  bool fz_copied = false;  // the step being enqueued made its hook's copy
  hipStream_t side = nullptr;
  hipEvent_t ev_fused = nullptr, ev_copy = nullptr, ev_fs = nullptr;
  float *fz_depth = nullptr;          // two depth slots
  nrf_dualquat *T_snap = nullptr, *c_snap = nullptr;  // two snapshot slots of max_nodes entries
  const float *fz_src = nullptr;      // the depth of the step being enqueued
  long fz_steps = 0;                  // overlapped steps so far (slot parity)
  int fz_slot = 0, fz_snap = 0;       // the deferred fusion's depth and snapshot slots
  unsigned fz_ticket = 0;             // the ticket of the step's first PCG launch
  hipStream_t fz_stream = nullptr;    // the deferred step's stream (pipeline_flush)
  float fz_pose[16] = {};
  bool graph_step = false;    // the last step was a graph replay (stage events: whole step only)
  bool mesh_valid = false;    // verts / tris / counts hold the marching cubes of the current volume
  // tracking frame buffers (allocated by the first nrf_pipeline_track)
  struct {
    float *dnorm = nullptr, *warped = nullptr, *live_v = nullptr, *live_n = nullptr;
    unsigned long long *ztri = nullptr;
    int32_t *hit = nullptr, *opt = nullptr, *opt_frame = nullptr;
    nrf_corr *corrs = nullptr;
    double *part = nullptr;
    nrf_dualquat *Tsave = nullptr;
    TrackDev *dev = nullptr;
  } tk;
};

static void pipeline_free(nrf_pipeline *p) {
  void *ptrs[] = {p->vox, p->knn, p->knn_w, p->knn_rows, p->nodes, p->T, p->centered, p->delta_x, p->mc_ws, p->occ, p->dblk,
                  p->verts, p->tris, p->counts, p->updated, p->t16, p->w8, p->k16, p->kw16,
                  p->tk.dnorm, p->tk.warped, p->tk.live_v, p->tk.live_n, p->tk.ztri, p->tk.hit, p->tk.opt,
                  p->tk.opt_frame, p->tk.corrs, p->tk.part, p->tk.Tsave, p->tk.dev,
                  p->g_depth, p->g_corrs, p->g_canon, p->g_pose, p->fz_depth, p->T_snap, p->c_snap};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  for (auto &e : p->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto &e : p->prof)
    if (e) (void)hipEventDestroy(e);
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->cap) (void)hipStreamDestroy(p->cap);
  if (p->side) {
    (void)hipStreamSynchronize(p->side);
    (void)hipStreamDestroy(p->side);
  }
  for (hipEvent_t e : {p->ev_fused, p->ev_copy, p->ev_fs})
    if (e) (void)hipEventDestroy(e);
  if (p->solver) nrf_solver_destroy(p->solver);
  delete p;
}

static int pipeline_fusion(nrf_pipeline *p, const float *depth, const float camera_pose[16], hipStream_t st,
                           bool events, const CamXf *dev_pose, hipEvent_t pe_int = nullptr,
                           hipEvent_t pe_mc = nullptr, const nrf_dualquat *T = nullptr,
                           const nrf_dualquat *centered = nullptr);

// Enqueues the deferred fusion on `side`, reading its depth and snapshot slots.
static int pipeline_fuse(nrf_pipeline *p) {
  if (!p->fz_pending) return NRF_OK;
  p->fz_pending = false;
  p->side_pending = true;  // (also on an error: what was enqueued is drained)
  const size_t npx = (size_t)p->cfg.camera.width * p->cfg.camera.height;
  const size_t nn = (size_t)std::max(1, p->cfg.max_nodes);
  NRF_HIP_CHECK(hipEventRecord(p->ev_fs, p->side));
  if (int rc = pipeline_fusion(p, p->fz_depth + (size_t)p->fz_slot * npx, p->fz_pose, p->side, true, nullptr,
                               nullptr, nullptr, p->T_snap + (size_t)p->fz_snap * nn,
                               p->c_snap + (size_t)p->fz_snap * nn))
    return rc;
  NRF_HIP_CHECK(hipEventRecord(p->ev_fused, p->side));
  p->fz_events = true;
  return NRF_OK;
}

// The side stream's part of step F: behind F's first PCG launch (`ticket`)
// or, for a solve without one, after `after` (an event on F's stream): the
// deferred fusion of step F-1 (its snapshot was taken by F's first kernel),
// then F's depth into its slot.
static int pipeline_side_step(nrf_pipeline *p, hipEvent_t after, unsigned ticket = 0) {
  if (after) NRF_HIP_CHECK(hipStreamWaitEvent(p->side, after, 0));
  p->side_pending = true;
  // behind the step's first PCG launch: its workgroups start first (they
  // need whole halves of CUs), and the step's work before it (the
  // transforms' snapshot, the caller's depth) has ended
  if (ticket)
    if (int rc = solver_pcg_start_wait(p->solver, ticket, p->side)) return rc;
  if (int rc = pipeline_fuse(p)) return rc;
  const size_t npx = (size_t)p->cfg.camera.width * p->cfg.camera.height;
  NRF_HIP_CHECK(hipMemcpyAsync(p->fz_depth + (size_t)(p->fz_steps & 1) * npx, p->fz_src, npx * sizeof(float),
                               hipMemcpyDeviceToDevice, p->side));
  NRF_HIP_CHECK(hipEventRecord(p->ev_copy, p->side));
  p->fz_copied = true;
  return NRF_OK;
}

// The solver's callback once a step's first PCG launch, which carries the
// step's ticket, is enqueued.
static int pipeline_fuse_hook(void *ctx, hipStream_t st) {
  nrf_pipeline *p = static_cast<nrf_pipeline *>(ctx);
  (void)st;
  return pipeline_side_step(p, nullptr, p->fz_ticket);
}

// A deferred fusion that no step takes: after the deferred step's work on
// its stream, the snapshot into the slot the next step's first kernel would
// write, then the fusion.
static int pipeline_flush(nrf_pipeline *p) {
  if (!p->fz_pending) return NRF_OK;
  // the end of the deferred step's work: an event recorded on its stream now
  NRF_HIP_CHECK(hipEventRecord(p->ev[1], p->fz_stream));
  NRF_HIP_CHECK(hipStreamWaitEvent(p->side, p->ev[1], 0));
  p->side_pending = true;
  const size_t nn = (size_t)std::max(1, p->cfg.max_nodes), bytes = sizeof(nrf_dualquat) * (size_t)p->N;
  p->fz_snap = (int)(p->fz_steps & 1);
  NRF_HIP_CHECK(hipMemcpyAsync(p->T_snap + (size_t)p->fz_snap * nn, p->T, bytes, hipMemcpyDeviceToDevice, p->side));
  NRF_HIP_CHECK(
      hipMemcpyAsync(p->c_snap + (size_t)p->fz_snap * nn, p->centered, bytes, hipMemcpyDeviceToDevice, p->side));
  p->fz_flushed = true;
  return pipeline_fuse(p);
}

// Every entry point other than the step itself first enqueues a deferred
// fusion and waits for it (host-synchronous; the step's callers that want to
// stay stream-ordered use nrf_pipeline_join).
static int pipeline_drain(nrf_pipeline *p) {
  int rc = pipeline_flush(p);
  if (p->side_pending) {
    NRF_HIP_CHECK(hipStreamSynchronize(p->side));
    p->side_pending = false;
    p->fz_flushed = false;
  }
  return rc;
}

extern "C" nrf_pipeline *nrf_pipeline_create(const nrf_pipeline_config *cfg) {
  if (!cfg || !volume_ok(&cfg->volume) || cfg->max_nodes <= 0 || cfg->max_corrs < 0 || cfg->max_triangles < 0)
    return nullptr;
  nrf_pipeline *p = new nrf_pipeline();
  p->cfg = *cfg;
  const size_t V = (size_t)volume_count(&cfg->volume);
  bool ok = true;
  ok = ok && hipMalloc(&p->vox, V * sizeof(nrf_voxel)) == hipSuccess;
  ok = ok && hipMalloc(&p->knn, V * 4 * sizeof(int32_t)) == hipSuccess;
  ok = ok && hipMalloc(&p->knn_w, V * 4 * sizeof(float)) == hipSuccess;
  const size_t nrun = (size_t)knn_row_words(&cfg->volume);
  ok = ok && hipMalloc(&p->knn_rows, nrun * sizeof(uint32_t)) == hipSuccess;
  ok = ok && hipMalloc(&p->nodes, (size_t)cfg->max_nodes * sizeof(nrf_node)) == hipSuccess;
  ok = ok && hipMalloc(&p->T, (size_t)cfg->max_nodes * sizeof(nrf_dualquat)) == hipSuccess;
  ok = ok && hipMalloc(&p->centered, (size_t)cfg->max_nodes * sizeof(nrf_dualquat)) == hipSuccess;
  ok = ok && hipMalloc(&p->delta_x, (size_t)cfg->max_nodes * 6 * sizeof(float)) == hipSuccess;
  p->mc_bytes = nrf_mc_workspace_size(&cfg->volume);
  ok = ok && hipMalloc(&p->mc_ws, p->mc_bytes) == hipSuccess;
  const size_t nocc = mc_occupancy_bytes(&cfg->volume);
  ok = ok && hipMalloc(&p->occ, nocc) == hipSuccess;
  ok = ok && hipMalloc(&p->dblk, std::max<size_t>(1, depth_block_count(&cfg->camera)) * sizeof(float)) == hipSuccess;
  const size_t T = (size_t)std::max(1, cfg->max_triangles);
  ok = ok && hipMalloc(&p->verts, T * 9 * sizeof(float)) == hipSuccess;
  ok = ok && hipMalloc(&p->tris, T * 3 * sizeof(int32_t)) == hipSuccess;
  ok = ok && hipMalloc(&p->counts, 4 * sizeof(int32_t)) == hipSuccess;
  ok = ok && hipMalloc(&p->updated, 4 * sizeof(int32_t)) == hipSuccess;
  if (cfg->storage16) {
    ok = ok && hipMalloc(&p->t16, V * sizeof(uint16_t)) == hipSuccess;
    ok = ok && hipMalloc(&p->w8, V * sizeof(uint8_t)) == hipSuccess;
    ok = ok && hipMalloc(&p->k16, V * 4 * sizeof(uint16_t)) == hipSuccess;
    ok = ok && hipMalloc(&p->kw16, V * 4 * sizeof(uint16_t)) == hipSuccess;
    if (ok) ok = nrf_tsdf16_reset(&cfg->volume, p->t16, p->w8, nullptr) == NRF_OK;
  }
  for (auto &e : p->ev) ok = ok && hipEventCreate(&e) == hipSuccess;
  ok = ok && hipStreamCreateWithFlags(&p->cap, hipStreamNonBlocking) == hipSuccess;
  if (ok) {
    p->solver = nrf_solver_create(&cfg->solver, cfg->max_nodes);
    ok = p->solver != nullptr;
  }
  if (ok) ok = nrf_tsdf_reset(&cfg->volume, p->vox, nullptr) == NRF_OK;
  if (ok) ok = hipMemset(p->knn, 0xff, V * 4 * sizeof(int32_t)) == hipSuccess;
  if (ok) ok = hipMemset(p->knn_w, 0, V * 4 * sizeof(float)) == hipSuccess;
  if (ok) ok = hipMemset(p->knn_rows, 0, nrun * sizeof(uint32_t)) == hipSuccess;  // no id >= 0 yet
  if (ok) ok = hipMemset(p->occ, 0, nocc) == hipSuccess;  // every weight is 0 after the reset
  if (ok) ok = hipMemset(p->counts, 0, 4 * sizeof(int32_t)) == hipSuccess;
  if (ok) ok = hipMemset(p->updated, 0, 4 * sizeof(int32_t)) == hipSuccess;
  if (ok) ok = hipDeviceSynchronize() == hipSuccess;
  if (!ok) {
    pipeline_free(p);
    return nullptr;
  }
  return p;
}

extern "C" void nrf_pipeline_destroy(nrf_pipeline *p) {
  if (p) pipeline_free(p);
}

// The volume and mesh getters wait for an overlapped step's fusion.
extern "C" nrf_voxel *nrf_pipeline_voxels(nrf_pipeline *p) { return p && !pipeline_drain(p) ? p->vox : nullptr; }
extern "C" nrf_dualquat *nrf_pipeline_transforms(nrf_pipeline *p) { return p ? p->T : nullptr; }
extern "C" nrf_node *nrf_pipeline_nodes(nrf_pipeline *p) { return p ? p->nodes : nullptr; }
extern "C" int32_t *nrf_pipeline_knn(nrf_pipeline *p) { return p ? p->knn : nullptr; }
extern "C" float *nrf_pipeline_knn_w(nrf_pipeline *p) { return p ? p->knn_w : nullptr; }
extern "C" float *nrf_pipeline_mesh(nrf_pipeline *p, int32_t **counts) {
  if (!p || pipeline_drain(p)) return nullptr;
  if (counts) *counts = p->counts;
  return p->verts;
}
extern "C" nrf_solver *nrf_pipeline_solver(nrf_pipeline *p) { return p ? p->solver : nullptr; }
extern "C" int nrf_pipeline_volume16(nrf_pipeline *p, uint16_t **tsdf16, uint8_t **weight8, uint16_t **knn16,
                                     uint16_t **knn_w16) {
  if (!p) return NRF_ERR_INVALID_ARGUMENT;
  if (int rc = pipeline_drain(p)) return rc;
  if (tsdf16) *tsdf16 = p->t16;
  if (weight8) *weight8 = p->w8;
  if (knn16) *knn16 = p->k16;
  if (knn_w16) *knn_w16 = p->kw16;
  return NRF_OK;
}

extern "C" int nrf_pipeline_integrate_rigid(nrf_pipeline *p, const float *depth, const float camera_pose[16],
                                            void *stream) {
  if (!p || !depth || !camera_pose) return NRF_ERR_INVALID_ARGUMENT;
  if (int rc0 = pipeline_drain(p)) return rc0;
  p->mesh_valid = false;
  return tsdf_integrate(&p->cfg.volume, p->vox, depth, &p->cfg.camera, camera_pose, nullptr, nullptr, 0, nullptr,
                        nullptr, nullptr, nullptr, nullptr, 0, nullptr, as_stream(stream), nullptr, p->occ, p->dblk);
}

extern "C" int nrf_pipeline_set_nodes(nrf_pipeline *p, const nrf_node *host_nodes, int num_nodes) {
  if (!p || num_nodes < 0 || num_nodes > p->cfg.max_nodes || (num_nodes > 0 && !host_nodes))
    return NRF_ERR_INVALID_ARGUMENT;
  if (int rc0 = pipeline_drain(p)) return rc0;
  p->N = num_nodes;
  if (num_nodes == 0) return NRF_OK;
  NRF_HIP_CHECK(hipMemcpy(p->nodes, host_nodes, sizeof(nrf_node) * num_nodes, hipMemcpyHostToDevice));
  std::vector<nrf_dualquat> id(num_nodes);
  for (auto &q : id) {
    q.real[0] = q.real[1] = q.real[2] = 0.f;
    q.real[3] = 1.f;
    q.dual[0] = q.dual[1] = q.dual[2] = q.dual[3] = 0.f;
  }
  NRF_HIP_CHECK(hipMemcpy(p->T, id.data(), sizeof(nrf_dualquat) * num_nodes, hipMemcpyHostToDevice));
  int rc = nrf_warp_centered_table(p->nodes, p->T, num_nodes, p->centered, nullptr);
  if (rc) return rc;
  rc = nrf_solver_set_pattern(p->solver, host_nodes, num_nodes);
  if (rc) return rc;
  rc = solver_reserve(p->solver, p->cfg.max_corrs);
  if (rc) return rc;
  rc = nrf_voxel_knn(&p->cfg.volume, p->vox, p->nodes, num_nodes, p->knn, p->knn_w, nullptr);
  if (rc) return rc;
  rc = knn_row_flags(&p->cfg.volume, p->knn, p->knn_rows, nullptr);
  if (rc) return rc;
  if (p->cfg.storage16) {  // the FP32 state so far, rounded once into the compact form
    rc = nrf_tsdf16_from_voxels(&p->cfg.volume, p->vox, p->t16, p->w8, nullptr);
    if (rc) return rc;
    rc = nrf_knn_compact(&p->cfg.volume, p->knn, p->knn_w, p->k16, p->kw16, nullptr);
    if (rc) return rc;
  }
  NRF_HIP_CHECK(hipDeviceSynchronize());
  if (p->exec) {
    (void)hipGraphExecDestroy(p->exec);
    p->exec = nullptr;
  }
  return NRF_OK;
}

// the marching cubes of the current volume into the pipeline's mesh
static int pipeline_mc(nrf_pipeline *p, hipStream_t st) {
  int rc;
  if (p->cfg.storage16)
    rc = mc16_extract(&p->cfg.volume, p->t16, p->w8, p->occ, p->mc_ws, p->mc_bytes, p->verts, p->tris,
                      p->cfg.max_triangles, p->counts, st);
  else
    rc = mc_extract(&p->cfg.volume, p->vox, p->occ, p->mc_ws, p->mc_bytes, p->verts, nullptr, p->tris,
                    p->cfg.max_triangles, p->counts, st);
  if (rc == NRF_OK) p->mesh_valid = true;
  return rc;
}

// The warped integration and marching cubes of a step on `st` (stage events
// ev[2] / ev[3] when `events`; per-kernel profiling events pe_int / pe_mc),
// with the node transforms T / centered (null: the live tables).
static int pipeline_fusion(nrf_pipeline *p, const float *depth, const float camera_pose[16], hipStream_t st,
                           bool events, const CamXf *dev_pose, hipEvent_t pe_int, hipEvent_t pe_mc,
                           const nrf_dualquat *T, const nrf_dualquat *centered) {
  if (!T) T = p->T;
  const nrf_dualquat *C = centered ? centered : p->centered;
  int rc;
  if (p->cfg.storage16)
    rc = tsdf16_integrate(&p->cfg.volume, p->t16, p->w8, depth, &p->cfg.camera, camera_pose, C, p->N, p->k16,
                          p->kw16, p->knn_rows, p->updated, st, nullptr, 0, dev_pose, p->occ, p->dblk);
  else
    rc = tsdf_integrate(&p->cfg.volume, p->vox, depth, &p->cfg.camera, camera_pose, p->nodes, T, p->N, C, p->knn,
                        p->knn_w, p->knn_rows, nullptr, p->cfg.min_opt_count, p->updated, st, dev_pose, p->occ,
                        p->dblk);
  if (rc) return rc;
  if (events) NRF_HIP_CHECK(hipEventRecord(p->ev[2], st));
  if (pe_int) NRF_HIP_CHECK(hipEventRecord(pe_int, st));
  if ((rc = pipeline_mc(p, st))) return rc;
  if (events) NRF_HIP_CHECK(hipEventRecord(p->ev[3], st));
  if (pe_mc) NRF_HIP_CHECK(hipEventRecord(pe_mc, st));
  return NRF_OK;
}

// `events`: record the stage events (not inside a capture).  dev_pose: the
// integration reads its camera transform there (a captured step).
static int pipeline_enqueue(nrf_pipeline *p, const float *depth, const float camera_pose[16],
                            const nrf_corr *corrs, const float *canonical, int num_corrs, void *stream,
                            bool profile, bool events = true, const CamXf *dev_pose = nullptr) {
  hipStream_t st = as_stream(stream);
  int rc;
  hipEvent_t *pe = nullptr;
  // overlapped fusion: eager steps with stage events only (not captured, not
  // under the per-kernel event ring, whose integration and marching-cubes
  // events sit on the caller's stream)
  const bool ovl = p->overlap && events && !dev_pose && !(profile && !p->prof_pcg);
  if (!ovl && (rc = pipeline_drain(p))) return rc;
  if (ovl) {
    // a flushed fusion read the snapshot slot this step writes
    if (p->fz_flushed) NRF_HIP_CHECK(hipStreamWaitEvent(st, p->ev_fused, 0));
    p->fz_flushed = false;
    // the deferred fusion's snapshot in this solve's first kernel; at its
    // first PCG launch that fusion and this step's depth copy
    if (p->fz_pending) {
      const size_t nn = (size_t)std::max(1, p->cfg.max_nodes);
      p->fz_snap = (int)(p->fz_steps & 1);
      solver_set_snapshot(p->solver, p->T_snap + (size_t)p->fz_snap * nn, p->c_snap + (size_t)p->fz_snap * nn);
    }
    p->fz_src = depth;
    p->fz_copied = false;
    if (++p->fz_ticket == 0) p->fz_ticket = 1;
    solver_set_pcg_ticket(p->solver, p->fz_ticket);
    solver_set_pcg_hook(p->solver, pipeline_fuse_hook, p, nullptr);
  }
  if (profile) {
    pe = &p->prof[(size_t)(p->prof_steps % p->prof_ring) * p->prof_per];
    if (p->prof_pcg) {
      // the PCG launch of GN iteration (step mod G): sink points 2 + 4 g, 3 + 4 g
      solver_set_event_sink(p->solver, pe, 2, 2 + 4 * (int)(p->prof_steps % p->prof_G));
    } else {
      NRF_HIP_CHECK(hipEventRecord(pe[0], st));
      solver_set_event_sink(p->solver, pe + 1, p->prof_per - 3);
    }
  }
  if (events && !ovl) NRF_HIP_CHECK(hipEventRecord(p->ev[0], st));  // (overlapped: no solve stage event)
  rc = nrf_solver_solve_gn(p->solver, corrs, canonical, num_corrs, p->nodes, p->T, p->centered, p->N, &p->cfg.gn,
                           p->delta_x, nullptr, stream);
  if (profile) solver_set_event_sink(p->solver, nullptr, 0);
  // all consumed by the solve (or dropped with an error)
  solver_set_pcg_wait(p->solver, nullptr);
  solver_set_pcg_hook(p->solver, nullptr, nullptr, nullptr);
  solver_set_pcg_ticket(p->solver, 0);
  solver_set_snapshot(p->solver, nullptr, nullptr);
  if (rc) return rc;
  if (events && !ovl) NRF_HIP_CHECK(hipEventRecord(p->ev[1], st));
  if (ovl) {
    // a solve without a PCG launch (which updates no transform, so its
    // snapshot holds what the deferred fusion needs) leaves the hook's part
    // to be done here
    if (!p->fz_copied) {
      NRF_HIP_CHECK(hipEventRecord(p->ev[1], st));
      if ((rc = pipeline_side_step(p, p->ev[1]))) return rc;
    }
    p->fz_stream = st;  // a flush orders the fusion after this step's work there
    // this step's fusion is deferred; the caller's stream waits for the copy
    // of its depth
    NRF_HIP_CHECK(hipStreamWaitEvent(st, p->ev_copy, 0));
    std::memcpy(p->fz_pose, camera_pose, sizeof(p->fz_pose));
    p->fz_slot = (int)(p->fz_steps & 1);
    ++p->fz_steps;
    p->fz_pending = true;
    if (profile) ++p->prof_steps;
    return NRF_OK;
  }
  p->fz_events = false;
  const bool ring = profile && !p->prof_pcg;
  if ((rc = pipeline_fusion(p, depth, camera_pose, st, events, dev_pose, ring ? pe[p->prof_per - 2] : nullptr,
                            ring ? pe[p->prof_per - 1] : nullptr)))
    return rc;
  if (profile) ++p->prof_steps;
  return NRF_OK;
}

extern "C" int nrf_pipeline_step(nrf_pipeline *p, const float *depth, const float camera_pose[16],
                                 const nrf_corr *corrs, const float *canonical, int num_corrs, void *stream) {
  if (!p || !depth || !camera_pose || num_corrs < 0 || num_corrs > p->cfg.max_corrs || (num_corrs > 0 && !corrs))
    return NRF_ERR_INVALID_ARGUMENT;
  if (p->N <= 0) return NRF_ERR_NOT_READY;
  if (!p->cfg.use_graph || p->prof_ring > 0) {
    p->graph_step = false;
    const bool sampled = p->prof_ring > 0 && p->prof_calls++ % p->prof_every == 0;
    return pipeline_enqueue(p, depth, camera_pose, corrs, canonical, num_corrs, stream, sampled);
  }
  if (int rc0 = pipeline_drain(p)) return rc0;
  // the frame into the captured step's input buffers (stream-ordered, ahead
  // of the replay on the same stream)
  hipStream_t cs = as_stream(stream);
  const nrf_camera &cam = p->cfg.camera;
  const size_t npx = (size_t)cam.width * cam.height;
  if (!p->g_depth) {
    NRF_HIP_CHECK(hipMalloc(&p->g_depth, std::max<size_t>(npx, 1) * sizeof(float)));
    NRF_HIP_CHECK(hipMalloc(&p->g_corrs, (size_t)std::max(1, p->cfg.max_corrs) * sizeof(nrf_corr)));
    NRF_HIP_CHECK(hipMalloc(&p->g_canon, (size_t)std::max(1, p->cfg.max_corrs) * 3 * sizeof(float)));
    NRF_HIP_CHECK(hipMalloc(&p->g_pose, sizeof(CamXf)));
  }
  if (depth != p->g_depth)
    NRF_HIP_CHECK(hipMemcpyAsync(p->g_depth, depth, npx * sizeof(float), hipMemcpyDeviceToDevice, cs));
  if (num_corrs > 0 && corrs != p->g_corrs)
    NRF_HIP_CHECK(hipMemcpyAsync(p->g_corrs, corrs, (size_t)num_corrs * sizeof(nrf_corr), hipMemcpyDeviceToDevice, cs));
  if (num_corrs > 0 && canonical && canonical != p->g_canon)
    NRF_HIP_CHECK(
        hipMemcpyAsync(p->g_canon, canonical, (size_t)num_corrs * 3 * sizeof(float), hipMemcpyDeviceToDevice, cs));
  int rc0 = set_device_pose(p->g_pose, camera_pose, cs);
  if (rc0) return rc0;
  const bool has_can = canonical != nullptr;
  const bool same = p->exec && p->g_n == num_corrs && p->g_has_can == has_can &&
                    p->g_gen == solver_state_generation(p->solver);
  if (!same) {
    if (p->exec) {
      (void)hipGraphExecDestroy(p->exec);
      p->exec = nullptr;
    }
    if (p->graph) {
      (void)hipGraphDestroy(p->graph);
      p->graph = nullptr;
    }
    hipStream_t st = p->cap;
    NRF_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = pipeline_enqueue(p, p->g_depth, camera_pose, p->g_corrs, has_can ? p->g_canon : nullptr, num_corrs, st,
                              false, false, p->g_pose);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(st, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) {
      set_last_error("hipStreamEndCapture", e);
      return NRF_ERR_HIP;
    }
    p->graph = g;
    NRF_HIP_CHECK(hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0));
    p->g_n = num_corrs;
    p->g_has_can = has_can;
    p->g_gen = solver_state_generation(p->solver);
    ++p->g_captures;
  }
  hipStream_t st = as_stream(stream);
  NRF_HIP_CHECK(hipEventRecord(p->ev[0], st));
  NRF_HIP_CHECK(hipGraphLaunch(p->exec, st));
  NRF_HIP_CHECK(hipEventRecord(p->ev[3], st));
  p->graph_step = true;
  return NRF_OK;
}

extern "C" long long nrf_pipeline_graph_captures(nrf_pipeline *p) { return p ? p->g_captures : -1; }

extern "C" int nrf_pipeline_set_overlap(nrf_pipeline *p, int enable) {
  if (!p || (enable != 0 && enable != 1)) return NRF_ERR_INVALID_ARGUMENT;
  if (int rc = pipeline_drain(p)) return rc;
  if (enable && !p->side) {
    NRF_HIP_CHECK(hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking));
    for (hipEvent_t *e : {&p->ev_fused, &p->ev_copy})
      NRF_HIP_CHECK(hipEventCreateWithFlags(e, hipEventDisableTiming));
    NRF_HIP_CHECK(hipEventCreate(&p->ev_fs));
    const size_t npx = (size_t)p->cfg.camera.width * p->cfg.camera.height;
    const size_t nn = (size_t)std::max(1, p->cfg.max_nodes);
    NRF_HIP_CHECK(hipMalloc(&p->fz_depth, 2 * std::max<size_t>(npx, 1) * sizeof(float)));
    NRF_HIP_CHECK(hipMalloc(&p->T_snap, 2 * nn * sizeof(nrf_dualquat)));
    NRF_HIP_CHECK(hipMalloc(&p->c_snap, 2 * nn * sizeof(nrf_dualquat)));
  }
  p->overlap = enable != 0;
  return NRF_OK;
}

extern "C" int nrf_pipeline_join(nrf_pipeline *p, void *stream) {
  if (!p) return NRF_ERR_INVALID_ARGUMENT;
  if (int rc = pipeline_flush(p)) return rc;
  if (p->side_pending) NRF_HIP_CHECK(hipStreamWaitEvent(as_stream(stream), p->ev_fused, 0));
  return NRF_OK;
}

extern "C" int nrf_pipeline_stage_ms(nrf_pipeline *p, float out_ms[3]) {
  if (!p || !out_ms) return NRF_ERR_INVALID_ARGUMENT;
  if (p->graph_step) {  // a replayed graph: the whole step, no stage split
    NRF_HIP_CHECK(hipEventElapsedTime(&out_ms[0], p->ev[0], p->ev[3]));
    out_ms[1] = out_ms[2] = -1.f;
    return NRF_OK;
  }
  if (int rc = pipeline_drain(p)) return rc;
  // an overlapped step: no solve stage, its fusion from its start on the
  // side stream
  if (p->fz_events)
    out_ms[0] = -1.f;
  else
    NRF_HIP_CHECK(hipEventElapsedTime(&out_ms[0], p->ev[0], p->ev[1]));
  NRF_HIP_CHECK(hipEventElapsedTime(&out_ms[1], p->fz_events ? p->ev_fs : p->ev[1], p->ev[2]));
  NRF_HIP_CHECK(hipEventElapsedTime(&out_ms[2], p->ev[2], p->ev[3]));
  return NRF_OK;
}

static int pipeline_profile_setup(nrf_pipeline *p, int ring_steps, int pcg_only) {
  if (!p || ring_steps < 0) return NRF_ERR_INVALID_ARGUMENT;
  for (auto &e : p->prof)
    if (e) (void)hipEventDestroy(e);
  p->prof.clear();
  p->prof_ring = 0;
  p->prof_steps = 0;
  p->prof_calls = 0;
  p->prof_every = 1;
  p->prof_pcg = 0;
  if (ring_steps == 0) return NRF_OK;
  const int G = std::max(1, p->cfg.gn.gn_iterations);
  p->prof_G = G;
  p->prof_pcg = pcg_only;
  p->prof_per = pcg_only ? 2 : 1 + (1 + 4 * G) + 2;
  p->prof.assign((size_t)ring_steps * p->prof_per, nullptr);
  for (auto &e : p->prof) NRF_HIP_CHECK(hipEventCreate(&e));
  p->prof_ring = ring_steps;
  return NRF_OK;
}
extern "C" int nrf_pipeline_profile(nrf_pipeline *p, int ring_steps) {
  return pipeline_profile_setup(p, ring_steps, 0);
}
extern "C" int nrf_pipeline_profile_pcg(nrf_pipeline *p, int ring_steps) {
  return pipeline_profile_setup(p, ring_steps, 1);
}

extern "C" int nrf_pipeline_profile_every(nrf_pipeline *p, int k) {
  if (!p || k < 1) return NRF_ERR_INVALID_ARGUMENT;
  p->prof_every = k;
  p->prof_calls = 0;
  return NRF_OK;
}

extern "C" int nrf_pipeline_profile_read(nrf_pipeline *p, int step_back, float out_ms[6]) {
  if (!p || !out_ms || p->prof_ring == 0 || step_back < 0 || step_back >= p->prof_ring ||
      step_back >= p->prof_steps)
    return NRF_ERR_INVALID_ARGUMENT;
  const long step = p->prof_steps - 1 - step_back;
  hipEvent_t *e = &p->prof[(size_t)(step % p->prof_ring) * p->prof_per];
  auto ms = [&](int a, int b, float &acc) -> int {
    float v = 0;
    NRF_HIP_CHECK(hipEventElapsedTime(&v, e[a], e[b]));
    acc += v;
    return NRF_OK;
  };
  for (int k = 0; k < 6; ++k) out_ms[k] = 0;
  int rc;
  if (p->prof_pcg) return ms(0, 1, out_ms[2]);  // one PCG launch
  if ((rc = ms(0, 1, out_ms[0]))) return rc;  // membership build
  int prev = 1;
  for (int g = 0; g < p->prof_G; ++g) {
    const int a = 2 + 4 * g;  // after assembly, before PCG, after PCG, after update
    if ((rc = ms(prev, a, out_ms[1]))) return rc;
    if ((rc = ms(a + 1, a + 2, out_ms[2]))) return rc;
    if ((rc = ms(a + 2, a + 3, out_ms[3]))) return rc;
    prev = a + 3;
  }
  if ((rc = ms(prev, p->prof_per - 2, out_ms[4]))) return rc;
  if ((rc = ms(p->prof_per - 2, p->prof_per - 1, out_ms[5]))) return rc;
  return NRF_OK;
}

// ----------------------------------------------------------- tracking frame

static int track_alloc(nrf_pipeline *p) {
  const nrf_camera &cam = p->cfg.camera;
  const size_t npx = (size_t)cam.width * cam.height;
  const size_t V = (size_t)volume_count(&p->cfg.volume);
  const size_t T = (size_t)std::max(1, p->cfg.max_triangles);
  auto &k = p->tk;
  NRF_HIP_CHECK(hipMalloc(&k.dnorm, npx * 3 * sizeof(float)));
  NRF_HIP_CHECK(hipMalloc(&k.warped, T * 9 * sizeof(float)));
  NRF_HIP_CHECK(hipMalloc(&k.live_v, npx * 3 * sizeof(float)));
  NRF_HIP_CHECK(hipMalloc(&k.live_n, npx * 3 * sizeof(float)));
  NRF_HIP_CHECK(hipMalloc(&k.ztri, npx * sizeof(unsigned long long)));
  NRF_HIP_CHECK(hipMalloc(&k.hit, npx * sizeof(int32_t)));
  NRF_HIP_CHECK(hipMalloc(&k.corrs, npx * sizeof(nrf_corr)));
  NRF_HIP_CHECK(hipMalloc(&k.part, (size_t)track_rmse_blocks((int)npx) * 2 * sizeof(double)));
  NRF_HIP_CHECK(hipMalloc(&k.Tsave, (size_t)p->cfg.max_nodes * sizeof(nrf_dualquat)));
  NRF_HIP_CHECK(hipMalloc(&k.dev, sizeof(TrackDev)));
  NRF_HIP_CHECK(hipMalloc(&k.opt, V * sizeof(int32_t)));
  NRF_HIP_CHECK(hipMalloc(&k.opt_frame, V * sizeof(int32_t)));
  // d_voxel_opt_counts_ / d_voxel_opt_frame_ start at zero (fusion_app.h:143-146)
  NRF_HIP_CHECK(hipMemset(k.opt, 0, V * sizeof(int32_t)));
  NRF_HIP_CHECK(hipMemset(k.opt_frame, 0, V * sizeof(int32_t)));
  NRF_HIP_CHECK(hipMemset(k.dev, 0, sizeof(TrackDev)));
  NRF_HIP_CHECK(hipMemset(k.corrs, 0, npx * sizeof(nrf_corr)));
  NRF_HIP_CHECK(hipMemset(k.hit, 0xff, npx * sizeof(int32_t)));
  int rc = solver_reserve(p->solver, std::max(p->cfg.max_corrs, (int)npx));
  if (rc) return rc;
  NRF_HIP_CHECK(hipDeviceSynchronize());
  return NRF_OK;
}

// live surface prediction + association at the current transforms, each
// kernel gated by the stop flag
static int track_associate(nrf_pipeline *p, const float *depth, const float camera_pose[16],
                           const nrf_assoc_params *ap, const int32_t *stop, hipStream_t st) {
  auto &k = p->tk;
  const nrf_camera *cam = &p->cfg.camera;
  const int maxT = std::max(1, p->cfg.max_triangles);
  const int32_t *ntri = p->counts + 1;
  int rc;
  if ((rc = skin_vertices(p->verts, 3 * maxT, &p->cfg.volume, p->centered, p->N, p->knn, p->knn_w, k.warped, st,
                          ntri, stop)))
    return rc;
  if ((rc = rasterize_mesh(p->verts, k.warped, p->tris, 3 * maxT, maxT, cam, camera_pose, &p->cfg.volume, k.ztri,
                           k.live_v, k.live_n, k.hit, nullptr, st, ntri, stop)))
    return rc;
  if ((rc = track_count_reset(k.dev, stop, st))) return rc;
  return find_correspondences(k.live_v, k.live_n, depth, k.dnorm, cam, camera_pose, nullptr, nullptr, k.hit, p->knn,
                              p->knn_w, ap, k.corrs, &k.dev->counter, nullptr, stop, st);
}

extern "C" int nrf_pipeline_track(nrf_pipeline *p, const float *depth, const float camera_pose[16],
                                  const nrf_track_params *tp, int frame_id, void *stream) {
  if (!p || !depth || !camera_pose || !tp) return NRF_ERR_INVALID_ARGUMENT;
  if (p->N <= 0) return NRF_ERR_NOT_READY;
  const nrf_gn_params &gn = p->cfg.gn;
  const int G = std::max(1, gn.gn_iterations);
  if (G > NRF_TRACK_MAX_GN || p->cfg.camera.width <= 0 || p->cfg.camera.height <= 0)
    return NRF_ERR_INVALID_ARGUMENT;
  int rc;
  if ((rc = pipeline_drain(p))) return rc;  // the mesh of an overlapped step feeds the prediction
  if (!p->tk.dev && (rc = track_alloc(p))) return rc;
  hipStream_t st = as_stream(stream);
  auto &k = p->tk;
  const nrf_camera *cam = &p->cfg.camera;
  const int npx = cam->width * cam->height;
  const bool ls = tp->line_search != 0;
  p->graph_step = false;
  if (!p->mesh_valid && (rc = pipeline_mc(p, st))) return rc;
  if ((rc = nrf_depth_normals(depth, cam, k.dnorm, st))) return rc;
  if ((rc = solver_gn_begin(p->solver, st))) return rc;
  int32_t *stop = solver_stop_flag(p->solver);
  if ((rc = track_begin(k.dev, st))) return rc;
  NRF_HIP_CHECK(hipEventRecord(p->ev[0], st));
  nrf_gn_params g1 = gn;
  g1.gn_iterations = 1;
  // fusion_app.h:500-511, the products in FP32 as there
  const float us = gn.update_scale;
  const float trials[10] = {us, 0.5f * us, 0.25f * us, 0.125f * us, 0.0625f * us,
                            -0.0625f * us, -0.125f * us, -0.25f * us, -0.5f * us, -us};
  for (int g = 0; g < G; ++g) {
    if ((rc = track_associate(p, depth, camera_pose, &tp->assoc, stop, st))) return rc;
    if ((rc = track_gate(k.dev, g, tp->min_valid_corrs, stop, st))) return rc;
    if (!ls) {
      if ((rc = solver_solve_gn(p->solver, k.corrs, nullptr, npx, p->nodes, p->T, p->centered, p->N, &g1,
                                p->delta_x, nullptr, st, false)))
        return rc;
      continue;
    }
    if ((rc = track_rmse(k.dev, k.corrs, npx, k.part, g, false, 0.f, 0, stop, st))) return rc;
    if ((rc = solver_gn_linearise(p->solver, k.corrs, npx, p->nodes, p->T, p->centered, p->N, &gn, p->delta_x, st)))
      return rc;
    if ((rc = track_ls_begin(k.dev, g, us, p->T, k.Tsave, p->N, stop, st))) return rc;
    for (float s : trials) {
      if ((rc = track_ls_apply(k.dev, p->delta_x, k.Tsave, p->T, p->centered, p->nodes, p->N, s, false,
                               gn.max_update_rot, gn.max_update_trans, stop, st)))
        return rc;
      if ((rc = track_associate(p, depth, camera_pose, &tp->assoc, stop, st))) return rc;
      if ((rc = track_rmse(k.dev, k.corrs, npx, k.part, g, true, s, tp->min_valid_corrs, stop, st))) return rc;
    }
    if ((rc = track_ls_apply(k.dev, p->delta_x, k.Tsave, p->T, p->centered, p->nodes, p->N, 0.f, true,
                             gn.max_update_rot, gn.max_update_trans, stop, st)))
      return rc;
    if ((rc = track_ls_commit(k.dev, g, stop, st))) return rc;
    // the association after the applied step (:636-639); the next
    // iteration's first association would repeat it on the same transforms
    if (g == G - 1 && (rc = track_associate(p, depth, camera_pose, &tp->assoc, stop, st))) return rc;
  }
  if ((rc = track_end(k.dev, solver_result_device(p->solver), ls, stop, st))) return rc;
  NRF_HIP_CHECK(hipEventRecord(p->ev[1], st));
  if ((rc = nrf_mark_optimized_voxels(k.corrs, k.hit, npx, &p->cfg.volume, frame_id, k.opt, k.opt_frame, st)))
    return rc;
  if (p->cfg.storage16)
    rc = tsdf16_integrate(&p->cfg.volume, p->t16, p->w8, depth, cam, camera_pose, p->centered, p->N, p->k16, p->kw16,
                          p->knn_rows, p->updated, st, k.opt, p->cfg.min_opt_count, nullptr, p->occ, p->dblk);
  else
    rc = tsdf_integrate(&p->cfg.volume, p->vox, depth, cam, camera_pose, p->nodes, p->T, p->N, p->centered, p->knn,
                        p->knn_w, p->knn_rows, k.opt, p->cfg.min_opt_count, p->updated, st, nullptr, p->occ, p->dblk);
  if (rc) return rc;
  NRF_HIP_CHECK(hipEventRecord(p->ev[2], st));
  if ((rc = pipeline_mc(p, st))) return rc;
  NRF_HIP_CHECK(hipEventRecord(p->ev[3], st));
  return NRF_OK;
}

extern "C" int nrf_pipeline_track_stats(nrf_pipeline *p, nrf_track_stats *out) {
  if (!p || !out) return NRF_ERR_INVALID_ARGUMENT;
  if (!p->tk.dev) return NRF_ERR_NOT_READY;
  NRF_HIP_CHECK(hipDeviceSynchronize());
  NRF_HIP_CHECK(hipMemcpy(out, &p->tk.dev->stats, sizeof(nrf_track_stats), hipMemcpyDeviceToHost));
  return NRF_OK;
}

extern "C" int nrf_pipeline_track_maps(nrf_pipeline *p, float **live_vertices, float **live_normals,
                                       int32_t **hit_voxel, nrf_corr **corrs, int32_t **voxel_opt_counts,
                                       float **depth_normals) {
  if (!p) return NRF_ERR_INVALID_ARGUMENT;
  if (live_vertices) *live_vertices = p->tk.live_v;
  if (live_normals) *live_normals = p->tk.live_n;
  if (hit_voxel) *hit_voxel = p->tk.hit;
  if (corrs) *corrs = p->tk.corrs;
  if (voxel_opt_counts) *voxel_opt_counts = p->tk.opt;
  if (depth_normals) *depth_normals = p->tk.dnorm;
  return NRF_OK;
}

extern "C" long long nrf_pipeline_updated_voxels(nrf_pipeline *p) {
  if (!p) return -1;
  if (pipeline_drain(p)) return -1;
  int32_t v = 0;
  if (hipMemcpy(&v, p->updated, sizeof(v), hipMemcpyDeviceToHost) != hipSuccess) return -1;
  return v;
}

// ------------------------------------------------------------------ batch

static_assert(NRF_MAX_BATCH == kMaxBatch, "batch size limits must agree");

struct nrf_batch {
  std::vector<nrf_pipeline *> p;
  nrf_batch_tuning tune{};
  SolverBatch *sb = nullptr;
  IntegrateBatch *ib = nullptr;
  int use_graph = 0;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  std::vector<nrf_batch_frame> g_frames;
  void *g_stream = nullptr;
  std::vector<unsigned> g_gen;
  std::vector<unsigned> sb_gen;  // member solver generations the device records were built at
  hipStream_t cap = nullptr;     // capture stream
  bool graph_step = false;
};

static std::vector<unsigned> batch_generations(const nrf_batch *b) {
  std::vector<unsigned> g;
  for (const nrf_pipeline *p : b->p) g.push_back(solver_state_generation(p->solver));
  return g;
}

static void batch_drop_graph(nrf_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

static void batch_free(nrf_batch *b) {
  batch_drop_graph(b);
  solver_batch_destroy(b->sb);
  integrate_batch_destroy(b->ib);
  for (auto &e : b->ev)
    if (e) (void)hipEventDestroy(e);
  if (b->cap) (void)hipStreamDestroy(b->cap);
  delete b;
}

static int batch_build_integrate(const std::vector<nrf_pipeline *> &p, int z_chunk, nrf_voxel *const *vox,
                                 int32_t *updated, IntegrateBatch **out) {
  std::vector<IntegrateBatchInstance> ii(p.size());
  for (size_t k = 0; k < p.size(); ++k) {
    IntegrateBatchInstance &q = ii[k];
    q.vp = p[k]->cfg.volume;
    q.cam = p[k]->cfg.camera;
    q.vox = vox ? vox[k] : p[k]->vox;
    q.nodes = p[k]->nodes;
    q.T = p[k]->T;
    q.centered = p[k]->centered;
    q.N = p[k]->N;
    q.knn = p[k]->knn;
    q.knn_w = p[k]->knn_w;
    q.knn_rows = p[k]->knn_rows;
    q.updated = updated ? updated : p[k]->updated;
    q.occ = vox ? nullptr : p[k]->occ;  // the tuner's scratch copies mark nothing
  }
  return integrate_batch_create(ii.data(), (int)p.size(), z_chunk, out);
}

static int batch_build_solver(const std::vector<nrf_pipeline *> &p, int lanes, int wg_nodes, SolverBatch **out) {
  std::vector<SolverBatchInstance> si(p.size());
  for (size_t k = 0; k < p.size(); ++k) {
    SolverBatchInstance &q = si[k];
    q.solver = p[k]->solver;
    q.nodes = p[k]->nodes;
    q.T = p[k]->T;
    q.centered = p[k]->centered;
    q.N = p[k]->N;
    q.gn = p[k]->cfg.gn;
    q.delta_x = p[k]->delta_x;
    q.max_corrs = p[k]->cfg.max_corrs;
  }
  return solver_batch_create(si.data(), (int)p.size(), lanes, wg_nodes, out);
}

static int batch_apply_tuning(nrf_batch *b, const nrf_batch_tuning &t) {
  // the rebuild rewrites the members' partition tables: no step may be in flight
  NRF_HIP_CHECK(hipDeviceSynchronize());
  batch_drop_graph(b);
  solver_batch_destroy(b->sb);
  b->sb = nullptr;
  integrate_batch_destroy(b->ib);
  b->ib = nullptr;
  int rc = batch_build_solver(b->p, t.lanes_per_row, t.wg_nodes, &b->sb);
  if (!rc) rc = batch_build_integrate(b->p, t.z_chunk, nullptr, nullptr, &b->ib);
  if (rc) return rc;
  b->tune.lanes_per_row = t.lanes_per_row;
  b->tune.wg_nodes = t.wg_nodes;
  b->tune.z_chunk = t.z_chunk;
  b->tune.launch_groups = solver_batch_launch_groups(b->sb);
  b->sb_gen = batch_generations(b);
  return NRF_OK;
}

extern "C" nrf_batch *nrf_batch_create(nrf_pipeline *const *pipelines, int count, const nrf_batch_tuning *tuning,
                                       int use_graph) {
  if (!pipelines || count < 1 || count > kMaxBatch) return nullptr;
  for (int k = 0; k < count; ++k) {
    nrf_pipeline *q = pipelines[k];
    if (!q || q->N <= 0 || q->cfg.storage16 || q->overlap) return nullptr;  // batch members step unoverlapped
    for (int j = 0; j < k; ++j)
      if (pipelines[j] == q) return nullptr;
  }
  nrf_batch *b = new nrf_batch();
  b->p.assign(pipelines, pipelines + count);
  b->use_graph = use_graph ? 1 : 0;
  nrf_batch_tuning t{};
  t.lanes_per_row = 4;
  t.wg_nodes = 32;
  t.z_chunk = 4;
  if (tuning) t = *tuning;
  bool ok = true;
  for (auto &e : b->ev) ok = ok && hipEventCreate(&e) == hipSuccess;
  ok = ok && hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking) == hipSuccess;
  if (!ok || batch_apply_tuning(b, t) != NRF_OK) {
    batch_free(b);
    return nullptr;
  }
  return b;
}

extern "C" void nrf_batch_destroy(nrf_batch *b) {
  if (b) batch_free(b);
}

extern "C" int nrf_batch_set_tuning(nrf_batch *b, const nrf_batch_tuning *tuning) {
  if (!b || !tuning) return NRF_ERR_INVALID_ARGUMENT;
  const nrf_batch_tuning old = b->tune;
  int rc = batch_apply_tuning(b, *tuning);
  if (rc && batch_apply_tuning(b, old) != NRF_OK) return NRF_ERR_HIP;
  return rc;
}

extern "C" int nrf_batch_get_tuning(const nrf_batch *b, nrf_batch_tuning *tuning) {
  if (!b || !tuning) return NRF_ERR_INVALID_ARGUMENT;
  *tuning = b->tune;
  return NRF_OK;
}

static int batch_frames(const nrf_batch *b, const nrf_batch_frame *frames, BatchCorrs &fc, BatchDepth &fd) {
  memset(&fc, 0, sizeof(fc));
  memset(&fd, 0, sizeof(fd));
  for (size_t k = 0; k < b->p.size(); ++k) {
    const nrf_batch_frame &f = frames[k];
    const nrf_pipeline *q = b->p[k];
    if (!f.depth || f.num_corrs < 0 || f.num_corrs > q->cfg.max_corrs || (f.num_corrs > 0 && !f.corrs))
      return NRF_ERR_INVALID_ARGUMENT;
    fc.corrs[k] = f.corrs;
    fc.canonical[k] = f.canonical;
    fc.n[k] = f.num_corrs;
    fc.max_n = std::max(fc.max_n, f.num_corrs);
    fc.cbase[k + 1] = fc.cbase[k] + q->N + (4 * f.num_corrs + kAsmChunk - 1) / kAsmChunk;
    fd.depth[k] = f.depth;
    fd.xf[k] = invert_pose(f.camera_pose);
  }
  return NRF_OK;
}

static int batch_enqueue(nrf_batch *b, const BatchCorrs &fc, const BatchDepth &fd, hipStream_t st, bool events) {
  int rc;
  if (events) NRF_HIP_CHECK(hipEventRecord(b->ev[0], st));
  if ((rc = solver_batch_solve_gn(b->sb, fc, st))) return rc;
  if (events) NRF_HIP_CHECK(hipEventRecord(b->ev[1], st));
  if ((rc = integrate_batch_run(b->ib, fd, st))) return rc;
  if (events) NRF_HIP_CHECK(hipEventRecord(b->ev[2], st));
  for (nrf_pipeline *p : b->p) {
    rc = mc_extract(&p->cfg.volume, p->vox, p->occ, p->mc_ws, p->mc_bytes, p->verts, nullptr, p->tris,
                    p->cfg.max_triangles, p->counts, st);
    if (rc) return rc;
  }
  if (events) NRF_HIP_CHECK(hipEventRecord(b->ev[3], st));
  return NRF_OK;
}

extern "C" int nrf_batch_step(nrf_batch *b, const nrf_batch_frame *frames, void *stream) {
  if (!b || !frames) return NRF_ERR_INVALID_ARGUMENT;
  BatchCorrs fc;
  BatchDepth fd;
  int rc = batch_frames(b, frames, fc, fd);
  if (rc) return rc;
  hipStream_t st = as_stream(stream);
  // a member's partition or solver buffers changed since the records were built
  if (b->sb_gen != batch_generations(b) && (rc = batch_apply_tuning(b, b->tune))) return rc;
  if (!b->use_graph) {
    b->graph_step = false;
    return batch_enqueue(b, fc, fd, st, true);
  }
  const size_t n = b->p.size();
  bool same = b->exec && b->g_stream == stream && b->g_frames.size() == n &&
              memcmp(b->g_frames.data(), frames, n * sizeof(nrf_batch_frame)) == 0 &&
              b->g_gen == batch_generations(b);
  if (!same) {
    batch_drop_graph(b);
    NRF_HIP_CHECK(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
    rc = batch_enqueue(b, fc, fd, b->cap, false);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->cap, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) {
      set_last_error("hipStreamEndCapture", e);
      return NRF_ERR_HIP;
    }
    b->graph = g;
    NRF_HIP_CHECK(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    b->g_frames.assign(frames, frames + n);
    b->g_stream = stream;
    b->g_gen = batch_generations(b);
  }
  NRF_HIP_CHECK(hipEventRecord(b->ev[0], st));
  NRF_HIP_CHECK(hipGraphLaunch(b->exec, st));
  NRF_HIP_CHECK(hipEventRecord(b->ev[3], st));
  b->graph_step = true;
  return NRF_OK;
}

extern "C" int nrf_batch_stage_ms(nrf_batch *b, float out_ms[3]) {
  if (!b || !out_ms) return NRF_ERR_INVALID_ARGUMENT;
  if (b->graph_step) {
    NRF_HIP_CHECK(hipEventElapsedTime(&out_ms[0], b->ev[0], b->ev[3]));
    out_ms[1] = out_ms[2] = -1.f;
    return NRF_OK;
  }
  for (int k = 0; k < 3; ++k) NRF_HIP_CHECK(hipEventElapsedTime(&out_ms[k], b->ev[k], b->ev[k + 1]));
  return NRF_OK;
}

// ---------------------------------------------------------------- tuner

static std::mutex g_tune_mu;
static std::map<std::string, nrf_batch_tuning> g_tune_cache;

static std::string batch_shape_key(const nrf_batch *b) {
  std::string key;
  char buf[160];
  for (const nrf_pipeline *p : b->p) {
    const nrf_volume_params &v = p->cfg.volume;
    snprintf(buf, sizeof(buf), "%d/%d/%d.%d.%d/%dx%d/%d;", p->N, solver_rmax(p->solver), v.dims[0], v.dims[1],
             v.dims[2], p->cfg.camera.width, p->cfg.camera.height, p->cfg.gn.gn_iterations);
    key += buf;
  }
  return key;
}

// Mean device time of `reps` runs of fn after one untimed run.
template <typename F>
static int time_ms(hipStream_t st, hipEvent_t e0, hipEvent_t e1, int reps, F fn, float &ms) {
  int rc;
  if ((rc = fn())) return rc;
  NRF_HIP_CHECK(hipEventRecord(e0, st));
  for (int r = 0; r < reps; ++r)
    if ((rc = fn())) return rc;
  NRF_HIP_CHECK(hipEventRecord(e1, st));
  NRF_HIP_CHECK(hipEventSynchronize(e1));
  NRF_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  ms /= (float)reps;
  return NRF_OK;
}

static int batch_tune(nrf_batch *b, const BatchCorrs &fc, const BatchDepth &fd, hipStream_t st,
                      nrf_batch_tuning &best) {
  const int kReps = 3;
  int rc = NRF_OK;
  batch_drop_graph(b);
  // systems of this frame at the current transforms (results and totals saved)
  if ((rc = solver_batch_prepare(b->sb, fc, st))) return rc;
  NRF_HIP_CHECK(hipStreamSynchronize(st));
  SolverBatch *prep = b->sb;
  b->sb = nullptr;
  struct Cand {
    int lanes, nodes;
  };
  const Cand cands[] = {{4, 24}, {4, 32}, {2, 48}, {2, 64}, {2, 80}, {1, 96}, {1, 128}, {1, 170}};
  best.pcg_ms = -1.f;
  for (const Cand &c : cands) {
    SolverBatch *sb = nullptr;
    if (batch_build_solver(b->p, c.lanes, c.nodes, &sb) != NRF_OK) {
      (void)hipGetLastError();
      continue;  // does not fit this layout
    }
    float ms = 0.f;
    rc = time_ms(st, b->ev[0], b->ev[1], kReps, [&]() { return solver_batch_pcg(sb, st); }, ms);
    solver_batch_destroy(sb);
    if (rc) break;
    if (best.pcg_ms < 0.f || ms < best.pcg_ms) {
      best.pcg_ms = ms;
      best.lanes_per_row = c.lanes;
      best.wg_nodes = c.nodes;
    }
  }
  int rc2 = solver_batch_restore(prep, st);
  if (!rc) rc = rc2;
  rc2 = hipStreamSynchronize(st) == hipSuccess ? NRF_OK : NRF_ERR_HIP;
  if (!rc) rc = rc2;
  solver_batch_destroy(prep);
  if (!rc && best.pcg_ms < 0.f) rc = NRF_ERR_TOO_LARGE;
  if (rc) return rc;
  // integration on scratch copies of the volumes
  const size_t n = b->p.size();
  std::vector<nrf_voxel *> copies(n, nullptr);
  int32_t *upd = nullptr;
  bool ok = hipMalloc(&upd, 4 * sizeof(int32_t)) == hipSuccess;
  for (size_t k = 0; k < n && ok; ++k) {
    const size_t bytes = (size_t)volume_count(&b->p[k]->cfg.volume) * sizeof(nrf_voxel);
    ok = hipMalloc(&copies[k], bytes) == hipSuccess &&
         hipMemcpyAsync(copies[k], b->p[k]->vox, bytes, hipMemcpyDeviceToDevice, st) == hipSuccess;
  }
  if (!ok) {
    set_last_error("autotune scratch volumes", hipGetLastError());
    rc = NRF_ERR_HIP;
  }
  best.integrate_ms = -1.f;
  const int zcs[] = {1, 2, 4, 8, 16};
  for (int zc : zcs) {
    if (rc) break;
    IntegrateBatch *ib = nullptr;
    if ((rc = batch_build_integrate(b->p, zc, copies.data(), upd, &ib))) break;
    float ms = 0.f;
    rc = time_ms(st, b->ev[0], b->ev[1], kReps, [&]() { return integrate_batch_run(ib, fd, st); }, ms);
    integrate_batch_destroy(ib);
    if (!rc && (best.integrate_ms < 0.f || ms < best.integrate_ms)) {
      best.integrate_ms = ms;
      best.z_chunk = zc;
    }
  }
  (void)hipStreamSynchronize(st);
  for (nrf_voxel *c : copies)
    if (c) (void)hipFree(c);
  if (upd) (void)hipFree(upd);
  return rc;
}

extern "C" int nrf_batch_autotune(nrf_batch *b, const nrf_batch_frame *frames, int use_cache,
                                  nrf_batch_tuning *out, void *stream) {
  if (!b || !frames) return NRF_ERR_INVALID_ARGUMENT;
  BatchCorrs fc;
  BatchDepth fd;
  int rc = batch_frames(b, frames, fc, fd);
  if (rc) return rc;
  const std::string key = batch_shape_key(b);
  nrf_batch_tuning best = b->tune;
  bool cached = false;
  if (use_cache) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      best = it->second;
      cached = true;
    }
  }
  hipStream_t st = as_stream(stream);
  if (!cached) {
    rc = batch_tune(b, fc, fd, st, best);
    if (rc) {
      // leave the batch usable with its previous layout
      if (batch_apply_tuning(b, b->tune) != NRF_OK) return NRF_ERR_HIP;
      return rc;
    }
  }
  if ((rc = batch_apply_tuning(b, best))) return rc;
  b->tune.pcg_ms = best.pcg_ms;
  b->tune.integrate_ms = best.integrate_ms;
  if (!cached) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[key] = b->tune;
  }
  if (out) *out = b->tune;
  return NRF_OK;
}
