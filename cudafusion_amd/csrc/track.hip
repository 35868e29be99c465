// track.hip - the device-side control of the tracking loop for gfx950.
//
// DynamicFusionPipeline::process_frame (include/fusion_app.h:253-669) runs
// its GN loop on the host: after each association it reads the valid count
// (break below min_valid_corrs, :342-343), downloads the correspondences for
// the RMSE (correspondence_rmse, :1331-1351) and, per GN iteration, tries
// ten step scales (:494-641), each with a fresh live surface and
// association, keeping the one of lowest RMSE.  Here every decision is a
// one-workgroup kernel on the stream, so the loop runs with no host
// synchronisation: a decision that ends the loop raises the solver's stop
// flag, and every later kernel of the frame (association, solve, trial,
// update) reads the flag first and returns.  The flag is written only by
// one-workgroup kernels, so no multi-workgroup kernel sees it change.
//
// The RMSE is summed in FP64 (as the reference's host loop), by fixed-order
// per-workgroup partial sums and a fixed-order final sum: deterministic, so
// equal correspondence sets give equal RMSEs and the reference's strict <
// keeps the first of equal trials.
#include <algorithm>
#include <cmath>

#include "nrf_internal.h"
#include "nrf_math.h"

namespace nrf {
namespace {

constexpr int kRmseThreads = 256;

// the one-workgroup entry test: every lane reads the flag before any lane
// of this workgroup may write it
__device__ inline bool stopped_uniform(const int32_t *stop) {
  __shared__ int s_stop;
  if (threadIdx.x == 0) s_stop = *stop;
  __syncthreads();
  const bool r = s_stop != 0;
  __syncthreads();
  return r;
}

__global__ void track_begin_kernel(TrackDev *t) {
  if (threadIdx.x != 0) return;
  nrf_track_stats &s = t->stats;
  s.gn_run = 0;
  s.gn_applied = 0;
  s.stop_reason = 0;
  s.last_valid = 0;
  for (int g = 0; g < NRF_TRACK_MAX_GN; ++g) {
    s.num_valid[g] = 0;
    s.best_scale[g] = 0.f;
    s.base_rmse[g] = 0.0;
    s.best_rmse[g] = 0.0;
  }
  t->counter = 0;
  t->best_scale = 0.f;
  t->best_rmse = 0.0;
}

__global__ void count_reset_kernel(TrackDev *t, const int32_t *stop) {
  if (threadIdx.x == 0 && !*stop) t->counter = 0;
}

// fusion_app.h:342-343: fewer than min_valid valid correspondences end the loop
__global__ void gate_kernel(TrackDev *t, int g, int min_valid, int32_t *stop) {
  if (threadIdx.x != 0 || *stop) return;
  const int nv = t->counter;
  t->stats.num_valid[g] = nv;
  t->stats.gn_run = g + 1;
  if (nv < min_valid) {
    *stop = 1;
    t->stats.stop_reason = 1;
  }
}

// correspondence_rmse (fusion_app.h:1331-1351): over the valid slots with
// weight > 0, r = n . (src - dst) with the difference in FP32 and the dot in
// FP64, w = max(weight, 0); per-workgroup sums of w r^2 and w.
__global__ __launch_bounds__(kRmseThreads) void rmse_partial_kernel(const nrf_corr *__restrict__ corrs, int n,
                                                                    double *__restrict__ part,
                                                                    const int32_t *__restrict__ stop) {
  if (*stop) return;
  const int i = blockIdx.x * kRmseThreads + threadIdx.x;
  double sw = 0.0, sr = 0.0;
  if (i < n) {
    const nrf_corr &c = corrs[i];
    if (c.valid && c.weight > 0.0f) {
      const float dx = c.src[0] - c.dst[0], dy = c.src[1] - c.dst[1], dz = c.src[2] - c.dst[2];
      const double r = (double)c.normal[0] * dx + (double)c.normal[1] * dy + (double)c.normal[2] * dz;
      const double w = (double)fmaxf(0.0f, c.weight);
      sr = w * r * r;
      sw = w;
    }
  }
  __shared__ double s_r[kRmseThreads], s_w[kRmseThreads];
  s_r[threadIdx.x] = sr;
  s_w[threadIdx.x] = sw;
  __syncthreads();
  for (int h = kRmseThreads / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      s_r[threadIdx.x] += s_r[threadIdx.x + h];
      s_w[threadIdx.x] += s_w[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s_r[0];
    part[2 * blockIdx.x + 1] = s_w[0];
  }
}

// the RMSE from the partial sums, by one workgroup in a fixed order
__device__ double rmse_from_parts(const double *part, int nblk) {
  __shared__ double s_r[kRmseThreads], s_w[kRmseThreads];
  double sr = 0.0, sw = 0.0;
  for (int b = threadIdx.x; b < nblk; b += kRmseThreads) {
    sr += part[2 * b];
    sw += part[2 * b + 1];
  }
  s_r[threadIdx.x] = sr;
  s_w[threadIdx.x] = sw;
  __syncthreads();
  for (int h = kRmseThreads / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      s_r[threadIdx.x] += s_r[threadIdx.x + h];
      s_w[threadIdx.x] += s_w[threadIdx.x + h];
    }
    __syncthreads();
  }
  const double r = s_w[0] > 0.0 ? sqrt(s_r[0] / s_w[0]) : (double)INFINITY;
  __syncthreads();
  return r;
}

// the base RMSE of GN iteration g (fusion_app.h:344-346)
__global__ __launch_bounds__(kRmseThreads) void ls_base_kernel(TrackDev *t, const double *part, int nblk, int g,
                                                               const int32_t *stop) {
  if (stopped_uniform(stop)) return;
  const double r = rmse_from_parts(part, nblk);
  if (threadIdx.x == 0) t->stats.base_rmse[g] = r;
}

// After the solve and its validation (a rejected step has raised the flag):
// the reference's break on update_scale <= 0 or a non-finite base RMSE
// (:496-497), then save_transforms (:499) and the search state.
__global__ __launch_bounds__(256) void ls_begin_kernel(TrackDev *t, int g, float update_scale,
                                                       const DQ *__restrict__ T, DQ *__restrict__ Tsave, int N,
                                                       int32_t *stop) {
  if (stopped_uniform(stop)) return;
  for (int i = threadIdx.x; i < N; i += 256) Tsave[i] = T[i];
  if (threadIdx.x == 0) {
    const double base = t->stats.base_rmse[g];
    t->best_rmse = base;
    t->best_scale = 0.f;
    if (update_scale <= 0.0f || !isfinite(base)) {
      *stop = 1;
      t->stats.stop_reason = 4;
    }
  }
}

// restore_transforms + upload_scaled_delta_x(scale) + apply_twist_increment
// (:505-536): T = step(Tsave, dx, scale), and the centred table.  use_best:
// the scale is the search's choice, and a choice of 0 restores Tsave
// (:590-597).
__global__ __launch_bounds__(256) void ls_apply_kernel(TrackDev *t, const float *__restrict__ dx,
                                                       const DQ *__restrict__ Tsave, DQ *__restrict__ T,
                                                       DQ *__restrict__ centered, const nrf_node *__restrict__ nodes,
                                                       int N, float scale, int use_best, float max_rot,
                                                       float max_trans, const int32_t *__restrict__ stop) {
  if (*stop) return;
  const float s = use_best ? t->best_scale : scale;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < N; i += gridDim.x * 256) {
    const DQ Tn = (use_best && s == 0.0f) ? Tsave[i] : scaled_twist_update(dx + 6 * i, s, max_rot, max_trans, Tsave[i]);
    T[i] = Tn;
    centered[i] = dq_centered(Tn, f3(nodes[i].pos));
  }
}

// The trial test (:572-579): enough valid correspondences (at least
// min_valid and 75 % of the iteration's), a finite RMSE below the best.
__global__ __launch_bounds__(kRmseThreads) void ls_trial_kernel(TrackDev *t, const double *part, int nblk, int g,
                                                                float scale, int min_valid, const int32_t *stop) {
  if (stopped_uniform(stop)) return;
  const double r = rmse_from_parts(part, nblk);
  if (threadIdx.x == 0) {
    const int checked = t->counter, num_corrs = t->stats.num_valid[g];
    if (checked >= min_valid && checked >= (int)(0.75f * (float)num_corrs) && isfinite(r) && r < t->best_rmse) {
      t->best_rmse = r;
      t->best_scale = scale;
    }
  }
}

// After the trials: record the choice; a choice of 0 ends the loop (:592-597).
__global__ void ls_commit_kernel(TrackDev *t, int g, int32_t *stop) {
  if (threadIdx.x != 0 || *stop) return;
  t->stats.best_scale[g] = t->best_scale;
  t->stats.best_rmse[g] = t->best_rmse;
  if (t->best_scale == 0.0f) {
    *stop = 1;
    t->stats.stop_reason = 3;
  } else {
    t->stats.gn_applied += 1;
  }
}

__global__ void track_end_kernel(TrackDev *t, const nrf_pcg_result *res, int line_search, const int32_t *stop) {
  if (threadIdx.x != 0) return;
  t->stats.last_valid = t->counter;
  if (!line_search) t->stats.gn_applied = res->gn_done;
  if (*stop && t->stats.stop_reason == 0) t->stats.stop_reason = 2;  // the validation's rejection
}

}  // namespace

int track_begin(TrackDev *t, hipStream_t st) {
  track_begin_kernel<<<1, 64, 0, st>>>(t);
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}
int track_count_reset(TrackDev *t, const int32_t *stop, hipStream_t st) {
  count_reset_kernel<<<1, 64, 0, st>>>(t, stop);
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}
int track_gate(TrackDev *t, int g, int min_valid, int32_t *stop, hipStream_t st) {
  gate_kernel<<<1, 64, 0, st>>>(t, g, min_valid, stop);
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}
int track_rmse_blocks(int n) { return std::max(1, (n + kRmseThreads - 1) / kRmseThreads); }
int track_rmse(TrackDev *t, const nrf_corr *corrs, int n, double *part, int g, bool trial, float scale,
               int min_valid, const int32_t *stop, hipStream_t st) {
  const int nblk = track_rmse_blocks(n);
  rmse_partial_kernel<<<nblk, kRmseThreads, 0, st>>>(corrs, n, part, stop);
  NRF_LAUNCH_CHECK();
  if (trial)
    ls_trial_kernel<<<1, kRmseThreads, 0, st>>>(t, part, nblk, g, scale, min_valid, stop);
  else
    ls_base_kernel<<<1, kRmseThreads, 0, st>>>(t, part, nblk, g, stop);
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}
int track_ls_begin(TrackDev *t, int g, float update_scale, const nrf_dualquat *T, nrf_dualquat *Tsave, int N,
                   int32_t *stop, hipStream_t st) {
  ls_begin_kernel<<<1, 256, 0, st>>>(t, g, update_scale, reinterpret_cast<const DQ *>(T),
                                     reinterpret_cast<DQ *>(Tsave), N, stop);
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}
int track_ls_apply(TrackDev *t, const float *dx, const nrf_dualquat *Tsave, nrf_dualquat *T, nrf_dualquat *centered,
                   const nrf_node *nodes, int N, float scale, bool use_best, float max_rot, float max_trans,
                   const int32_t *stop, hipStream_t st) {
  if (N <= 0) return NRF_OK;
  ls_apply_kernel<<<(N + 255) / 256, 256, 0, st>>>(t, dx, reinterpret_cast<const DQ *>(Tsave),
                                                   reinterpret_cast<DQ *>(T), reinterpret_cast<DQ *>(centered),
                                                   nodes, N, scale, use_best ? 1 : 0, max_rot, max_trans, stop);
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}
int track_ls_commit(TrackDev *t, int g, int32_t *stop, hipStream_t st) {
  ls_commit_kernel<<<1, 64, 0, st>>>(t, g, stop);
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}
int track_end(TrackDev *t, const nrf_pcg_result *res, bool line_search, const int32_t *stop, hipStream_t st) {
  track_end_kernel<<<1, 64, 0, st>>>(t, res, line_search ? 1 : 0, stop);
  NRF_LAUNCH_CHECK();
  return NRF_OK;
}

}  // namespace nrf
