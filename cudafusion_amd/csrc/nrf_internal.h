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
void solver_set_event_sink(nrf_solver *s, hipEvent_t *events, int count);
int device_max_lds_per_block();
bool device_cooperative();

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
