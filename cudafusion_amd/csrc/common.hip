// common.hip - status strings, error state, device queries, host helpers.
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "nrf_internal.h"

namespace nrf {

static thread_local std::string g_last_error;

void set_last_error(const char *where, hipError_t e) {
  g_last_error = std::string(where) + ": " + hipGetErrorString(e);
}
void set_last_error_msg(const std::string &msg) { g_last_error = msg; }

CamXf invert_pose(const float pose[16]) {
  // src/tsdf_volume.cu:36-49: rigid inverse R^T, -R^T t, FP32.
  CamXf T;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) T.m[i][j] = pose[j * 4 + i];
  for (int i = 0; i < 3; ++i)
    T.m[i][3] = -(T.m[i][0] * pose[0 * 4 + 3] + T.m[i][1] * pose[1 * 4 + 3] + T.m[i][2] * pose[2 * 4 + 3]);
  return T;
}

namespace {
struct DevInfo {
  int cus = 0, lds = 65536;
};
DevInfo query() {
  DevInfo d;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return d;
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess) d.cus = v;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess) d.lds = v;
  return d;
}
const DevInfo &info() {
  static DevInfo d = query();
  return d;
}
}  // namespace

int device_cu_count() { return info().cus > 0 ? info().cus : 1; }
int device_max_lds_per_block() {
  // gfx950 has 160 KiB of LDS per CU; dynamic LDS above 64 KiB needs the
  // per-kernel opt-in attribute (set by the callers that use it).
  return info().lds > 0 ? info().lds : 65536;
}

}  // namespace nrf

extern "C" const char *nrf_status_string(int status) {
  switch (status) {
    case NRF_OK: return "ok";
    case NRF_ERR_INVALID_ARGUMENT: return "invalid argument";
    case NRF_ERR_TOO_LARGE: return "size exceeds a documented capacity";
    case NRF_ERR_HIP: return "HIP runtime error";
    case NRF_ERR_WORKSPACE: return "workspace too small";
    case NRF_ERR_NOT_READY: return "pattern or graph not set";
    case NRF_ERR_BARRIER_TIMEOUT: return "persistent kernel barrier timed out";
    case NRF_ERR_NO_KERNEL: return "no gfx950 kernels in the code object";
    default: return "unknown status";
  }
}

extern "C" const char *nrf_last_error(void) { return nrf::g_last_error.c_str(); }

extern "C" const char *nrf_build_info(void) {
  return "nrf gfx950; device code -O3 -ffp-contract=off, correctly rounded fp32 div/sqrt";
}
