// Error reporting shared by every entry point of the library.
#include "common.h"
#include "seistorch_amd.h"

#include <vector>

namespace seis {

static thread_local char g_last_error[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}

// Launch timing with HIP events on the launching stream (off by default).
namespace {
struct Profile {
  bool on = false;
  std::vector<hipEvent_t> ev[SEIS_PROFILE_KERNELS];
  size_t used[SEIS_PROFILE_KERNELS] = {};
};
Profile g_prof;
}  // namespace

int prof_mark(int kernel, hipStream_t s) {
  if (!g_prof.on || kernel < 0 || kernel >= SEIS_PROFILE_KERNELS) return 0;
  auto& v = g_prof.ev[kernel];
  size_t& u = g_prof.used[kernel];
  if (u == v.size()) {
    hipEvent_t e;
    SEIS_CHECK_HIP(hipEventCreate(&e));
    v.push_back(e);
  }
  SEIS_CHECK_HIP(hipEventRecord(v[u++], s));
  return 0;
}

}  // namespace seis

extern "C" {

int seis_profile_enable(int32_t on) {
  seis::g_prof.on = on != 0;
  return 0;
}

int seis_profile_collect(int32_t kernel, double* total_ms, int64_t* launches) {
  SEIS_CHECK_ARG(kernel >= 0 && kernel < SEIS_PROFILE_KERNELS, "unknown kernel id %d", kernel);
  SEIS_CHECK_ARG(total_ms && launches, "null pointer argument");
  auto& v = seis::g_prof.ev[kernel];
  size_t& u = seis::g_prof.used[kernel];
  double tot = 0.0;
  if (u >= 2) SEIS_CHECK_HIP(hipEventSynchronize(v[u - 1]));
  for (size_t i = 0; i + 1 < u; i += 2) {
    float ms = 0.0f;
    SEIS_CHECK_HIP(hipEventElapsedTime(&ms, v[i], v[i + 1]));
    tot += ms;
  }
  *total_ms = tot;
  *launches = (int64_t)(u / 2);
  u = 0;
  return 0;
}

const char* seis_last_error(void) { return seis::g_last_error; }

int32_t seis_abi_version(void) { return SEIS_ABI_VERSION; }

}  // extern "C"
