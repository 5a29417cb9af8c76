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

// Streaming copy for the HBM calibration: float4 loads and stores, a
// grid-stride loop over all CUs with several blocks each; every thread keeps
// four loads in flight before its stores.
__global__ __launch_bounds__(256) void k_stream_copy(const float4* __restrict__ a,
                                                     float4* __restrict__ b, int64_t n4) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < n4; i += 4 * stride) {
    const float4 v0 = a[i], v1 = a[i + stride], v2 = a[i + 2 * stride], v3 = a[i + 3 * stride];
    b[i] = v0;
    b[i + stride] = v1;
    b[i + 2 * stride] = v2;
    b[i + 3 * stride] = v3;
  }
  for (; i < n4; i += stride) b[i] = a[i];
}

}  // namespace seis

extern "C" {

int seis_stream_copy(const float* src, float* dst, int64_t n, int32_t blocks, void* stream) {
  SEIS_CHECK_ARG(src && dst && src != dst, "null or aliased pointer argument");
  SEIS_CHECK_ARG(n > 0 && n % 4 == 0, "n must be a positive multiple of 4, got %lld",
                 (long long)n);
  SEIS_CHECK_ARG((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) % 16 == 0,
                 "pointers must be 16-byte aligned");
  SEIS_CHECK_ARG(blocks >= 1 && blocks <= 1 << 20, "blocks %d outside [1, 2^20]", blocks);
  hipStream_t s = (hipStream_t)stream;
  SEIS_PROF(SEIS_KERNEL_STREAM_COPY, s);
  seis::k_stream_copy<<<blocks, 256, 0, s>>>(reinterpret_cast<const float4*>(src),
                                              reinterpret_cast<float4*>(dst), n / 4);
  SEIS_CHECK_LAUNCH();
  SEIS_PROF(SEIS_KERNEL_STREAM_COPY, s);
  return 0;
}

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
