// Shared helpers for the seistorch_amd HIP library (gfx950 / CDNA4).
//
// Conventions used by every kernel in this library:
//   * wavefields are FP32, shot-major (B, nz, nx) with x the contiguous axis,
//     exactly the tensors the reference's time loop hands to an equation
//     backend (seistorch/rnn.py:116-123);
//   * models and damping arrays are (nz, nx) and shared by all shots;
//   * every entry point validates its arguments on the host, launches on the
//     caller's stream and returns 0, or a nonzero code whose message is
//     available from seis_last_error().
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstdarg>

namespace seis {

void set_error(const char* fmt, ...);

// Records a timing event for `kernel` (a SEIS_KERNEL_* id) on stream s when
// profiling is enabled; call it right before and right after a launch.
int prof_mark(int kernel, hipStream_t s);

// Host-side FP32 time constants, formed with the same FP32 operations the
// reference applies to its float32 dt tensor (acoustic.py:81-84):
// dt**-2 -> 1/(dt*dt), dt**-1 -> 1/dt, 2/dt**2 -> 2/(dt*dt).
struct TimeConsts {
  float idt2;      // dt^-2
  float idt;       // dt^-1
  float two_idt2;  // 2 / dt^2
  float ih2;       // h^-2
  float dt2;       // dt^2
};

inline TimeConsts make_consts(float dt, float h) {
  TimeConsts c;
  volatile float dtdt = dt * dt;
  volatile float hh = h * h;
  c.idt2 = 1.0f / dtdt;
  c.idt = 1.0f / dt;
  c.two_idt2 = 2.0f / dtdt;
  c.ih2 = 1.0f / hh;
  c.dt2 = dtdt;
  return c;
}

}  // namespace seis

#define SEIS_CHECK_ARG(cond, ...)      \
  do {                                 \
    if (!(cond)) {                     \
      seis::set_error(__VA_ARGS__);    \
      return 1;                        \
    }                                  \
  } while (0)

#define SEIS_CHECK_HIP(expr)                                                  \
  do {                                                                        \
    hipError_t _e = (expr);                                                   \
    if (_e != hipSuccess) {                                                   \
      seis::set_error("%s failed: %s", #expr, hipGetErrorString(_e));         \
      return 2;                                                               \
    }                                                                         \
  } while (0)

#define SEIS_CHECK_LAUNCH() SEIS_CHECK_HIP(hipGetLastError())

#define SEIS_PROF(kernel, stream)                              \
  do {                                                         \
    if (int _pe = seis::prof_mark((kernel), (stream))) return _pe; \
  } while (0)
