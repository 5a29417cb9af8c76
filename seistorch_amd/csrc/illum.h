// Source illumination of the forward loops (rnn.py:127-128, 204-205): per
// time level, illum += the sum over shots of the squared post-source field.
#pragma once

#include "common.h"

namespace {

// illum += sum_b u[b]^2 per cell, the shots summed in order into one FP32
// value, then added: torch.sum(u**2, 0) followed by +=.
__global__ __launch_bounds__(256) void k_illum(const float* __restrict__ u, int B, int64_t plane,
                                               float* __restrict__ illum) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= plane) return;
  float acc = 0.0f;
  for (int b = 0; b < B; ++b) {
    const float v = u[b * plane + c];
    acc = acc + v * v;
  }
  illum[c] = illum[c] + acc;
}

// no-op when illum is NULL
inline int illuminate(const float* u, int B, int64_t plane, float* illum, hipStream_t s) {
  if (!illum) return 0;
  k_illum<<<(unsigned)((plane + 255) / 256), 256, 0, s>>>(u, B, plane, illum);
  SEIS_CHECK_LAUNCH();
  return 0;
}

}  // namespace
