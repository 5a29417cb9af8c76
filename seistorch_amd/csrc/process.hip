// Gradient post-processing for gfx950: Gaussian smoothing, seabed cut and
// illumination preconditioning of a model gradient between the adjoint and
// the optimizer step.
//
// seistorch applies, in this order (seistorch_dist.py:270-280):
//   smooth_gradient (process.py:66-112): if the seabed cut is on, g *= mask;
//     then `counts` times a 1-D Gaussian along z, then along x
//     (signal.py:247-304);
//   cut_gradient (process.py:43-47): g *= mask;
//   precondition (process.py:114-117): g /= illumination.
// seis_smooth_gradient runs any of these in one sequence of passes: the
// pre-mask rides on the first pass's loads, the cut and the division on the
// last pass's store.
// The 1-D Gaussian builds a kernel of ks taps (ks = 2r+1 for even radius r,
// 2r for odd r), normalised to sum 1 in FP32, reflect-pads the field by
// ks/2 cells on both axes (numpy 'reflect': the edge cell is not repeated),
// correlates along one axis with zero padding ks/2 and crops ks/2 cells from
// both ends of both axes.  Written out, with refl() the reflection and
// r = ks/2:
//   out[i] = sum_{k=0}^{ks-1} w[k] * x[refl(i + k - r)],  i < n + 2r - ks + 1
// so an odd radius (even ks) lengthens the smoothed axis by one cell.
//
// A C2-sized gradient (900 x 2500) is 9 MB: each pass is a few microseconds,
// negligible next to a gradient.
#include "common.h"
#include "seistorch_amd.h"

#include <mutex>

// Workspace of seis_smooth_gradient_noalloc: one FP32 plane of at most
// SEIS_NOALLOC_MAX_CELLS cells (16 MiB), part of the library's code object, so
// each device that runs the library holds its own copy until the process
// ends.  Host code never uses the name `scratch` as an address (on the host
// it names a stand-in, not device memory): scratch_ptr() resolves the
// current device's copy once with hipGetSymbolAddress and keeps it.  Every
// pass writes every cell of its destination before the next pass reads it,
// so nothing here has to start at zero; calls that overlap on one device
// share the plane and race.
static __device__ __align__(256) float scratch[SEIS_NOALLOC_MAX_CELLS];

namespace {

constexpr int kMaxTaps = 129;  // radius <= 64

struct Taps {
  float w[kMaxTaps];
};

__device__ __forceinline__ int64_t reflect(int64_t j, int64_t n) {
  return j < 0 ? -j : (j >= n ? 2 * (n - 1) - j : j);
}

// Tiles: the z pass gives a block 64 columns x kTz output rows (4 waves, one
// column per lane, kTz/4 rows per thread); the x pass 256 columns x kTxRows
// rows (one column per thread).  A block first stages its input window,
// reflected and pre-multiplied, in LDS with coalesced loads (each input value
// is read from HBM once per tile instead of once per tap), then sums the
// taps from LDS in tap order.
constexpr int kLanes = 64;  // z pass: one column per lane of a wave
constexpr int kTz = 32;
constexpr int kTxRows = 4;
constexpr int kBatch = 8;

__device__ __forceinline__ int64_t window(int64_t j, int64_t n, int r) {
  // rows/columns past n - 1 + r only feed outputs past the end (never
  // stored); clamping keeps the reflection inside [0, n) for r < n
  return reflect(j < n - 1 + r ? j : n - 1 + r, n);
}

// One 1-D pass along AXIS (0 = z, 1 = x) from an (nz, nx) field to an
// (oz, ox) one.  Every input value is multiplied by pre first (when given);
// the result is multiplied by post (when given), then divided by illum (when
// given).
template <int AXIS>
__global__ void __launch_bounds__(256)
    k_gauss(const float* __restrict__ in, float* __restrict__ out, const float* __restrict__ pre,
            const float* __restrict__ post, const float* __restrict__ illum, int64_t nz,
            int64_t nx, int64_t oz, int64_t ox, Taps t, int ks, int r) {
  extern __shared__ float win[];
  const int tid = threadIdx.x;
  if (AXIS == 0) {
    const int64_t x0 = (int64_t)blockIdx.x * kLanes, z0 = (int64_t)blockIdx.y * kTz;
    const int lane = tid % kLanes, wv = tid / kLanes;
    const int64_t x = x0 + lane;
    const int rows = kTz + ks - 1;
    if (x < nx) {
      // batches of kBatch rows per wave: all loads of a batch are in flight
      // before the first LDS store
      for (int q0 = wv; q0 < rows; q0 += 4 * kBatch) {
        float v[kBatch], m[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
          const int q = q0 + 4 * u;
          if (q < rows) {
            const int64_t j = window(z0 + q - r, nz, r) * nx + x;
            v[u] = in[j];
            m[u] = pre ? pre[j] : 1.0f;
          }
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
          const int q = q0 + 4 * u;
          if (q < rows) win[q * kLanes + lane] = pre ? v[u] * m[u] : v[u];
        }
      }
    }
    __syncthreads();
    if (x >= ox) return;
    // taps outer, rows inner: one scalar load of a tap feeds kPer
    // independent accumulators (each output still sums its taps in order)
    constexpr int kPer = kTz / 4;
    float acc[kPer];
#pragma unroll
    for (int i = 0; i < kPer; ++i) acc[i] = 0.0f;
    const float* col = win + wv * kPer * kLanes + lane;
    for (int k = 0; k < ks; ++k) {
      const float wk = t.w[k];
#pragma unroll
      for (int i = 0; i < kPer; ++i) acc[i] = acc[i] + wk * col[(i + k) * kLanes];
    }
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      const int64_t z = z0 + wv * kPer + i;
      if (z >= oz) break;
      const int64_t c = z * ox + x;
      float a = acc[i];
      if (post) a = a * post[c];
      if (illum) a = a / illum[c];
      out[c] = a;
    }
  } else {
    const int64_t x0 = (int64_t)blockIdx.x * 256, z0 = (int64_t)blockIdx.y * kTxRows;
    const int cols = 256 + ks - 1;
    // the window is kTxRows rows of cols <= 384 values: at most 2 per thread
    // and row, all loads in flight before the first LDS store
    float v[kTxRows][2], m[kTxRows][2];
#pragma unroll
    for (int rr = 0; rr < kTxRows; ++rr) {
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int q = tid + 256 * u;
        if (z0 + rr < nz && q < cols) {
          const int64_t j = (z0 + rr) * nx + window(x0 + q - r, nx, r);
          v[rr][u] = in[j];
          m[rr][u] = pre ? pre[j] : 1.0f;
        }
      }
    }
#pragma unroll
    for (int rr = 0; rr < kTxRows; ++rr) {
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int q = tid + 256 * u;
        if (z0 + rr < nz && q < cols) win[rr * cols + q] = pre ? v[rr][u] * m[rr][u] : v[rr][u];
      }
    }
    __syncthreads();
    const int64_t x = x0 + tid;
    if (x >= ox) return;
    float acc[kTxRows];
#pragma unroll
    for (int rr = 0; rr < kTxRows; ++rr) acc[rr] = 0.0f;
    for (int k = 0; k < ks; ++k) {
      const float wk = t.w[k];
#pragma unroll
      for (int rr = 0; rr < kTxRows; ++rr) acc[rr] = acc[rr] + wk * win[rr * cols + tid + k];
    }
#pragma unroll
    for (int rr = 0; rr < kTxRows; ++rr) {
      const int64_t z = z0 + rr;
      if (z >= oz) break;
      const int64_t c = z * ox + x;
      float a = acc[rr];
      if (post) a = a * post[c];
      if (illum) a = a / illum[c];
      out[c] = a;
    }
  }
}

void launch_gauss(int axis, const float* in, float* out, const float* pre, const float* post,
                  const float* illum, int64_t nz, int64_t nx, int64_t oz, int64_t ox,
                  const Taps& t, int ks, hipStream_t s) {
  if (axis == 0) {
    const dim3 grid((unsigned)((ox + kLanes - 1) / kLanes), (unsigned)((oz + kTz - 1) / kTz));
    const size_t lds = sizeof(float) * (size_t)(kTz + ks - 1) * kLanes;
    k_gauss<0><<<grid, 256, lds, s>>>(in, out, pre, post, illum, nz, nx, oz, ox, t, ks, ks / 2);
  } else {
    const dim3 grid((unsigned)((ox + 255) / 256), (unsigned)((oz + kTxRows - 1) / kTxRows));
    const size_t lds = sizeof(float) * (size_t)kTxRows * (256 + ks - 1);
    k_gauss<1><<<grid, 256, lds, s>>>(in, out, pre, post, illum, nz, nx, oz, ox, t, ks, ks / 2);
  }
}

int smooth_gradient(const float* grad, float* out, float* work, int64_t nz, int64_t nx,
                    const float* wz, int32_t kz, const float* wx, int32_t kx, int32_t counts,
                    const float* pre_mask, const float* post_mask, const float* illum,
                    hipStream_t s) {
  SEIS_CHECK_ARG(grad && out, "null pointer argument");
  SEIS_CHECK_ARG(out != grad, "out must not alias grad");
  SEIS_CHECK_ARG(nz >= 1 && nx >= 1, "empty field (%lld x %lld)", (long long)nz, (long long)nx);
  SEIS_CHECK_ARG(counts >= 0, "counts must be >= 0, got %d", counts);
  Taps one{};
  one.w[0] = 1.0f;
  if (counts == 0) {
    // masks and preconditioning only: a one-tap pass (1 * g is exact)
    launch_gauss(1, grad, out, pre_mask, post_mask, illum, nz, nx, nz, nx, one, 1, s);
    SEIS_CHECK_LAUNCH();
    return 0;
  }
  SEIS_CHECK_ARG(work && wz && wx, "null pointer argument");
  SEIS_CHECK_ARG(work != grad && work != out, "work must not alias grad or out");
  SEIS_CHECK_ARG(kz >= 1 && kz <= kMaxTaps && kx >= 1 && kx <= kMaxTaps,
                 "kernel sizes %d, %d outside [1, %d]", kz, kx, kMaxTaps);
  SEIS_CHECK_ARG(kz % 2 == 1 && kx % 2 == 1,
                 "kernel sizes must be odd (even radius): an even size (odd radius) lengthens "
                 "the gradient by one cell per pass; got %d, %d", kz, kx);
  SEIS_CHECK_ARG(kz / 2 < nz && kx / 2 < nx, "radius must be smaller than the axis length");
  Taps tz{}, tx{};
  for (int k = 0; k < kz; ++k) tz.w[k] = wz[k];
  for (int k = 0; k < kx; ++k) tx.w[k] = wx[k];
  // 2*counts passes alternate between work and out, ending in out
  const int passes = 2 * counts;
  const float* src = grad;
  for (int p = 0; p < passes; ++p) {
    float* dst = ((passes - 1 - p) % 2 == 0) ? out : work;
    const int axis = p % 2;
    const bool last = p == passes - 1;
    launch_gauss(axis, src, dst, p == 0 ? pre_mask : nullptr, last ? post_mask : nullptr,
                 last ? illum : nullptr, nz, nx, nz, nx, axis == 0 ? tz : tx, axis == 0 ? kz : kx,
                 s);
    src = dst;
  }
  SEIS_CHECK_LAUNCH();
  return 0;
}

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_dev[kMaxDevices] = {};

int scratch_ptr(float** out) {
  int dev = -1;
  SEIS_CHECK_HIP(hipGetDevice(&dev));
  SEIS_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "device ordinal %d outside [0, %d)", dev,
                 kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    SEIS_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    SEIS_CHECK_ARG(p != nullptr, "hipGetSymbolAddress returned no address for the workspace");
    g_scratch_dev[dev] = static_cast<float*>(p);
  }
  *out = g_scratch_dev[dev];
  return 0;
}

}  // namespace

extern "C" {

int seis_gaussian_filter(const float* x, float* y, int64_t nz, int64_t nx, int32_t axis,
                         const float* w, int32_t ks, void* stream) {
  SEIS_CHECK_ARG(x && y && w, "null pointer argument");
  SEIS_CHECK_ARG(x != y, "y must not alias x");
  SEIS_CHECK_ARG(axis == 0 || axis == 1, "axis must be 0 (z) or 1 (x), got %d", axis);
  SEIS_CHECK_ARG(ks >= 1 && ks <= kMaxTaps, "kernel size %d outside [1, %d]", ks, kMaxTaps);
  SEIS_CHECK_ARG(nz >= 1 && nx >= 1, "empty field (%lld x %lld)", (long long)nz, (long long)nx);
  const int64_t n = axis == 0 ? nz : nx;
  SEIS_CHECK_ARG(ks / 2 < n, "half kernel size %d must be smaller than the axis length %lld",
                 ks / 2, (long long)n);
  Taps t{};
  for (int k = 0; k < ks; ++k) t.w[k] = w[k];
  const int64_t grow = 2 * (ks / 2) - ks + 1;
  const int64_t oz = axis == 0 ? nz + grow : nz, ox = axis == 1 ? nx + grow : nx;
  launch_gauss(axis, x, y, nullptr, nullptr, nullptr, nz, nx, oz, ox, t, ks, (hipStream_t)stream);
  SEIS_CHECK_LAUNCH();
  return 0;
}

int seis_smooth_gradient(const float* grad, float* out, float* work, int64_t nz, int64_t nx,
                         const float* wz, int32_t kz, const float* wx, int32_t kx, int32_t counts,
                         const float* pre_mask, const float* post_mask, const float* illum,
                         void* stream) {
  return smooth_gradient(grad, out, work, nz, nx, wz, kz, wx, kx, counts, pre_mask, post_mask,
                         illum, (hipStream_t)stream);
}

int seis_smooth_gradient_noalloc(const float* grad, float* out, int64_t nz, int64_t nx,
                                 const float* wz, int32_t kz, const float* wx, int32_t kx,
                                 int32_t counts, const float* pre_mask, const float* post_mask,
                                 const float* illum, void* stream) {
  SEIS_CHECK_ARG(nz >= 1 && nx >= 1, "empty field (%lld x %lld)", (long long)nz, (long long)nx);
  SEIS_CHECK_ARG(nz <= SEIS_NOALLOC_MAX_CELLS / nx,
                 "seis_smooth_gradient_noalloc: %lld x %lld = %lld cells exceed the library's "
                 "static workspace of %lld cells; use seis_smooth_gradient with a work buffer",
                 (long long)nz, (long long)nx, (long long)(nz * nx),
                 (long long)SEIS_NOALLOC_MAX_CELLS);
  float* work = nullptr;
  if (int rc = scratch_ptr(&work)) return rc;
  return smooth_gradient(grad, out, work, nz, nx, wz, kz, wx, kx, counts, pre_mask, post_mask,
                         illum, (hipStream_t)stream);
}

}  // extern "C"
