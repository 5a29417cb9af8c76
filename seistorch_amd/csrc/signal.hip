// Zero-phase IIR filtering of seismic traces for gfx950 (multiscale FWI).
//
// seistorch filters synthetic and observed records band by band inside the
// FWI loop (seistorch_dist.py:222-225) through seistorch/signal.py:91-96:
// Butterworth (b, a) from scipy.signal.butter, applied with
// torchaudio.functional.filtfilt(x.double(), a, b, clamp=False).float() over
// the time axis of every trace.  That filtfilt is an lfilter with zero
// initial state forward in time, then the same lfilter over the reversed
// result, reversed back (no edge padding), in FP64:
//   w[t] = (sum_k b[k] x[t-k]) / a[0],   y[t] = w[t] - sum_{k>=1} (a[k]/a[0]) y[t-k]
// (x[t < 0] = y[t < 0] = 0).  The map is self-adjoint (the transpose of a
// causal zero-state recursion is the same recursion backward in time), so
// its gradient is the same filter applied to the cotangent.
//
// Layout: time-major (nt, ntr), traces contiguous, i.e. the records layout
// (nt, R, nch) with ntr = R*nch.  One thread owns one trace and runs both
// passes; the forward pass keeps its output in an FP64 work array (the
// reference's intermediate precision), the backward pass reads it in reverse
// and writes FP32.  Loads run one chunk of kChunk samples ahead of the
// recursion (double-buffered in registers), so that a wave keeps 2*kChunk
// HBM requests in flight with few waves resident (one per 64 traces).
#include "common.h"
#include "seistorch_amd.h"

namespace {

constexpr int kMaxCoef = 32;
constexpr int kChunk = 16;

struct Iir {
  double b[kMaxCoef];   // b[k] / 1 (the FIR part is divided by a[0] afterwards)
  double an[kMaxCoef];  // a[k] / a[0]
  double a0;
};

// One recursion step for output t; xh[k] = x[t-k] (k = 0..N-1), yh[j] =
// y[t-1-j] (j = 0..N-2).  The FIR and feedback sums run oldest term first.
template <int N>
__device__ __forceinline__ double iir_step(const Iir& f, const double (&xh)[N],
                                           const double (&yh)[N]) {
  double w = f.b[N - 1] * xh[N - 1];
#pragma unroll
  for (int k = N - 2; k >= 0; --k) w = w + f.b[k] * xh[k];
  w = w / f.a0;
  double d = 0.0;
  if (N > 1) {
    d = f.an[N - 1] * yh[N - 2];
#pragma unroll
    for (int k = N - 2; k >= 1; --k) d = d + f.an[k] * yh[k - 1];
  }
  return w - d;
}

template <int N>
__device__ __forceinline__ void push(double (&h)[N], double v) {
#pragma unroll
  for (int k = N - 1; k >= 1; --k) h[k] = h[k - 1];
  h[0] = v;
}

template <int N>
__global__ __launch_bounds__(256) void k_filtfilt(const float* x, float* y,
                                                  double* __restrict__ work, int64_t nt,
                                                  int64_t ntr, Iir f) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= ntr) return;
  double xh[N], yh[N];
#pragma unroll
  for (int k = 0; k < N; ++k) { xh[k] = 0.0; yh[k] = 0.0; }
  // forward in time: x -> work; the next chunk's loads are in flight while
  // the recursion runs over the current one
  {
    float cur[kChunk], nxt[kChunk];
#pragma unroll
    for (int u = 0; u < kChunk; ++u) cur[u] = u < nt ? x[u * ntr + r] : 0.0f;
    for (int64_t t0 = 0; t0 < nt; t0 += kChunk) {
      const int64_t t1 = t0 + kChunk;
#pragma unroll
      for (int u = 0; u < kChunk; ++u) nxt[u] = t1 + u < nt ? x[(t1 + u) * ntr + r] : 0.0f;
#pragma unroll
      for (int u = 0; u < kChunk; ++u) {
        if (t0 + u >= nt) break;
        push<N>(xh, (double)cur[u]);
        const double o = iir_step<N>(f, xh, yh);
        push<N>(yh, o);
        work[(t0 + u) * ntr + r] = o;
      }
#pragma unroll
      for (int u = 0; u < kChunk; ++u) cur[u] = nxt[u];
    }
  }
#pragma unroll
  for (int k = 0; k < N; ++k) { xh[k] = 0.0; yh[k] = 0.0; }
  // backward in time: work -> y (each work value was written by this thread)
  {
    double cur[kChunk], nxt[kChunk];
#pragma unroll
    for (int u = 0; u < kChunk; ++u) cur[u] = nt - 1 - u >= 0 ? work[(nt - 1 - u) * ntr + r] : 0.0;
    for (int64_t t0 = nt - 1; t0 >= 0; t0 -= kChunk) {
      const int64_t t1 = t0 - kChunk;
#pragma unroll
      for (int u = 0; u < kChunk; ++u) nxt[u] = t1 - u >= 0 ? work[(t1 - u) * ntr + r] : 0.0;
#pragma unroll
      for (int u = 0; u < kChunk; ++u) {
        if (t0 - u < 0) break;
        push<N>(xh, cur[u]);
        const double o = iir_step<N>(f, xh, yh);
        push<N>(yh, o);
        y[(t0 - u) * ntr + r] = (float)o;
      }
#pragma unroll
      for (int u = 0; u < kChunk; ++u) cur[u] = nxt[u];
    }
  }
}

template <int N>
void launch_filtfilt(const float* x, float* y, double* work, int64_t nt, int64_t ntr,
                     const Iir& f, hipStream_t s) {
  const unsigned blocks = (unsigned)((ntr + 255) / 256);
  k_filtfilt<N><<<blocks, 256, 0, s>>>(x, y, work, nt, ntr, f);
}

template <int N>
int dispatch(int n, const float* x, float* y, double* work, int64_t nt, int64_t ntr,
             const Iir& f, hipStream_t s) {
  if constexpr (N > kMaxCoef) {
    return -1;
  } else {
    if (n == N) {
      launch_filtfilt<N>(x, y, work, nt, ntr, f, s);
      return 0;
    }
    return dispatch<N + 1>(n, x, y, work, nt, ntr, f, s);
  }
}

}  // namespace

extern "C" {

int seis_filtfilt(const float* x, float* y, double* work, int64_t nt, int64_t ntr,
                  const double* b, const double* a, int32_t ncoef, void* stream) {
  SEIS_CHECK_ARG(x && y && work && b && a, "null pointer argument");
  SEIS_CHECK_ARG(nt >= 1 && ntr >= 1, "need nt >= 1 and ntr >= 1 (got %lld, %lld)",
                 (long long)nt, (long long)ntr);
  SEIS_CHECK_ARG(ncoef >= 1 && ncoef <= kMaxCoef, "ncoef %d outside [1, %d]", ncoef, kMaxCoef);
  SEIS_CHECK_ARG(a[0] != 0.0, "a[0] must be nonzero");
  Iir f{};
  f.a0 = a[0];
  for (int k = 0; k < ncoef; ++k) {
    f.b[k] = b[k];
    f.an[k] = a[k] / a[0];
  }
  if (dispatch<1>(ncoef, x, y, work, nt, ntr, f, (hipStream_t)stream))
    SEIS_CHECK_ARG(false, "no kernel for ncoef %d", ncoef);
  SEIS_CHECK_LAUNCH();
  return 0;
}

}  // extern "C"
