// Kernels of the learned gradient preconditioner (a small U-Net) for gfx950.
//
// Activations are NHWC with channels contiguous; weights and activations are
// bfloat16 (stored as raw uint16 bit patterns), accumulation is FP32 on the
// matrix cores (v_mfma_f32_16x16x32_bf16), and gradients of activations are
// kept in FP32.  A 3x3 convolution (zero padding 1, stride 1) is an implicit
// GEMM: rows = output pixels, columns = output channels, reduction index
// k = tap * Cin + ci over the 9 taps (tap = 3*(dy+1) + (dx+1)).  Weights are
// stored [Cout][9*Cin] so that one lane's 8 consecutive k are one 16-byte load.
//
// MFMA operand layout (16x16x32 bf16, one wave): lane l holds
//   A[m = l%16][k = 8*(l/16) + i], B[k = 8*(l/16) + i][n = l%16], i = 0..7,
// and the result D[m = 4*(l/16) + i][n = l%16], i = 0..3.
//
// There is no reference counterpart: the network and its training are this
// project's own (the plain-gradient FWI step of the reference,
// seistorch/fwi.py, is its fallback).
#include "common.h"
#include "seistorch_amd.h"

#include <cstdint>

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kWave = 64;

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// FP32 -> bf16, round to nearest even (NaN kept quiet).
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ bf16x8 load8(const uint16_t* p) {
  const uint4 v = *reinterpret_cast<const uint4*>(p);
  return __builtin_bit_cast(bf16x8, v);
}

__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, make_uint4(0, 0, 0, 0)); }

constexpr float kInvSqrt2 = 0.70710678118654752f;
constexpr float kInvSqrt2Pi = 0.39894228040143268f;

__device__ __forceinline__ float gelu(float v) { return 0.5f * v * (1.0f + erff(v * kInvSqrt2)); }

__device__ __forceinline__ float gelu_grad(float v) {
  return 0.5f * (1.0f + erff(v * kInvSqrt2)) + v * kInvSqrt2Pi * expf(-0.5f * v * v);
}

// ---------------------------------------------------------------- conv 3x3

// One block = 64 consecutive pixels of one image row x NT output channels.
// The K loop runs over chunks of 32 reduction indices: the block's 64 x 32
// input slice (gathered from the 9 taps with zero padding) and NT x 32 weight
// slice are loaded once, 16 bytes per thread, into a double-buffered LDS tile
// (rows padded to 40 elements against bank conflicts); each wave then reads
// its 16-pixel A fragment and NT/16 B fragments from LDS and issues NT/16
// MFMAs, while the next chunk's global loads are in flight.
// OUT 0: z = acc + bias (bf16, optional) and y = act(z) (bf16);
// OUT 1: y = acc (+ bias) in FP32 (the data gradient of a convolution).
constexpr int kCvPx = 64;
constexpr int kCvK = 32;
constexpr int kCvRow = kCvK + 8;  // padded LDS row (elements)

template <int NT>
__device__ __forceinline__ void conv_fetch(const uint16_t* __restrict__ x, int n, int H, int W,
                                           int Cin, int yrow, int px0,
                                           const uint16_t* __restrict__ w, int co0, int Cout,
                                           int K, int kc, uint4& ra, uint4& rb) {
  const int t = threadIdx.x;
  // A: pixel t/4, 8 reduction indices starting at (t%4)*8
  {
    const int p = t / 4, ks = (t % 4) * 8;
    const int kb = kc * kCvK + ks;
    const int tap = kb / Cin;
    const int ci = kb - tap * Cin;
    ra = make_uint4(0u, 0u, 0u, 0u);
    if (tap < 9) {
      const int yy = yrow + tap / 3 - 1;
      const int xx = px0 + p + tap % 3 - 1;
      if (px0 + p < W && yy >= 0 && yy < H && xx >= 0 && xx < W)
        ra = *reinterpret_cast<const uint4*>(x + (((int64_t)n * H + yy) * W + xx) * Cin + ci);
    }
  }
  // B: output channel t/4 (threads beyond NT*4 load nothing)
  rb = make_uint4(0u, 0u, 0u, 0u);
  if (t < NT * 4) {
    const int co = t / 4, ks = (t % 4) * 8;
    const int kb = kc * kCvK + ks;
    if (kb < K && co0 + co < Cout)
      rb = *reinterpret_cast<const uint4*>(w + (int64_t)(co0 + co) * K + kb);
  }
}

template <int NT, int OUT>
__global__ __launch_bounds__(256) void k_conv3x3(const uint16_t* __restrict__ x, int N, int H,
                                                 int W, int Cin, const uint16_t* __restrict__ w,
                                                 const float* __restrict__ bias, int Cout, int act,
                                                 uint16_t* __restrict__ z, uint16_t* __restrict__ y,
                                                 float* __restrict__ yf) {
  __shared__ __align__(16) uint16_t sA[2][kCvPx][kCvRow];
  __shared__ __align__(16) uint16_t sB[2][NT][kCvRow];
  const int segs = (W + kCvPx - 1) / kCvPx;
  const int seg = blockIdx.x % segs;
  const int row = blockIdx.x / segs;  // n*H + yrow
  const int yrow = row % H;
  const int n = row / H;
  const int px0 = seg * kCvPx;
  const int co0 = blockIdx.y * NT;
  const int wave = threadIdx.x / kWave;
  const int l = threadIdx.x % kWave;
  const int m = l % 16, kq = l / 16;
  const int K = 9 * Cin;
  const int nk = (K + kCvK - 1) / kCvK;
  const int t = threadIdx.x;
  floatx4 acc[NT / 16];
#pragma unroll
  for (int j = 0; j < NT / 16; ++j) acc[j] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
  uint4 ra, rb;
  conv_fetch<NT>(x, n, H, W, Cin, yrow, px0, w, co0, Cout, K, 0, ra, rb);
  *reinterpret_cast<uint4*>(&sA[0][t / 4][(t % 4) * 8]) = ra;
  if (t < NT * 4) *reinterpret_cast<uint4*>(&sB[0][t / 4][(t % 4) * 8]) = rb;
  __syncthreads();
  for (int kc = 0; kc < nk; ++kc) {
    const int buf = kc & 1;
    const bool more = kc + 1 < nk;
    if (more) conv_fetch<NT>(x, n, H, W, Cin, yrow, px0, w, co0, Cout, K, kc + 1, ra, rb);
    const bf16x8 a = __builtin_bit_cast(
        bf16x8, *reinterpret_cast<const uint4*>(&sA[buf][wave * 16 + m][kq * 8]));
#pragma unroll
    for (int j = 0; j < NT / 16; ++j) {
      const bf16x8 b = __builtin_bit_cast(
          bf16x8, *reinterpret_cast<const uint4*>(&sB[buf][j * 16 + m][kq * 8]));
      acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[j], 0, 0, 0);
    }
    if (more) {
      // the other buffer was last read in iteration kc-1, before its barrier
      *reinterpret_cast<uint4*>(&sA[buf ^ 1][t / 4][(t % 4) * 8]) = ra;
      if (t < NT * 4) *reinterpret_cast<uint4*>(&sB[buf ^ 1][t / 4][(t % 4) * 8]) = rb;
    }
    __syncthreads();
  }
  const int pbase = px0 + wave * 16 + kq * 4;
#pragma unroll
  for (int j = 0; j < NT / 16; ++j) {
    const int co = co0 + j * 16 + m;
    if (co >= Cout) continue;
    const float bv = bias ? bias[co] : 0.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int p = pbase + i;
      if (p >= W) continue;
      const int64_t o = (((int64_t)n * H + yrow) * W + p) * Cout + co;
      const float v = acc[j][i] + bv;
      if (OUT == 0) {
        if (z) z[o] = f2bf(v);
        y[o] = f2bf(act ? gelu(v) : v);
      } else {
        yf[o] = v;
      }
    }
  }
}

// Weights of the data gradient: wt[ci][t][co] = w[co][8-t][ci] (taps flipped,
// channel roles swapped), so the data gradient is k_conv3x3 of the output
// gradient with wt.
__global__ __launch_bounds__(256) void k_flip_weights(const uint16_t* __restrict__ w, int Cout,
                                                      int Cin, uint16_t* __restrict__ wt) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)Cout * 9 * Cin;
  if (t >= total) return;
  const int co = (int)(t % Cout);
  const int tap = (int)((t / Cout) % 9);
  const int ci = (int)(t / ((int64_t)Cout * 9));
  wt[t] = w[((int64_t)co * 9 + (8 - tap)) * Cin + ci];
}

// Weight gradient dw[co][tap*Cin + ci] += sum_p dz[p][co] * x[p + tap][ci].
// A block owns CIT input x COT output channels and loops over 32-pixel row
// segments: the segment's dz (transposed to [co][x]) and its 9 shifted input
// windows (transposed to [tap][ci][x]) are staged in LDS, then the 4 waves
// split the 9 * (CIT/16) * (COT/16) output tiles, accumulating in registers
// over the segments; FP32 atomics add the block's sums at the end.
constexpr int kWgCIT = 32;
constexpr int kWgCOT = 64;
constexpr int kWgSeg = 32;
constexpr int kWgTiles = 9 * (kWgCIT / 16) * (kWgCOT / 16);  // 72
constexpr int kWgPerWave = kWgTiles / 4;                       // 18

__global__ __launch_bounds__(256) void k_conv3x3_wgrad(const uint16_t* __restrict__ x,
                                                       const uint16_t* __restrict__ dz, int N,
                                                       int H, int W, int Cin, int Cout,
                                                       float* __restrict__ dw) {
  __shared__ __align__(16) uint16_t sdz[kWgCOT][kWgSeg];
  __shared__ __align__(16) uint16_t sx[9][kWgCIT][kWgSeg];
  const int ntci = (Cin + kWgCIT - 1) / kWgCIT;
  const int ci0 = (blockIdx.x % ntci) * kWgCIT;
  const int co0 = (blockIdx.x / ntci) * kWgCOT;
  const int wave = threadIdx.x / kWave;
  const int l = threadIdx.x % kWave;
  const int m = l % 16, kq = l / 16;
  const int segs_row = (W + kWgSeg - 1) / kWgSeg;
  const int64_t nseg = (int64_t)N * H * segs_row;
  floatx4 acc[kWgPerWave];
#pragma unroll
  for (int s = 0; s < kWgPerWave; ++s) acc[s] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
  for (int64_t sg = blockIdx.y; sg < nseg; sg += gridDim.y) {
    const int xs0 = (int)(sg % segs_row) * kWgSeg;
    const int64_t row = sg / segs_row;
    const int yrow = (int)(row % H);
    const int n = (int)(row / H);
    __syncthreads();  // the previous segment's fragments have been read
    // dz: 32 pixels x 64 channels, 8 channels (16 bytes) per load, written
    // transposed to [co][x]
    {
      const int xo = threadIdx.x / (kWgCOT / 8), c8 = (threadIdx.x % (kWgCOT / 8)) * 8;
      const int px = xs0 + xo;
      uint4 v = make_uint4(0u, 0u, 0u, 0u);
      if (px < W && co0 + c8 < Cout)
        v = *reinterpret_cast<const uint4*>(dz + (((int64_t)n * H + yrow) * W + px) * Cout +
                                            co0 + c8);
      const uint16_t* e = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
      for (int j = 0; j < 8; ++j) sdz[c8 + j][xo] = e[j];
    }
    // input window: rows yrow-1..yrow+1, x from xs0-1 to xs0+32, 32 channels,
    // 8 channels per load; each load feeds the three horizontal taps
    for (int e = threadIdx.x; e < 3 * (kWgSeg + 2) * (kWgCIT / 8); e += 256) {
      const int c8 = (e % (kWgCIT / 8)) * 8;
      const int xw = (e / (kWgCIT / 8)) % (kWgSeg + 2);
      const int dy = e / ((kWgCIT / 8) * (kWgSeg + 2)) - 1;
      const int yy = yrow + dy, xx = xs0 + xw - 1;
      uint4 v = make_uint4(0u, 0u, 0u, 0u);
      if (yy >= 0 && yy < H && xx >= 0 && xx < W && ci0 + c8 < Cin)
        v = *reinterpret_cast<const uint4*>(x + (((int64_t)n * H + yy) * W + xx) * Cin + ci0 +
                                            c8);
      const uint16_t* ev = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
      for (int dx = -1; dx <= 1; ++dx) {
        const int xo = xw - dx - 1;
        if (xo < 0 || xo >= kWgSeg) continue;
        const int tap = 3 * (dy + 1) + (dx + 1);
#pragma unroll
        for (int j = 0; j < 8; ++j) sx[tap][c8 + j][xo] = ev[j];
      }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < kWgPerWave; ++s) {
      const int t = wave + 4 * s;  // tile: (tap, ci sub-tile, co sub-tile)
      const int cos = t % (kWgCOT / 16);
      const int cis = (t / (kWgCOT / 16)) % (kWgCIT / 16);
      const int tap = t / ((kWgCOT / 16) * (kWgCIT / 16));
      const bf16x8 a = __builtin_bit_cast(
          bf16x8, *reinterpret_cast<const uint4*>(&sdz[cos * 16 + m][kq * 8]));
      const bf16x8 b = __builtin_bit_cast(
          bf16x8, *reinterpret_cast<const uint4*>(&sx[tap][cis * 16 + m][kq * 8]));
      acc[s] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[s], 0, 0, 0);
    }
  }
  const int K = 9 * Cin;
#pragma unroll
  for (int s = 0; s < kWgPerWave; ++s) {
    const int t = wave + 4 * s;
    const int cos = t % (kWgCOT / 16);
    const int cis = (t / (kWgCOT / 16)) % (kWgCIT / 16);
    const int tap = t / ((kWgCOT / 16) * (kWgCIT / 16));
    const int ci = ci0 + cis * 16 + m;
    if (ci >= Cin) continue;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int co = co0 + cos * 16 + kq * 4 + i;
      if (co < Cout) atomicAdd(dw + (int64_t)co * K + tap * Cin + ci, acc[s][i]);
    }
  }
}

// db[co] += sum_p dz[p][co]: one block per 256-pixel chunk, one thread per
// channel (Cout <= 1024), strided over the chunk.
__global__ __launch_bounds__(256) void k_bias_grad(const uint16_t* __restrict__ dz, int64_t P,
                                                   int Cout, float* __restrict__ db) {
  const int64_t p0 = (int64_t)blockIdx.x * 256;
  for (int co = threadIdx.x; co < Cout; co += blockDim.x) {
    float s = 0.0f;
    const int64_t p1 = p0 + 256 < P ? p0 + 256 : P;
    for (int64_t p = p0; p < p1; ++p) s += bf2f(dz[p * Cout + co]);
    atomicAdd(db + co, s);
  }
}

// ----------------------------------------------------------- elementwise

// dz = dy * gelu'(z), rounded to bf16 (the operand of the next GEMMs).
__global__ __launch_bounds__(256) void k_gelu_bwd(const float* __restrict__ dy,
                                                  const uint16_t* __restrict__ z, int64_t n,
                                                  uint16_t* __restrict__ dz) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dz[t] = f2bf(dy[t] * gelu_grad(bf2f(z[t])));
}

__global__ __launch_bounds__(256) void k_to_bf16(const float* __restrict__ x, int64_t n,
                                                 uint16_t* __restrict__ y) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) y[t] = f2bf(x[t]);
}

// 2x2 average pooling, (N, H, W, C) -> (N, H/2, W/2, C); H, W even.
__global__ __launch_bounds__(256) void k_avgpool2(const uint16_t* __restrict__ x, int N, int H,
                                                  int W, int C, uint16_t* __restrict__ y) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int Ho = H / 2, Wo = W / 2;
  if (t >= (int64_t)N * Ho * Wo * C) return;
  const int c = (int)(t % C);
  const int64_t q = t / C;
  const int xo = (int)(q % Wo);
  const int yo = (int)((q / Wo) % Ho);
  const int n = (int)(q / ((int64_t)Wo * Ho));
  const int64_t b = (((int64_t)n * H + 2 * yo) * W + 2 * xo) * C + c;
  const float s = (bf2f(x[b]) + bf2f(x[b + C])) + (bf2f(x[b + (int64_t)W * C]) +
                                                   bf2f(x[b + (int64_t)W * C + C]));
  y[t] = f2bf(0.25f * s);
}

// dx (N, H, W, C) = 0.25 * dy[pixel / 2] (+= dx if accumulate).
__global__ __launch_bounds__(256) void k_avgpool2_bwd(const float* __restrict__ dy, int N, int H,
                                                      int W, int C, float* __restrict__ dx,
                                                      int accumulate) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)N * H * W * C) return;
  const int c = (int)(t % C);
  const int64_t q = t / C;
  const int xx = (int)(q % W);
  const int yy = (int)((q / W) % H);
  const int n = (int)(q / ((int64_t)W * H));
  const float g = 0.25f * dy[(((int64_t)n * (H / 2) + yy / 2) * (W / 2) + xx / 2) * C + c];
  dx[t] = accumulate ? dx[t] + g : g;
}

// out (N, H, W, C1 + C2) = [nearest 2x upsampling of low (N, H/2, W/2, C1), skip].
__global__ __launch_bounds__(256) void k_up_concat(const uint16_t* __restrict__ low,
                                                   const uint16_t* __restrict__ skip, int N,
                                                   int H, int W, int C1, int C2,
                                                   uint16_t* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int C = C1 + C2;
  if (t >= (int64_t)N * H * W * C) return;
  const int c = (int)(t % C);
  const int64_t q = t / C;  // pixel
  const int xx = (int)(q % W);
  const int yy = (int)((q / W) % H);
  const int n = (int)(q / ((int64_t)W * H));
  out[t] = c < C1 ? low[(((int64_t)n * (H / 2) + yy / 2) * (W / 2) + xx / 2) * C1 + c]
                  : skip[q * C2 + (c - C1)];
}

// Gradient of k_up_concat: dlow = sum of the 2x2 children; dskip (+)= its slice.
__global__ __launch_bounds__(256) void k_up_concat_bwd_low(const float* __restrict__ dout,
                                                           int N, int H, int W, int C1, int C2,
                                                           float* __restrict__ dlow) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int Ho = H / 2, Wo = W / 2, C = C1 + C2;
  if (t >= (int64_t)N * Ho * Wo * C1) return;
  const int c = (int)(t % C1);
  const int64_t q = t / C1;
  const int xo = (int)(q % Wo);
  const int yo = (int)((q / Wo) % Ho);
  const int n = (int)(q / ((int64_t)Wo * Ho));
  const int64_t b = (((int64_t)n * H + 2 * yo) * W + 2 * xo) * C + c;
  dlow[t] = (dout[b] + dout[b + C]) + (dout[b + (int64_t)W * C] + dout[b + (int64_t)W * C + C]);
}

__global__ __launch_bounds__(256) void k_up_concat_bwd_skip(const float* __restrict__ dout,
                                                            int64_t P, int C1, int C2,
                                                            float* __restrict__ dskip,
                                                            int accumulate) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P * C2) return;
  const int64_t q = t / C2;
  const int c = (int)(t % C2);
  const float g = dout[q * (C1 + C2) + C1 + c];
  dskip[t] = accumulate ? dskip[t] + g : g;
}

// Network input: channel 0 = grad * gscale, channel 1 = (model - mshift) * mscale,
// channels 2..7 zero; (N, H, W) FP32 planes -> (N, H, W, 8) bf16.
__global__ __launch_bounds__(256) void k_pack_input(const float* __restrict__ g,
                                                    const float* __restrict__ mdl, int64_t P,
                                                    const float* __restrict__ scales,
                                                    int64_t plane, uint16_t* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P) return;
  const int64_t n = t / plane;
  const float gs = scales[3 * n], ms = scales[3 * n + 1], mm = scales[3 * n + 2];
  const uint32_t lo = (uint32_t)f2bf(g[t] * gs) | ((uint32_t)f2bf((mdl[t] - mm) * ms) << 16);
  *reinterpret_cast<uint4*>(out + t * 8) = make_uint4(lo, 0u, 0u, 0u);
}

// 1x1 head to one channel: out[p] = sum_c x[p][c] * w[c] + b (FP32).
__global__ __launch_bounds__(256) void k_head_fwd(const uint16_t* __restrict__ x, int64_t P, int C,
                                                  const float* __restrict__ w,
                                                  const float* __restrict__ b,
                                                  float* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P) return;
  float s = 0.0f;
  for (int c = 0; c < C; ++c) s += bf2f(x[t * C + c]) * w[c];
  out[t] = s + b[0];
}

// Head gradient: dx[p][c] = dout[p] * w[c]; dw[c] += sum_p dout[p] x[p][c];
// db += sum_p dout[p].  One block per 256 pixels; thread c < C sums its channel.
__global__ __launch_bounds__(256) void k_head_bwd(const uint16_t* __restrict__ x,
                                                  const float* __restrict__ dout, int64_t P,
                                                  int C, const float* __restrict__ w,
                                                  float* __restrict__ dx, float* __restrict__ dw,
                                                  float* __restrict__ db) {
  const int64_t p0 = (int64_t)blockIdx.x * 256;
  const int64_t p1 = p0 + 256 < P ? p0 + 256 : P;
  for (int64_t e = threadIdx.x; e < (p1 - p0) * C; e += blockDim.x) {
    const int64_t p = p0 + e / C;
    const int c = (int)(e % C);
    dx[p * C + c] = dout[p] * w[c];
  }
  if (threadIdx.x < C) {
    const int c = threadIdx.x;
    float s = 0.0f;
    for (int64_t p = p0; p < p1; ++p) s += dout[p] * bf2f(x[p * C + c]);
    atomicAdd(dw + c, s);
  } else if (threadIdx.x == 255) {
    float s = 0.0f;
    for (int64_t p = p0; p < p1; ++p) s += dout[p];
    atomicAdd(db, s);
  }
}

// Masked mean squared error: loss += sum mask*(pred-target)^2 / count,
// dpred = 2*mask*(pred-target)/count.
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ pred,
                                             const float* __restrict__ target,
                                             const float* __restrict__ mask, int64_t n,
                                             float inv_count, float* __restrict__ loss,
                                             float* __restrict__ dpred) {
  __shared__ float red[256];
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float e = 0.0f;
  if (t < n) {
    const float mk = mask ? mask[t] : 1.0f;
    const float r = mk * (pred[t] - target[t]);
    dpred[t] = 2.0f * r * inv_count;
    e = r * r * inv_count;
  }
  red[threadIdx.x] = e;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

// AdamW (decoupled weight decay) on FP32 master weights; refreshes the bf16 copy.
__global__ __launch_bounds__(256) void k_adamw(float* __restrict__ p, const float* __restrict__ g,
                                               float* __restrict__ m, float* __restrict__ v,
                                               int64_t n, float lr, float b1, float b2, float eps,
                                               float wd, float bc1, float bc2,
                                               uint16_t* __restrict__ pb) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  float w = p[t];
  w -= lr * wd * w;
  const float gt = g[t];
  const float mt = b1 * m[t] + (1.0f - b1) * gt;
  const float vt = b2 * v[t] + (1.0f - b2) * gt * gt;
  m[t] = mt;
  v[t] = vt;
  w -= lr * (mt / bc1) / (sqrtf(vt / bc2) + eps);
  p[t] = w;
  if (pb) pb[t] = f2bf(w);
}

__global__ __launch_bounds__(256) void k_fill_f32(float* __restrict__ p, int64_t n, float v) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

}  // namespace

extern "C" {

int seis_unet_conv3x3(const uint16_t* x, int32_t N, int32_t H, int32_t W, int32_t Cin,
                      const uint16_t* w, const float* bias, int32_t Cout, int32_t act,
                      uint16_t* z, uint16_t* y, float* y_f32, void* stream) {
  SEIS_CHECK_ARG(N >= 1 && H >= 1 && W >= 1, "empty image batch");
  SEIS_CHECK_ARG(Cin >= 8 && Cin % 8 == 0 && Cin <= 4096, "Cin %d must be a multiple of 8", Cin);
  SEIS_CHECK_ARG(Cout >= 1 && Cout <= 4096, "Cout %d outside [1, 4096]", Cout);
  SEIS_CHECK_ARG(x && w && aligned16(x) && aligned16(w), "x and w must be 16-byte aligned");
  SEIS_CHECK_ARG((y != nullptr) != (y_f32 != nullptr), "exactly one of y (bf16) and y_f32");
  SEIS_CHECK_ARG(y || (!z && !act), "z and act apply to the bf16 output only");
  SEIS_CHECK_ARG((int64_t)N * H <= (int64_t)1 << 24, "too many image rows");
  hipStream_t s = (hipStream_t)stream;
  const int segs = (W + kCvPx - 1) / kCvPx;
  const bool wide = Cout > 32;
  const dim3 grid((unsigned)(segs * N * H), (unsigned)((Cout + (wide ? 63 : 31)) / (wide ? 64 : 32)));
  if (y) {
    if (wide) k_conv3x3<64, 0><<<grid, 256, 0, s>>>(x, N, H, W, Cin, w, bias, Cout, act, z, y, nullptr);
    else k_conv3x3<32, 0><<<grid, 256, 0, s>>>(x, N, H, W, Cin, w, bias, Cout, act, z, y, nullptr);
  } else {
    if (wide) k_conv3x3<64, 1><<<grid, 256, 0, s>>>(x, N, H, W, Cin, w, bias, Cout, 0, nullptr, nullptr, y_f32);
    else k_conv3x3<32, 1><<<grid, 256, 0, s>>>(x, N, H, W, Cin, w, bias, Cout, 0, nullptr, nullptr, y_f32);
  }
  SEIS_CHECK_LAUNCH();
  return 0;
}

int seis_unet_flip_weights(const uint16_t* w, int32_t Cout, int32_t Cin, uint16_t* wt,
                           void* stream) {
  SEIS_CHECK_ARG(w && wt && Cout >= 1 && Cin >= 1, "bad arguments");
  const int64_t n = (int64_t)Cout * 9 * Cin;
  k_flip_weights<<<blocks_for(n), 256, 0, (hipStream_t)stream>>>(w, Cout, Cin, wt);
  SEIS_CHECK_LAUNCH();
  return 0;
}

int seis_unet_conv3x3_wgrad(const uint16_t* x, const uint16_t* dz, int32_t N, int32_t H,
                            int32_t W, int32_t Cin, int32_t Cout, float* dw, float* db,
                            void* stream) {
  SEIS_CHECK_ARG(N >= 1 && H >= 1 && W >= 1, "empty image batch");
  SEIS_CHECK_ARG(Cin >= 8 && Cout >= 8 && Cin <= 4096 && Cout <= 4096 && Cin % 8 == 0 &&
                     Cout % 8 == 0,
                 "channel counts (%d, %d) must be multiples of 8", Cin, Cout);
  SEIS_CHECK_ARG(x && dz && dw, "null pointer argument");
  SEIS_CHECK_ARG(aligned16(x) && aligned16(dz), "x and dz must be 16-byte aligned");
  hipStream_t s = (hipStream_t)stream;
  const int tiles = ((Cin + kWgCIT - 1) / kWgCIT) * ((Cout + kWgCOT - 1) / kWgCOT);
  const int64_t nseg = (int64_t)N * H * ((W + kWgSeg - 1) / kWgSeg);
  // enough blocks to fill the chip, each summing many segments before its atomics
  int64_t gy = (4096 + tiles - 1) / tiles;
  if (gy > nseg) gy = nseg;
  if (gy < 1) gy = 1;
  k_conv3x3_wgrad<<<dim3((unsigned)tiles, (unsigned)gy), 256, 0, s>>>(x, dz, N, H, W, Cin, Cout,
                                                                     dw);
  SEIS_CHECK_LAUNCH();
  if (db) {
    const int64_t P = (int64_t)N * H * W;
    k_bias_grad<<<blocks_for(P), 256, 0, s>>>(dz, P, Cout, db);
    SEIS_CHECK_LAUNCH();
  }
  return 0;
}

int seis_unet_gelu_bwd(const float* dy, const uint16_t* z, int64_t n, uint16_t* dz, void* stream) {
  SEIS_CHECK_ARG(dy && z && dz && n >= 0, "bad arguments");
  if (n == 0) return 0;
  k_gelu_bwd<<<blocks_for(n), 256, 0, (hipStream_t)stream>>>(dy, z, n, dz);
  SEIS_CHECK_LAUNCH();
  return 0;
}

int seis_unet_to_bf16(const float* x, int64_t n, uint16_t* y, void* stream) {
  SEIS_CHECK_ARG(x && y && n >= 0, "bad arguments");
  if (n == 0) return 0;
  k_to_bf16<<<blocks_for(n), 256, 0, (hipStream_t)stream>>>(x, n, y);
  SEIS_CHECK_LAUNCH();
  return 0;
}

int seis_unet_avgpool2(const uint16_t* x, int32_t N, int32_t H, int32_t W, int32_t C, uint16_t* y,
                       void* stream) {
  SEIS_CHECK_ARG(x && y && N >= 1 && C >= 1 && H >= 2 && W >= 2 && H % 2 == 0 && W % 2 == 0,
                 "avgpool2 needs even H, W");
  const int64_t n = (int64_t)N * (H / 2) * (W / 2) * C;
  k_avgpool2<<<blocks_for(n), 256, 0, (hipStream_t)stream>>>(x, N, H, W, C, y);
  SEIS_CHECK_LAUNCH();
  return 0;
}

int seis_unet_avgpool2_bwd(const float* dy, int32_t N, int32_t H, int32_t W, int32_t C, float* dx,
                           int32_t accumulate, void* stream) {
  SEIS_CHECK_ARG(dy && dx && N >= 1 && C >= 1 && H % 2 == 0 && W % 2 == 0 && H >= 2 && W >= 2,
                 "avgpool2 needs even H, W");
  const int64_t n = (int64_t)N * H * W * C;
  k_avgpool2_bwd<<<blocks_for(n), 256, 0, (hipStream_t)stream>>>(dy, N, H, W, C, dx, accumulate);
  SEIS_CHECK_LAUNCH();
  return 0;
}

int seis_unet_up_concat(const uint16_t* low, const uint16_t* skip, int32_t N, int32_t H,
                        int32_t W, int32_t C1, int32_t C2, uint16_t* out, void* stream) {
  SEIS_CHECK_ARG(low && skip && out && N >= 1 && H % 2 == 0 && W % 2 == 0 && C1 >= 1 && C2 >= 1,
                 "bad arguments");
  const int64_t n = (int64_t)N * H * W * (C1 + C2);
  k_up_concat<<<blocks_for(n), 256, 0, (hipStream_t)stream>>>(low, skip, N, H, W, C1, C2, out);
  SEIS_CHECK_LAUNCH();
  return 0;
}

int seis_unet_up_concat_bwd(const float* dout, int32_t N, int32_t H, int32_t W, int32_t C1,
                            int32_t C2, float* dlow, float* dskip, int32_t accumulate_skip,
                            void* stream) {
  SEIS_CHECK_ARG(dout && dlow && dskip && N >= 1 && H % 2 == 0 && W % 2 == 0, "bad arguments");
  hipStream_t s = (hipStream_t)stream;
  const int64_t nl = (int64_t)N * (H / 2) * (W / 2) * C1;
  k_up_concat_bwd_low<<<blocks_for(nl), 256, 0, s>>>(dout, N, H, W, C1, C2, dlow);
  SEIS_CHECK_LAUNCH();
  const int64_t P = (int64_t)N * H * W;
  k_up_concat_bwd_skip<<<blocks_for(P * C2), 256, 0, s>>>(dout, P, C1, C2, dskip,
                                                          accumulate_skip);
  SEIS_CHECK_LAUNCH();
  return 0;
}

int seis_unet_pack_input(const float* grad, const float* model, int32_t N, int32_t H, int32_t W,
                         const float* scales, uint16_t* out, void* stream) {
  SEIS_CHECK_ARG(grad && model && scales && out && aligned16(out), "bad arguments");
  const int64_t P = (int64_t)N * H * W;
  k_pack_input<<<blocks_for(P), 256, 0, (hipStream_t)stream>>>(grad, model, P, scales,
                                                              (int64_t)H * W, out);
  SEIS_CHECK_LAUNCH();
  return 0;
}

int seis_unet_head(const uint16_t* x, int64_t P, int32_t C, const float* w, const float* b,
                   float* out, void* stream) {
  SEIS_CHECK_ARG(x && w && b && out && C >= 1, "bad arguments");
  k_head_fwd<<<blocks_for(P), 256, 0, (hipStream_t)stream>>>(x, P, C, w, b, out);
  SEIS_CHECK_LAUNCH();
  return 0;
}

int seis_unet_head_bwd(const uint16_t* x, const float* dout, int64_t P, int32_t C, const float* w,
                       float* dx, float* dw, float* db, void* stream) {
  SEIS_CHECK_ARG(x && dout && w && dx && dw && db && C >= 1 && C <= 255, "bad arguments");
  k_head_bwd<<<blocks_for(P), 256, 0, (hipStream_t)stream>>>(x, dout, P, C, w, dx, dw, db);
  SEIS_CHECK_LAUNCH();
  return 0;
}

int seis_unet_mse(const float* pred, const float* target, const float* mask, int64_t n,
                  float inv_count, float* loss, float* dpred, void* stream) {
  SEIS_CHECK_ARG(pred && target && loss && dpred && n >= 1, "bad arguments");
  hipStream_t s = (hipStream_t)stream;
  k_fill_f32<<<1, 256, 0, s>>>(loss, 1, 0.0f);
  k_mse<<<blocks_for(n), 256, 0, s>>>(pred, target, mask, n, inv_count, loss, dpred);
  SEIS_CHECK_LAUNCH();
  return 0;
}

int seis_unet_adamw(float* param, const float* grad, float* m, float* v, int64_t n, float lr,
                    float beta1, float beta2, float eps, float weight_decay, int32_t step,
                    uint16_t* param_bf16, void* stream) {
  SEIS_CHECK_ARG(param && grad && m && v && n >= 1 && step >= 1, "bad arguments");
  const float bc1 = 1.0f - powf(beta1, (float)step);
  const float bc2 = 1.0f - powf(beta2, (float)step);
  k_adamw<<<blocks_for(n), 256, 0, (hipStream_t)stream>>>(param, grad, m, v, n, lr, beta1, beta2,
                                                          eps, weight_decay, bc1, bc2, param_bf16);
  SEIS_CHECK_LAUNCH();
  return 0;
}

int seis_unet_fill(float* p, int64_t n, float v, void* stream) {
  SEIS_CHECK_ARG(p && n >= 0, "bad arguments");
  if (n == 0) return 0;
  k_fill_f32<<<blocks_for(n), 256, 0, (hipStream_t)stream>>>(p, n, v);
  SEIS_CHECK_LAUNCH();
  return 0;
}

}  // extern "C"
