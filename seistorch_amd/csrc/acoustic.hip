// Constant-density acoustic PML kernels for gfx950.
//
// Forward step (seistorch/equations2d/acoustic.py:73-86):
//   y = A*(2/dt^2*u - B*u_prev + c^2*L u),  A = 1/(dt^-2 + d/dt),
//   B = dt^-2 - d/dt, L = zero-padded 5-point Laplacian / h^2.
// Loop order per step (seistorch/rnn.py:178-202): step, raw add of the
// wavelet sample at each shot's source cell, record the receivers.
//
// Adjoint (exact transpose of that loop, mu = A*lambda):
//   mu^k  = A*(s^k + 2/dt^2*mu^{k+1} - B*mu^{k+2} + L(c^2*mu^{k+1}))
//   dJ/dc = sum_n 2c*(L u^n)*mu^{n+1}
// The forward field is rebuilt backwards in time on the undamped interior
// [npml+1, n-npml-1) from the 1-cell ring stored at index npml during the
// forward (seistorch/equations2d/utils.py:50-90 stores the same ring), so the
// gradient is exact on that interior and left at zero in the damped strips.
//
// Work decomposition: a 64x4 thread tile covers 64 contiguous x by 4 z rows;
// each thread advances its cell for a group of S shots so that the model,
// the damping and (in the adjoint) the per-group gradient are read/written
// once per S shot-updates.  Neighbour reads hit L1/L2.
#include "common.h"
#include "seistorch_amd.h"
#include "illum.h"
#include "tiles.h"

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

using seis::Grid;
using seis::lap_taps;
using seis::acoustic_update;
using seis::lap5;
using namespace seis::tiles;

namespace {

constexpr int kTX = 64;
constexpr int kTZ = 4;



// ---------------------------------------------------------------- per step

__global__ __launch_bounds__(256) void k_step(const float* __restrict__ vp,
                                              const float* __restrict__ d,
                                              const float* __restrict__ h1,
                                              const float* __restrict__ h2,
                                              float* __restrict__ y, Grid g,
                                              seis::TimeConsts k) {
  const int x = blockIdx.x * kTX + threadIdx.x;
  const int z = blockIdx.y * kTZ + threadIdx.y;
  if (x >= g.nx || z >= g.nz) return;
  const int64_t off = (int64_t)blockIdx.z * g.plane;
  const int i = z * g.nx + x;
  const float c = vp[i];
  const float dd = d[i];
  const float A = 1.0f / (k.idt2 + dd * k.idt);
  const float Bc = k.idt2 - dd * k.idt;
  const float L = lap5(h1 + off, z, x, i, g, k.ih2);
  y[off + i] = acoustic_update(A, Bc, k.two_idt2, h1[off + i], h2[off + i], c * c, L);
}

// d/d(h1), d/d(h2) of one step: gh1 = 2/dt^2*A*gy + L(c^2*A*gy) + pass,
// gh2 = -B*A*gy.
__global__ __launch_bounds__(256) void k_step_vjp_fields(
    const float* __restrict__ vp, const float* __restrict__ d, const float* __restrict__ gy,
    const float* __restrict__ pass, float* __restrict__ gh1, float* __restrict__ gh2, Grid g,
    seis::TimeConsts k) {
  const int x = blockIdx.x * kTX + threadIdx.x;
  const int z = blockIdx.y * kTZ + threadIdx.y;
  if (x >= g.nx || z >= g.nz) return;
  const int64_t off = (int64_t)blockIdx.z * g.plane;
  const float* gb = gy + off;
  const int i = z * g.nx + x;
  auto q = [&](int j) {
    const float c = vp[j];
    const float A = 1.0f / (k.idt2 + d[j] * k.idt);
    return c * c * A * gb[j];
  };
  const float dd = d[i];
  const float A = 1.0f / (k.idt2 + dd * k.idt);
  const float Bc = k.idt2 - dd * k.idt;
  const float qc = q(i);
  const float qzm = z > 0 ? q(i - g.nx) : 0.0f;
  const float qzp = z < g.nz - 1 ? q(i + g.nx) : 0.0f;
  const float qxm = x > 0 ? q(i - 1) : 0.0f;
  const float qxp = x < g.nx - 1 ? q(i + 1) : 0.0f;
  const float Lq = lap_taps(qzm, qxm, qc, qxp, qzp, k.ih2);
  const float ag = A * gb[i];
  float v = k.two_idt2 * ag + Lq;
  if (pass) v += pass[off + i];
  gh1[off + i] = v;
  gh2[off + i] = -Bc * ag;
}

// d/dc of one step summed over shots: 2c*A*(L h1)*gy.
__global__ __launch_bounds__(256) void k_step_vjp_model(
    const float* __restrict__ vp, const float* __restrict__ d, const float* __restrict__ h1,
    const float* __restrict__ gy, float* __restrict__ gvp, int B, Grid g, seis::TimeConsts k) {
  const int x = blockIdx.x * kTX + threadIdx.x;
  const int z = blockIdx.y * kTZ + threadIdx.y;
  if (x >= g.nx || z >= g.nz) return;
  const int i = z * g.nx + x;
  const float c = vp[i];
  const float A = 1.0f / (k.idt2 + d[i] * k.idt);
  float acc = 0.0f;
  for (int b = 0; b < B; ++b) {
    const int64_t off = (int64_t)b * g.plane;
    acc += lap5(h1 + off, z, x, i, g, k.ih2) * gy[off + i];
  }
  gvp[i] = 2.0f * c * A * acc;
}

// ------------------------------------------------------------ fused loop

// u^{n+1} = step(u^n, u^{n-1}) + w[n] at the source cell, written in place
// over u^{n-1} (each cell of u^{n-1} is read only by the thread that
// overwrites it).  Same Laplacian and update helpers as the per-step kernel.
template <int S, int RZ, int VX>
__device__ __forceinline__ void fwd_tile(
    int tx, int tzi, int grp, const float* __restrict__ vp, const float* __restrict__ d,
    const float* __restrict__ u, float* __restrict__ up, const int32_t* __restrict__ src,
    const float* __restrict__ wn, int B, Grid g, const seis::TimeConsts& k) {
  const int x0 = (tx * kLanes + (int)(threadIdx.x % kLanes)) * VX;
  const int zb = (tzi * kRowsOfThreads + (int)(threadIdx.x / kLanes)) * RZ;
  if (x0 >= g.nx || zb >= g.nz) return;
  const int b0 = grp * S;
  const int nb = min(S, B - b0);
  const float w = *wn;
  float A[RZ][VX], Bc[RZ][VX], c2[RZ][VX];
#pragma unroll
  for (int r = 0; r < RZ; ++r) {
    const int z = min(zb + r, g.nz - 1);
    float c[VX], dd[VX];
    vload<VX>(vp + z * g.nx + x0, c);
    vload<VX>(d + z * g.nx + x0, dd);
#pragma unroll
    for (int e = 0; e < VX; ++e) {
      A[r][e] = 1.0f / (k.idt2 + dd[e] * k.idt);
      Bc[r][e] = k.idt2 - dd[e] * k.idt;
      c2[r][e] = c[e] * c[e];
    }
  }
  for (int s = 0; s < nb; ++s) {
    const int b = b0 + s;
    const float* f = u + (int64_t)b * g.plane;
    float* o = up + (int64_t)b * g.plane;
    const int sz = src[2 * b], sx = src[2 * b + 1];
    float col[RZ + 2][VX];
#pragma unroll
    for (int r = 0; r < RZ + 2; ++r) {
      const int z = zb - 1 + r;
      if (z >= 0 && z < g.nz) vload<VX>(f + z * g.nx + x0, col[r]);
      else vzero<VX>(col[r]);
    }
#pragma unroll
    for (int r = 0; r < RZ; ++r) {
      const int z = zb + r;
      if (z < g.nz) {
        const int i = z * g.nx + x0;
        const float xm = x0 > 0 ? f[i - 1] : 0.0f;
        const float xp = x0 + VX < g.nx ? f[i + VX] : 0.0f;
        float prev[VX], out[VX];
        vload<VX>(o + i, prev);
#pragma unroll
        for (int e = 0; e < VX; ++e) {
          const float left = e == 0 ? xm : col[r + 1][e - 1];
          const float right = e == VX - 1 ? xp : col[r + 1][e + 1];
          const float L = lap_taps(col[r][e], left, col[r + 1][e], right, col[r + 2][e], k.ih2);
          float v = acoustic_update(A[r][e], Bc[r][e], k.two_idt2, col[r + 1][e], prev[e],
                                    c2[r][e], L);
          if (z == sz && x0 + e == sx) v += w;
          out[e] = v;
        }
        vstore<VX>(o + i, out);
      }
    }
  }
}

// u^{n+1} of one cell from u^n (f) and u^{n-1} (fp) at (z, x), with the zero
// padding of the Laplacian and the source sample; the same operands and
// helpers as fwd_tile, so the value is bit-identical to the one the cell's
// own tile computes.
__device__ __forceinline__ float step_cell(const float* __restrict__ vp,
                                           const float* __restrict__ d,
                                           const float* __restrict__ f,
                                           const float* __restrict__ fp, int z, int x, Grid g,
                                           const seis::TimeConsts& k, int sz, int sx, float w) {
  const int i = z * g.nx + x;
  const float A = 1.0f / (k.idt2 + d[i] * k.idt);
  const float Bc = k.idt2 - d[i] * k.idt;
  const float c = vp[i];
  const float zm = z > 0 ? f[i - g.nx] : 0.0f;
  const float zp = z < g.nz - 1 ? f[i + g.nx] : 0.0f;
  const float xm = x > 0 ? f[i - 1] : 0.0f;
  const float xp = x < g.nx - 1 ? f[i + 1] : 0.0f;
  const float L = lap_taps(zm, xm, f[i], xp, zp, k.ih2);
  float v = acoustic_update(A, Bc, k.two_idt2, f[i], fp[i], c * c, L);
  if (z == sz && x == sx) v += w;
  return v;
}

// Two forward steps per launch: u^{n+1} = step(u^n, u^{n-1}) + w[n] at the
// source, u^{n+2} = step(u^{n+1}, u^n) + w[n+1].  Reads ua = u^n, ub =
// u^{n-1}; writes uc = u^{n+1}, ud = u^{n+2} (four distinct buffers, since
// neighbouring tiles still read u^n and u^{n-1} around this one).  A wave is
// one row of 64 lanes; each lane computes u^{n+1} on its RZ rows plus the row
// above and below (recomputed from u^n and u^{n-1} with the same operands as
// the neighbouring thread row, so bit-identical) and takes its x-neighbours
// from the adjacent lanes by cross-lane shuffles.  Tiles overlap by one lane
// vector on each side: lanes 0 and 63 only supply u^{n+1} to lanes 1 and 62,
// and the 62 inner lanes write both steps (kFwd2Out * VX columns per tile).
// No LDS, no barriers; each field crosses HBM once per two steps.
constexpr int kFwd2Out = kLanes - 2;

//
// ILL (the source illumination of rnn.py:204-205): each inner lane also sums
// the squares of both new levels over its S shots and adds them to its
// shot group's plane of ipart, (ipart + sum_s (u^{n+1})^2) + sum_s (u^{n+2})^2
// (the host adds the groups' planes at the end).
template <int S, int RZ, int VX, bool ILL>
__device__ __forceinline__ void fwd2_tile(
    const float* __restrict__ vp, const float* __restrict__ d, const float* __restrict__ ua,
    const float* __restrict__ ub, float* __restrict__ uc, float* __restrict__ ud,
    const int32_t* __restrict__ src, const float* __restrict__ wn, int B, const Grid& g,
    const seis::TimeConsts& k, int tx, int tzi, int grp, float* __restrict__ ipart, bool two) {
  const int lane = (int)(threadIdx.x % kLanes);
  const int x0 = (tx * kFwd2Out + lane - 1) * VX;  // lane 0 starts one vector left
  const int zb = (tzi * kRowsOfThreads + (int)(threadIdx.x / kLanes)) * RZ;
  if (zb >= g.nz) return;  // uniform over the wave
  const bool xin = x0 >= 0 && x0 < g.nx;  // nx is a multiple of VX
  const bool inner = lane > 0 && lane < kLanes - 1;
  const float w0 = wn[0], w1 = two ? wn[1] : 0.0f;  // wn[1] exists only for two steps
  // coefficients on the rows zb-1 .. zb+RZ (index q = z - zb + 1)
  float A[RZ + 2][VX], Bc[RZ + 2][VX], c2[RZ + 2][VX];
#pragma unroll
  for (int q = 0; q < RZ + 2; ++q) {
    const int z = min(max(zb - 1 + q, 0), g.nz - 1);
    float c[VX], dd[VX];
    if (xin) {
      vload<VX>(vp + z * g.nx + x0, c);
      vload<VX>(d + z * g.nx + x0, dd);
    } else {
      vzero<VX>(c);
      vzero<VX>(dd);
    }
#pragma unroll
    for (int e = 0; e < VX; ++e) {
      A[q][e] = 1.0f / (k.idt2 + dd[e] * k.idt);
      Bc[q][e] = k.idt2 - dd[e] * k.idt;
      c2[q][e] = c[e] * c[e];
    }
  }
  const int b0 = grp * S;
  const int nb = min(S, B - b0);
  float il1[ILL ? RZ : 1][VX], il2[ILL ? RZ : 1][VX];
  if constexpr (ILL) {
#pragma unroll
    for (int r = 0; r < RZ; ++r) {
      vzero<VX>(il1[r]);
      vzero<VX>(il2[r]);
    }
  }
  for (int s = 0; s < nb; ++s) {
    const int b = b0 + s;
    const int64_t off = (int64_t)b * g.plane;
    const float* f = ua + off;
    const float* fp = ub + off;
    const int sz = src[2 * b], sx = src[2 * b + 1];
    // u^n on rows zb-2 .. zb+RZ+1 (index p = z - zb + 2), zero outside
    float un[RZ + 4][VX];
#pragma unroll
    for (int p = 0; p < RZ + 4; ++p) {
      const int z = zb - 2 + p;
      if (xin && z >= 0 && z < g.nz) vload<VX>(f + z * g.nx + x0, un[p]);
      else vzero<VX>(un[p]);
    }
    // step 1 on rows zb-1 .. zb+RZ
    float u1[RZ + 2][VX];
#pragma unroll
    for (int q = 0; q < RZ + 2; ++q) {
      const int z = zb - 1 + q;
      const bool zok = z >= 0 && z < g.nz;
      float xm = __shfl_up(un[q + 1][VX - 1], 1, kLanes);
      float xp = __shfl_down(un[q + 1][0], 1, kLanes);
      // lanes 0 and 63 pass on only their inner-side vector element, which
      // needs the outer neighbour only when a vector is a single cell
      if (VX == 1) {
        if (lane == 0) xm = (xin && zok && x0 > 0) ? f[z * g.nx + x0 - 1] : 0.0f;
        if (lane == kLanes - 1)
          xp = (xin && zok && x0 + VX < g.nx) ? f[z * g.nx + x0 + VX] : 0.0f;
      }
      if (xin && zok) {
        float prev[VX];
        vload<VX>(fp + z * g.nx + x0, prev);
#pragma unroll
        for (int e = 0; e < VX; ++e) {
          const float left = e == 0 ? xm : un[q + 1][e - 1];
          const float right = e == VX - 1 ? xp : un[q + 1][e + 1];
          const float L = lap_taps(un[q][e], left, un[q + 1][e], right, un[q + 2][e], k.ih2);
          float v = acoustic_update(A[q][e], Bc[q][e], k.two_idt2, un[q + 1][e], prev[e],
                                    c2[q][e], L);
          if (z == sz && x0 + e == sx) v += w0;
          u1[q][e] = v;
        }
      } else {
        vzero<VX>(u1[q]);  // outside the grid: the Laplacian's zero padding
      }
    }
    if (inner && xin) {
#pragma unroll
      for (int r = 0; r < RZ; ++r) {
        const int z = zb + r;
        if (z < g.nz) vstore<VX>(uc + off + z * g.nx + x0, u1[r + 1]);
        if constexpr (ILL) {
#pragma unroll
          for (int e = 0; e < VX; ++e) il1[r][e] = il1[r][e] + u1[r + 1][e] * u1[r + 1][e];
        }
      }
    }
    if (!two) continue;  // uniform over the block (one instance)
    // step 2 on rows zb .. zb+RZ-1 (inner lanes)
#pragma unroll
    for (int r = 0; r < RZ; ++r) {
      const int z = zb + r;
      const int q = r + 1;
      const float xm = __shfl_up(u1[q][VX - 1], 1, kLanes);
      const float xp = __shfl_down(u1[q][0], 1, kLanes);
      if (!inner || !xin || z >= g.nz) continue;
      float out[VX];
#pragma unroll
      for (int e = 0; e < VX; ++e) {
        const float left = e == 0 ? xm : u1[q][e - 1];
        const float right = e == VX - 1 ? xp : u1[q][e + 1];
        const float L = lap_taps(u1[q - 1][e], left, u1[q][e], right, u1[q + 1][e], k.ih2);
        float v = acoustic_update(A[q][e], Bc[q][e], k.two_idt2, u1[q][e], un[q + 1][e],
                                  c2[q][e], L);
        if (z == sz && x0 + e == sx) v += w1;
        out[e] = v;
        if constexpr (ILL) il2[r][e] = il2[r][e] + v * v;
      }
      vstore<VX>(ud + off + z * g.nx + x0, out);
    }
  }
  if constexpr (ILL) {
    if (!inner || !xin) return;
    float* P = ipart + (int64_t)grp * g.plane;
#pragma unroll
    for (int r = 0; r < RZ; ++r) {
      const int z = zb + r;
      if (z >= g.nz) continue;
      float v[VX];
      vload<VX>(P + z * g.nx + x0, v);
#pragma unroll
      for (int e = 0; e < VX; ++e) v[e] = (v[e] + il1[r][e]) + il2[r][e];
      vstore<VX>(P + z * g.nx + x0, v);
    }
  }
}

template <int S, int RZ, int VX>
__global__ __launch_bounds__(kBlock) void k_fwd2(
    const float* __restrict__ vp, const float* __restrict__ d, const float* __restrict__ ua,
    const float* __restrict__ ub, float* __restrict__ uc, float* __restrict__ ud,
    const int32_t* __restrict__ src, const float* __restrict__ wn, int B, Grid g,
    seis::TimeConsts k, TileMap tm) {
  int tx, tzi, grp;
  if (!tile_of(tm, tx, tzi, grp)) return;  // uniform over the block
  fwd2_tile<S, RZ, VX, false>(vp, d, ua, ub, uc, ud, src, wn, B, g, k, tx, tzi, grp, nullptr,
                              true);
}

// k_fwd2 that also accumulates the source illumination into ipart
template <int S, int RZ, int VX>
__global__ __launch_bounds__(kBlock) void k_fwd2_illum(
    const float* __restrict__ vp, const float* __restrict__ d, const float* __restrict__ ua,
    const float* __restrict__ ub, float* __restrict__ uc, float* __restrict__ ud,
    const int32_t* __restrict__ src, const float* __restrict__ wn, int B, Grid g,
    seis::TimeConsts k, TileMap tm, float* __restrict__ ipart) {
  int tx, tzi, grp;
  if (!tile_of(tm, tx, tzi, grp)) return;  // uniform over the block
  fwd2_tile<S, RZ, VX, true>(vp, d, ua, ub, uc, ud, src, wn, B, g, k, tx, tzi, grp, ipart, true);
}

// Receivers of the new field into one record row, and its ring into one slot.
// Additional point sources of one step (source encoding, codingfwi.py:184-334:
// many sources with their own encoded wavelets in one field; each adds its
// sample at its cell, source.py:54-55 with the time loop's dt = 1):
// u[b, z, x] += sign * wav[t*nt + n] for each entry t = (b, z, x) of esrc.
// Sources that share a cell add.
__global__ __launch_bounds__(256) void k_extra_src(float* __restrict__ u,
                                                   const int32_t* __restrict__ esrc,
                                                   const float* __restrict__ ewav, int64_t nt,
                                                   int64_t n, int64_t n_extra, float sign,
                                                   Grid g) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_extra) return;
  const int64_t cell = (int64_t)esrc[3 * t] * g.plane + (int64_t)esrc[3 * t + 1] * g.nx +
                       esrc[3 * t + 2];
  atomicAdd(u + cell, sign * ewav[t * nt + n]);
}

int extra_sources(float* u, const int32_t* esrc, const float* ewav, int64_t nt, int64_t n,
                  int64_t n_extra, float sign, const Grid& g, hipStream_t s) {
  if (n_extra <= 0) return 0;
  k_extra_src<<<(unsigned)((n_extra + 255) / 256), 256, 0, s>>>(u, esrc, ewav, nt, n, n_extra,
                                                                  sign, g);
  SEIS_CHECK_LAUNCH();
  return 0;
}

__global__ __launch_bounds__(256) void k_record(const float* __restrict__ u,
                                                const int64_t* __restrict__ rec, int64_t R,
                                                float* __restrict__ rec_row,
                                                float* __restrict__ ring_slot, int64_t ring_len,
                                                int B, Grid g) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < R) {
    rec_row[t] = u[rec[t]];
    return;
  }
  t -= R;
  if (ring_slot && t < (int64_t)B * ring_len) {
    const int64_t b = t / ring_len;
    const int64_t j = t - b * ring_len;
    ring_slot[t] = u[b * g.plane + ring_cell(j, g)];
  }
}

// One reverse iteration k (see file header): mu^k over the whole grid into
// m2 (was mu^{k+2}); on the interior the gradient term 2c*L(u^{k-1})*mu^k and
// the rebuilt u^{k-2} = (2/dt^2 u^{k-1} + c^2 L u^{k-1} - (u^k - w[k-1]*src)/A)/B into
// uc (was u^k).  recon: 0 none, 1 rebuild, 2 write zeros (u^0).  Tiles as
// k_fwd; u^{k-1} and u^k are only touched on vectors that reach the interior
// and on the interior's rows and ring rows.  Vectors straddling the interior
// edge write their non-interior cells back unchanged.
//
// The reconstruction inverts the forward's own FP32 update u^k = A*(2/dt^2
// u^{k-1} - B u^{k-2} + c^2 L u^{k-1}) with true divisions by its A and B: the
// shortcut 2u^{k-1} - u^k + c^2 dt^2 L assumes A*B == 1, which FP32 misses by
// ~1 ulp and which then drifts coherently over the reversed steps.
template <int S, int RZ, int VX>
__device__ __forceinline__ void bwd_tile(
    int tx, int tzi, int grp, const float* __restrict__ vp, const float* __restrict__ d,
    const float* __restrict__ um, float* __restrict__ uc, const float* __restrict__ m1,
    float* __restrict__ m2, float* __restrict__ gpart, const int32_t* __restrict__ src,
    const float* __restrict__ wn, int recon, int B, Grid g, const seis::TimeConsts& k) {
  const int x0 = (tx * kLanes + (int)(threadIdx.x % kLanes)) * VX;
  const int zb = (tzi * kRowsOfThreads + (int)(threadIdx.x / kLanes)) * RZ;
  if (x0 >= g.nx || zb >= g.nz) return;
  const int b0 = grp * S;
  const int nb = min(S, B - b0);
  const float w = *wn;
  bool xin[VX];
  bool anyx = false;
#pragma unroll
  for (int e = 0; e < VX; ++e) {
    xin[e] = x0 + e > g.npml && x0 + e < g.nx - 1 - g.npml;
    anyx = anyx || xin[e];
  }
  // rows of u^{k-1} that are read: the interior's and its ring rows
  const int urow0 = max(g.ztop, 0), urow1 = g.nz - 1 - g.npml;
  // model: c^2 on the column rows zb-1 .. zb+RZ and at the vector's two x-ends
  float c2col[RZ + 2][VX], c2L[RZ], c2R[RZ], cc[RZ][VX], A[RZ][VX], Bc[RZ][VX];
#pragma unroll
  for (int r = 0; r < RZ + 2; ++r) {
    const int z = zb - 1 + r;
    if (z >= 0 && z < g.nz) vload<VX>(vp + z * g.nx + x0, c2col[r]);
    else vzero<VX>(c2col[r]);
#pragma unroll
    for (int e = 0; e < VX; ++e) c2col[r][e] *= c2col[r][e];
  }
#pragma unroll
  for (int r = 0; r < RZ; ++r) {
    const int z = min(zb + r, g.nz - 1);
    const int i = z * g.nx + x0;
    const float cl = x0 > 0 ? vp[i - 1] : 0.0f;
    const float cr = x0 + VX < g.nx ? vp[i + VX] : 0.0f;
    c2L[r] = cl * cl;
    c2R[r] = cr * cr;
    float dd[VX];
    vload<VX>(vp + i, cc[r]);
    vload<VX>(d + i, dd);
#pragma unroll
    for (int e = 0; e < VX; ++e) {
      A[r][e] = 1.0f / (k.idt2 + dd[e] * k.idt);
      Bc[r][e] = k.idt2 - dd[e] * k.idt;
    }
  }
  float gsum[RZ][VX];
#pragma unroll
  for (int r = 0; r < RZ; ++r) vzero<VX>(gsum[r]);
  for (int s = 0; s < nb; ++s) {
    const int b = b0 + s;
    const int64_t off = (int64_t)b * g.plane;
    const float* M = m1 + off;
    float* M2 = m2 + off;
    const float* U = um + off;
    float* UC = uc + off;
    const int sz = src[2 * b], sx = src[2 * b + 1];
    float mcol[RZ + 2][VX], ucol[RZ + 2][VX];
#pragma unroll
    for (int r = 0; r < RZ + 2; ++r) {
      const int z = zb - 1 + r;
      if (z >= 0 && z < g.nz) vload<VX>(M + z * g.nx + x0, mcol[r]);
      else vzero<VX>(mcol[r]);
      if (anyx && z >= urow0 && z <= urow1) vload<VX>(U + z * g.nx + x0, ucol[r]);
      else vzero<VX>(ucol[r]);
    }
#pragma unroll
    for (int r = 0; r < RZ; ++r) {
      const int z = zb + r;
      if (z < g.nz) {
        const int i = z * g.nx + x0;
        const float qL = c2L[r] * (x0 > 0 ? M[i - 1] : 0.0f);
        const float qR = c2R[r] * (x0 + VX < g.nx ? M[i + VX] : 0.0f);
        float mu[VX];
        vload<VX>(M2 + i, mu);
#pragma unroll
        for (int e = 0; e < VX; ++e) {
          const float left = e == 0 ? qL : c2col[r + 1][e - 1] * mcol[r + 1][e - 1];
          const float right = e == VX - 1 ? qR : c2col[r + 1][e + 1] * mcol[r + 1][e + 1];
          const float Lm = lap_taps(c2col[r][e] * mcol[r][e], left, c2col[r + 1][e] * mcol[r + 1][e],
                                    right, c2col[r + 2][e] * mcol[r + 2][e], k.ih2);
          mu[e] = acoustic_update(A[r][e], Bc[r][e], k.two_idt2, mcol[r + 1][e], mu[e], 1.0f, Lm);
        }
        vstore<VX>(M2 + i, mu);
        if (anyx && interior_row(z, g)) {
          const float uL = x0 > 0 ? U[i - 1] : 0.0f;
          const float uR = x0 + VX < g.nx ? U[i + VX] : 0.0f;
          float uv[VX];
          vload<VX>(UC + i, uv);
#pragma unroll
          for (int e = 0; e < VX; ++e) {
            if (xin[e]) {
              const float left = e == 0 ? uL : ucol[r + 1][e - 1];
              const float right = e == VX - 1 ? uR : ucol[r + 1][e + 1];
              const float Lu = lap_taps(ucol[r][e], left, ucol[r + 1][e], right, ucol[r + 2][e],
                                        k.ih2);
              gsum[r][e] += Lu * mu[e];
              if (recon == 1) {
                float un = uv[e];
                if (z == sz && x0 + e == sx) un -= w;
                uv[e] = ((k.two_idt2 * ucol[r + 1][e] + c2col[r + 1][e] * Lu) - un / A[r][e]) /
                        Bc[r][e];
              } else if (recon == 2) {
                uv[e] = 0.0f;
              }
            }
          }
          if (recon != 0) vstore<VX>(UC + i, uv);
        }
      }
    }
  }
  float* G = gpart + (int64_t)grp * g.plane;
#pragma unroll
  for (int r = 0; r < RZ; ++r) {
    const int z = zb + r;
    if (anyx && interior_row(z, g)) {
      float gv[VX];
      vload<VX>(G + z * g.nx + x0, gv);
#pragma unroll
      for (int e = 0; e < VX; ++e)
        if (xin[e]) gv[e] += 2.0f * cc[r][e] * gsum[r][e];
      vstore<VX>(G + z * g.nx + x0, gv);
    }
  }
}

template <int S, int RZ, int VX>
__global__ __launch_bounds__(kBlock) void k_fwd(
    const float* __restrict__ vp, const float* __restrict__ d, const float* __restrict__ u,
    float* __restrict__ up, const int32_t* __restrict__ src, const float* __restrict__ wn, int B,
    Grid g, seis::TimeConsts k, TileMap tm) {
  int tx, tzi, grp;
  if (!tile_of(tm, tx, tzi, grp)) return;
  fwd_tile<S, RZ, VX>(tx, tzi, grp, vp, d, u, up, src, wn, B, g, k);
}

template <int S, int RZ, int VX>
__global__ __launch_bounds__(kBlock) void k_bwd(
    const float* __restrict__ vp, const float* __restrict__ d, const float* __restrict__ um,
    float* __restrict__ uc, const float* __restrict__ m1, float* __restrict__ m2,
    float* __restrict__ gpart, const int32_t* __restrict__ src, const float* __restrict__ wn,
    int recon, int B, Grid g, seis::TimeConsts k, TileMap tm) {
  int tx, tzi, grp;
  if (!tile_of(tm, tx, tzi, grp)) return;
  bwd_tile<S, RZ, VX>(tx, tzi, grp, vp, d, um, uc, m1, m2, gpart, src, wn, recon, B, g, k);
}

// Two reverse iterations per launch, kk and kk-1 (both rebuilding, kk >= 4),
// over rotating buffers (neighbouring tiles still read the inputs around
// this one): reads ua = u^{kk-1}, ub = u^{kk}, ma = mu^{kk+1}, mb = mu^{kk+2};
// writes mc = mu^{kk} (with the receivers' A*s^{kk}), md = mu^{kk-1} (its
// receivers are added afterwards by k_bwd_post), uc = u^{kk-2} (interior
// rebuilt, ring from ring_slot = the ring of u^{kk-2}, elsewhere 0) and ud =
// u^{kk-3} on the interior; adds both iterations' gradient terms to gpart.
// The adjoint source of iteration kk enters mu^{kk} inside the launch (the
// second iteration reads it): rmap (nz x nx) holds each cell's receiver
// index within a shot (-1 if none), shot b's cotangent at rg[b*Rs + r]; all
// shots share the receiver cells.  Layout as k_fwd2: a lane computes the
// first iteration on its RZ rows and the rows above and below, x-neighbours
// come by shuffles, lanes 0 and 63 load their outer neighbours and only the
// 62 inner lanes write.  Every value uses bwd_tile's operands and order, so
// fields and gradients equal two k_bwd iterations bit for bit, except the
// gradient at receiver cells, where the receivers' gradient shares are added
// after both iterations' terms instead of between them.
constexpr int kBwd2MaxVX = 4;

template <int S, int RZ, int VX>
__global__ __launch_bounds__(kBlock) void k_bwd2(
    const float* __restrict__ vp, const float* __restrict__ d, const float* __restrict__ ua,
    const float* __restrict__ ub, float* __restrict__ uc, float* __restrict__ ud,
    const float* __restrict__ ma, const float* __restrict__ mb, float* __restrict__ mc,
    float* __restrict__ md, float* __restrict__ gpart, const int32_t* __restrict__ src,
    const float* __restrict__ wn, const int32_t* __restrict__ rmap, const float* __restrict__ rg,
    int Rs, const float* __restrict__ ring_slot, int64_t ring_len, int B, Grid g,
    seis::TimeConsts k, TileMap tm) {
  int tx, tzi, grp;
  if (!tile_of(tm, tx, tzi, grp)) return;  // uniform over the block
  const int lane = (int)(threadIdx.x % kLanes);
  const int x0 = (tx * kFwd2Out + lane - 1) * VX;
  const int zb = (tzi * kRowsOfThreads + (int)(threadIdx.x / kLanes)) * RZ;
  if (zb >= g.nz) return;  // uniform over the wave
  const bool xin = x0 >= 0 && x0 < g.nx;  // nx is a multiple of VX
  const bool inner = lane > 0 && lane < kLanes - 1;
  const float wA = wn[1], wB = wn[0];  // w[kk-1], w[kk-2]
  const int ur0 = max(g.ztop, 0), ur1 = g.nz - 1 - g.npml;  // rows of u that are read
  bool xint[VX];
  bool anyi = false, anyu = false;
#pragma unroll
  for (int e = 0; e < VX; ++e) {
    const int x = x0 + e;
    xint[e] = xin && x > g.npml && x < g.nx - 1 - g.npml;
    anyi = anyi || xint[e];
    anyu = anyu || (xin && x >= g.npml && x <= g.nx - 1 - g.npml);
  }
  // c^2 on rows zb-2 .. zb+RZ+1 (index p = z - zb + 2); A, Bc and the
  // receiver map on rows zb-1 .. zb+RZ (index q = z - zb + 1)
  float c2[RZ + 4][VX], A[RZ + 2][VX], Bc[RZ + 2][VX], cc[RZ][VX];
  int rm[RZ + 2][VX];
#pragma unroll
  for (int p = 0; p < RZ + 4; ++p) {
    const int z = zb - 2 + p;
    if (xin && z >= 0 && z < g.nz) vload<VX>(vp + z * g.nx + x0, c2[p]);
    else vzero<VX>(c2[p]);
#pragma unroll
    for (int e = 0; e < VX; ++e) c2[p][e] *= c2[p][e];
  }
#pragma unroll
  for (int q = 0; q < RZ + 2; ++q) {
    const int z = min(max(zb - 1 + q, 0), g.nz - 1);
    const int i = z * g.nx + (xin ? x0 : 0);
    float dd[VX];
    vload<VX>(d + i, dd);
#pragma unroll
    for (int e = 0; e < VX; ++e) {
      A[q][e] = 1.0f / (k.idt2 + dd[e] * k.idt);
      Bc[q][e] = k.idt2 - dd[e] * k.idt;
      rm[q][e] = (rmap && xin && zb - 1 + q >= 0 && zb - 1 + q < g.nz) ? rmap[i + e] : -1;
    }
  }
#pragma unroll
  for (int r = 0; r < RZ; ++r) {
    const int z = min(zb + r, g.nz - 1);
    if (xin) vload<VX>(vp + z * g.nx + x0, cc[r]);
    else vzero<VX>(cc[r]);
  }
  float gA[RZ][VX], gB[RZ][VX];
#pragma unroll
  for (int r = 0; r < RZ; ++r) {
    vzero<VX>(gA[r]);
    vzero<VX>(gB[r]);
  }
  const int b0 = grp * S;
  const int nb = min(S, B - b0);
  for (int s = 0; s < nb; ++s) {
    const int b = b0 + s;
    const int64_t off = (int64_t)b * g.plane;
    const int sz = src[2 * b], sx = src[2 * b + 1];
    // mu^{kk+1} and u^{kk-1} on rows p
    float M1[RZ + 4][VX], U1[RZ + 4][VX];
#pragma unroll
    for (int p = 0; p < RZ + 4; ++p) {
      const int z = zb - 2 + p;
      if (xin && z >= 0 && z < g.nz) vload<VX>(ma + off + z * g.nx + x0, M1[p]);
      else vzero<VX>(M1[p]);
      if (anyu && z >= ur0 && z <= ur1) vload<VX>(ua + off + z * g.nx + x0, U1[p]);
      else vzero<VX>(U1[p]);
    }
    // iteration kk on rows q: mu^{kk} (MU, with the receivers' source; the
    // own rows also without it, for the gradient) and u^{kk-2} (UA)
    float MU[RZ + 2][VX], MUn[RZ][VX], UA[RZ + 2][VX];
#pragma unroll
    for (int q = 0; q < RZ + 2; ++q) {
      const int z = zb - 1 + q;
      const bool zok = z >= 0 && z < g.nz;
      float pl = __shfl_up(c2[q + 1][VX - 1] * M1[q + 1][VX - 1], 1, kLanes);
      float pr = __shfl_down(c2[q + 1][0] * M1[q + 1][0], 1, kLanes);
      float ul = __shfl_up(U1[q + 1][VX - 1], 1, kLanes);
      float ur = __shfl_down(U1[q + 1][0], 1, kLanes);
      const bool urow = z >= ur0 && z <= ur1;
      // lanes 0 and 63 pass on only their inner-side vector element, which
      // needs the outer neighbour only when a vector is a single cell
      if (VX == 1 && lane == 0) {
        const bool ok = xin && zok && x0 > 0;
        const int j = z * g.nx + x0 - 1;
        const float cl = ok ? vp[j] : 0.0f;
        pl = cl * cl * (ok ? ma[off + j] : 0.0f);
        ul = (ok && urow) ? ua[off + j] : 0.0f;
      }
      if (VX == 1 && lane == kLanes - 1) {
        const bool ok = xin && zok && x0 + VX < g.nx;
        const int j = z * g.nx + x0 + VX;
        const float cr = ok ? vp[j] : 0.0f;
        pr = cr * cr * (ok ? ma[off + j] : 0.0f);
        ur = (ok && urow) ? ua[off + j] : 0.0f;
      }
      if (!(xin && zok)) {
        vzero<VX>(MU[q]);
        vzero<VX>(UA[q]);
        if (q >= 1 && q <= RZ) vzero<VX>(MUn[q - 1]);
        continue;
      }
      const int i = z * g.nx + x0;
      float m2v[VX];
      vload<VX>(mb + off + i, m2v);
      const bool zi = interior_row(z, g);
      float ukk[VX];
      if (anyi && zi) vload<VX>(ub + off + i, ukk);
      else vzero<VX>(ukk);
#pragma unroll
      for (int e = 0; e < VX; ++e) {
        const float left = e == 0 ? pl : c2[q + 1][e - 1] * M1[q + 1][e - 1];
        const float right = e == VX - 1 ? pr : c2[q + 1][e + 1] * M1[q + 1][e + 1];
        const float Lm = lap_taps(c2[q][e] * M1[q][e], left, c2[q + 1][e] * M1[q + 1][e], right,
                                  c2[q + 2][e] * M1[q + 2][e], k.ih2);
        float mu = acoustic_update(A[q][e], Bc[q][e], k.two_idt2, M1[q + 1][e], m2v[e], 1.0f, Lm);
        if (q >= 1 && q <= RZ) MUn[q - 1][e] = mu;
        if (rm[q][e] >= 0) mu = mu + A[q][e] * rg[(int64_t)b * Rs + rm[q][e]];
        MU[q][e] = mu;
        float uo = 0.0f;
        if (zi && xint[e]) {
          const float l = e == 0 ? ul : U1[q + 1][e - 1];
          const float r = e == VX - 1 ? ur : U1[q + 1][e + 1];
          const float Lu = lap_taps(U1[q][e], l, U1[q + 1][e], r, U1[q + 2][e], k.ih2);
          if (q >= 1 && q <= RZ) gA[q - 1][e] += Lu * MUn[q - 1][e];
          float un = ukk[e];
          if (z == sz && x0 + e == sx) un -= wA;
          uo = ((k.two_idt2 * U1[q + 1][e] + c2[q + 1][e] * Lu) - un / A[q][e]) / Bc[q][e];
        } else if (urow) {
          const int64_t jr = ring_index(z, x0 + e, g);
          if (jr >= 0) uo = ring_slot[(int64_t)b * ring_len + jr];
        }
        UA[q][e] = uo;
      }
    }
    // iteration kk-1 on the own rows (inner lanes)
#pragma unroll
    for (int r = 0; r < RZ; ++r) {
      const int q = r + 1;
      const int z = zb + r;
      const float pl = __shfl_up(c2[q + 1][VX - 1] * MU[q][VX - 1], 1, kLanes);
      const float pr = __shfl_down(c2[q + 1][0] * MU[q][0], 1, kLanes);
      const float ul = __shfl_up(UA[q][VX - 1], 1, kLanes);
      const float ur = __shfl_down(UA[q][0], 1, kLanes);
      if (!inner || !xin || z >= g.nz) continue;
      const int i = z * g.nx + x0;
      const bool zi = interior_row(z, g);
      float mo[VX], uo[VX];
#pragma unroll
      for (int e = 0; e < VX; ++e) {
        const float left = e == 0 ? pl : c2[q + 1][e - 1] * MU[q][e - 1];
        const float right = e == VX - 1 ? pr : c2[q + 1][e + 1] * MU[q][e + 1];
        const float Lm = lap_taps(c2[q][e] * MU[q - 1][e], left, c2[q + 1][e] * MU[q][e], right,
                                  c2[q + 2][e] * MU[q + 1][e], k.ih2);
        mo[e] = acoustic_update(A[q][e], Bc[q][e], k.two_idt2, MU[q][e], M1[q + 1][e], 1.0f, Lm);
        uo[e] = 0.0f;
        if (zi && xint[e]) {
          const float l = e == 0 ? ul : UA[q][e - 1];
          const float rr = e == VX - 1 ? ur : UA[q][e + 1];
          const float Lu = lap_taps(UA[q - 1][e], l, UA[q][e], rr, UA[q + 1][e], k.ih2);
          gB[r][e] += Lu * mo[e];
          float un = U1[q + 1][e];
          if (z == sz && x0 + e == sx) un -= wB;
          uo[e] = ((k.two_idt2 * UA[q][e] + c2[q + 1][e] * Lu) - un / A[q][e]) / Bc[q][e];
        }
      }
      vstore<VX>(mc + off + i, MU[q]);
      vstore<VX>(md + off + i, mo);
      vstore<VX>(uc + off + i, UA[q]);
      vstore<VX>(ud + off + i, uo);
    }
  }
  if (!inner || !anyi) return;
  float* G = gpart + (int64_t)grp * g.plane;
#pragma unroll
  for (int r = 0; r < RZ; ++r) {
    const int z = zb + r;
    if (interior_row(z, g)) {
      float gv[VX];
      vload<VX>(G + z * g.nx + x0, gv);
#pragma unroll
      for (int e = 0; e < VX; ++e) {
        if (!xint[e]) continue;
        gv[e] += 2.0f * cc[r][e] * gA[r][e];
        gv[e] += 2.0f * cc[r][e] * gB[r][e];
      }
      vstore<VX>(G + z * g.nx + x0, gv);
    }
  }
}

// After k_bwd: inject the adjoint source A*s^k at the receivers of mu^k (and
// its gradient share 2c*L(u^{k-1})*A*s^k), restore the ring of u^{k-2}.
// m2 == nullptr: the source is already in mu^k (k_bwd2), only the gradient
// share is added.
__global__ __launch_bounds__(256) void k_bwd_post(
    const float* __restrict__ vp, const float* __restrict__ d, const float* __restrict__ um,
    float* __restrict__ uc, float* __restrict__ m2, float* __restrict__ gpart,
    const int64_t* __restrict__ rec, const float* __restrict__ rgrad, int64_t R,
    const float* __restrict__ ring_slot, int64_t ring_len, int recon, int B, int S, Grid g,
    seis::TimeConsts k) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < R) {
    const int64_t f = rec[t];
    const int64_t b = f / g.plane;
    const int i = (int)(f - b * g.plane);
    const int z = i / g.nx;
    const int x = i - z * g.nx;
    const float A = 1.0f / (k.idt2 + d[i] * k.idt);
    const float as = A * rgrad[t];
    if (m2) atomicAdd(m2 + f, as);
    if (interior(z, x, g)) {
      const float Lu = lap5(um + b * g.plane, z, x, i, g, k.ih2);
      atomicAdd(gpart + (b / S) * g.plane + i, 2.0f * vp[i] * Lu * as);
    }
    return;
  }
  t -= R;
  if (recon != 0 && t < (int64_t)B * ring_len) {
    const int64_t b = t / ring_len;
    const int64_t j = t - b * ring_len;
    uc[b * g.plane + ring_cell(j, g)] = recon == 1 ? ring_slot[t] : 0.0f;
  }
}

__global__ __launch_bounds__(256) void k_fill(float* __restrict__ p, int64_t n, float v) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

// Two equal-size buffers copied in one launch (snapshot store / reload).
__global__ __launch_bounds__(256) void k_copy2(const float* __restrict__ a, float* __restrict__ ao,
                                               const float* __restrict__ b, float* __restrict__ bo,
                                               int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) {
    ao[t] = a[t];
    bo[t] = b[t];
  }
}

__global__ __launch_bounds__(256) void k_sum_groups(const float* __restrict__ gpart, int G,
                                                    float* __restrict__ out, int64_t plane) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= plane) return;
  float s = 0.0f;
  for (int q = 0; q < G; ++q) s += gpart[(int64_t)q * plane + t];
  out[t] = s;
}

inline int copy2(const float* a, float* ao, const float* b, float* bo, int64_t n, hipStream_t s) {
  k_copy2<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(a, ao, b, bo, n);
  SEIS_CHECK_LAUNCH();
  return 0;
}

// ------------------------------------------------- uneven multi-instance batch
//
// Many independent problems (own grid, shots, nt, dt, h, receivers) advance in
// one launch per time step.  Instance q is one row of an int64 table on the
// device (columns below).  The rows are sorted by nt, longest first, and every
// instance starts at global step 0 (the reverse loop at its own last step), so
// the instances still running at global step j are a prefix of the table and
// each launch covers exactly their tiles: global tile t belongs to the last
// row whose first tile is <= t.  Offsets index the concatenated arenas.
enum BatchCol {
  kCNz = 0, kCNx, kCB, kCNt, kCR, kCNtx, kCNtz,
  kCTile0,     // first global tile of the row (prefix over the sorted rows)
  kCAux0,      // first global index of its receiver + ring work
  kCField,     // (B, nz, nx) fields in the field arenas
  kCModel,     // (nz, nx) in the model arenas
  kCSrc,       // shot index into the (sum B, 2) source arena
  kCRec,       // receiver index into the rec arena (flat into its own fields)
  kCOut,       // (nt, R) records
  kCWav,       // wavelet samples
  kCRing,      // (nt-1, B, ring_len) rings
  kCRingLen,
  kCGpart,     // (groups, nz, nx) gradient planes
  kCIdt2, kCIdt, kCTwoIdt2, kCIh2,  // FP32 time constants as bit patterns
  kCTile2,     // first global tile of the row in the two-step forward's tiling
  kBatchCols
};
static_assert(kBatchCols == SEIS_BATCH_COLS, "batch table width");

// Last row of the active prefix [0, n_active) whose column `col` is <= t.
__device__ __forceinline__ int find_row(const int64_t* __restrict__ tab, int n_active, int col,
                                        int64_t t) {
  int lo = 0, hi = n_active - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tab[(int64_t)mid * kBatchCols + col] <= t) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ Grid row_grid(const int64_t* r, int npml) {
  return seis::make_grid(r[kCNz], r[kCNx], npml, false);
}

__device__ __forceinline__ seis::TimeConsts row_consts(const int64_t* r) {
  seis::TimeConsts c;
  c.idt2 = __int_as_float((int)r[kCIdt2]);
  c.idt = __int_as_float((int)r[kCIdt]);
  c.two_idt2 = __int_as_float((int)r[kCTwoIdt2]);
  c.ih2 = __int_as_float((int)r[kCIh2]);
  c.dt2 = 0.0f;
  return c;
}

// u^m of an instance lives in field arena a iff m + nt is even (so u^nt ends
// in a, as in the single-instance loop); mu^m lives in a iff m + nt + 1 is even.
__host__ __device__ __forceinline__ bool in_a(int64_t m, int64_t nt) {
  return ((m + nt) & 1) == 0;
}

// The two-step forward keeps four levels in four arenas (a, c, d, b): u^m of
// an instance lives in arena (m - nt) mod 4 of that order, so u^nt ends in a
// and u^{nt-1} in b, where the one-step loop leaves them.
__host__ __device__ __forceinline__ float* level_arena(int64_t m, int64_t nt, float* fa,
                                                       float* fb, float* fc, float* fd) {
  switch ((int)(((m - nt) % 4 + 4) % 4)) {
    case 0: return fa;
    case 1: return fc;
    case 2: return fd;
    default: return fb;
  }
}

// Row of global tile t (XCD-aware order over the active tiles) and its local
// tile coordinates.
__device__ __forceinline__ const int64_t* batch_tile(const int64_t* __restrict__ tab,
                                                     int n_active, int64_t tiles, int per,
                                                     int& tx, int& tzi, int& grp) {
  const int w = blockIdx.x;
  const int64_t t = (int64_t)(w % kXcds) * per + w / kXcds;
  if (t >= tiles) return nullptr;
  const int64_t* r = tab + (int64_t)find_row(tab, n_active, kCTile0, t) * kBatchCols;
  const int64_t local = t - r[kCTile0];
  const int64_t ntx = r[kCNtx], ntz = r[kCNtz];
  tx = (int)(local % ntx);
  tzi = (int)((local / ntx) % ntz);
  grp = (int)(local / (ntx * ntz));
  return r;
}

template <int S, int RZ, int VX>
__global__ __launch_bounds__(kBlock) void k_fwd_batch(
    const int64_t* __restrict__ tab, int n_active, int64_t tiles, int per, int npml,
    const float* __restrict__ vp, const float* __restrict__ d, float* __restrict__ fa,
    float* __restrict__ fb, const int32_t* __restrict__ src, const float* __restrict__ wav,
    int64_t n) {
  int tx, tzi, grp;
  const int64_t* r = batch_tile(tab, n_active, tiles, per, tx, tzi, grp);
  if (!r) return;
  const int64_t fo = r[kCField];
  const bool a = in_a(n, r[kCNt]);
  fwd_tile<S, RZ, VX>(tx, tzi, grp, vp + r[kCModel], d + r[kCModel], (a ? fa : fb) + fo,
                      (a ? fb : fa) + fo, src + 2 * r[kCSrc], wav + r[kCWav] + n, (int)r[kCB],
                      row_grid(r, npml), row_consts(r));
}

// Global steps n and n+1 in one launch (fwd2_tile): every active row (nt > n)
// writes u^{n+1}, and u^{n+2} where nt > n+1; tiles of kFwd2Out lane vectors
// in x, numbered by kCTile2.
template <int S, int RZ, int VX>
__global__ __launch_bounds__(kBlock) void k_fwd2_batch(
    const int64_t* __restrict__ tab, int n_active, int64_t tiles, int per, int npml,
    const float* __restrict__ vp, const float* __restrict__ d, float* __restrict__ fa,
    float* __restrict__ fb, float* __restrict__ fc, float* __restrict__ fd,
    const int32_t* __restrict__ src, const float* __restrict__ wav, int64_t n) {
  const int w = blockIdx.x;
  const int64_t t = (int64_t)(w % kXcds) * per + w / kXcds;
  if (t >= tiles) return;
  const int64_t* r = tab + (int64_t)find_row(tab, n_active, kCTile2, t) * kBatchCols;
  const int64_t local = t - r[kCTile2];
  const int64_t nx = r[kCNx], ntz = r[kCNtz], nt = r[kCNt];
  const int64_t ntx2 = (nx + kFwd2Out * VX - 1) / (kFwd2Out * VX);
  const int tx = (int)(local % ntx2);
  const int tzi = (int)((local / ntx2) % ntz);
  const int grp = (int)(local / (ntx2 * ntz));
  const int64_t fo = r[kCField];
  fwd2_tile<S, RZ, VX, false>(
      vp + r[kCModel], d + r[kCModel], level_arena(n, nt, fa, fb, fc, fd) + fo,
      level_arena(n - 1, nt, fa, fb, fc, fd) + fo, level_arena(n + 1, nt, fa, fb, fc, fd) + fo,
      level_arena(n + 2, nt, fa, fb, fc, fd) + fo, src + 2 * r[kCSrc], wav + r[kCWav] + n,
      (int)r[kCB], row_grid(r, npml), row_consts(r), tx, tzi, grp, nullptr, n + 2 <= nt);
}

// Receivers of u^{n+1} into records row n, and its ring into slot n (fc null:
// the one-step loop's two arenas; else the two-step forward's four).
__global__ __launch_bounds__(256) void k_record_batch(
    const int64_t* __restrict__ tab, int n_active, int64_t work, int npml,
    float* __restrict__ fa, float* __restrict__ fb, float* __restrict__ fc,
    float* __restrict__ fd, const int64_t* __restrict__ rec, float* __restrict__ out,
    float* __restrict__ ring, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= work) return;
  const int64_t* r = tab + (int64_t)find_row(tab, n_active, kCAux0, t) * kBatchCols;
  const int64_t local = t - r[kCAux0];
  const int64_t R = r[kCR], nt = r[kCNt];
  const float* u = (fc ? level_arena(n + 1, nt, fa, fb, fc, fd)
                       : (in_a(n + 1, nt) ? fa : fb)) + r[kCField];
  if (local < R) {
    out[r[kCOut] + n * R + local] = u[rec[r[kCRec] + local]];
    return;
  }
  const int64_t j0 = local - R;
  const int64_t L = r[kCRingLen];
  if (ring && n + 1 < nt && j0 < r[kCB] * L) {
    const int64_t b = j0 / L;
    const Grid g = row_grid(r, npml);
    ring[r[kCRing] + n * r[kCB] * L + j0] = u[b * g.plane + seis::ring_cell(j0 - b * L, g)];
  }
}

// Reverse global iteration j: each active row runs its own iteration kk = nt - j.
template <int S, int RZ, int VX>
__global__ __launch_bounds__(kBlock) void k_bwd_batch(
    const int64_t* __restrict__ tab, int n_active, int64_t tiles, int per, int npml,
    const float* __restrict__ vp, const float* __restrict__ d, float* __restrict__ fa,
    float* __restrict__ fb, float* __restrict__ ma, float* __restrict__ mb,
    float* __restrict__ gpart, const int32_t* __restrict__ src, const float* __restrict__ wav,
    int64_t j) {
  int tx, tzi, grp;
  const int64_t* r = batch_tile(tab, n_active, tiles, per, tx, tzi, grp);
  if (!r) return;
  const int64_t nt = r[kCNt], kk = nt - j, fo = r[kCField];
  const int recon = kk >= 3 ? 1 : (kk == 2 ? 2 : 0);
  bwd_tile<S, RZ, VX>(tx, tzi, grp, vp + r[kCModel], d + r[kCModel],
                      (in_a(kk - 1, nt) ? fa : fb) + fo, (in_a(kk, nt) ? fa : fb) + fo,
                      (in_a(kk + 2, nt) ? ma : mb) + fo, (in_a(kk + 3, nt) ? ma : mb) + fo,
                      gpart + r[kCGpart], src + 2 * r[kCSrc], wav + r[kCWav] + (kk - 1), recon,
                      (int)r[kCB], row_grid(r, npml), row_consts(r));
}

// k_bwd_post for every active row.
__global__ __launch_bounds__(256) void k_post_batch(
    const int64_t* __restrict__ tab, int n_active, int64_t work, int npml, int S,
    const float* __restrict__ vp, const float* __restrict__ d, float* __restrict__ fa,
    float* __restrict__ fb, float* __restrict__ ma, float* __restrict__ mb,
    float* __restrict__ gpart, const int64_t* __restrict__ rec, const float* __restrict__ rgrad,
    const float* __restrict__ ring, int64_t j) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= work) return;
  const int64_t* r = tab + (int64_t)find_row(tab, n_active, kCAux0, t) * kBatchCols;
  const int64_t local = t - r[kCAux0];
  const int64_t R = r[kCR], nt = r[kCNt], kk = nt - j, fo = r[kCField], mo = r[kCModel];
  const Grid g = row_grid(r, npml);
  if (local < R) {
    const seis::TimeConsts k = row_consts(r);
    const float* X = (in_a(kk - 1, nt) ? fa : fb) + fo;
    float* M2 = (in_a(kk + 3, nt) ? ma : mb) + fo;
    const int64_t f = rec[r[kCRec] + local];
    const int64_t b = f / g.plane;
    const int i = (int)(f - b * g.plane);
    const int z = i / g.nx, x = i - z * g.nx;
    const float A = 1.0f / (k.idt2 + d[mo + i] * k.idt);
    const float as = A * rgrad[r[kCOut] + (kk - 1) * R + local];
    atomicAdd(M2 + f, as);
    if (seis::interior(z, x, g)) {
      const float Lu = lap5(X + b * g.plane, z, x, i, g, k.ih2);
      atomicAdd(gpart + r[kCGpart] + (b / S) * g.plane + i, 2.0f * vp[mo + i] * Lu * as);
    }
    return;
  }
  const int recon = kk >= 3 ? 1 : (kk == 2 ? 2 : 0);
  const int64_t j0 = local - R;
  const int64_t L = r[kCRingLen];
  if (recon != 0 && j0 < r[kCB] * L) {
    float* Y = (in_a(kk, nt) ? fa : fb) + fo;
    const int64_t b = j0 / L;
    Y[b * g.plane + seis::ring_cell(j0 - b * L, g)] =
        recon == 1 ? ring[r[kCRing] + (kk - 3) * r[kCB] * L + j0] : 0.0f;
  }
}

// Instantiated launch shapes: shots per thread S, rows per thread RZ, cells
// per lane VX.
#define SEIS_RZ_VX(K, S_)                                                              \
  case S_ * 256 + 1 * 16 + 1: K<S_, 1, 1><<<grid, kBlock, 0, s>>>(a...); break;        \
  case S_ * 256 + 2 * 16 + 1: K<S_, 2, 1><<<grid, kBlock, 0, s>>>(a...); break;        \
  case S_ * 256 + 4 * 16 + 1: K<S_, 4, 1><<<grid, kBlock, 0, s>>>(a...); break;        \
  case S_ * 256 + 1 * 16 + 2: K<S_, 1, 2><<<grid, kBlock, 0, s>>>(a...); break;        \
  case S_ * 256 + 2 * 16 + 2: K<S_, 2, 2><<<grid, kBlock, 0, s>>>(a...); break;        \
  case S_ * 256 + 4 * 16 + 2: K<S_, 4, 2><<<grid, kBlock, 0, s>>>(a...); break;        \
  case S_ * 256 + 1 * 16 + 4: K<S_, 1, 4><<<grid, kBlock, 0, s>>>(a...); break;        \
  case S_ * 256 + 2 * 16 + 4: K<S_, 2, 4><<<grid, kBlock, 0, s>>>(a...); break;        \
  case S_ * 256 + 4 * 16 + 4: K<S_, 4, 4><<<grid, kBlock, 0, s>>>(a...); break;      

inline bool valid_shape(int S, int RZ, int VX) {
  return (S == 1 || S == 2 || S == 4 || S == 8) && (RZ == 1 || RZ == 2 || RZ == 4) &&
         (VX == 1 || VX == 2 || VX == 4);
}

#define SEIS_DEFINE_LAUNCH(NAME, K)                                                   \
  template <typename... Args>                                                        \
  int NAME(int S, int RZ, int VX, int64_t blocks, hipStream_t s, Args... a) {      \
    const dim3 grid((unsigned)blocks);                                               \
    switch (S * 256 + RZ * 16 + VX) {                                                \
      SEIS_RZ_VX(K, 1)                                                               \
      SEIS_RZ_VX(K, 2)                                                               \
      SEIS_RZ_VX(K, 4)                                                               \
      SEIS_RZ_VX(K, 8)                                                               \
      default:                                                                       \
        seis::set_error("unsupported launch shape (%d, %d, %d)", S, RZ, VX);         \
        return 1;                                                                    \
    }                                                                                \
    SEIS_CHECK_LAUNCH();                                                             \
    return 0;                                                                        \
  }

SEIS_DEFINE_LAUNCH(launch_fwd, k_fwd)
SEIS_DEFINE_LAUNCH(launch_bwd, k_bwd)
SEIS_DEFINE_LAUNCH(launch_fwd2, k_fwd2)
SEIS_DEFINE_LAUNCH(launch_fwd2_illum, k_fwd2_illum)
SEIS_DEFINE_LAUNCH(launch_bwd2, k_bwd2)
SEIS_DEFINE_LAUNCH(launch_fwd_batch, k_fwd_batch)
SEIS_DEFINE_LAUNCH(launch_fwd2_batch, k_fwd2_batch)
SEIS_DEFINE_LAUNCH(launch_bwd_batch, k_bwd_batch)

inline dim3 tile_grid(const Grid& g, int zdim) {
  return dim3((g.nx + kTX - 1) / kTX, (g.nz + kTZ - 1) / kTZ, zdim);
}

inline int fill(float* p, int64_t n, float v, hipStream_t s) {
  if (n <= 0) return 0;
  k_fill<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, n, v);
  SEIS_CHECK_LAUNCH();
  return 0;
}

inline bool aligned(const void* p, int bytes) {
  return p == nullptr || (reinterpret_cast<uintptr_t>(p) % (uintptr_t)bytes) == 0;
}

int check_grid(int64_t B, int64_t nz, int64_t nx, float dt, float h) {
  SEIS_CHECK_ARG(B >= 1 && B <= 65535, "batch size %lld outside [1, 65535]", (long long)B);
  SEIS_CHECK_ARG(nz >= 1 && nx >= 1, "empty grid %lld x %lld", (long long)nz, (long long)nx);
  SEIS_CHECK_ARG(nz * nx < (int64_t)1 << 31, "grid %lld x %lld too large", (long long)nz,
                 (long long)nx);
  SEIS_CHECK_ARG(dt > 0.0f && h > 0.0f, "dt and h must be positive (dt=%g, h=%g)", dt, h);
  return 0;
}

// Forward loop with two steps per launch (k_fwd2) over four rotating
// buffers; a last single step (k_fwd) when nt is odd.  Records, rings and
// snapshots are taken from the buffers that hold each time level, exactly as
// in the one-step loop; u^{nt} and u^{nt-1} end in u_a and u_b.
int forward_two_step(const float* vp, const float* d, const float* wavelet, int64_t nt,
                     const int32_t* src_zx, const int64_t* rec_flat, int64_t R, float* rec_out,
                     float* ring, float* snaps, int32_t snap_every, float* illum, float* ipart,
                     float* u_a, float* u_b,
                     float* u_c, float* u_d, int B, const Grid& g, const seis::TimeConsts& k,
                     int S, int RZ, int VX, const TileMap& tm, int64_t ring_len, int64_t aux,
                     hipStream_t s) {
  const int64_t field = (int64_t)B * g.plane;
  // A launch maps (cur, prev, n1, n2) -> (n2, n1, cur, prev), a period-4
  // rotation, and an odd tail swaps cur and prev; the starting assignment is
  // chosen so that u^{nt} ends in u_a and u^{nt-1} in u_b without copies.
  const int64_t q = nt / 2;
  float* want_cur = (nt % 2 == 0) ? u_a : u_b;
  float* want_prev = (nt % 2 == 0) ? u_b : u_a;
  float* P[4];
  switch (q % 4) {
    case 0: P[0] = want_cur; P[1] = want_prev; P[2] = u_c; P[3] = u_d; break;
    case 1: P[3] = want_cur; P[2] = want_prev; P[0] = u_c; P[1] = u_d; break;
    case 2: P[1] = want_cur; P[0] = want_prev; P[2] = u_c; P[3] = u_d; break;
    default: P[2] = want_cur; P[3] = want_prev; P[0] = u_c; P[1] = u_d; break;
  }
  float* cur = P[0];   // u^n
  float* prev = P[1];  // u^{n-1}
  float* n1 = P[2];    // u^{n+1}
  float* n2 = P[3];    // u^{n+2}
  if (int e = fill(cur, field, 0.0f, s)) return e;
  if (int e = fill(prev, field, 0.0f, s)) return e;
  // k_fwd2 tiles: kFwd2Out lane vectors of output per tile in x
  TileMap tm2 = tm;
  tm2.ntx = (g.nx + kFwd2Out * VX - 1) / (kFwd2Out * VX);
  tm2.per = (int)(((int64_t)tm2.ntx * tm2.ntz * tm2.groups + kXcds - 1) / kXcds);
  // with ipart the two-step launches accumulate the illumination themselves
  auto record = [&](const float* u, int64_t n, bool ill) -> int {  // u = u^{n+1}
    if (ill)
      if (int e = illuminate(u, B, g.plane, illum, s)) return e;
    float* slot = (ring && n + 1 < nt) ? ring + n * B * ring_len : nullptr;
    const int64_t work = R + (slot ? B * ring_len : 0);
    if (work > 0 && aux > 0) {
      k_record<<<(unsigned)((work + 255) / 256), 256, 0, s>>>(
          u, rec_flat, R, R ? rec_out + n * R : nullptr, slot, ring_len, B, g);
      SEIS_CHECK_LAUNCH();
    }
    return 0;
  };
  auto snapshot = [&](const float* um, const float* um1, int64_t m) -> int {
    if (snaps && m >= snap_every && m % snap_every == 0 && m <= nt - 2) {
      float* slot_p = snaps + (m / snap_every - 1) * 2 * field;
      return copy2(um, slot_p, um1, slot_p + field, field, s);
    }
    return 0;
  };
  const bool fold = illum && ipart;
  if (fold)
    if (int e = fill(ipart, (int64_t)tm.groups * g.plane, 0.0f, s)) return e;
  int64_t n = 0;
  for (; n + 1 < nt; n += 2) {
    SEIS_PROF(SEIS_KERNEL_ACOUSTIC_FWD2, s);
    if (fold) {
      if (int e = launch_fwd2_illum(S, RZ, VX, (int64_t)tm2.per * kXcds, s, vp, d, cur, prev, n1,
                                    n2, src_zx, wavelet + n, B, g, k, tm2, ipart))
        return e;
    } else if (int e = launch_fwd2(S, RZ, VX, (int64_t)tm2.per * kXcds, s, vp, d, cur, prev, n1,
                                   n2, src_zx, wavelet + n, B, g, k, tm2)) {
      return e;
    }
    SEIS_PROF(SEIS_KERNEL_ACOUSTIC_FWD2, s);
    if (int e = record(n1, n, !fold)) return e;
    if (int e = record(n2, n + 1, !fold)) return e;
    if (int e = snapshot(cur, n1, n)) return e;
    if (int e = snapshot(n1, n2, n + 1)) return e;
    float* t0 = cur; float* t1 = prev;
    cur = n2; prev = n1; n1 = t0; n2 = t1;
  }
  if (fold) {  // the groups' planes, then the odd tail's level below
    k_sum_groups<<<(unsigned)((g.plane + 255) / 256), 256, 0, s>>>(ipart, tm.groups, illum,
                                                                    g.plane);
    SEIS_CHECK_LAUNCH();
  }
  if (n < nt) {  // odd nt: one more single step, u^{nt} written over u^{nt-2}
    SEIS_PROF(SEIS_KERNEL_ACOUSTIC_FWD, s);
    if (int e = launch_fwd(S, RZ, VX, (int64_t)tm.per * kXcds, s, vp, d, cur, prev, src_zx,
                           wavelet + n, B, g, k, tm))
      return e;
    SEIS_PROF(SEIS_KERNEL_ACOUSTIC_FWD, s);
    if (int e = record(prev, n, true)) return e;
    if (int e = snapshot(cur, prev, n)) return e;
    float* t = cur; cur = prev; prev = t;
  }
  // by the choice of P, u^{nt} is now in u_a and u^{nt-1} in u_b
  return 0;
}


// ------------------------------------------------- batch plan (host side)

// Float bit pattern as a table entry.
inline int64_t fbits(float v) {
  int32_t b;
  std::memcpy(&b, &v, sizeof b);
  return (int64_t)b;
}

struct BatchPlan {
  int n = 0;
  std::vector<int64_t> tab;        // sorted rows, kBatchCols each
  std::vector<int64_t> nt_sorted;  // nt of the sorted rows
  std::vector<int64_t> tile_end;   // tiles of the first k sorted rows, k = 0..n
  std::vector<int64_t> tile2_end;  // the same in the two-step forward's tiling
  std::vector<int64_t> aux_end;    // receiver + ring work of the first k sorted rows
  std::vector<int64_t> snap_off;   // per sorted row: offset into the snapshot arena
  std::vector<int64_t> groups;     // per sorted row: shot groups
  int64_t nt_max = 0, fields = 0, gparts = 0;
};

// Validates every instance and lays out the arenas in the caller's order;
// rows are then sorted by nt (longest first, stable).
int make_plan(BatchPlan& p, int64_t n, const int64_t* shape, const float* dt_h, int npml, int S,
              int RZ, int VX, int32_t snap_every, bool need_ring) {
  SEIS_CHECK_ARG(n >= 1 && n <= (1 << 20), "instance count %lld outside [1, 2^20]",
                 (long long)n);
  SEIS_CHECK_ARG(shape && dt_h, "null host pointer argument");
  SEIS_CHECK_ARG(valid_shape(S, RZ, VX),
                 "launch shape (shots_per_group %d, rows_per_thread %d, vec_x %d) not in "
                 "{1,2,4,8} x {1,2,4} x {1,2,4}",
                 S, RZ, VX);
  p.n = (int)n;
  std::vector<int64_t> rows((size_t)n * kBatchCols, 0);
  int64_t o_model = 0, o_wav = 0, o_src = 0, o_rec = 0, o_out = 0, o_ring = 0, o_snap = 0,
          o_field = 0, o_gpart = 0;
  std::vector<int64_t> snap_off(n), groups(n);
  for (int64_t q = 0; q < n; ++q) {
    const int64_t nz = shape[5 * q], nx = shape[5 * q + 1], B = shape[5 * q + 2],
                  nt = shape[5 * q + 3], R = shape[5 * q + 4];
    const float dt = dt_h[2 * q], h = dt_h[2 * q + 1];
    if (int e = check_grid(B, nz, nx, dt, h)) {
      const std::string msg = seis_last_error();
      seis::set_error("instance %lld: %s", (long long)q, msg.c_str());
      return e;
    }
    SEIS_CHECK_ARG(nt >= 1 && R >= 0, "instance %lld: nt %lld, R %lld", (long long)q,
                   (long long)nt, (long long)R);
    SEIS_CHECK_ARG(nx % VX == 0, "instance %lld: vec_x %d does not divide nx = %lld",
                   (long long)q, VX, (long long)nx);
    const int64_t L = seis_ring_len(nz, nx, npml, 0);
    SEIS_CHECK_ARG(!need_ring || L > 0, "instance %lld: grid %lld x %lld has no interior "
                   "inside npml=%d", (long long)q, (long long)nz, (long long)nx, npml);
    const int64_t plane = nz * nx;
    const int64_t G = (B + S - 1) / S;
    const Grid g = seis::make_grid(nz, nx, npml, false);
    const TileMap tm = make_tiles(g, VX, RZ, (int)G);
    const seis::TimeConsts k = seis::make_consts(dt, h);
    int64_t* r = rows.data() + q * kBatchCols;
    r[kCNz] = nz; r[kCNx] = nx; r[kCB] = B; r[kCNt] = nt; r[kCR] = R;
    r[kCNtx] = tm.ntx; r[kCNtz] = tm.ntz;
    r[kCField] = o_field; r[kCModel] = o_model; r[kCSrc] = o_src; r[kCRec] = o_rec;
    r[kCOut] = o_out; r[kCWav] = o_wav; r[kCRing] = o_ring; r[kCRingLen] = need_ring && L > 0 ? L : 0;
    r[kCGpart] = o_gpart;
    r[kCIdt2] = fbits(k.idt2); r[kCIdt] = fbits(k.idt); r[kCTwoIdt2] = fbits(k.two_idt2);
    r[kCIh2] = fbits(k.ih2);
    snap_off[q] = o_snap;
    groups[q] = G;
    o_model += plane; o_wav += nt; o_src += B; o_rec += R; o_out += nt * R;
    if (need_ring && nt > 1) o_ring += (nt - 1) * B * L;
    o_snap += seis_snapshot_count(nt, snap_every) * 2 * B * plane;
    o_field += B * plane; o_gpart += G * plane;
  }
  std::vector<int> order(n);
  for (int q = 0; q < (int)n; ++q) order[q] = q;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return rows[(size_t)a * kBatchCols + kCNt] > rows[(size_t)b * kBatchCols + kCNt];
  });
  p.tab.assign((size_t)n * kBatchCols, 0);
  p.nt_sorted.resize(n); p.snap_off.resize(n); p.groups.resize(n);
  p.tile_end.assign(n + 1, 0); p.tile2_end.assign(n + 1, 0); p.aux_end.assign(n + 1, 0);
  for (int k = 0; k < (int)n; ++k) {
    int64_t* r = p.tab.data() + (size_t)k * kBatchCols;
    std::memcpy(r, rows.data() + (size_t)order[k] * kBatchCols, sizeof(int64_t) * kBatchCols);
    r[kCTile0] = p.tile_end[k];
    r[kCTile2] = p.tile2_end[k];
    r[kCAux0] = p.aux_end[k];
    p.tile_end[k + 1] = p.tile_end[k] + r[kCNtx] * r[kCNtz] * groups[order[k]];
    const int64_t ntx2 = (r[kCNx] + kFwd2Out * VX - 1) / (kFwd2Out * VX);
    p.tile2_end[k + 1] = p.tile2_end[k] + ntx2 * r[kCNtz] * groups[order[k]];
    p.aux_end[k + 1] = p.aux_end[k] + r[kCR] + r[kCB] * r[kCRingLen];
    p.nt_sorted[k] = r[kCNt];
    p.snap_off[k] = snap_off[order[k]];
    p.groups[k] = groups[order[k]];
  }
  SEIS_CHECK_ARG(p.tile2_end[n] / kXcds < (int64_t)1 << 30, "too many tiles in one batch");
  p.nt_max = p.nt_sorted[0];
  p.fields = o_field;
  p.gparts = o_gpart;
  return 0;
}

int upload_plan(const BatchPlan& p, int64_t* table, hipStream_t s) {
  SEIS_CHECK_HIP(hipMemcpyAsync(table, p.tab.data(), p.tab.size() * sizeof(int64_t),
                                hipMemcpyHostToDevice, s));
  // the host rows are released when the caller's plan goes out of scope
  SEIS_CHECK_HIP(hipStreamSynchronize(s));
  return 0;
}

// Number of sorted rows still running at global step j (nt > j).
inline int active_rows(const BatchPlan& p, int64_t j) {
  return (int)(std::partition_point(p.nt_sorted.begin(), p.nt_sorted.end(),
                                    [j](int64_t nt) { return nt > j; }) -
               p.nt_sorted.begin());
}

struct BatchArgs {
  const float *vp, *d, *wavelet;
  const int32_t* src;
  const int64_t* rec;
  float* out;           // forward records
  const float* rgrad;   // backward record cotangents
  float* ring;
  float* snaps;
  int32_t K;
  float *ua, *ub, *ma, *mb, *gpart, *grad;
  const int64_t* table;
  int npml, S, RZ, VX;
  float *uc = nullptr, *ud = nullptr;  // two-step forward: the third and fourth arenas
};

// Two steps per launch (k_fwd2_batch) over four arenas; each instance's
// levels rotate through them by level_arena, so u^nt and u^{nt-1} end in ua
// and ub.  Records, rings and snapshots are taken per level as in the
// one-step loop.
int enqueue_forward_batch2(const BatchPlan& p, const BatchArgs& a, hipStream_t s, bool prof) {
  for (float* f : {a.ua, a.ub, a.uc, a.ud})
    if (int e = fill(f, p.fields, 0.0f, s)) return e;
  auto record = [&](int64_t m) -> int {  // u^{m+1} of the rows with nt > m
    const int na = active_rows(p, m);
    const int64_t work = p.aux_end[na];
    if (work > 0) {
      k_record_batch<<<(unsigned)((work + 255) / 256), 256, 0, s>>>(
          a.table, na, work, a.npml, a.ua, a.ub, a.uc, a.ud, a.rec, a.out, a.ring, m);
      SEIS_CHECK_LAUNCH();
    }
    return 0;
  };
  auto snapshot = [&](int64_t m) -> int {  // (u^m, u^{m+1}) at m = K, 2K, ...
    if (!a.snaps || m < a.K || m % a.K != 0) return 0;
    const int na = active_rows(p, m);
    for (int k = 0; k < na; ++k) {
      const int64_t* r = p.tab.data() + (size_t)k * kBatchCols;
      const int64_t nt = r[kCNt];
      if (m > nt - 2) continue;
      const int64_t sz = r[kCB] * r[kCNz] * r[kCNx];
      float* cur = level_arena(m, nt, a.ua, a.ub, a.uc, a.ud) + r[kCField];
      float* nxt = level_arena(m + 1, nt, a.ua, a.ub, a.uc, a.ud) + r[kCField];
      float* slot = a.snaps + p.snap_off[k] + (m / a.K - 1) * 2 * sz;
      if (int e = copy2(cur, slot, nxt, slot + sz, sz, s)) return e;
    }
    return 0;
  };
  for (int64_t n = 0; n < p.nt_max; n += 2) {
    const int na = active_rows(p, n);
    const int64_t tiles = p.tile2_end[na];
    const int64_t per = (tiles + kXcds - 1) / kXcds;
    if (prof) SEIS_PROF(SEIS_KERNEL_BATCH_FWD, s);
    if (int e = launch_fwd2_batch(a.S, a.RZ, a.VX, per * kXcds, s, a.table, na, tiles, (int)per,
                                  a.npml, a.vp, a.d, a.ua, a.ub, a.uc, a.ud, a.src, a.wavelet,
                                  n))
      return e;
    if (prof) SEIS_PROF(SEIS_KERNEL_BATCH_FWD, s);
    if (int e = record(n)) return e;
    if (int e = snapshot(n)) return e;
    if (n + 1 < p.nt_max) {
      if (int e = record(n + 1)) return e;
      if (int e = snapshot(n + 1)) return e;
    }
  }
  return 0;
}

int enqueue_forward_batch(const BatchPlan& p, const BatchArgs& a, hipStream_t s, bool prof) {
  if (int e = fill(a.ua, p.fields, 0.0f, s)) return e;
  if (int e = fill(a.ub, p.fields, 0.0f, s)) return e;
  for (int64_t n = 0; n < p.nt_max; ++n) {
    const int na = active_rows(p, n);
    const int64_t tiles = p.tile_end[na];
    const int64_t per = (tiles + kXcds - 1) / kXcds;
    if (prof) SEIS_PROF(SEIS_KERNEL_BATCH_FWD, s);
    if (int e = launch_fwd_batch(a.S, a.RZ, a.VX, per * kXcds, s, a.table, na, tiles, (int)per,
                                 a.npml, a.vp, a.d, a.ua, a.ub, a.src, a.wavelet, n))
      return e;
    if (prof) SEIS_PROF(SEIS_KERNEL_BATCH_FWD, s);
    const int64_t work = p.aux_end[na];
    if (work > 0) {
      k_record_batch<<<(unsigned)((work + 255) / 256), 256, 0, s>>>(
          a.table, na, work, a.npml, a.ua, a.ub, nullptr, nullptr, a.rec, a.out, a.ring, n);
      SEIS_CHECK_LAUNCH();
    }
    if (a.snaps && n >= a.K && n % a.K == 0) {
      for (int k = 0; k < na; ++k) {
        const int64_t* r = p.tab.data() + (size_t)k * kBatchCols;
        const int64_t nt = r[kCNt];
        if (n > nt - 2) continue;
        const int64_t sz = r[kCB] * r[kCNz] * r[kCNx];
        float* cur = (((n + nt) & 1) == 0 ? a.ua : a.ub) + r[kCField];
        float* nxt = (((n + nt) & 1) == 0 ? a.ub : a.ua) + r[kCField];
        float* slot = a.snaps + p.snap_off[k] + (n / a.K - 1) * 2 * sz;
        if (int e = copy2(cur, slot, nxt, slot + sz, sz, s)) return e;
      }
    }
  }
  return 0;
}

int enqueue_backward_batch(const BatchPlan& p, const BatchArgs& a, hipStream_t s, bool prof) {
  if (int e = fill(a.ma, p.fields, 0.0f, s)) return e;
  if (int e = fill(a.mb, p.fields, 0.0f, s)) return e;
  if (int e = fill(a.gpart, p.gparts, 0.0f, s)) return e;
  for (int64_t j = 0; j < p.nt_max; ++j) {
    const int na = active_rows(p, j);
    const int64_t tiles = p.tile_end[na];
    const int64_t per = (tiles + kXcds - 1) / kXcds;
    if (prof) SEIS_PROF(SEIS_KERNEL_BATCH_BWD, s);
    if (int e = launch_bwd_batch(a.S, a.RZ, a.VX, per * kXcds, s, a.table, na, tiles, (int)per,
                                 a.npml, a.vp, a.d, a.ua, a.ub, a.ma, a.mb, a.gpart, a.src,
                                 a.wavelet, j))
      return e;
    if (prof) SEIS_PROF(SEIS_KERNEL_BATCH_BWD, s);
    const int64_t work = p.aux_end[na];
    if (work > 0) {
      k_post_batch<<<(unsigned)((work + 255) / 256), 256, 0, s>>>(
          a.table, na, work, a.npml, a.S, a.vp, a.d, a.ua, a.ub, a.ma, a.mb, a.gpart, a.rec,
          a.rgrad, a.ring, j);
      SEIS_CHECK_LAUNCH();
    }
    if (a.snaps) {
      for (int k = 0; k < na; ++k) {
        const int64_t* r = p.tab.data() + (size_t)k * kBatchCols;
        const int64_t nt = r[kCNt], m = nt - j - 2;
        if (m < a.K || m % a.K != 0 || m > nt - 2) continue;
        const int64_t sz = r[kCB] * r[kCNz] * r[kCNx];
        // (u^m, u^{m+1}) -> the buffers of u^{kk-2} = u^m and u^{kk-1}
        float* um = (((m + nt) & 1) == 0 ? a.ua : a.ub) + r[kCField];
        float* um1 = (((m + nt) & 1) == 0 ? a.ub : a.ua) + r[kCField];
        const float* slot = a.snaps + p.snap_off[k] + (m / a.K - 1) * 2 * sz;
        if (int e = copy2(slot, um, slot + sz, um1, sz, s)) return e;
      }
    }
  }
  for (int k = 0; k < p.n; ++k) {
    const int64_t* r = p.tab.data() + (size_t)k * kBatchCols;
    const int64_t plane = r[kCNz] * r[kCNx];
    k_sum_groups<<<(unsigned)((plane + 255) / 256), 256, 0, s>>>(
        a.gpart + r[kCGpart], (int)p.groups[k], a.grad + r[kCModel], plane);
    SEIS_CHECK_LAUNCH();
  }
  return 0;
}

// Runs `enqueue` directly on s, or records it into a graph on a private
// capture stream and launches that graph on s (then waits for it).
template <typename F>
int run_maybe_graph(bool use_graph, hipStream_t s, F enqueue) {
  if (!use_graph) return enqueue(s, true);
  hipStream_t cs = nullptr;
  SEIS_CHECK_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  int rc = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipError_t he = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (he != hipSuccess) {
    seis::set_error("hipStreamBeginCapture failed: %s", hipGetErrorString(he));
    rc = 2;
  } else {
    rc = enqueue(cs, false);
    he = hipStreamEndCapture(cs, &graph);
    if (rc == 0 && he != hipSuccess) {
      seis::set_error("hipStreamEndCapture failed: %s", hipGetErrorString(he));
      rc = 2;
    }
  }
  if (rc == 0) {
    he = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (he == hipSuccess) he = hipGraphLaunch(exec, s);
    if (he == hipSuccess) he = hipStreamSynchronize(s);
    if (he != hipSuccess) {
      seis::set_error("batch graph failed: %s", hipGetErrorString(he));
      rc = 2;
    }
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  (void)hipStreamDestroy(cs);
  return rc;
}

}  // namespace

extern "C" {

int64_t seis_ring_len(int64_t nz, int64_t nx, int32_t npml, int32_t multiple) {
  const int64_t nxi = nx - 2 * (int64_t)npml - 2;
  const int64_t nzi = multiple ? nz - (int64_t)npml - 1 : nz - 2 * (int64_t)npml - 2;
  if (npml < 0 || nxi < 1 || nzi < 1) return -1;
  return (multiple ? 1 : 2) * nxi + 2 * nzi;
}

int64_t seis_snapshot_count(int64_t nt, int32_t snap_every) {
  if (snap_every <= 0 || nt < 2) return 0;
  return (nt - 2) / snap_every;
}

int seis_acoustic_step(const float* vp, const float* d, const float* h1, const float* h2,
                       float* y, int64_t B, int64_t nz, int64_t nx, float dt, float h,
                       void* stream) {
  if (int e = check_grid(B, nz, nx, dt, h)) return e;
  SEIS_CHECK_ARG(vp && d && h1 && h2 && y, "null pointer argument");
  const Grid g = seis::make_grid(nz, nx, 0, false);
  hipStream_t s = (hipStream_t)stream;
  k_step<<<tile_grid(g, (int)B), dim3(kTX, kTZ), 0, s>>>(vp, d, h1, h2, y, g,
                                                         seis::make_consts(dt, h));
  SEIS_CHECK_LAUNCH();
  return 0;
}

int seis_acoustic_step_vjp(const float* vp, const float* d, const float* h1, const float* gy,
                           const float* gh1_pass, float* gh1, float* gh2, float* gvp, int64_t B,
                           int64_t nz, int64_t nx, float dt, float h, void* stream) {
  if (int e = check_grid(B, nz, nx, dt, h)) return e;
  SEIS_CHECK_ARG(vp && d && h1 && gy && gh1 && gh2, "null pointer argument");
  const Grid g = seis::make_grid(nz, nx, 0, false);
  const seis::TimeConsts k = seis::make_consts(dt, h);
  hipStream_t s = (hipStream_t)stream;
  k_step_vjp_fields<<<tile_grid(g, (int)B), dim3(kTX, kTZ), 0, s>>>(vp, d, gy, gh1_pass, gh1,
                                                                    gh2, g, k);
  SEIS_CHECK_LAUNCH();
  if (gvp) {
    k_step_vjp_model<<<tile_grid(g, 1), dim3(kTX, kTZ), 0, s>>>(vp, d, h1, gy, gvp, (int)B, g,
                                                                k);
    SEIS_CHECK_LAUNCH();
  }
  return 0;
}

int seis_acoustic_forward(const float* vp, const float* d, const float* wavelet, int64_t nt,
                          const int32_t* src_zx, const int32_t* extra_src,
                          const float* extra_wav, int64_t n_extra, const int64_t* rec_flat,
                          int64_t R, float* rec_out, float* ring, float* snaps, int32_t snap_every,
                          float* illum, float* illum_part, float* u_a, float* u_b, float* u_c,
                          float* u_d, int64_t B, int64_t nz, int64_t nx, float dt, float h,
                          int32_t npml, int32_t multiple, int32_t shots_per_group,
                          int32_t rows_per_thread, int32_t vec_x, void* stream) {
  if (int e = check_grid(B, nz, nx, dt, h)) return e;
  SEIS_CHECK_ARG(valid_shape(shots_per_group, rows_per_thread, vec_x),
                 "launch shape (shots_per_group %d, rows_per_thread %d, vec_x %d) not in "
                 "{1,2,4,8} x {1,2,4} x {1,2,4}",
                 shots_per_group, rows_per_thread, vec_x);
  SEIS_CHECK_ARG(nx % vec_x == 0, "vec_x %d does not divide nx = %lld", vec_x, (long long)nx);
  SEIS_CHECK_ARG(vp && d && wavelet && src_zx && u_a && u_b, "null pointer argument");
  SEIS_CHECK_ARG(aligned(vp, 4 * vec_x) && aligned(d, 4 * vec_x) && aligned(u_a, 4 * vec_x) &&
                     aligned(u_b, 4 * vec_x),
                 "field and model pointers must be %d-byte aligned for vec_x %d", 4 * vec_x,
                 vec_x);
  SEIS_CHECK_ARG(nt >= 1, "nt must be >= 1");
  SEIS_CHECK_ARG(R >= 0 && (R == 0 || (rec_flat && rec_out)), "bad receiver arguments");
  SEIS_CHECK_ARG(shots_per_group >= 1, "shots_per_group must be >= 1");
  const int64_t ring_len = ring ? seis_ring_len(nz, nx, npml, multiple) : 0;
  SEIS_CHECK_ARG(!ring || ring_len > 0, "grid %lld x %lld has no interior inside npml=%d",
                 (long long)nz, (long long)nx, npml);
  SEIS_CHECK_ARG(!snaps || snap_every > 0, "snapshots need snap_every > 0");
  SEIS_CHECK_ARG(n_extra >= 0 && (n_extra == 0 || (extra_src && extra_wav)),
                 "bad extra-source arguments");
  SEIS_CHECK_ARG(n_extra == 0 || !(u_c && u_d),
                 "extra sources need one time step per launch (u_c = u_d = NULL)");
  const Grid g = seis::make_grid(nz, nx, npml, multiple != 0);
  const seis::TimeConsts k = seis::make_consts(dt, h);
  hipStream_t s = (hipStream_t)stream;
  const int S = shots_per_group;
  const int G = (int)((B + S - 1) / S);
  // After nt steps u^{nt} must sit in u_a: step n writes into the buffer that
  // holds u^{n-1}, so u^{n} lives in the step-0 "current" buffer iff n is even.
  const TileMap tm = make_tiles(g, vec_x, rows_per_thread, G);
  const int64_t aux = R + ((ring && nt > 1) ? B * ring_len : 0);
  if (illum) {
    SEIS_CHECK_ARG(illum != u_a && illum != u_b && illum != u_c && illum != u_d,
                   "illum must not alias a field buffer");
    SEIS_CHECK_ARG(!illum_part || (illum_part != illum && illum_part != u_a &&
                                   illum_part != u_b && illum_part != u_c && illum_part != u_d),
                   "illum_part must not alias illum or a field buffer");
    if (int e = fill(illum, g.plane, 0.0f, s)) return e;
  }
  if (u_c && u_d) {
    SEIS_CHECK_ARG(aligned(u_c, 4 * vec_x) && aligned(u_d, 4 * vec_x),
                   "u_c and u_d must be %d-byte aligned for vec_x %d", 4 * vec_x, vec_x);
    SEIS_CHECK_ARG(u_c != u_a && u_c != u_b && u_d != u_a && u_d != u_b && u_c != u_d,
                   "the four field buffers must be distinct");
    return forward_two_step(vp, d, wavelet, nt, src_zx, rec_flat, R, rec_out, ring, snaps,
                            snap_every, illum, illum ? illum_part : nullptr, u_a, u_b, u_c, u_d,
                            (int)B, g, k, S, rows_per_thread,
                            vec_x, tm, ring_len, aux, s);
  }
  float* cur = (nt % 2 == 0) ? u_a : u_b;
  float* prev = (nt % 2 == 0) ? u_b : u_a;
  if (int e = fill(u_a, B * g.plane, 0.0f, s)) return e;
  if (int e = fill(u_b, B * g.plane, 0.0f, s)) return e;
  for (int64_t n = 0; n < nt; ++n) {
    SEIS_PROF(SEIS_KERNEL_ACOUSTIC_FWD, s);
    if (int e = launch_fwd(S, rows_per_thread, vec_x, (int64_t)tm.per * kXcds, s, vp, d, cur,
                           prev, src_zx, wavelet + n, (int)B, g, k, tm))
      return e;
    SEIS_PROF(SEIS_KERNEL_ACOUSTIC_FWD, s);
    if (int e = extra_sources(prev, extra_src, extra_wav, nt, n, n_extra, 1.0f, g, s)) return e;
    if (int e = illuminate(prev, (int)B, g.plane, illum, s)) return e;
    float* slot = (ring && n + 1 < nt) ? ring + n * B * ring_len : nullptr;
    const int64_t work = R + (slot ? B * ring_len : 0);
    if (work > 0 && aux > 0) {
      k_record<<<(unsigned)((work + 255) / 256), 256, 0, s>>>(
          prev, rec_flat, R, R ? rec_out + n * R : nullptr, slot, ring_len, (int)B, g);
      SEIS_CHECK_LAUNCH();
    }
    // snapshot at time m = n: the pair (u^m, u^{m+1}), slot m/snap_every - 1
    if (snaps && n >= snap_every && n % snap_every == 0 && n <= nt - 2) {
      float* slot_p = snaps + (n / snap_every - 1) * 2 * B * g.plane;
      if (int e = copy2(cur, slot_p, prev, slot_p + B * g.plane, B * g.plane, s)) return e;
    }
    float* t = cur; cur = prev; prev = t;
  }
  return 0;
}

int seis_acoustic_backward(const float* vp, const float* d, const float* wavelet, int64_t nt,
                           const int32_t* src_zx, const int32_t* extra_src,
                           const float* extra_wav, int64_t n_extra, const int64_t* rec_flat,
                           int64_t R,
                           const float* rec_grad, const float* ring, const float* snaps,
                           int32_t snap_every, float* u_a, float* u_b,
                           float* mu_a, float* mu_b, float* u_c, float* u_d, float* mu_c,
                           float* mu_d, const int32_t* rec_map, float* grad_part, float* grad,
                           int64_t B, int64_t nz, int64_t nx, float dt, float h, int32_t npml,
                           int32_t multiple, int32_t shots_per_group, int32_t rows_per_thread,
                           int32_t vec_x, void* stream) {
  if (int e = check_grid(B, nz, nx, dt, h)) return e;
  SEIS_CHECK_ARG(valid_shape(shots_per_group, rows_per_thread, vec_x),
                 "launch shape (shots_per_group %d, rows_per_thread %d, vec_x %d) not in "
                 "{1,2,4,8} x {1,2,4} x {1,2,4}",
                 shots_per_group, rows_per_thread, vec_x);
  SEIS_CHECK_ARG(nx % vec_x == 0, "vec_x %d does not divide nx = %lld", vec_x, (long long)nx);
  SEIS_CHECK_ARG(vp && d && wavelet && src_zx && u_a && u_b && mu_a && mu_b && grad_part && grad,
                 "null pointer argument");
  SEIS_CHECK_ARG(aligned(vp, 4 * vec_x) && aligned(d, 4 * vec_x) && aligned(u_a, 4 * vec_x) &&
                     aligned(u_b, 4 * vec_x) && aligned(mu_a, 4 * vec_x) &&
                     aligned(mu_b, 4 * vec_x) && aligned(grad_part, 4 * vec_x),
                 "field and model pointers must be %d-byte aligned for vec_x %d", 4 * vec_x,
                 vec_x);
  SEIS_CHECK_ARG(nt >= 1, "nt must be >= 1");
  SEIS_CHECK_ARG(R >= 0 && (R == 0 || (rec_flat && rec_grad)), "bad receiver arguments");
  SEIS_CHECK_ARG(shots_per_group >= 1, "shots_per_group must be >= 1");
  const int64_t ring_len = seis_ring_len(nz, nx, npml, multiple);
  SEIS_CHECK_ARG(ring_len > 0, "grid %lld x %lld has no interior inside npml=%d", (long long)nz,
                 (long long)nx, npml);
  SEIS_CHECK_ARG(ring || nt <= 2, "ring storage is required for nt > 2");
  SEIS_CHECK_ARG(!snaps || snap_every > 0, "snapshots need snap_every > 0");
  SEIS_CHECK_ARG(n_extra >= 0 && (n_extra == 0 || (extra_src && extra_wav)),
                 "bad extra-source arguments");
  const Grid g = seis::make_grid(nz, nx, npml, multiple != 0);
  const seis::TimeConsts k = seis::make_consts(dt, h);
  hipStream_t s = (hipStream_t)stream;
  const int S = shots_per_group;
  const int G = (int)((B + S - 1) / S);
  if (int e = fill(mu_a, B * g.plane, 0.0f, s)) return e;
  if (int e = fill(mu_b, B * g.plane, 0.0f, s)) return e;
  if (int e = fill(grad_part, (int64_t)G * g.plane, 0.0f, s)) return e;
  const TileMap tm = make_tiles(g, vec_x, rows_per_thread, G);
  float* X = u_b;   // u^{k-1}
  float* Y = u_a;   // u^k  -> u^{k-2}
  float* M1 = mu_a; // mu^{k+1}
  float* M2 = mu_b; // mu^{k+2} -> mu^k
  // two reverse iterations per launch (k_bwd2) where possible
  const bool two = u_c && u_d && mu_c && mu_d && ring && (R == 0 || rec_map) && n_extra == 0;
  if (two) {
    SEIS_CHECK_ARG(aligned(u_c, 4 * vec_x) && aligned(u_d, 4 * vec_x) &&
                       aligned(mu_c, 4 * vec_x) && aligned(mu_d, 4 * vec_x),
                   "field pointers must be %d-byte aligned for vec_x %d", 4 * vec_x, vec_x);
    SEIS_CHECK_ARG(R % B == 0, "rec_map needs the same receiver count in every shot");
  }
  const int Rs = (int)(R / B);
  float* XC = u_c;  // free buffers for the pair's outputs
  float* XD = u_d;
  float* MC = mu_c;
  float* MD = mu_d;
  // k_bwd2 keeps two iterations' fields of its rows in registers: one row per
  // thread and at most kBwd2MaxVX cells per lane (vec_x divides nx, so does
  // any smaller power of two)
  const int rz2 = 1;
  const int vx2 = vec_x < kBwd2MaxVX ? vec_x : kBwd2MaxVX;
  TileMap tm2 = make_tiles(g, vx2, rz2, G);
  tm2.ntx = (g.nx + kFwd2Out * vx2 - 1) / (kFwd2Out * vx2);
  tm2.per = (int)(((int64_t)tm2.ntx * tm2.ntz * tm2.groups + kXcds - 1) / kXcds);
  auto snap_at = [&](int64_t m) {
    return snaps && m >= snap_every && m % snap_every == 0 && m <= nt - 2;
  };
  int64_t kk = nt;
  while (kk >= 1) {
    if (two && kk >= 4 && !snap_at(kk - 2)) {
      // iterations kk and kk-1: X = u^{kk-1}, Y = u^kk, M1 = mu^{kk+1}, M2 =
      // mu^{kk+2} -> XC = u^{kk-2}, XD = u^{kk-3}, MC = mu^kk, MD = mu^{kk-1}
      SEIS_PROF(SEIS_KERNEL_ACOUSTIC_BWD2, s);
      if (int e = launch_bwd2(shots_per_group, rz2, vx2, (int64_t)tm2.per * kXcds, s, vp, d, X, Y, XC, XD, M1, M2, MC, MD,
                              grad_part, src_zx, wavelet + (kk - 2), R ? rec_map : nullptr,
                              R ? rec_grad + (kk - 1) * R : nullptr, Rs,
                              ring + (kk - 3) * B * ring_len, ring_len, (int)B, g, k, tm2))
        return e;
      SEIS_PROF(SEIS_KERNEL_ACOUSTIC_BWD2, s);
      if (R > 0) {  // gradient share of iteration kk's receivers (already in mu^kk)
        k_bwd_post<<<(unsigned)((R + 255) / 256), 256, 0, s>>>(
            vp, d, X, nullptr, nullptr, grad_part, rec_flat, rec_grad + (kk - 1) * R, R,
            nullptr, ring_len, 0, (int)B, S, g, k);
        SEIS_CHECK_LAUNCH();
      }
      {  // iteration kk-1's receivers and the ring of u^{kk-3}
        const int64_t work = R + B * ring_len;
        k_bwd_post<<<(unsigned)((work + 255) / 256), 256, 0, s>>>(
            vp, d, XC, XD, MD, grad_part, rec_flat, R ? rec_grad + (kk - 2) * R : nullptr, R,
            ring + (kk - 4) * B * ring_len, ring_len, 1, (int)B, S, g, k);
        SEIS_CHECK_LAUNCH();
      }
      if (snap_at(kk - 3)) {  // reload the stored pair (u^{kk-3}, u^{kk-2})
        const float* slot_p = snaps + ((kk - 3) / snap_every - 1) * 2 * B * g.plane;
        if (int e = copy2(slot_p, XD, slot_p + B * g.plane, XC, B * g.plane, s)) return e;
      }
      float* t0 = X; float* t1 = Y;
      X = XD; Y = XC; XC = t0; XD = t1;
      t0 = M1; t1 = M2;
      M1 = MD; M2 = MC; MC = t0; MD = t1;
      kk -= 2;
      continue;
    }
    const int recon = kk >= 3 ? 1 : (kk == 2 ? 2 : 0);
    // the rebuild of u^{kk-2} takes u^kk without its sources of step kk-1;
    // Y's value of u^kk is not read again
    if (recon == 1)
      if (int e = extra_sources(Y, extra_src, extra_wav, nt, kk - 1, n_extra, -1.0f, g, s))
        return e;
    SEIS_PROF(SEIS_KERNEL_ACOUSTIC_BWD, s);
    if (int e = launch_bwd(S, rows_per_thread, vec_x, (int64_t)tm.per * kXcds, s, vp, d, X, Y, M1,
                           M2, grad_part, src_zx, wavelet + (kk - 1), recon, (int)B, g, k, tm))
      return e;
    SEIS_PROF(SEIS_KERNEL_ACOUSTIC_BWD, s);
    const int64_t work = R + (recon ? B * ring_len : 0);
    if (work > 0) {
      const float* slot = recon == 1 ? ring + (kk - 3) * B * ring_len : nullptr;
      k_bwd_post<<<(unsigned)((work + 255) / 256), 256, 0, s>>>(
          vp, d, X, Y, M2, grad_part, rec_flat, R ? rec_grad + (kk - 1) * R : nullptr, R, slot,
          ring_len, recon, (int)B, S, g, k);
      SEIS_CHECK_LAUNCH();
    }
    // reload the stored pair (u^m, u^{m+1}) at snapshot times m = kk-2
    const int64_t m = kk - 2;
    if (snaps && m >= snap_every && m % snap_every == 0 && m <= nt - 2) {
      const float* slot_p = snaps + (m / snap_every - 1) * 2 * B * g.plane;
      if (int e = copy2(slot_p, Y, slot_p + B * g.plane, X, B * g.plane, s)) return e;
    }
    float* t = X; X = Y; Y = t;
    t = M1; M1 = M2; M2 = t;
    --kk;
  }
  k_sum_groups<<<(unsigned)((g.plane + 255) / 256), 256, 0, s>>>(grad_part, G, grad, g.plane);
  SEIS_CHECK_LAUNCH();
  return 0;
}

int seis_acoustic_forward_batch(int64_t n, const int64_t* shape, const float* dt_h,
                                const float* vp, const float* d, const float* wavelet,
                                const int32_t* src_zx, const int64_t* rec_flat, float* rec_out,
                                float* ring, float* snaps, int32_t snap_every, float* u_a,
                                float* u_b, float* u_c, float* u_d, int64_t* table, int32_t npml,
                                int32_t shots_per_group, int32_t rows_per_thread, int32_t vec_x,
                                int32_t use_graph, void* stream) {
  BatchPlan p;
  if (int e = make_plan(p, n, shape, dt_h, npml, shots_per_group, rows_per_thread, vec_x,
                        snaps ? snap_every : 0, ring != nullptr))
    return e;
  SEIS_CHECK_ARG(vp && d && wavelet && src_zx && u_a && u_b && table, "null pointer argument");
  bool any_r = false;
  for (int k = 0; k < p.n; ++k) any_r = any_r || p.tab[(size_t)k * kBatchCols + kCR] > 0;
  SEIS_CHECK_ARG(!any_r || (rec_flat && rec_out), "receivers need rec_flat and rec_out");
  SEIS_CHECK_ARG(!snaps || snap_every > 0, "snapshots need snap_every > 0");
  SEIS_CHECK_ARG(aligned(vp, 4 * vec_x) && aligned(d, 4 * vec_x) && aligned(u_a, 4 * vec_x) &&
                     aligned(u_b, 4 * vec_x),
                 "field and model pointers must be %d-byte aligned for vec_x %d", 4 * vec_x,
                 vec_x);
  SEIS_CHECK_ARG(!u_c == !u_d, "u_c and u_d are given together (two steps per launch) or not");
  if (u_c) {
    SEIS_CHECK_ARG(aligned(u_c, 4 * vec_x) && aligned(u_d, 4 * vec_x),
                   "u_c and u_d must be %d-byte aligned for vec_x %d", 4 * vec_x, vec_x);
    SEIS_CHECK_ARG(u_c != u_a && u_c != u_b && u_d != u_a && u_d != u_b && u_c != u_d,
                   "the four field buffers must be distinct");
  }
  hipStream_t s = (hipStream_t)stream;
  if (int e = upload_plan(p, table, s)) return e;
  BatchArgs a{vp, d, wavelet, src_zx, rec_flat, rec_out, nullptr, ring, snaps, snap_every,
              u_a, u_b, nullptr, nullptr, nullptr, nullptr, table, npml, shots_per_group,
              rows_per_thread, vec_x, u_c, u_d};
  return run_maybe_graph(use_graph != 0, s, [&](hipStream_t st, bool prof) {
    return u_c ? enqueue_forward_batch2(p, a, st, prof) : enqueue_forward_batch(p, a, st, prof);
  });
}

int seis_acoustic_backward_batch(int64_t n, const int64_t* shape, const float* dt_h,
                                 const float* vp, const float* d, const float* wavelet,
                                 const int32_t* src_zx, const int64_t* rec_flat,
                                 const float* rec_grad, const float* ring, const float* snaps,
                                 int32_t snap_every, float* u_a, float* u_b, float* mu_a,
                                 float* mu_b, float* grad_part, float* grad, int64_t* table,
                                 int32_t npml, int32_t shots_per_group, int32_t rows_per_thread,
                                 int32_t vec_x, int32_t use_graph, void* stream) {
  BatchPlan p;
  if (int e = make_plan(p, n, shape, dt_h, npml, shots_per_group, rows_per_thread, vec_x,
                        snaps ? snap_every : 0, true))
    return e;
  SEIS_CHECK_ARG(vp && d && wavelet && src_zx && u_a && u_b && mu_a && mu_b && grad_part &&
                     grad && table,
                 "null pointer argument");
  SEIS_CHECK_ARG(ring || p.nt_max <= 2, "ring storage is required for nt > 2");
  bool any_r = false;
  for (int k = 0; k < p.n; ++k) any_r = any_r || p.tab[(size_t)k * kBatchCols + kCR] > 0;
  SEIS_CHECK_ARG(!any_r || (rec_flat && rec_grad), "receivers need rec_flat and rec_grad");
  SEIS_CHECK_ARG(!snaps || snap_every > 0, "snapshots need snap_every > 0");
  SEIS_CHECK_ARG(aligned(vp, 4 * vec_x) && aligned(d, 4 * vec_x) && aligned(u_a, 4 * vec_x) &&
                     aligned(u_b, 4 * vec_x) && aligned(mu_a, 4 * vec_x) &&
                     aligned(mu_b, 4 * vec_x) && aligned(grad_part, 4 * vec_x),
                 "field and model pointers must be %d-byte aligned for vec_x %d", 4 * vec_x,
                 vec_x);
  hipStream_t s = (hipStream_t)stream;
  if (int e = upload_plan(p, table, s)) return e;
  BatchArgs a{vp, d, wavelet, src_zx, rec_flat, nullptr, rec_grad, const_cast<float*>(ring),
              const_cast<float*>(snaps), snap_every, u_a, u_b, mu_a, mu_b, grad_part, grad,
              table, npml, shots_per_group, rows_per_thread, vec_x};
  return run_maybe_graph(use_graph != 0, s, [&](hipStream_t st, bool prof) {
    return enqueue_backward_batch(p, a, st, prof);
  });
}

}  // extern "C"
