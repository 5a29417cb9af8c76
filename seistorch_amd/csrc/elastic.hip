// 2D isotropic elastic velocity-stress kernels for gfx950.
//
// One time step (seistorch/equations2d/elastic.py:7-37) is two linear stages:
//   T (stresses from velocities)
//     txx' = a*(s*dt/h*((lam+2mu) Dx- vx + lam Dz+ vz) + b*txx)
//     tzz' = a*(s*dt/h*((lam+2mu) Dz+ vz + lam Dx- vx) + b*tzz)
//     txz' = a*(s*(dt*mu)/h*(Dx+ vz + Dz- vx) + b*txz)
//   V (velocities from the new stresses)
//     vx' = a*(s*(dt/rho)/h*(Dx+ txx' + Dz+ txz') + b*vx)
//     vz' = a*(s*(dt/rho)/h*(Dx- txz' + Dz- tzz') + b*vz)
// with a = 1/(1+c), b = 1-c, c = dt*d/2, D- / D+ the zero-ended differences
// of seistorch/equations2d/utils.py:3-48, s = +1.  The forward step is T then
// V; the reference's reverse step (elastic.py:39-122) is V then T with s = -1
// on the interior region.  Each stage kernel works in place on a region.
//
// adjT / adjV are the exact transposes of the stages (out of place), with
// optional accumulation of the gradients w.r.t. the per-cell coefficients
// (lam+2mu, lam, dt*mu/h, dt/(rho h)) and, in the fused adjoint loop, the
// reverse-time rebuild of the forward fields on the undamped interior from
// the boundary-saving rings.  k_el_chain maps the coefficient gradients to
// vp, vs and rho.
//
// The coefficients are precomputed per cell with the reference's FP32
// operation order (k_el_coef), and the stage kernels round op by op in the
// reference's expression order, so the forward matches its FP32 path.
#include "common.h"
#include "illum.h"
#include "seistorch_amd.h"
#include "tiles.h"

#include <initializer_list>

using seis::Grid;
using namespace seis::tiles;

namespace {

// coefficient planes, each (nz, nx)
enum { kL2M = 0, kLam, kMc, kRc, kAinv, kBm, kNumCoef };

struct Region {
  int z0, z1, x0, x1;  // cells z0 <= z < z1, x0 <= x < x1
};

__device__ __forceinline__ bool in_region(int z, int x, const Region r) {
  return z >= r.z0 && z < r.z1 && x >= r.x0 && x < r.x1;
}

struct Fields {
  float* f[5];  // vx, vz, txx, tzz, txz, each B x (nz, nx)
};

struct CFields {
  const float* f[5];
};

__global__ __launch_bounds__(256) void k_el_coef(const float* __restrict__ vp,
                                                 const float* __restrict__ vs,
                                                 const float* __restrict__ rho,
                                                 const float* __restrict__ d,
                                                 float* __restrict__ cf, float dt, float ih,
                                                 int64_t plane) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= plane) return;
  const float p = vp[i], s = vs[i], r = rho[i];
  const float lam = r * (p * p - 2.0f * (s * s));   // rho*(vp.pow(2)-2*vs.pow(2))
  const float mu = r * (s * s);                      // rho*(vs.pow(2))
  const float c = (0.5f * dt) * d[i];                // 0.5*dt*d
  cf[kL2M * plane + i] = lam + 2.0f * mu;
  cf[kLam * plane + i] = lam;
  cf[kMc * plane + i] = (dt * mu) * ih;              // dt*lame_mu*h.pow(-1)
  cf[kRc * plane + i] = (dt * (1.0f / r)) * ih;      // dt*rho.pow(-1)*h.pow(-1)
  cf[kAinv * plane + i] = 1.0f / (1.0f + c);         // (1+c)**-1
  cf[kBm * plane + i] = 1.0f - c;                    // (1-c)
}

// Lane cells of the vectorised stage kernels: VX consecutive x cells of one
// row (float/float2/float4 access), S shots per thread, XCD-aware order.
template <int VX>
__device__ __forceinline__ bool vcell_of(const TileMap tm, const Grid g, int& z, int& x0,
                                         int& grp) {
  int tx, tzi;
  if (!tile_of(tm, tx, tzi, grp)) return false;
  x0 = (tx * kLanes + (int)(threadIdx.x % kLanes)) * VX;
  z = tzi * kRowsOfThreads + (int)(threadIdx.x / kLanes);
  return x0 < g.nx && z < g.nz;
}

// Stage T in place over the region: stresses from velocities.  Each lane
// updates VX cells; cells outside the region keep their values.  The
// arithmetic per cell is the scalar expression order of elastic.py:16-24.
template <int S, int VX>
__global__ __launch_bounds__(256) void k_el_T(const float* __restrict__ cf, float sdti,
                                              float sgn, Region rg, Fields F, int B, Grid g,
                                              TileMap tm) {
  int z, x0, grp;
  if (!vcell_of<VX>(tm, g, z, x0, grp)) return;
  if (z < rg.z0 || z >= rg.z1 || x0 + VX <= rg.x0 || x0 >= rg.x1) return;
  const int i = z * g.nx + x0;
  const int64_t P = g.plane;
  float l2m[VX], lam[VX], mc[VX], a[VX], bm[VX];
  vload<VX>(cf + kL2M * P + i, l2m);
  vload<VX>(cf + kLam * P + i, lam);
  vload<VX>(cf + kMc * P + i, mc);
  vload<VX>(cf + kAinv * P + i, a);
  vload<VX>(cf + kBm * P + i, bm);
  bool in[VX];
#pragma unroll
  for (int e = 0; e < VX; ++e) in[e] = x0 + e >= rg.x0 && x0 + e < rg.x1;
  const bool zl = z > 0, zr = z < g.nz - 1;
  const int b0 = grp * S, nb = min(S, B - b0);
  for (int s = 0; s < nb; ++s) {
    const int64_t off = (int64_t)(b0 + s) * P + i;
    const float* vx = F.f[0] + off;
    const float* vz = F.f[1] + off;
    float vxc[VX], vzc[VX], vxu[VX], vzd[VX];
    vload<VX>(vx, vxc);
    vload<VX>(vz, vzc);
    if (zl) vload<VX>(vx - g.nx, vxu); else vzero<VX>(vxu);
    if (zr) vload<VX>(vz + g.nx, vzd); else vzero<VX>(vzd);
    const float vxl = x0 > 0 ? vx[-1] : 0.0f;
    const float vzr = x0 + VX < g.nx ? vz[VX] : 0.0f;
    float txx[VX], tzz[VX], txz[VX];
    vload<VX>(F.f[2] + off, txx);
    vload<VX>(F.f[3] + off, tzz);
    vload<VX>(F.f[4] + off, txz);
#pragma unroll
    for (int e = 0; e < VX; ++e) {
      const int x = x0 + e;
      const float vx_x = x > 0 ? vxc[e] - (e == 0 ? vxl : vxc[e - 1]) : 0.0f;
      const float vz_z = zr ? vzd[e] - vzc[e] : 0.0f;
      const float vx_z = zl ? vxc[e] - vxu[e] : 0.0f;
      const float vz_x = x < g.nx - 1 ? (e == VX - 1 ? vzr : vzc[e + 1]) - vzc[e] : 0.0f;
      if (in[e]) {
        txx[e] = a[e] * (sdti * (l2m[e] * vx_x + lam[e] * vz_z) + bm[e] * txx[e]);
        tzz[e] = a[e] * (sdti * (l2m[e] * vz_z + lam[e] * vx_x) + bm[e] * tzz[e]);
        txz[e] = a[e] * ((sgn * mc[e]) * (vz_x + vx_z) + bm[e] * txz[e]);
      }
    }
    vstore<VX>(F.f[2] + off, txx);
    vstore<VX>(F.f[3] + off, tzz);
    vstore<VX>(F.f[4] + off, txz);
  }
}

// Stage V in place over the region: velocities from stresses (elastic.py:26-35).
template <int S, int VX>
__global__ __launch_bounds__(256) void k_el_V(const float* __restrict__ cf, float sgn, Region rg,
                                              Fields F, int B, Grid g, TileMap tm) {
  int z, x0, grp;
  if (!vcell_of<VX>(tm, g, z, x0, grp)) return;
  if (z < rg.z0 || z >= rg.z1 || x0 + VX <= rg.x0 || x0 >= rg.x1) return;
  const int i = z * g.nx + x0;
  const int64_t P = g.plane;
  float rc[VX], a[VX], bm[VX];
  vload<VX>(cf + kRc * P + i, rc);
  vload<VX>(cf + kAinv * P + i, a);
  vload<VX>(cf + kBm * P + i, bm);
  bool in[VX];
#pragma unroll
  for (int e = 0; e < VX; ++e) in[e] = x0 + e >= rg.x0 && x0 + e < rg.x1;
  const bool zl = z > 0, zr = z < g.nz - 1;
  const int b0 = grp * S, nb = min(S, B - b0);
  for (int s = 0; s < nb; ++s) {
    const int64_t off = (int64_t)(b0 + s) * P + i;
    const float* txx = F.f[2] + off;
    const float* tzz = F.f[3] + off;
    const float* txz = F.f[4] + off;
    float txxc[VX], tzzc[VX], txzc[VX], tzzu[VX], txzd[VX];
    vload<VX>(txx, txxc);
    vload<VX>(tzz, tzzc);
    vload<VX>(txz, txzc);
    if (zl) vload<VX>(tzz - g.nx, tzzu); else vzero<VX>(tzzu);
    if (zr) vload<VX>(txz + g.nx, txzd); else vzero<VX>(txzd);
    const float txxr = x0 + VX < g.nx ? txx[VX] : 0.0f;
    const float txzl = x0 > 0 ? txz[-1] : 0.0f;
    float vx[VX], vz[VX];
    vload<VX>(F.f[0] + off, vx);
    vload<VX>(F.f[1] + off, vz);
#pragma unroll
    for (int e = 0; e < VX; ++e) {
      const int x = x0 + e;
      const float txx_x = x < g.nx - 1 ? (e == VX - 1 ? txxr : txxc[e + 1]) - txxc[e] : 0.0f;
      const float txz_z = zr ? txzd[e] - txzc[e] : 0.0f;
      const float tzz_z = zl ? tzzc[e] - tzzu[e] : 0.0f;
      const float txz_x = x > 0 ? txzc[e] - (e == 0 ? txzl : txzc[e - 1]) : 0.0f;
      if (in[e]) {
        const float src = sgn * rc[e];
        vx[e] = a[e] * (src * (txx_x + txz_z) + bm[e] * vx[e]);
        vz[e] = a[e] * (src * (txz_x + tzz_z) + bm[e] * vz[e]);
      }
    }
    vstore<VX>(F.f[0] + off, vx);
    vstore<VX>(F.f[1] + off, vz);
  }
}

// One whole forward step (stage T then stage V, elastic.py:16-35) in one
// pass, reading state n from Fi and writing state n+1 to Fo (distinct
// buffers: neighbouring tiles still read state n around this one).  A wave
// is one row of 64 lanes; each lane computes the stresses of its RZ rows and
// of the rows above and below (recomputed with the same operands as the
// neighbouring thread row, so bit-identical), then the velocities of its own
// rows, taking x-neighbours from the adjacent lanes by cross-lane shuffles.
// Tiles overlap by one lane vector on each side: lanes 0 and 63 only supply
// neighbour values, the 62 inner lanes write.  Coefficients are re-read per
// shot (L1/L2 hits) to keep the register footprint small.
constexpr int kElOut = kLanes - 2;

template <int S, int RZ, int VX>
__device__ __forceinline__ void el_step_body(const float* __restrict__ cf, float dti, CFields Fi,
                                             Fields Fo, int B, Grid g, TileMap tm) {
  int tx, tzi, grp;
  if (!tile_of(tm, tx, tzi, grp)) return;  // uniform over the block
  const int lane = (int)(threadIdx.x % kLanes);
  const int trow = (int)(threadIdx.x / kLanes);
  const int x0 = (tx * kElOut + lane - 1) * VX;
  const int zb0 = tzi * kRowsOfThreads * RZ;  // the block's first row
  const int zb = zb0 + trow * RZ;
  const bool xin = x0 >= 0 && x0 < g.nx;  // nx is a multiple of VX
  const bool inner = lane > 0 && lane < kLanes - 1;
  const int64_t P = g.plane;
  // the block's coefficient planes in LDS, rows zb0-1 .. zb0+4*RZ (local
  // row z - zb0 + 1): read from HBM once per block, not once per shot.
  // Rows outside the grid are never read.
  constexpr int kCR = kRowsOfThreads * RZ + 2;
  __shared__ __align__(16) float lcf[kNumCoef][kCR][kLanes * VX];
  for (int q = trow; q < kCR; q += kRowsOfThreads) {
    const int z = zb0 - 1 + q;
    if (!xin || z < 0 || z >= g.nz) continue;
#pragma unroll
    for (int c = 0; c < kNumCoef; ++c) {
      float t[VX];
      vload<VX>(cf + c * P + (int64_t)z * g.nx + x0, t);
      vstore<VX>(&lcf[c][q][lane * VX], t);
    }
  }
  __syncthreads();
  if (zb >= g.nz) return;  // uniform over the wave, after the block's barrier
  // coefficient c on row z
  auto cload = [&](int c, int z, float(&o)[VX]) {
    vload<VX>(&lcf[c][z - zb0 + 1][lane * VX], o);
  };
  const int b0 = grp * S, nb = min(S, B - b0);
  for (int s = 0; s < nb; ++s) {
    const int64_t off = (int64_t)(b0 + s) * P;
    const float* vx = Fi.f[0] + off;
    const float* vz = Fi.f[1] + off;
    // vx on rows zb-2 .. zb+RZ (index z - zb + 2), vz on rows zb-1 .. zb+RZ+1
    // (index z - zb + 1); zero outside the grid
    float VXr[RZ + 3][VX], VZr[RZ + 3][VX];
#pragma unroll
    for (int p = 0; p < RZ + 3; ++p) {
      const int za = zb - 2 + p, zc = zb - 1 + p;
      if (xin && za >= 0 && za < g.nz) vload<VX>(vx + za * g.nx + x0, VXr[p]);
      else vzero<VX>(VXr[p]);
      if (xin && zc >= 0 && zc < g.nz) vload<VX>(vz + zc * g.nx + x0, VZr[p]);
      else vzero<VX>(VZr[p]);
    }
    // stage T on rows zb-1 .. zb+RZ (index q = z - zb + 1)
    float TXX[RZ + 2][VX], TZZ[RZ + 2][VX], TXZ[RZ + 2][VX];
#pragma unroll
    for (int q = 0; q < RZ + 2; ++q) {
      const int z = zb - 1 + q;
      const bool zok = z >= 0 && z < g.nz;
      // vx row z is VXr[q + 1], vz row z is VZr[q]
      // lane 0's stresses reach lane 1 only through its last txz, lane 63's
      // only through its first txx: neither uses the outer neighbour
      const float vxl = __shfl_up(VXr[q + 1][VX - 1], 1, kLanes);
      const float vzr = __shfl_down(VZr[q][0], 1, kLanes);
      if (xin && zok) {
        const int i = z * g.nx + x0;
        float l2m[VX], lam[VX], mc[VX], a[VX], bm[VX];
        cload(kL2M, z, l2m);
        cload(kLam, z, lam);
        cload(kMc, z, mc);
        cload(kAinv, z, a);
        cload(kBm, z, bm);
        vload<VX>(Fi.f[2] + off + i, TXX[q]);
        vload<VX>(Fi.f[3] + off + i, TZZ[q]);
        vload<VX>(Fi.f[4] + off + i, TXZ[q]);
        const bool zl = z > 0, zr = z < g.nz - 1;
#pragma unroll
        for (int e = 0; e < VX; ++e) {
          const int x = x0 + e;
          const float vx_x = x > 0 ? VXr[q + 1][e] - (e == 0 ? vxl : VXr[q + 1][e - 1]) : 0.0f;
          const float vz_z = zr ? VZr[q + 1][e] - VZr[q][e] : 0.0f;
          const float vx_z = zl ? VXr[q + 1][e] - VXr[q][e] : 0.0f;
          const float vz_x = x < g.nx - 1 ? (e == VX - 1 ? vzr : VZr[q][e + 1]) - VZr[q][e] : 0.0f;
          TXX[q][e] = a[e] * (dti * (l2m[e] * vx_x + lam[e] * vz_z) + bm[e] * TXX[q][e]);
          TZZ[q][e] = a[e] * (dti * (l2m[e] * vz_z + lam[e] * vx_x) + bm[e] * TZZ[q][e]);
          TXZ[q][e] = a[e] * ((1.0f * mc[e]) * (vz_x + vx_z) + bm[e] * TXZ[q][e]);
        }
      } else {
        vzero<VX>(TXX[q]);
        vzero<VX>(TZZ[q]);
        vzero<VX>(TXZ[q]);
      }
    }
    // stage V on rows zb .. zb+RZ-1 from the new stresses; write both
#pragma unroll
    for (int r = 0; r < RZ; ++r) {
      const int z = zb + r;
      const int q = r + 1;
      const float txxr = __shfl_down(TXX[q][0], 1, kLanes);
      const float txzl = __shfl_up(TXZ[q][VX - 1], 1, kLanes);
      if (!inner || !xin || z >= g.nz) continue;
      const int i = z * g.nx + x0;
      float rc[VX], a[VX], bm[VX], vxo[VX], vzo[VX];
      cload(kRc, z, rc);
      cload(kAinv, z, a);
      cload(kBm, z, bm);
      const bool zl = z > 0, zr = z < g.nz - 1;
#pragma unroll
      for (int e = 0; e < VX; ++e) {
        const int x = x0 + e;
        const float txx_x = x < g.nx - 1 ? (e == VX - 1 ? txxr : TXX[q][e + 1]) - TXX[q][e] : 0.0f;
        const float txz_z = zr ? TXZ[q + 1][e] - TXZ[q][e] : 0.0f;
        const float tzz_z = zl ? TZZ[q][e] - TZZ[q - 1][e] : 0.0f;
        const float txz_x = x > 0 ? TXZ[q][e] - (e == 0 ? txzl : TXZ[q][e - 1]) : 0.0f;
        const float src = 1.0f * rc[e];
        vxo[e] = a[e] * (src * (txx_x + txz_z) + bm[e] * VXr[q + 1][e]);
        vzo[e] = a[e] * (src * (txz_x + tzz_z) + bm[e] * VZr[q][e]);
      }
      vstore<VX>(Fo.f[0] + off + i, vxo);
      vstore<VX>(Fo.f[1] + off + i, vzo);
      vstore<VX>(Fo.f[2] + off + i, TXX[q]);
      vstore<VX>(Fo.f[3] + off + i, TZZ[q]);
      vstore<VX>(Fo.f[4] + off + i, TXZ[q]);
    }
  }
}

// Rows per thread of the fused step.  One: with the coefficients in LDS a
// block of one-row threads holds 36.9 KB (4-cell lanes) and keeps 4 waves per
// SIMD; two rows per thread would need 61 KB and halve that.  495 -> 410 us
// per step at S = 2 on the 700x1900, 32-shot grid against two rows without
// the staging.
constexpr int kElRZ = 1;

template <int S, int VX>
__global__ __launch_bounds__(256) void k_el_step_rz(const float* __restrict__ cf, float dti,
                                                    CFields Fi, Fields Fo, int B, Grid g,
                                                    TileMap tm) {
  el_step_body<S, kElRZ, VX>(cf, dti, Fi, Fo, B, g, tm);
}

// One whole reverse step of the fused adjoint loop in one pass: the transpose
// of stage V, then of stage T (k_el_adjV then k_el_adjT over the whole grid,
// sign +1, gradients on the interior), with the reverse-time rebuild of the
// forward state.  Reads the cotangent of state n+1 (Xi) and state n+1 (Fi,
// whose velocities on the ring already hold state n); writes the cotangent
// of state n (Xo) and state n (Fo: rebuilt on the interior, copied
// elsewhere; velocities zero when recon == 0, state 0) and adds the
// coefficient gradients of its cells to the group's planes of gco.  Wave
// layout as el_step_body: a lane recomputes what it needs on the rows above
// and below, x-neighbours come from the adjacent lanes by shuffles, and only
// the 62 inner lanes write (what they take from lanes 0 and 63 depends on
// no cell outside the tile).  Every value uses the operands and order of the
// two stage kernels, so the results are bit-identical to them.
template <int S, int RZ, int VX>
__device__ __forceinline__ void el_adj_body(const float* __restrict__ cf, float dti, CFields Xi,
                                            Fields Xo, CFields Fi, Fields Fo,
                                            float* __restrict__ gco, int recon, int B, Grid g,
                                            TileMap tm) {
  int tx, tzi, grp;
  if (!tile_of(tm, tx, tzi, grp)) return;  // uniform over the block
  const int lane = (int)(threadIdx.x % kLanes);
  const int trow = (int)(threadIdx.x / kLanes);
  const int x0 = (tx * kElOut + lane - 1) * VX;
  const int zb0 = tzi * kRowsOfThreads * RZ;  // the block's first row
  const int zb = zb0 + trow * RZ;
  const bool xin = x0 >= 0 && x0 < g.nx;  // nx is a multiple of VX
  const bool inner = lane > 0 && lane < kLanes - 1;
  const int64_t P = g.plane;
  // the block's coefficient planes in LDS, rows zb0-1 .. zb0+4*RZ (local
  // row q = z - zb0 + 1), one lane vector per lane: read from HBM once per
  // block instead of once per shot.  Rows outside the grid are never read.
  constexpr int kCR = kRowsOfThreads * RZ + 2;
  __shared__ __align__(16) float lcf[kNumCoef][kCR][kLanes * VX];
  for (int q = trow; q < kCR; q += kRowsOfThreads) {
    const int z = zb0 - 1 + q;
    if (!xin || z < 0 || z >= g.nz) continue;
#pragma unroll
    for (int c = 0; c < kNumCoef; ++c) {
      float t[VX];
      vload<VX>(cf + c * P + (int64_t)z * g.nx + x0, t);
      vstore<VX>(&lcf[c][q][lane * VX], t);
    }
  }
  __syncthreads();
  if (zb >= g.nz) return;  // uniform over the wave, after the block's barrier
  // coefficient c on local row j (z = zb - 1 + j)
  auto cload = [&](int c, int j, float(&o)[VX]) {
    vload<VX>(&lcf[c][trow * RZ + j][lane * VX], o);
  };
  const float sgn = 1.0f;
  const float sdti = dti;
  // row j of the local arrays is z = zb - 1 + j
  auto rok = [&](int j) { return xin && zb - 1 + j >= 0 && zb - 1 + j < g.nz; };
  auto idx = [&](int j) { return (zb - 1 + j) * g.nx + x0; };
  float G0[RZ][VX], G1[RZ][VX], G2[RZ][VX], G3[RZ][VX];
#pragma unroll
  for (int r = 0; r < RZ; ++r) {
    vzero<VX>(G0[r]);
    vzero<VX>(G1[r]);
    vzero<VX>(G2[r]);
    vzero<VX>(G3[r]);
  }
  const int b0 = grp * S, nb = min(S, B - b0);
  for (int s = 0; s < nb; ++s) {
    const int64_t off = (int64_t)(b0 + s) * P;
    // velocity cotangents of state n+1 and their products with the stage-V
    // weight c = s*a*Rc, rows zb-1 .. zb+RZ
    float XV[RZ + 2][VX], ZV[RZ + 2][VX], PX[RZ + 2][VX], PZ[RZ + 2][VX];
#pragma unroll
    for (int j = 0; j < RZ + 2; ++j) {
      if (rok(j)) {
        const int i = idx(j);
        float a[VX], rc[VX];
        vload<VX>(Xi.f[0] + off + i, XV[j]);
        vload<VX>(Xi.f[1] + off + i, ZV[j]);
        cload(kAinv, j, a);
        cload(kRc, j, rc);
#pragma unroll
        for (int e = 0; e < VX; ++e) {
          const float c = sgn * a[e] * rc[e];
          PX[j][e] = c * XV[j][e];
          PZ[j][e] = c * ZV[j][e];
        }
      } else {
        vzero<VX>(XV[j]);
        vzero<VX>(ZV[j]);
        vzero<VX>(PX[j]);
        vzero<VX>(PZ[j]);
      }
    }
    // stress cotangents after the stage-V transpose, times a:
    // U1 = a*txx, U2 = a*tzz on rows zb-1 .. zb+RZ-1 (index j),
    // U3 = a*txz on rows zb .. zb+RZ (index j-1)
    float U1[RZ + 1][VX], U2[RZ + 1][VX], U3[RZ + 1][VX];
#pragma unroll
    for (int j = 0; j <= RZ; ++j) {
      const float pxl = __shfl_up(PX[j][VX - 1], 1, kLanes);
      if (rok(j)) {
        const int i = idx(j);
        const int z = zb - 1 + j;
        const bool zl = z >= 1, zr = z <= g.nz - 2;
        float t2[VX], t3[VX], a[VX];
        vload<VX>(Xi.f[2] + off + i, t2);
        vload<VX>(Xi.f[3] + off + i, t3);
        cload(kAinv, j, a);
#pragma unroll
        for (int e = 0; e < VX; ++e) {
          const int x = x0 + e;
          const bool xl = x >= 1, xr = x <= g.nx - 2;
          const float txm = xl ? (e == 0 ? pxl : PX[j][e - 1]) : 0.0f;
          const float tzzp = zr ? PZ[j + 1][e] : 0.0f;
          t2[e] = t2[e] + ((xl ? txm : 0.0f) - (xr ? PX[j][e] : 0.0f));
          t3[e] = t3[e] + ((zl ? PZ[j][e] : 0.0f) - (zr ? tzzp : 0.0f));
          U1[j][e] = a[e] * t2[e];
          U2[j][e] = a[e] * t3[e];
        }
      } else {
        vzero<VX>(U1[j]);
        vzero<VX>(U2[j]);
      }
    }
#pragma unroll
    for (int j = 1; j <= RZ + 1; ++j) {
      const float pzr = __shfl_down(PZ[j][0], 1, kLanes);
      if (rok(j)) {
        const int i = idx(j);
        const int z = zb - 1 + j;
        const bool zl = z >= 1, zr = z <= g.nz - 2;
        float t4[VX], a[VX];
        vload<VX>(Xi.f[4] + off + i, t4);
        cload(kAinv, j, a);
#pragma unroll
        for (int e = 0; e < VX; ++e) {
          const int x = x0 + e;
          const bool xl = x >= 1, xr = x <= g.nx - 2;
          const float txzm = zl ? PX[j - 1][e] : 0.0f;
          const float tzxp = xr ? (e == VX - 1 ? pzr : PZ[j][e + 1]) : 0.0f;
          t4[e] = t4[e] + ((zl ? txzm : 0.0f) - (zr ? PX[j][e] : 0.0f)) +
                  ((xl ? PZ[j][e] : 0.0f) - (xr ? tzxp : 0.0f));
          U3[j - 1][e] = a[e] * t4[e];
        }
      } else {
        vzero<VX>(U3[j - 1]);
      }
    }
    // stage-T transpose: q1 = sdti*(l2m*U1 + lam*U2) on the own rows,
    // q2 = sdti*(lam*U1 + l2m*U2) on rows zb-1 .. zb+RZ-1, q3 = s*Mc*U3 on
    // rows zb .. zb+RZ; then the velocity cotangents of state n
    float Q2[RZ + 1][VX], Q3[RZ + 1][VX];
#pragma unroll
    for (int j = 0; j <= RZ; ++j) {
      if (rok(j)) {
        const int i = idx(j);
        float l2m[VX], lam[VX];
        cload(kL2M, j, l2m);
        cload(kLam, j, lam);
#pragma unroll
        for (int e = 0; e < VX; ++e) Q2[j][e] = sdti * (lam[e] * U1[j][e] + l2m[e] * U2[j][e]);
      } else {
        vzero<VX>(Q2[j]);
      }
      if (rok(j + 1)) {
        float mc[VX];
        cload(kMc, j + 1, mc);
#pragma unroll
        for (int e = 0; e < VX; ++e) Q3[j][e] = sgn * mc[e] * U3[j][e];
      } else {
        vzero<VX>(Q3[j]);
      }
    }
#pragma unroll
    for (int r = 0; r < RZ; ++r) {
      const int j = r + 1;
      float Q1[VX];
      if (rok(j)) {
        const int i = idx(j);
        float l2m[VX], lam[VX];
        cload(kL2M, j, l2m);
        cload(kLam, j, lam);
#pragma unroll
        for (int e = 0; e < VX; ++e) Q1[e] = sdti * (l2m[e] * U1[j][e] + lam[e] * U2[j][e]);
      } else {
        vzero<VX>(Q1);
      }
      const float q1r = __shfl_down(Q1[0], 1, kLanes);
      const float q3l = __shfl_up(Q3[r][VX - 1], 1, kLanes);
      if (!inner || !rok(j)) continue;
      const int i = idx(j);
      const int z = zb + r;
      const bool zl = z >= 1, zr = z <= g.nz - 2;
      float a[VX], bm[VX], o0[VX], o1[VX], o2[VX], o3[VX], o4[VX];
      cload(kAinv, j, a);
      cload(kBm, j, bm);
#pragma unroll
      for (int e = 0; e < VX; ++e) {
        const int x = x0 + e;
        const bool xl = x >= 1, xr = x <= g.nx - 2;
        const float own = a[e] * bm[e];
        const float q1xp = xr ? (e == VX - 1 ? q1r : Q1[e + 1]) : 0.0f;
        const float q3zp = zr ? Q3[r + 1][e] : 0.0f;
        const float q2zm = zl ? Q2[j - 1][e] : 0.0f;
        const float q3xm = xl ? (e == 0 ? q3l : Q3[r][e - 1]) : 0.0f;
        o0[e] = own * XV[j][e];
        o1[e] = own * ZV[j][e];
        o0[e] = o0[e] + ((xl ? Q1[e] : 0.0f) - q1xp) + ((zl ? Q3[r][e] : 0.0f) - q3zp);
        o1[e] = o1[e] + (q2zm - (zr ? Q2[j][e] : 0.0f)) + (q3xm - (xr ? Q3[r][e] : 0.0f));
        o2[e] = bm[e] * U1[j][e];
        o3[e] = bm[e] * U2[j][e];
        o4[e] = bm[e] * U3[r][e];
      }
      vstore_nt<VX>(Xo.f[0] + off + i, o0);
      vstore_nt<VX>(Xo.f[1] + off + i, o1);
      vstore_nt<VX>(Xo.f[2] + off + i, o2);
      vstore_nt<VX>(Xo.f[3] + off + i, o3);
      vstore_nt<VX>(Xo.f[4] + off + i, o4);
    }
    // state n+1 stresses on rows zb-1 .. zb+RZ; the stage-V sums
    // sx = Dx+ txx + Dz+ txz (rows zb-1 .. zb+RZ-1) and sz = Dx- txz + Dz- tzz
    // (rows zb .. zb+RZ); velocities of state n: vx on rows zb-1 .. zb+RZ-1
    // (index j), vz on rows zb .. zb+RZ (index j-1)
    float SXX[RZ + 2][VX], SZZ[RZ + 2][VX], SXZ[RZ + 2][VX];
#pragma unroll
    for (int j = 0; j < RZ + 2; ++j) {
      if (rok(j)) {
        const int i = idx(j);
        vload<VX>(Fi.f[2] + off + i, SXX[j]);
        vload<VX>(Fi.f[3] + off + i, SZZ[j]);
        vload<VX>(Fi.f[4] + off + i, SXZ[j]);
      } else {
        vzero<VX>(SXX[j]);
        vzero<VX>(SZZ[j]);
        vzero<VX>(SXZ[j]);
      }
    }
    float SX[RZ + 1][VX], SZ[RZ + 1][VX], VXN[RZ + 1][VX], VZN[RZ + 1][VX];
#pragma unroll
    for (int j = 0; j <= RZ; ++j) {
      const float txxr = __shfl_down(SXX[j][0], 1, kLanes);
      const int z = zb - 1 + j;
      const bool zr = z < g.nz - 1;
#pragma unroll
      for (int e = 0; e < VX; ++e) {
        const int x = x0 + e;
        SX[j][e] = (x < g.nx - 1 ? (e == VX - 1 ? txxr : SXX[j][e + 1]) - SXX[j][e] : 0.0f) +
                   (zr ? SXZ[j + 1][e] - SXZ[j][e] : 0.0f);
      }
      if (rok(j) && recon) {
        const int i = idx(j);
        float rc[VX];
        vload<VX>(Fi.f[0] + off + i, VXN[j]);
        cload(kRc, j, rc);
#pragma unroll
        for (int e = 0; e < VX; ++e)
          if (seis::interior(z, x0 + e, g)) VXN[j][e] = VXN[j][e] - rc[e] * SX[j][e];
      } else {
        vzero<VX>(VXN[j]);
      }
    }
#pragma unroll
    for (int j = 1; j <= RZ + 1; ++j) {
      const float txzl = __shfl_up(SXZ[j][VX - 1], 1, kLanes);
      const int z = zb - 1 + j;
      const bool zl = z >= 1;
#pragma unroll
      for (int e = 0; e < VX; ++e) {
        const int x = x0 + e;
        SZ[j - 1][e] = (x > 0 ? SXZ[j][e] - (e == 0 ? txzl : SXZ[j][e - 1]) : 0.0f) +
                       (zl ? SZZ[j][e] - SZZ[j - 1][e] : 0.0f);
      }
      if (rok(j) && recon) {
        const int i = idx(j);
        float rc[VX];
        vload<VX>(Fi.f[1] + off + i, VZN[j - 1]);
        cload(kRc, j, rc);
#pragma unroll
        for (int e = 0; e < VX; ++e)
          if (seis::interior(z, x0 + e, g)) VZN[j - 1][e] = VZN[j - 1][e] - rc[e] * SZ[j - 1][e];
      } else {
        vzero<VX>(VZN[j - 1]);
      }
    }
    // own rows: coefficient gradients, the stress rebuild, state n out
#pragma unroll
    for (int r = 0; r < RZ; ++r) {
      const int j = r + 1;
      const float vxl = __shfl_up(VXN[j][VX - 1], 1, kLanes);
      const float vzr = __shfl_down(VZN[r][0], 1, kLanes);
      if (!inner || !rok(j)) continue;
      const int i = idx(j);
      const int z = zb + r;
      const bool zl = z >= 1, zr = z <= g.nz - 2;
      float l2m[VX], lam[VX], mc[VX], a[VX];
      cload(kL2M, j, l2m);
      cload(kLam, j, lam);
      cload(kMc, j, mc);
      cload(kAinv, j, a);
      float txx[VX], tzz[VX], txz[VX];
#pragma unroll
      for (int e = 0; e < VX; ++e) {
        const int x = x0 + e;
        txx[e] = SXX[j][e];
        tzz[e] = SZZ[j][e];
        txz[e] = SXZ[j][e];
        if (!seis::interior(z, x, g)) continue;
        G3[r][e] += a[e] * sgn * (SX[j][e] * XV[j][e] + SZ[r][e] * ZV[j][e]);
        const float vx_x = x > 0 ? VXN[j][e] - (e == 0 ? vxl : VXN[j][e - 1]) : 0.0f;
        const float vz_z = zr ? VZN[r + 1][e] - VZN[r][e] : 0.0f;
        const float vx_z = zl ? VXN[j][e] - VXN[j - 1][e] : 0.0f;
        const float vz_x = x < g.nx - 1 ? (e == VX - 1 ? vzr : VZN[r][e + 1]) - VZN[r][e] : 0.0f;
        G0[r][e] += sdti * (vx_x * U1[j][e] + vz_z * U2[j][e]);
        G1[r][e] += sdti * (vz_z * U1[j][e] + vx_x * U2[j][e]);
        G2[r][e] += sgn * (vz_x + vx_z) * U3[r][e];
        if (recon) {
          txx[e] = txx[e] - sdti * (l2m[e] * vx_x + lam[e] * vz_z);
          tzz[e] = tzz[e] - sdti * (l2m[e] * vz_z + lam[e] * vx_x);
          txz[e] = txz[e] - sgn * mc[e] * (vz_x + vx_z);
        }
      }
      vstore_nt<VX>(Fo.f[0] + off + i, VXN[j]);
      vstore_nt<VX>(Fo.f[1] + off + i, VZN[r]);
      vstore_nt<VX>(Fo.f[2] + off + i, txx);
      vstore_nt<VX>(Fo.f[3] + off + i, tzz);
      vstore_nt<VX>(Fo.f[4] + off + i, txz);
    }
  }
#pragma unroll
  for (int r = 0; r < RZ; ++r) {
    const int j = r + 1;
    if (!inner || !rok(j)) continue;
    // whole lane vectors: G0..G3 hold exact zeros outside the interior, and
    // adding them leaves those cells of this group's planes as they are
    float* Gp = gco + (int64_t)grp * 4 * P + idx(j);
    float t0[VX], t1[VX], t2[VX], t3[VX];
    vload<VX>(Gp, t0);
    vload<VX>(Gp + P, t1);
    vload<VX>(Gp + 2 * P, t2);
    vload<VX>(Gp + 3 * P, t3);
#pragma unroll
    for (int e = 0; e < VX; ++e) {
      t0[e] += G0[r][e];
      t1[e] += G1[r][e];
      t2[e] += G2[r][e];
      t3[e] += G3[r][e];
    }
    vstore<VX>(Gp, t0);
    vstore<VX>(Gp + P, t1);
    vstore<VX>(Gp + 2 * P, t2);
    vstore<VX>(Gp + 3 * P, t3);
  }
}

constexpr int kElAdjRZ = 1;  // rows per thread of the fused reverse step

// At least 4 waves per SIMD: the 4-cell-lane instances fit 128 VGPRs (135
// unconstrained, 3 waves; with the gradient partials added per element they
// spilled 12 bytes): 952 -> 924 us per step at S = 8 on the 700x1900, 32-shot
// grid (981 -> 986 at S = 4).
template <int S, int VX>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_el_adj_rz(const float* __restrict__ cf, float dti,
                                                   CFields Xi, Fields Xo, CFields Fi, Fields Fo,
                                                   float* __restrict__ gco, int recon, int B,
                                                   Grid g, TileMap tm) {
  el_adj_body<S, kElAdjRZ, VX>(cf, dti, Xi, Xo, Fi, Fo, gco, recon, B, g, tm);
}

// Transpose of stage V.  in: cotangents of the stage outputs; out: of its
// inputs.  fw: the stage's input fields (stresses used for the Rc gradient);
// gmode 0 none, 1 every region cell, 2 interior cells only; with recon the
// velocities in fw are rebuilt in place on the interior (the reverse of V).
// Each lane handles VX consecutive cells of a row; every cell's arithmetic is
// the scalar expression order.
template <int S, int VX>
__global__ __launch_bounds__(256) void k_el_adjV(const float* __restrict__ cf, float sgn,
                                                 Region rg, CFields in, Fields out, Fields fw,
                                                 float* __restrict__ gRc, int gmode, int recon,
                                                 int B, Grid g, TileMap tm) {
  int z, x0, grp;
  if (!vcell_of<VX>(tm, g, z, x0, grp)) return;
  const int i = z * g.nx + x0;
  const int64_t P = g.plane;
  const float* A = cf + kAinv * P;
  const float* R = cf + kRc * P;
  // s*a*Rc at (zz, xx), 0 outside the region
  auto aR = [&](int zz, int xx) {
    return in_region(zz, xx, rg) ? sgn * A[zz * g.nx + xx] * R[zz * g.nx + xx] : 0.0f;
  };
  const bool zl = z >= 1, zr = z <= g.nz - 2;
  // c at x0-1 .. x0+VX of row z (index e+1), and rows z-1, z+1
  float crow[VX + 2], cup[VX], cdn[VX], own[VX], rc[VX], ai[VX];
  bool grad[VX], rebuild[VX];
  bool anyg = false;
#pragma unroll
  for (int e = -1; e <= VX; ++e) {
    const int x = x0 + e;
    crow[e + 1] = (x >= 0 && x < g.nx) ? aR(z, x) : 0.0f;
  }
#pragma unroll
  for (int e = 0; e < VX; ++e) {
    const int x = x0 + e;
    cup[e] = zl ? aR(z - 1, x) : 0.0f;
    cdn[e] = zr ? aR(z + 1, x) : 0.0f;
    const bool reg = in_region(z, x, rg);
    own[e] = reg ? A[i + e] * cf[kBm * P + i + e] : 1.0f;
    rc[e] = R[i + e];
    ai[e] = A[i + e];
    grad[e] = gRc && reg && (gmode == 1 || seis::interior(z, x, g));
    rebuild[e] = recon && seis::interior(z, x, g);
    anyg = anyg || grad[e] || rebuild[e];
  }
  float gacc[VX];
  vzero<VX>(gacc);
  const int b0 = grp * S, nb = min(S, B - b0);
  for (int s = 0; s < nb; ++s) {
    const int64_t off = (int64_t)(b0 + s) * P + i;
    const float* avx = in.f[0] + off;
    const float* avz = in.f[1] + off;
    float vxc[VX], vzc[VX], vxu[VX], vzd[VX];
    vload<VX>(avx, vxc);
    vload<VX>(avz, vzc);
    if (zl) vload<VX>(avx - g.nx, vxu); else vzero<VX>(vxu);
    if (zr) vload<VX>(avz + g.nx, vzd); else vzero<VX>(vzd);
    const float vxl = x0 > 0 ? avx[-1] : 0.0f;
    const float vzr = x0 + VX < g.nx ? avz[VX] : 0.0f;
    float i2[VX], i3[VX], i4[VX], o0[VX], o1[VX];
    vload<VX>(in.f[2] + off, i2);
    vload<VX>(in.f[3] + off, i3);
    vload<VX>(in.f[4] + off, i4);
#pragma unroll
    for (int e = 0; e < VX; ++e) {
      const int x = x0 + e;
      const bool xl = x >= 1, xr = x <= g.nx - 2;
      const float c0 = crow[e + 1];
      const float tx0 = c0 * vxc[e];
      const float tz0 = c0 * vzc[e];
      const float txm = xl ? crow[e] * (e == 0 ? vxl : vxc[e - 1]) : 0.0f;
      const float txzm = zl ? cup[e] * vxu[e] : 0.0f;
      const float tzxp = xr ? crow[e + 2] * (e == VX - 1 ? vzr : vzc[e + 1]) : 0.0f;
      const float tzzp = zr ? cdn[e] * vzd[e] : 0.0f;
      i2[e] = i2[e] + ((xl ? txm : 0.0f) - (xr ? tx0 : 0.0f));
      i4[e] = i4[e] + ((zl ? txzm : 0.0f) - (zr ? tx0 : 0.0f)) +
              ((xl ? tz0 : 0.0f) - (xr ? tzxp : 0.0f));
      i3[e] = i3[e] + ((zl ? tz0 : 0.0f) - (zr ? tzzp : 0.0f));
      o0[e] = own[e] * vxc[e];
      o1[e] = own[e] * vzc[e];
    }
    vstore<VX>(out.f[2] + off, i2);
    vstore<VX>(out.f[4] + off, i4);
    vstore<VX>(out.f[3] + off, i3);
    vstore<VX>(out.f[0] + off, o0);
    vstore<VX>(out.f[1] + off, o1);
    if (anyg) {
      const float* txx = fw.f[2] + off;
      const float* tzz = fw.f[3] + off;
      const float* txz = fw.f[4] + off;
      float txxc[VX], tzzc[VX], txzc[VX], tzzu[VX], txzd[VX];
      vload<VX>(txx, txxc);
      vload<VX>(tzz, tzzc);
      vload<VX>(txz, txzc);
      if (zl) vload<VX>(tzz - g.nx, tzzu); else vzero<VX>(tzzu);
      if (zr) vload<VX>(txz + g.nx, txzd); else vzero<VX>(txzd);
      const float txxr = x0 + VX < g.nx ? txx[VX] : 0.0f;
      const float txzl = x0 > 0 ? txz[-1] : 0.0f;
      float fvx[VX], fvz[VX];
      vload<VX>(fw.f[0] + off, fvx);
      vload<VX>(fw.f[1] + off, fvz);
      bool anyr = false;
#pragma unroll
      for (int e = 0; e < VX; ++e) {
        const int x = x0 + e;
        const float sx = (x < g.nx - 1 ? (e == VX - 1 ? txxr : txxc[e + 1]) - txxc[e] : 0.0f) +
                         (zr ? txzd[e] - txzc[e] : 0.0f);
        const float sz = (x > 0 ? txzc[e] - (e == 0 ? txzl : txzc[e - 1]) : 0.0f) +
                         (zl ? tzzc[e] - tzzu[e] : 0.0f);
        if (grad[e]) gacc[e] += ai[e] * sgn * (sx * vxc[e] + sz * vzc[e]);
        if (rebuild[e]) {
          // forward vx' = a*(Rc*sx + b*vx) with a = b = 1 on the interior
          fvx[e] = fvx[e] - rc[e] * sx;
          fvz[e] = fvz[e] - rc[e] * sz;
          anyr = true;
        }
      }
      if (anyr) {
        vstore<VX>(fw.f[0] + off, fvx);
        vstore<VX>(fw.f[1] + off, fvz);
      }
    }
  }
#pragma unroll
  for (int e = 0; e < VX; ++e)
    if (grad[e]) gRc[(int64_t)grp * 4 * P + i + e] += gacc[e];
}

// Transpose of stage T (see k_el_adjV).  fw: the stage's input velocities
// (for the coefficient gradients) and, with recon, the stresses to rebuild in
// place on the interior (the reverse of T).  gradients: planes 0..2 of the
// group's 4-plane block (lam+2mu, lam, Mc).
template <int S, int VX>
__global__ __launch_bounds__(256) void k_el_adjT(const float* __restrict__ cf, float sdti,
                                                 float sgn, Region rg, CFields in, Fields out,
                                                 Fields fw, float* __restrict__ gco, int gmode,
                                                 int recon, int B, Grid g, TileMap tm) {
  int z, x0, grp;
  if (!vcell_of<VX>(tm, g, z, x0, grp)) return;
  const int i = z * g.nx + x0;
  const int64_t P = g.plane;
  const float* A = cf + kAinv * P;
  // per-neighbour coefficients of q1 = sdti*(l2m*u1 + lam*u2), q2 = sdti*(lam*u1
  // + l2m*u2), q3 = s*Mc*u3 with u = a*cotangent inside the region, 0 outside
  struct C3 { float a, l2m, lam, mc; };
  auto co = [&](int zz, int xx) {
    C3 r{0.0f, 0.0f, 0.0f, 0.0f};
    if (zz >= 0 && zz < g.nz && xx >= 0 && xx < g.nx && in_region(zz, xx, rg)) {
      const int j = zz * g.nx + xx;
      r = C3{A[j], cf[kL2M * P + j], cf[kLam * P + j], cf[kMc * P + j]};
    }
    return r;
  };
  const bool zl = z >= 1, zr = z <= g.nz - 2;
  C3 krow[VX + 2], kup[VX], kdn[VX];
#pragma unroll
  for (int e = -1; e <= VX; ++e) krow[e + 1] = co(z, x0 + e);
  float bm[VX];
  bool reg[VX], grad[VX], rebuild[VX];
  bool anyg = false;
#pragma unroll
  for (int e = 0; e < VX; ++e) {
    const int x = x0 + e;
    kup[e] = co(z - 1, x);
    kdn[e] = co(z + 1, x);
    reg[e] = in_region(z, x, rg);
    bm[e] = reg[e] ? cf[kBm * P + i + e] : 1.0f;
    grad[e] = gco && reg[e] && (gmode == 1 || seis::interior(z, x, g));
    rebuild[e] = recon && seis::interior(z, x, g);
    anyg = anyg || grad[e] || rebuild[e];
  }
  float g_l2m[VX], g_lam[VX], g_mc[VX];
  vzero<VX>(g_l2m);
  vzero<VX>(g_lam);
  vzero<VX>(g_mc);
  const int b0 = grp * S, nb = min(S, B - b0);
  for (int s = 0; s < nb; ++s) {
    const int64_t off = (int64_t)(b0 + s) * P + i;
    const float* axx = in.f[2] + off;
    const float* azz = in.f[3] + off;
    const float* axz = in.f[4] + off;
    float xxc[VX], zzc[VX], xzc[VX], xxu[VX], zzu[VX], xzd[VX];
    vload<VX>(axx, xxc);
    vload<VX>(azz, zzc);
    vload<VX>(axz, xzc);
    if (zl) { vload<VX>(axx - g.nx, xxu); vload<VX>(azz - g.nx, zzu); }
    else { vzero<VX>(xxu); vzero<VX>(zzu); }
    if (zr) vload<VX>(axz + g.nx, xzd); else vzero<VX>(xzd);
    const bool right = x0 + VX < g.nx;
    const float xxr = right ? axx[VX] : 0.0f;
    const float zzr = right ? azz[VX] : 0.0f;
    const float xzl = x0 > 0 ? axz[-1] : 0.0f;
    float i0[VX], i1[VX], o2[VX], o3[VX], o4[VX], u1[VX], u2[VX], u3[VX];
    vload<VX>(in.f[0] + off, i0);
    vload<VX>(in.f[1] + off, i1);
#pragma unroll
    for (int e = 0; e < VX; ++e) {
      const int x = x0 + e;
      const bool xl = x >= 1, xr = x <= g.nx - 2;
      const C3 k0 = krow[e + 1], kxp = krow[e + 2], kxm = krow[e];
      u1[e] = k0.a * xxc[e];
      u2[e] = k0.a * zzc[e];
      u3[e] = k0.a * xzc[e];
      const float q1 = sdti * (k0.l2m * u1[e] + k0.lam * u2[e]);
      const float q2 = sdti * (k0.lam * u1[e] + k0.l2m * u2[e]);
      const float q3 = sgn * k0.mc * u3[e];
      float q1xp = 0.0f, q3zp = 0.0f, q2zm = 0.0f, q3xm = 0.0f;
      if (xr) {
        const float ax1 = e == VX - 1 ? xxr : xxc[e + 1];
        const float az1 = e == VX - 1 ? zzr : zzc[e + 1];
        q1xp = sdti * (kxp.l2m * (kxp.a * ax1) + kxp.lam * (kxp.a * az1));
      }
      if (zr) q3zp = sgn * kdn[e].mc * (kdn[e].a * xzd[e]);
      if (zl) q2zm = sdti * (kup[e].lam * (kup[e].a * xxu[e]) + kup[e].l2m * (kup[e].a * zzu[e]));
      if (xl) q3xm = sgn * kxm.mc * (kxm.a * (e == 0 ? xzl : xzc[e - 1]));
      i0[e] = i0[e] + ((xl ? q1 : 0.0f) - q1xp) + ((zl ? q3 : 0.0f) - q3zp);
      i1[e] = i1[e] + (q2zm - (zr ? q2 : 0.0f)) + (q3xm - (xr ? q3 : 0.0f));
      if (reg[e]) {
        o2[e] = bm[e] * u1[e];
        o3[e] = bm[e] * u2[e];
        o4[e] = bm[e] * u3[e];
      } else {
        o2[e] = xxc[e];
        o3[e] = zzc[e];
        o4[e] = xzc[e];
      }
    }
    vstore<VX>(out.f[0] + off, i0);
    vstore<VX>(out.f[1] + off, i1);
    vstore<VX>(out.f[2] + off, o2);
    vstore<VX>(out.f[3] + off, o3);
    vstore<VX>(out.f[4] + off, o4);
    if (anyg) {
      const float* vx = fw.f[0] + off;
      const float* vz = fw.f[1] + off;
      float vxc[VX], vzc[VX], vxu[VX], vzd[VX];
      vload<VX>(vx, vxc);
      vload<VX>(vz, vzc);
      if (zl) vload<VX>(vx - g.nx, vxu); else vzero<VX>(vxu);
      if (zr) vload<VX>(vz + g.nx, vzd); else vzero<VX>(vzd);
      const float vxl = x0 > 0 ? vx[-1] : 0.0f;
      const float vzr = x0 + VX < g.nx ? vz[VX] : 0.0f;
      float txx[VX], tzz[VX], txz[VX];
      bool anyr = false;
#pragma unroll
      for (int e = 0; e < VX; ++e) anyr = anyr || rebuild[e];
      if (anyr) {
        vload<VX>(fw.f[2] + off, txx);
        vload<VX>(fw.f[3] + off, tzz);
        vload<VX>(fw.f[4] + off, txz);
      }
#pragma unroll
      for (int e = 0; e < VX; ++e) {
        const int x = x0 + e;
        const float vx_x = x > 0 ? vxc[e] - (e == 0 ? vxl : vxc[e - 1]) : 0.0f;
        const float vz_z = zr ? vzd[e] - vzc[e] : 0.0f;
        const float vx_z = zl ? vxc[e] - vxu[e] : 0.0f;
        const float vz_x = x < g.nx - 1 ? (e == VX - 1 ? vzr : vzc[e + 1]) - vzc[e] : 0.0f;
        if (grad[e]) {
          g_l2m[e] += sdti * (vx_x * u1[e] + vz_z * u2[e]);
          g_lam[e] += sdti * (vz_z * u1[e] + vx_x * u2[e]);
          g_mc[e] += sgn * (vz_x + vx_z) * u3[e];
        }
        if (rebuild[e]) {
          // forward txx' = a*(dti*(...) + b*txx) with a = b = 1 on the interior
          const C3 k0 = krow[e + 1];
          txx[e] = txx[e] - sdti * (k0.l2m * vx_x + k0.lam * vz_z);
          tzz[e] = tzz[e] - sdti * (k0.l2m * vz_z + k0.lam * vx_x);
          txz[e] = txz[e] - sgn * k0.mc * (vz_x + vx_z);
        }
      }
      if (anyr) {
        vstore<VX>(fw.f[2] + off, txx);
        vstore<VX>(fw.f[3] + off, tzz);
        vstore<VX>(fw.f[4] + off, txz);
      }
    }
  }
#pragma unroll
  for (int e = 0; e < VX; ++e) {
    if (!grad[e]) continue;
    float* G = gco + (int64_t)grp * 4 * P + i + e;
    G[0] += g_l2m[e];
    G[P] += g_lam[e];
    G[2 * P] += g_mc[e];
  }
}

// Coefficient gradients (summed over the G group blocks) -> dJ/dvp, dvs, drho.
__global__ __launch_bounds__(256) void k_el_chain(const float* __restrict__ gco, int G,
                                                  const float* __restrict__ vp,
                                                  const float* __restrict__ vs,
                                                  const float* __restrict__ rho, float dti,
                                                  float* __restrict__ gvp, float* __restrict__ gvs,
                                                  float* __restrict__ grho, int64_t plane) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= plane) return;
  float a = 0.0f, l = 0.0f, m = 0.0f, r = 0.0f;
  for (int q = 0; q < G; ++q) {
    const float* b = gco + (int64_t)q * 4 * plane + i;
    a += b[0];
    l += b[plane];
    m += b[2 * plane];
    r += b[3 * plane];
  }
  const float p = vp[i], s = vs[i], rh = rho[i];
  // lam+2mu = rho vp^2, lam = rho (vp^2 - 2 vs^2), Mc = dt/h rho vs^2, Rc = dt/(h rho)
  gvp[i] = 2.0f * rh * p * (a + l);
  gvs[i] = -4.0f * rh * s * l + 2.0f * rh * s * dti * m;
  grho[i] = p * p * a + (p * p - 2.0f * s * s) * l + s * s * dti * m - dti / (rh * rh) * r;
}

// After stage V of forward step n: add w[n] to the source fields, then (a
// separate launch) record receivers and store the rings of all five fields.
__global__ __launch_bounds__(256) void k_el_source(Fields F, const int32_t* __restrict__ src,
                                                   const float* __restrict__ wn, float sign,
                                                   int src_mask, int B, Grid g) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const int64_t i = (int64_t)b * g.plane + (int64_t)src[2 * b] * g.nx + src[2 * b + 1];
  const float w = sign * *wn;
  for (int f = 0; f < 5; ++f)
    if (src_mask & (1 << f)) F.f[f][i] += w;
}

// Receivers: rec_flat[r] (into (B, nz, nx)); channel c reads field rec_field[c].
__global__ __launch_bounds__(256) void k_el_record(CFields F, const int64_t* __restrict__ rec,
                                                   int64_t R, const int32_t* __restrict__ rfield,
                                                   int nch, float* __restrict__ rec_row,
                                                   float* __restrict__ ring_slot,
                                                   int64_t ring_len, int B, Grid g) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < R * nch) {
    const int64_t r = t / nch;
    const int c = (int)(t - r * nch);
    rec_row[t] = F.f[rfield[c]][rec[r]];
    return;
  }
  t -= R * nch;
  if (ring_slot && t < 5 * (int64_t)B * ring_len) {
    const int64_t per = (int64_t)B * ring_len;
    const int f = (int)(t / per);
    const int64_t rem = t - f * per;
    const int64_t b = rem / ring_len;
    const int64_t j = rem - b * ring_len;
    ring_slot[t] = F.f[f][b * g.plane + seis::ring_cell(j, g)];
  }
}

// Restore the rings of the fields in `mask` from a slot (5 x B x ring_len).
__global__ __launch_bounds__(256) void k_el_ring_restore(Fields F, const float* __restrict__ slot,
                                                         int mask, int64_t ring_len, int B,
                                                         Grid g) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per = (int64_t)B * ring_len;
  if (t >= 5 * per) return;
  const int f = (int)(t / per);
  if (!(mask & (1 << f))) return;
  const int64_t rem = t - f * per;
  const int64_t b = rem / ring_len;
  const int64_t j = rem - b * ring_len;
  F.f[f][b * g.plane + seis::ring_cell(j, g)] = slot ? slot[t] : 0.0f;
}

// Adjoint source: cotangents of the records of step n into the field adjoints.
__global__ __launch_bounds__(256) void k_el_inject(Fields adj, const int64_t* __restrict__ rec,
                                                   int64_t R, const int32_t* __restrict__ rfield,
                                                   int nch, const float* __restrict__ rgrad) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= R * nch) return;
  const int64_t r = t / nch;
  const int c = (int)(t - r * nch);
  atomicAdd(adj.f[rfield[c]] + rec[r], rgrad[t]);
}

__global__ __launch_bounds__(256) void k_el_fill(float* __restrict__ p, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0.0f;
}

__global__ __launch_bounds__(256) void k_el_copy(const float* __restrict__ a,
                                                 float* __restrict__ b, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = a[t];
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

inline int zero(float* p, int64_t n, hipStream_t s) {
  if (n <= 0) return 0;
  k_el_fill<<<blocks_for(n), 256, 0, s>>>(p, n);
  SEIS_CHECK_LAUNCH();
  return 0;
}

inline bool valid_s(int S) { return S == 1 || S == 2 || S == 4 || S == 8; }

#define SEIS_EL_VX(K, S_, VX_, grid, s, ...)                                  \
  switch (VX_) {                                                               \
    case 1: K<S_, 1><<<grid, 256, 0, s>>>(__VA_ARGS__); break;                \
    case 2: K<S_, 2><<<grid, 256, 0, s>>>(__VA_ARGS__); break;                \
    default: K<S_, 4><<<grid, 256, 0, s>>>(__VA_ARGS__); break;               \
  }

// Launch of a vectorised stage kernel K<S, VX>.
#define SEIS_EL_LAUNCH_V(K, S, VX, grid, s, ...)                             \
  do {                                                                      \
    switch (S) {                                                            \
      case 1: SEIS_EL_VX(K, 1, VX, grid, s, __VA_ARGS__) break;            \
      case 2: SEIS_EL_VX(K, 2, VX, grid, s, __VA_ARGS__) break;            \
      case 4: SEIS_EL_VX(K, 4, VX, grid, s, __VA_ARGS__) break;            \
      default: SEIS_EL_VX(K, 8, VX, grid, s, __VA_ARGS__) break;           \
    }                                                                       \
    SEIS_CHECK_LAUNCH();                                                    \
  } while (0)

// Widest lane vector (4, 2, 1 cells) that divides nx with every given
// pointer (coefficients, field sets of five, an extra plane) aligned to it.
int stage_vx(int64_t nx, const float* coef, std::initializer_list<float* const*> sets,
             const float* extra = nullptr) {
  for (int v : {4, 2}) {
    if (nx % v) continue;
    auto al = [v](const void* p) { return (reinterpret_cast<uintptr_t>(p) % (4 * v)) == 0; };
    bool ok = al(coef) && al(extra);
    for (float* const* f : sets)
      for (int k = 0; k < 5; ++k) ok = ok && al(f[k]);
    if (ok) return v;
  }
  return 1;
}

int check(int64_t B, int64_t nz, int64_t nx, float dt, float h, int S) {
  SEIS_CHECK_ARG(B >= 1 && B <= 65535, "batch size %lld outside [1, 65535]", (long long)B);
  SEIS_CHECK_ARG(nz >= 2 && nx >= 2, "grid %lld x %lld too small", (long long)nz, (long long)nx);
  SEIS_CHECK_ARG(nz * nx < (int64_t)1 << 31, "grid too large");
  SEIS_CHECK_ARG(dt > 0.0f && h > 0.0f, "dt and h must be positive");
  SEIS_CHECK_ARG(valid_s(S), "shots_per_group %d not in {1, 2, 4, 8}", S);
  return 0;
}

Fields fields_of(float* const* p) { return Fields{{p[0], p[1], p[2], p[3], p[4]}}; }
CFields cfields_of(float* const* p) { return CFields{{p[0], p[1], p[2], p[3], p[4]}}; }

}  // namespace

extern "C" {

int seis_elastic_coefficients(const float* vp, const float* vs, const float* rho, const float* d,
                              float* coef, int64_t nz, int64_t nx, float dt, float h,
                              void* stream) {
  SEIS_CHECK_ARG(vp && vs && rho && d && coef, "null pointer argument");
  SEIS_CHECK_ARG(nz >= 1 && nx >= 1 && dt > 0.0f && h > 0.0f, "bad grid or time step");
  const int64_t plane = nz * nx;
  k_el_coef<<<blocks_for(plane), 256, 0, (hipStream_t)stream>>>(vp, vs, rho, d, coef, dt,
                                                                1.0f / h, plane);
  SEIS_CHECK_LAUNCH();
  return 0;
}

int seis_elastic_stage(int32_t stage, const float* coef, float* const* fields, float sign,
                       int32_t z0, int32_t z1, int32_t x0, int32_t x1, int64_t B, int64_t nz,
                       int64_t nx, float dt, float h, int32_t shots_per_group, void* stream) {
  if (int e = check(B, nz, nx, dt, h, shots_per_group)) return e;
  SEIS_CHECK_ARG(coef && fields, "null pointer argument");
  for (int f = 0; f < 5; ++f) SEIS_CHECK_ARG(fields[f], "null field pointer");
  SEIS_CHECK_ARG(stage == 0 || stage == 1, "stage must be 0 (stresses) or 1 (velocities)");
  const Grid g = seis::make_grid(nz, nx, 0, false);
  const int S = shots_per_group;
  const int VX = stage_vx(nx, coef, {fields});
  const TileMap tm = make_tiles(g, VX, 1, (int)((B + S - 1) / S));
  const Region rg{z0, z1, x0, x1};
  hipStream_t s = (hipStream_t)stream;
  const dim3 grid(tm.per * kXcds);
  const float dti = dt * (1.0f / h);
  if (stage == 0)
    SEIS_EL_LAUNCH_V(k_el_T, S, VX, grid, s, coef, sign * dti, sign, rg, fields_of(fields),
                     (int)B, g, tm);
  else
    SEIS_EL_LAUNCH_V(k_el_V, S, VX, grid, s, coef, sign, rg, fields_of(fields), (int)B, g, tm);
  return 0;
}

int seis_elastic_stage_vjp(int32_t stage, const float* coef, float* const* cot_in,
                           float* const* cot_out, float* const* fwd, float* grad_coef,
                           float sign, int32_t z0, int32_t z1, int32_t x0, int32_t x1, int64_t B,
                           int64_t nz, int64_t nx, float dt, float h, int32_t shots_per_group,
                           void* stream) {
  if (int e = check(B, nz, nx, dt, h, shots_per_group)) return e;
  SEIS_CHECK_ARG(coef && cot_in && cot_out && fwd, "null pointer argument");
  for (int f = 0; f < 5; ++f)
    SEIS_CHECK_ARG(cot_in[f] && cot_out[f] && fwd[f], "null field pointer");
  SEIS_CHECK_ARG(stage == 0 || stage == 1, "stage must be 0 (stresses) or 1 (velocities)");
  const Grid g = seis::make_grid(nz, nx, 0, false);
  const int S = shots_per_group;
  const int VX = stage_vx(nx, coef, {cot_in, cot_out, fwd}, grad_coef);
  const TileMap tm = make_tiles(g, VX, 1, (int)((B + S - 1) / S));
  const Region rg{z0, z1, x0, x1};
  hipStream_t s = (hipStream_t)stream;
  const dim3 grid(tm.per * kXcds);
  const float dti = dt * (1.0f / h);
  const int gmode = grad_coef ? 1 : 0;
  if (stage == 0)
    SEIS_EL_LAUNCH_V(k_el_adjT, S, VX, grid, s, coef, sign * dti, sign, rg, cfields_of(cot_in),
                     fields_of(cot_out), fields_of(fwd), grad_coef, gmode, 0, (int)B, g, tm);
  else
    SEIS_EL_LAUNCH_V(k_el_adjV, S, VX, grid, s, coef, sign, rg, cfields_of(cot_in),
                     fields_of(cot_out), fields_of(fwd),
                     grad_coef ? grad_coef + 3 * g.plane : nullptr, gmode, 0, (int)B, g, tm);
  return 0;
}

int seis_elastic_chain(const float* grad_coef, int32_t groups, const float* vp, const float* vs,
                       const float* rho, float* gvp, float* gvs, float* grho, int64_t nz,
                       int64_t nx, float dt, float h, void* stream) {
  SEIS_CHECK_ARG(grad_coef && vp && vs && rho && gvp && gvs && grho, "null pointer argument");
  SEIS_CHECK_ARG(groups >= 1, "groups must be >= 1");
  const int64_t plane = nz * nx;
  k_el_chain<<<blocks_for(plane), 256, 0, (hipStream_t)stream>>>(
      grad_coef, groups, vp, vs, rho, dt * (1.0f / h), gvp, gvs, grho, plane);
  SEIS_CHECK_LAUNCH();
  return 0;
}

int seis_elastic_forward(const float* coef, const float* wavelet, int64_t nt,
                         const int32_t* src_zx, int32_t src_mask, const int64_t* rec_flat,
                         int64_t R, const int32_t* rec_field, int32_t nch, float* rec_out,
                         float* ring, float* snaps, int32_t snap_every, float* illum,
                         int32_t illum_field, float* const* fields, float* const* scratch,
                         int64_t B, int64_t nz, int64_t nx, float dt, float h, int32_t npml,
                         int32_t multiple, int32_t shots_per_group, void* stream) {
  if (int e = check(B, nz, nx, dt, h, shots_per_group)) return e;
  SEIS_CHECK_ARG(coef && wavelet && src_zx && fields, "null pointer argument");
  SEIS_CHECK_ARG(!illum || (illum_field >= 0 && illum_field < 5),
                 "illum_field %d outside [0, 5)", illum_field);
  for (int f = 0; f < 5; ++f) SEIS_CHECK_ARG(fields[f], "null field pointer");
  SEIS_CHECK_ARG(nt >= 1, "nt must be >= 1");
  SEIS_CHECK_ARG(R >= 0 && nch >= 1 && (R == 0 || (rec_flat && rec_field && rec_out)),
                 "bad receiver arguments");
  const int64_t ring_len = ring ? seis_ring_len(nz, nx, npml, multiple) : 0;
  SEIS_CHECK_ARG(!ring || ring_len > 0, "grid has no interior inside npml=%d", npml);
  SEIS_CHECK_ARG(!snaps || snap_every > 0, "snapshots need snap_every > 0");
  const Grid g = seis::make_grid(nz, nx, npml, multiple != 0);
  const int S = shots_per_group;
  if (scratch)
    for (int f = 0; f < 5; ++f) SEIS_CHECK_ARG(scratch[f], "null scratch field pointer");
  const int VX = scratch ? stage_vx(nx, coef, {fields, scratch}) : stage_vx(nx, coef, {fields});
  const TileMap tm = make_tiles(g, VX, 1, (int)((B + S - 1) / S));
  const Region all{0, (int)nz, 0, (int)nx};
  hipStream_t s = (hipStream_t)stream;
  const dim3 grid(tm.per * kXcds);
  const float dti = dt * (1.0f / h);
  const int64_t field = B * g.plane;
  // fused steps ping-pong between fields and scratch; the start is chosen so
  // that the last step writes `fields`
  Fields F = fields_of(fields);
  Fields Fs = scratch ? fields_of(scratch) : F;
  TileMap tm2 = make_tiles(g, VX, kElRZ, (int)((B + S - 1) / S));
  tm2.ntx = (g.nx + kElOut * VX - 1) / (kElOut * VX);
  tm2.per = (int)(((int64_t)tm2.ntx * tm2.ntz * tm2.groups + kXcds - 1) / kXcds);
  if (scratch && nt % 2 == 1) { Fields t = F; F = Fs; Fs = t; }
  for (int f = 0; f < 5; ++f)
    if (int e = zero(F.f[f], field, s)) return e;
  if (illum)
    if (int e = zero(illum, g.plane, s)) return e;
  for (int64_t n = 0; n < nt; ++n) {
    SEIS_PROF(SEIS_KERNEL_ELASTIC_FWD, s);
    if (scratch) {
      SEIS_EL_LAUNCH_V(k_el_step_rz, S, VX, dim3(tm2.per * kXcds), s, coef, dti,
                       CFields{{F.f[0], F.f[1], F.f[2], F.f[3], F.f[4]}}, Fs, (int)B, g, tm2);
      Fields t = F; F = Fs; Fs = t;
    } else {
      SEIS_EL_LAUNCH_V(k_el_T, S, VX, grid, s, coef, dti, 1.0f, all, F, (int)B, g, tm);
      SEIS_EL_LAUNCH_V(k_el_V, S, VX, grid, s, coef, 1.0f, all, F, (int)B, g, tm);
    }
    SEIS_PROF(SEIS_KERNEL_ELASTIC_FWD, s);
    k_el_source<<<blocks_for(B), 256, 0, s>>>(F, src_zx, wavelet + n, 1.0f, src_mask, (int)B, g);
    SEIS_CHECK_LAUNCH();
    if (illum)
      if (int e = illuminate(F.f[illum_field], (int)B, g.plane, illum, s)) return e;
    float* slot = (ring && n + 1 < nt) ? ring + n * 5 * B * ring_len : nullptr;
    const int64_t work = R * nch + (slot ? 5 * B * ring_len : 0);
    if (work > 0) {
      k_el_record<<<blocks_for(work), 256, 0, s>>>(CFields{{F.f[0], F.f[1], F.f[2], F.f[3], F.f[4]}},
                                                   rec_flat, R, rec_field, nch,
                                                   R ? rec_out + n * R * nch : nullptr, slot,
                                                   ring_len, (int)B, g);
      SEIS_CHECK_LAUNCH();
    }
    // snapshot of state n+1 at n+1 = j*snap_every, slot j-1
    const int64_t m = n + 1;
    if (snaps && m % snap_every == 0 && m <= nt - 1) {
      float* sp = snaps + (m / snap_every - 1) * 5 * field;
      for (int f = 0; f < 5; ++f) {
        k_el_copy<<<blocks_for(field), 256, 0, s>>>(F.f[f], sp + f * field, field);
        SEIS_CHECK_LAUNCH();
      }
    }
  }
  return 0;
}

int seis_elastic_backward(const float* coef, const float* wavelet, int64_t nt,
                          const int32_t* src_zx, int32_t src_mask, const int64_t* rec_flat,
                          int64_t R, const int32_t* rec_field, int32_t nch,
                          const float* rec_grad, const float* ring, const float* snaps,
                          int32_t snap_every, float* const* fields, float* const* adj_a,
                          float* const* adj_b, float* const* scratch, float* grad_coef, int64_t B,
                          int64_t nz, int64_t nx, float dt, float h, int32_t npml,
                          int32_t multiple, int32_t shots_per_group, void* stream) {
  if (int e = check(B, nz, nx, dt, h, shots_per_group)) return e;
  SEIS_CHECK_ARG(coef && wavelet && src_zx && fields && adj_a && adj_b && grad_coef,
                 "null pointer argument");
  for (int f = 0; f < 5; ++f)
    SEIS_CHECK_ARG(fields[f] && adj_a[f] && adj_b[f], "null field pointer");
  if (scratch)
    for (int f = 0; f < 5; ++f) SEIS_CHECK_ARG(scratch[f], "null scratch field pointer");
  SEIS_CHECK_ARG(nt >= 1, "nt must be >= 1");
  SEIS_CHECK_ARG(R >= 0 && nch >= 1 && (R == 0 || (rec_flat && rec_field && rec_grad)),
                 "bad receiver arguments");
  const int64_t ring_len = seis_ring_len(nz, nx, npml, multiple);
  SEIS_CHECK_ARG(ring_len > 0, "grid has no interior inside npml=%d", npml);
  SEIS_CHECK_ARG(ring || nt <= 1, "ring storage is required for nt > 1");
  SEIS_CHECK_ARG(!snaps || snap_every > 0, "snapshots need snap_every > 0");
  const Grid g = seis::make_grid(nz, nx, npml, multiple != 0);
  const int S = shots_per_group;
  const int G = (int)((B + S - 1) / S);
  const int VX = scratch ? stage_vx(nx, coef, {fields, adj_a, adj_b, scratch}, grad_coef)
                        : stage_vx(nx, coef, {fields, adj_a, adj_b}, grad_coef);
  const TileMap tm = make_tiles(g, VX, 1, G);
  const Region all{0, (int)nz, 0, (int)nx};
  hipStream_t s = (hipStream_t)stream;
  const dim3 grid(tm.per * kXcds);
  const float dti = dt * (1.0f / h);
  const int64_t field = B * g.plane;
  Fields F = fields_of(fields);
  Fields X = fields_of(adj_a), Y = fields_of(adj_b);
  for (int f = 0; f < 5; ++f) {
    if (int e = zero(X.f[f], field, s)) return e;
    if (int e = zero(Y.f[f], field, s)) return e;
  }
  if (int e = zero(grad_coef, (int64_t)G * 4 * g.plane, s)) return e;
  // state nt: remove the last source sample, inject the last records' cotangent
  k_el_source<<<blocks_for(B), 256, 0, s>>>(F, src_zx, wavelet + (nt - 1), -1.0f, src_mask,
                                            (int)B, g);
  SEIS_CHECK_LAUNCH();
  if (R > 0) {
    k_el_inject<<<blocks_for(R * nch), 256, 0, s>>>(X, rec_flat, R, rec_field, nch,
                                                    rec_grad + (nt - 1) * R * nch);
    SEIS_CHECK_LAUNCH();
  }
  const int64_t rwork = 5 * B * ring_len;
  if (scratch) {
    // fused reverse steps: X/Y and fields/scratch ping-pong
    Fields Fn = fields_of(scratch);
    TileMap tm2 = make_tiles(g, VX, kElAdjRZ, G);
    tm2.ntx = (g.nx + kElOut * VX - 1) / (kElOut * VX);
    tm2.per = (int)(((int64_t)tm2.ntx * tm2.ntz * tm2.groups + kXcds - 1) / kXcds);
    for (int64_t n = nt - 1; n >= 0; --n) {
      const int recon = n >= 1 ? 1 : 0;
      // velocities of state n on the ring (state n+1's are not read there)
      if (recon) {
        k_el_ring_restore<<<blocks_for(rwork), 256, 0, s>>>(
            F, ring + (n - 1) * 5 * B * ring_len, 0x3, ring_len, (int)B, g);
        SEIS_CHECK_LAUNCH();
      }
      SEIS_PROF(SEIS_KERNEL_ELASTIC_BWD, s);
      SEIS_EL_LAUNCH_V(k_el_adj_rz, S, VX, dim3(tm2.per * kXcds), s, coef, dti,
                       CFields{{X.f[0], X.f[1], X.f[2], X.f[3], X.f[4]}}, Y,
                       CFields{{F.f[0], F.f[1], F.f[2], F.f[3], F.f[4]}}, Fn, grad_coef, recon,
                       (int)B, g, tm2);
      SEIS_PROF(SEIS_KERNEL_ELASTIC_BWD, s);
      { Fields t = X; X = Y; Y = t; }
      { Fields t = F; F = Fn; Fn = t; }
      if (n >= 1) {
        k_el_ring_restore<<<blocks_for(rwork), 256, 0, s>>>(
            F, ring + (n - 1) * 5 * B * ring_len, 0x1c, ring_len, (int)B, g);
        SEIS_CHECK_LAUNCH();
        if (snaps && n % snap_every == 0 && n <= nt - 1) {
          const float* sp = snaps + (n / snap_every - 1) * 5 * field;
          for (int f = 0; f < 5; ++f) {
            k_el_copy<<<blocks_for(field), 256, 0, s>>>(sp + f * field, F.f[f], field);
            SEIS_CHECK_LAUNCH();
          }
        }
        k_el_source<<<blocks_for(B), 256, 0, s>>>(F, src_zx, wavelet + (n - 1), -1.0f,
                                                  src_mask, (int)B, g);
        SEIS_CHECK_LAUNCH();
        if (R > 0) {
          k_el_inject<<<blocks_for(R * nch), 256, 0, s>>>(X, rec_flat, R, rec_field, nch,
                                                          rec_grad + (n - 1) * R * nch);
          SEIS_CHECK_LAUNCH();
        }
      }
    }
    return 0;
  }
  for (int64_t n = nt - 1; n >= 0; --n) {
    // reverse of forward step n: fields hold state n+1 (stresses pre-source);
    // X holds the cotangent of state n+1.  recon rebuilds state n.
    const int recon = n >= 1 ? 1 : 0;
    SEIS_PROF(SEIS_KERNEL_ELASTIC_BWD, s);
    SEIS_EL_LAUNCH_V(k_el_adjV, S, VX, grid, s, coef, 1.0f, all,
                     CFields{{X.f[0], X.f[1], X.f[2], X.f[3], X.f[4]}}, Y, F,
                     grad_coef + 3 * g.plane, 2, recon, (int)B, g, tm);
    if (n == 0) {  // state 0 is zero: the stage-T gradient sees zero velocities
      if (int e = zero(F.f[0], field, s)) return e;
      if (int e = zero(F.f[1], field, s)) return e;
    }
    if (recon) {
      k_el_ring_restore<<<blocks_for(rwork), 256, 0, s>>>(F, ring + (n - 1) * 5 * B * ring_len,
                                                         0x3, ring_len, (int)B, g);
      SEIS_CHECK_LAUNCH();
    }
    SEIS_EL_LAUNCH_V(k_el_adjT, S, VX, grid, s, coef, dti, 1.0f, all,
                     CFields{{Y.f[0], Y.f[1], Y.f[2], Y.f[3], Y.f[4]}}, X, F, grad_coef, 2, recon,
                     (int)B, g, tm);
    SEIS_PROF(SEIS_KERNEL_ELASTIC_BWD, s);
    if (n >= 1) {
      k_el_ring_restore<<<blocks_for(rwork), 256, 0, s>>>(F, ring + (n - 1) * 5 * B * ring_len,
                                                         0x1c, ring_len, (int)B, g);
      SEIS_CHECK_LAUNCH();
      // reload the stored state n at snapshot times (bounds FP32 drift)
      if (snaps && n % snap_every == 0 && n <= nt - 1) {
        const float* sp = snaps + (n / snap_every - 1) * 5 * field;
        for (int f = 0; f < 5; ++f) {
          k_el_copy<<<blocks_for(field), 256, 0, s>>>(sp + f * field, F.f[f], field);
          SEIS_CHECK_LAUNCH();
        }
      }
      // state n stresses back to pre-source of step n-1, cotangent of records n-1
      k_el_source<<<blocks_for(B), 256, 0, s>>>(F, src_zx, wavelet + (n - 1), -1.0f, src_mask,
                                                (int)B, g);
      SEIS_CHECK_LAUNCH();
      if (R > 0) {
        k_el_inject<<<blocks_for(R * nch), 256, 0, s>>>(X, rec_flat, R, rec_field, nch,
                                                        rec_grad + (n - 1) * R * nch);
        SEIS_CHECK_LAUNCH();
      }
    }
  }
  return 0;
}

}  // extern "C"
