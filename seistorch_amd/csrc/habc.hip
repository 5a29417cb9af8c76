// Acoustic HABC (hybrid absorbing boundary) kernels for gfx950.
//
// One step (seistorch/equations2d/acoustic_habc.py:206-221):
//   y = (1/dt^-2)(2/dt^2 u - dt^-2 p + c^2 L u)   (u = u^n, p = u^{n-1})
// then in each 50-cell boundary strip s (top, bottom, left, right) with
// distance i from its outer edge and inward unit step e:
//   u1_s = (2-l-m) u + (l+2m) u[+e] - m u[+2e] + (l-1) p - l p[+e]
//   y_s  = b u1_s + (1-b) y,   l = 2c dt/h, m = c^2 dt^2/h^2   (:79-101)
// The cell takes y, then y_top, y_bottom, y_left, y_right where that strip's
// trapezoid mask holds (:167-183, masks of seistorch/habc.py:4-40), then the
// four diagonals take 0.5 y_s1 + 0.5 y_s2 (:185-202).  The reference works on
// 51-row strips shifted with torch.roll, so at strip row 49 the "u[+2e]" term
// is the strip's outer-edge cell; that matters where the blend weight of a
// row-49 cell is nonzero (the far diagonal ends) and is reproduced here.
//
// The adjoint is the exact transpose of that linear map in gather form: each
// cell collects the undamped-stencil terms weighted by the blend's (1-b) and
// the one-way terms from the up-to-three strip cells whose one-way stencil
// reaches it.  The forward uses three rotating field buffers (the blend reads
// u^{n-1} one cell inward, so it cannot overwrite it in place); the adjoint
// three rotating cotangent buffers for the same reason.
#include <cmath>

#include "common.h"
#include "illum.h"
#include "seistorch_amd.h"
#include "tiles.h"

#include <initializer_list>

using seis::Grid;
using seis::lap_taps;
using seis::acoustic_update;
using seis::lap5;
using namespace seis::tiles;

namespace {

constexpr int kW = 50;

struct HabcCtx {
  Grid g;
  int multiple;
  float dt, h;
};

__device__ __forceinline__ bool m_top(int z, int x, const HabcCtx& c) {
  return !c.multiple && z < kW && x >= z && x <= c.g.nx - 1 - z;
}
__device__ __forceinline__ bool m_bottom(int z, int x, const HabcCtx& c) {
  const int i = c.g.nz - 1 - z;
  return z >= c.g.nz - kW && x >= i && x <= c.g.nx - 1 - i;
}
__device__ __forceinline__ bool m_left(int z, int x, const HabcCtx& c) {
  return x < kW && ((z >= x && z <= c.g.nz - 1 - x) || (c.multiple && z < kW));
}
__device__ __forceinline__ bool m_right(int z, int x, const HabcCtx& c) {
  const int i = c.g.nx - 1 - x;
  return x >= c.g.nx - kW && ((z >= i && z <= c.g.nz - 1 - i) || (c.multiple && z < kW));
}

// strip s: 0 top, 1 bottom, 2 left, 3 right
__device__ __forceinline__ int sdist(int s, int z, int x, const Grid& g) {
  return s == 0 ? z : s == 1 ? g.nz - 1 - z : s == 2 ? x : g.nx - 1 - x;
}
__device__ __forceinline__ int sstep(int s, const Grid& g) {
  return s == 0 ? g.nx : s == 1 ? -g.nx : s == 2 ? 1 : -1;
}
__device__ __forceinline__ int sedge(int s, int z, int x, const Grid& g) {
  return s == 0 ? x : s == 1 ? (g.nz - 1) * g.nx + x : s == 2 ? z * g.nx : z * g.nx + g.nx - 1;
}
__device__ __forceinline__ bool near_edge(int z, int x, const Grid& g, int margin) {
  return z < margin || z >= g.nz - margin || x < margin || x >= g.nx - margin;
}

struct OwCoef {
  float k0, k1, k2, k3, k4;
};

// one-way coefficients with the reference's FP32 order: l = 2*c*dt/dh,
// m = c**2*dt**2/(dh**2), (2-l-m), (l+2m), -m, (l-1), -l
__device__ __forceinline__ OwCoef ow_coef(float c, const HabcCtx& x) {
  const float lam = ((2.0f * c) * x.dt) / x.h;
  const float mu = ((c * c) * (x.dt * x.dt)) / (x.h * x.h);
  return OwCoef{(2.0f - lam) - mu, lam + 2.0f * mu, -mu, lam - 1.0f, -lam};
}

// blended strip value y_s at cell i (flat) for fields u, p
__device__ __forceinline__ float strip_value(int s, int z, int x, int i, const float* u,
                                             const float* p, const OwCoef& k, float b, float y,
                                             const Grid& g) {
  const int e = sstep(s, g);
  const float u2 = sdist(s, z, x, g) == kW - 1 ? u[sedge(s, z, x, g)] : u[i + 2 * e];
  const float t1 = k.k0 * u[i];
  const float t2 = k.k1 * u[i + e];
  const float t3 = k.k2 * u2;
  const float t4 = k.k3 * p[i];
  const float t5 = k.k4 * p[i + e];
  const float one = (((t1 + t2) + t3) + t4) + t5;
  return one * b + (1.0f - b) * y;
}

// Linear weights of the final assignment at a cell: value = alpha*y +
// sum_s beta_s*u1_s.
struct HabcW {
  float alpha;
  float beta[4];
};

__device__ __forceinline__ HabcW habc_weights(int z, int x, const HabcCtx& c, float b) {
  HabcW w{1.0f, {0.0f, 0.0f, 0.0f, 0.0f}};
  const Grid& g = c.g;
  if (!near_edge(z, x, g, kW)) return w;
  int last = -1;
  if (m_top(z, x, c)) last = 0;
  if (m_bottom(z, x, c)) last = 1;
  if (m_left(z, x, c)) last = 2;
  if (m_right(z, x, c)) last = 3;
  int d1 = -1, d2 = -1;
  if (!c.multiple && z < kW && x == z) { d1 = 0; d2 = 2; }
  if (!c.multiple && z < kW && x == g.nx - kW + z) { d1 = 0; d2 = 3; }
  if (z >= g.nz - kW && x == z - (g.nz - kW)) { d1 = 1; d2 = 2; }
  if (z >= g.nz - kW && x == g.nx - kW + (z - (g.nz - kW))) { d1 = 1; d2 = 3; }
  if (d1 >= 0) {
    w.alpha = 0.5f * (1.0f - b) + 0.5f * (1.0f - b);
    w.beta[d1] = 0.5f * b;
    w.beta[d2] = 0.5f * b;
  } else if (last >= 0) {
    w.alpha = 1.0f - b;
    w.beta[last] = b;
  }
  return w;
}

// Forward step for S shots per thread: out = HABC(u, p) (+ w at the source
// cells when src != nullptr).  Cells are 64 x 4 tiles in XCD-aware order.
// Cells z in [z0, z1), x in [x0, x1) handled by a vectorised launch (see
// k_habc_fwd_deep / k_habc_adj_deep); the per-cell kernels skip them.
struct Deep {
  int z0, z1, x0, x1;
};

__device__ __forceinline__ bool in_deep(int z, int x, const Deep& dp) {
  return z >= dp.z0 && z < dp.z1 && x >= dp.x0 && x < dp.x1;
}

// Shot groups per frame thread: the per-cell set-up of the strips runs once
// per kFrameGroups groups of S shots.
constexpr int kFrameGroups = 8;

// Shot groups per thread of the adjoint's top and bottom bands.
constexpr int kBandGroups = 8;

// Cells per lane on the adjoint's bands: one, which divides every Deep lane
// width (Deep's column bounds are multiples of it).  Two and four cells per
// lane took 120 and 161 VGPRs in the fused kernel against 98.
constexpr int kBandVX = 1;

// Rows per thread of the vectorised adjoint on Deep in the fused kernel.  One: the deep and
// frame block ranges share one kernel, whose register count (and so the
// frame threads' occupancy) follows the larger branch; two rows per thread
// took 146 VGPRs against 91 and made the launch 20 % slower.
constexpr int kHabcRZ = 1;
// Shot groups per frame thread of the variable-density adjoint (its per-cell
// set-up, up to 16 neighbours' blend weights and one-way terms, runs once
// for them all).
constexpr int kRhoFrameGroups = 64;

// The cells outside Deep as up to four rectangles (top, bottom, left, right),
// tiled 64 x 4 each (one cell per thread) and numbered consecutively, so the
// per-cell launches only cover the strips.
struct Frame {
  int n;
  int z0[4], z1[4], x0[4], x1[4], ntx[4], ntz[4];
  int64_t t0[5];  // first global tile of each rectangle (prefix over the groups too)
  int groups, per;
};

// -> false if workgroup w (numbered within the frame's blocks) has no cell
__device__ __forceinline__ bool frame_cell(const Frame& f, int w, int& z, int& x, int& grp) {
  const int64_t t = (int64_t)(w % kXcds) * f.per + w / kXcds;
  if (t >= f.t0[f.n]) return false;
  int r = 0;
  while (r + 1 < f.n && t >= f.t0[r + 1]) ++r;
  const int64_t local = t - f.t0[r];
  const int tx = (int)(local % f.ntx[r]);
  const int tzi = (int)((local / f.ntx[r]) % f.ntz[r]);
  grp = (int)(local / ((int64_t)f.ntx[r] * f.ntz[r]));
  x = f.x0[r] + tx * kLanes + (int)(threadIdx.x % kLanes);
  z = f.z0[r] + tzi * kRowsOfThreads + (int)(threadIdx.x / kLanes);
  return x < f.x1[r] && z < f.z1[r];
}

template <int S>
__device__ __forceinline__ void habc_fwd_cell(int z, int x, int grp, const float* __restrict__ vp,
                                              const float* __restrict__ bw,
                                              const float* __restrict__ u,
                                              const float* __restrict__ p,
                                              float* __restrict__ out,
                                              const int32_t* __restrict__ src,
                                              const float* __restrict__ wn, int B,
                                              const HabcCtx& hc, const seis::TimeConsts& k,
                                              int ng = 1) {
  const Grid& g = hc.g;
  const int i = z * g.nx + x;
  const float c = vp[i];
  const float A0 = 1.0f / k.idt2;
  const float B0 = k.idt2;
  const bool edge = near_edge(z, x, g, kW);
  const float b = edge ? bw[i] : 0.0f;
  const OwCoef kc = ow_coef(c, hc);
  const bool mt = m_top(z, x, hc), mb = m_bottom(z, x, hc), ml = m_left(z, x, hc),
             mr = m_right(z, x, hc);
  int d1 = -1, d2 = -1;
  if (!hc.multiple && z < kW && x == z) { d1 = 0; d2 = 2; }
  if (!hc.multiple && z < kW && x == g.nx - kW + z) { d1 = 0; d2 = 3; }
  if (z >= g.nz - kW && x == z - (g.nz - kW)) { d1 = 1; d2 = 2; }
  if (z >= g.nz - kW && x == g.nx - kW + (z - (g.nz - kW))) { d1 = 1; d2 = 3; }
  const float w = src ? *wn : 0.0f;
  // groups grp .. grp+ng-1 of S shots
  const int b0 = grp * S, nb = min(S * ng, B - b0);
  for (int s = 0; s < nb; ++s) {
    const int bb = b0 + s;
    const int64_t off = (int64_t)bb * g.plane;
    const float* U = u + off;
    const float* P = p + off;
    const float L = lap5(U, z, x, i, g, k.ih2);
    const float y = acoustic_update(A0, B0, k.two_idt2, U[i], P[i], c * c, L);
    float v = y;
    if (edge) {
      if (mt) v = strip_value(0, z, x, i, U, P, kc, b, y, g);
      if (mb) v = strip_value(1, z, x, i, U, P, kc, b, y, g);
      if (ml) v = strip_value(2, z, x, i, U, P, kc, b, y, g);
      if (mr) v = strip_value(3, z, x, i, U, P, kc, b, y, g);
      if (d1 >= 0)
        v = 0.5f * strip_value(d1, z, x, i, U, P, kc, b, y, g) +
            0.5f * strip_value(d2, z, x, i, U, P, kc, b, y, g);
    }
    if (src && z == src[2 * bb] && x == src[2 * bb + 1]) v += w;
    out[off + i] = v;
  }
}

template <int S>
__global__ __launch_bounds__(256) void k_habc_fwd(const float* __restrict__ vp,
                                                  const float* __restrict__ bw,
                                                  const float* __restrict__ u,
                                                  const float* __restrict__ p,
                                                  float* __restrict__ out,
                                                  const int32_t* __restrict__ src,
                                                  const float* __restrict__ wn, int B,
                                                  HabcCtx hc, seis::TimeConsts k, TileMap tm) {
  int tx, tzi, grp;
  if (!tile_of(tm, tx, tzi, grp)) return;
  const Grid& g = hc.g;
  const int x = tx * kLanes + (int)(threadIdx.x % kLanes);
  const int z = tzi * kRowsOfThreads + (int)(threadIdx.x / kLanes);
  if (x >= g.nx || z >= g.nz) return;
  habc_fwd_cell<S>(z, x, grp, vp, bw, u, p, out, src, wn, B, hc, k);
}

// Variable-density acoustic step with the HABC strips
// (acoustic_rho_habc.py:31-57, Geiger and Daley 2003): with b = 1/rho and
// the face averages b_pos(x) = (b[x] + b[x+1]) / 2, b_neg(x) = (b[x-1] +
// b[x]) / 2 (per axis, b and u zero outside the grid),
//   y = (2u - p) + ((C * rho) * c^2) * (((bxp (u[x+1] - u) - bxn (u - u[x-1]))
//                                        + bzp (u[z+1] - u)) - bzn (u - u[z-1]))
// with C = dt^2 h^-2 formed in double on the host and rounded to FP32 (the
// reference's scalars are Python floats), then the same strip blend as
// habc_fwd_cell (blend = 0 skips it: the reverse step of
// acoustic_rho_habc.py:59-89).  One cell per thread for spt shots (a runtime
// count): the face averages, the strip set-up and their divisions are formed
// once per cell for all of them.
// 1/rho at (z, x), zero outside the grid; from the precomputed plane when
// the caller gives one (the same correctly rounded division)
__device__ __forceinline__ float buoyancy(const float* __restrict__ rho,
                                          const float* __restrict__ buoy, int z, int x,
                                          const Grid& g) {
  if (z < 0 || z >= g.nz || x < 0 || x >= g.nx) return 0.0f;
  const int j = z * g.nx + x;
  return buoy ? buoy[j] : 1.0f / rho[j];
}

__global__ __launch_bounds__(256) void k_reciprocal(const float* __restrict__ a,
                                                    float* __restrict__ r, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = 1.0f / a[i];
}

__global__ __launch_bounds__(256) void k_rho_habc_fwd(
    const float* __restrict__ vp, const float* __restrict__ rho, const float* __restrict__ bw,
    const float* __restrict__ buoy, const float* __restrict__ u, const float* __restrict__ p,
    float* __restrict__ out, const int32_t* __restrict__ src, const float* __restrict__ wn, int B,
    int spt, HabcCtx hc, float C, int blend, TileMap tm) {
  int tx, tzi, grp;
  if (!tile_of(tm, tx, tzi, grp)) return;
  const Grid& g = hc.g;
  const int x = tx * kLanes + (int)(threadIdx.x % kLanes);
  const int z = tzi * kRowsOfThreads + (int)(threadIdx.x / kLanes);
  if (x >= g.nx || z >= g.nz) return;
  const int i = z * g.nx + x;
  const float c = vp[i];
  const float bc = buoyancy(rho, buoy, z, x, g);
  const float bxp = (bc + buoyancy(rho, buoy, z, x + 1, g)) / 2.0f;
  const float bxn = (buoyancy(rho, buoy, z, x - 1, g) + bc) / 2.0f;
  const float bzp = (bc + buoyancy(rho, buoy, z + 1, x, g)) / 2.0f;
  const float bzn = (buoyancy(rho, buoy, z - 1, x, g) + bc) / 2.0f;
  const float coef = (C * rho[i]) * (c * c);
  const bool edge = blend && near_edge(z, x, g, kW);
  // the strip set-up only on the frame of strips
  float b = 0.0f;
  OwCoef kc{0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  bool mt = false, mb = false, ml = false, mr = false;
  int d1 = -1, d2 = -1;
  if (edge) {
    b = bw[i];
    kc = ow_coef(c, hc);
    mt = m_top(z, x, hc);
    mb = m_bottom(z, x, hc);
    ml = m_left(z, x, hc);
    mr = m_right(z, x, hc);
    if (!hc.multiple && z < kW && x == z) { d1 = 0; d2 = 2; }
    if (!hc.multiple && z < kW && x == g.nx - kW + z) { d1 = 0; d2 = 3; }
    if (z >= g.nz - kW && x == z - (g.nz - kW)) { d1 = 1; d2 = 2; }
    if (z >= g.nz - kW && x == g.nx - kW + (z - (g.nz - kW))) { d1 = 1; d2 = 3; }
  }
  const float w = src ? *wn : 0.0f;
  const int b0 = grp * spt, nb = min(spt, B - b0);
  for (int s = 0; s < nb; ++s) {
    const int bb = b0 + s;
    const int64_t off = (int64_t)bb * g.plane;
    const float* U = u + off;
    const float* P = p + off;
    const float uc = U[i];
    const float pxp = (x + 1 < g.nx ? U[i + 1] : 0.0f) - uc;
    const float pxn = uc - (x > 0 ? U[i - 1] : 0.0f);
    const float pzp = (z + 1 < g.nz ? U[i + g.nx] : 0.0f) - uc;
    const float pzn = uc - (z > 0 ? U[i - g.nx] : 0.0f);
    const float br = ((bxp * pxp - bxn * pxn) + bzp * pzp) - bzn * pzn;
    const float y = (2.0f * uc - P[i]) + coef * br;
    float v = y;
    if (edge) {
      if (mt) v = strip_value(0, z, x, i, U, P, kc, b, y, g);
      if (mb) v = strip_value(1, z, x, i, U, P, kc, b, y, g);
      if (ml) v = strip_value(2, z, x, i, U, P, kc, b, y, g);
      if (mr) v = strip_value(3, z, x, i, U, P, kc, b, y, g);
      if (d1 >= 0)
        v = 0.5f * strip_value(d1, z, x, i, U, P, kc, b, y, g) +
            0.5f * strip_value(d2, z, x, i, U, P, kc, b, y, g);
    }
    if (src && z == src[2 * bb] && x == src[2 * bb + 1]) v += w;
    out[off + i] = v;
  }
}

// The fused loop's HABC step in one launch of two block ranges: on Deep (at
// least kW cells from every edge, outside every strip) the undamped update
// only, vectorised over VX-cell lanes with the operands and order of
// k_habc_fwd's per-cell path (bit-identical); on the frame around Deep the
// per-cell step, one cell per thread.
template <int S, int VX>
__global__ __launch_bounds__(256) void k_habc_fwd_deep(const float* __restrict__ vp,
                                                       const float* __restrict__ u,
                                                       const float* __restrict__ p,
                                                       float* __restrict__ out,
                                                       const int32_t* __restrict__ src,
                                                       const float* __restrict__ wn, int B,
                                                       HabcCtx hc, seis::TimeConsts k, Deep dp,
                                                       TileMap tm, const float* __restrict__ bw,
                                                       Frame f, int ngf, int G) {
  // blocks [0, f.per*8) run the frame's cells, one cell per thread for all G
  // shot groups (the per-cell set-up once per step), the rest Deep's lane
  // vectors, so both kinds of work share the machine within one launch
  const int nf = f.per * kXcds;
  if ((int)blockIdx.x < nf) {
    int zc, xc, gc;
    if (frame_cell(f, (int)blockIdx.x, zc, xc, gc))
      habc_fwd_cell<S>(zc, xc, gc * ngf, vp, bw, u, p, out, src, wn, B, hc, k,
                       min(ngf, G - gc * ngf));
    return;
  }
  int tx, tzi, grp;
  if (!tile_of_w(tm, (int)blockIdx.x - nf, tx, tzi, grp)) return;
  const Grid& g = hc.g;
  const int x0 = dp.x0 + (tx * kLanes + (int)(threadIdx.x % kLanes)) * VX;
  const int z = dp.z0 + tzi * kRowsOfThreads + (int)(threadIdx.x / kLanes);
  if (x0 >= dp.x1 || z >= dp.z1) return;
  const int i = z * g.nx + x0;
  const float A0 = 1.0f / k.idt2;
  const float B0 = k.idt2;
  float c2[VX];
  vload<VX>(vp + i, c2);
#pragma unroll
  for (int e = 0; e < VX; ++e) c2[e] = c2[e] * c2[e];
  const float w = src ? *wn : 0.0f;
  const int b0 = grp * S, nb = min(S, B - b0);
  for (int s = 0; s < nb; ++s) {
    const int bb = b0 + s;
    const int64_t off = (int64_t)bb * g.plane;
    const float* U = u + off;
    float uu[VX], uup[VX], udn[VX], pp[VX], o[VX];
    vload<VX>(U + i, uu);
    vload<VX>(U + i - g.nx, uup);
    vload<VX>(U + i + g.nx, udn);
    vload<VX>(p + off + i, pp);
    const float ul = U[i - 1], ur = U[i + VX];
    const int sz = src ? src[2 * bb] : -1, sx = src ? src[2 * bb + 1] : -1;
#pragma unroll
    for (int e = 0; e < VX; ++e) {
      const float L = lap_taps(uup[e], e == 0 ? ul : uu[e - 1], uu[e],
                               e == VX - 1 ? ur : uu[e + 1], udn[e], k.ih2);
      float v = acoustic_update(A0, B0, k.two_idt2, uu[e], pp[e], c2[e], L);
      if (z == sz && x0 + e == sx) v += w;
      o[e] = v;
    }
    vstore<VX>(out + off + i, o);
  }
}

// One-way terms of the transposed strip blend that reach cell i: the strip
// cells n whose one-way stencil reads i, with the coefficients on u^n (c1)
// and on u^{n-1} (c2) times the blend weight beta of n.  At most 16.
__device__ __forceinline__ int oneway_terms(int z, int x, int i, const float* __restrict__ vp,
                                            const float* __restrict__ bw, const HabcCtx& hc,
                                            int* on, float* oc1, float* oc2) {
  const Grid& g = hc.g;
  int no = 0;
  for (int s = 0; s < 4; ++s) {
    const int dj = sdist(s, z, x, g);
    const int e = sstep(s, g);
    for (int m = 0; m < 3; ++m) {
      const int dn = dj - m;
      if (dn < 0 || dn > kW - 1 || (m == 2 && dn == kW - 1)) continue;
      const int n = i - m * e;
      const int nz_ = n / g.nx, nx_ = n - nz_ * g.nx;
      const float be = habc_weights(nz_, nx_, hc, bw[n]).beta[s];
      if (be == 0.0f) continue;
      const OwCoef kc = ow_coef(vp[n], hc);
      const float c1 = m == 0 ? kc.k0 : m == 1 ? kc.k1 : kc.k2;
      const float c2 = m == 0 ? kc.k3 : m == 1 ? kc.k4 : 0.0f;
      on[no] = n; oc1[no] = be * c1; oc2[no] = be * c2; ++no;
    }
    if (dj == 0) {  // the edge cell is the "+2" neighbour of the row-49 cell (roll wrap)
      const int n = i + (kW - 1) * e;
      const int nz_ = n / g.nx, nx_ = n - nz_ * g.nx;
      const float be = habc_weights(nz_, nx_, hc, bw[n]).beta[s];
      if (be != 0.0f) {
        on[no] = n; oc1[no] = be * ow_coef(vp[n], hc).k2; oc2[no] = 0.0f; ++no;
      }
    }
  }
  return no;
}

// Transpose of the step: out = H_u^T y1 + H_p^T y2 (either may be null).
// With recon/grad (fused adjoint, y1 = lambda^{k+1}, y2 = lambda^{k+2}):
// gradient 2c*A0*L(u^{k-1})*lambda^k on the interior (lambda^k without the
// adjoint source, which k_habc_post adds with its gradient share) and the
// interior rebuild of u^{k-2} over uc.  recon: 0 none, 1 rebuild, 2 zeros.
template <int S>
__device__ __forceinline__ void habc_adj_cell(
    int z, int x, int grp, const float* __restrict__ vp, const float* __restrict__ bw,
    const float* __restrict__ y1, const float* __restrict__ y2, float* __restrict__ out,
    const float* __restrict__ um, float* __restrict__ uc, float* __restrict__ gpart,
    const int32_t* __restrict__ src, const float* __restrict__ wn, int recon, int B,
    const HabcCtx& hc, const seis::TimeConsts& k, int ng = 1) {
  const Grid& g = hc.g;
  const int i = z * g.nx + x;
  const float A0 = 1.0f / k.idt2;
  const float B0 = k.idt2;
  const float c = vp[i];
  // undamped-stencil weights q(n) = c(n)^2*A0*alpha(n) at the cell and its neighbours
  const bool deep = !near_edge(z, x, g, kW + 3);
  auto qw = [&](int zz, int xx) {
    const int j = zz * g.nx + xx;
    const float cn = vp[j];
    const float al = deep ? 1.0f : habc_weights(zz, xx, hc, bw[j]).alpha;
    return (cn * cn) * A0 * al;
  };
  const float q0 = qw(z, x);
  const float qzm = z > 0 ? qw(z - 1, x) : 0.0f;
  const float qzp = z < g.nz - 1 ? qw(z + 1, x) : 0.0f;
  const float qxm = x > 0 ? qw(z, x - 1) : 0.0f;
  const float qxp = x < g.nx - 1 ? qw(z, x + 1) : 0.0f;
  const float al0 = deep ? 1.0f : habc_weights(z, x, hc, bw[i]).alpha;
  // one-way contributions: (flat offset, coefficient on y1, coefficient on y2)
  int on[16];
  float oc1[16], oc2[16];
  const int no = deep ? 0 : oneway_terms(z, x, i, vp, bw, hc, on, oc1, oc2);
  const bool in = interior(z, x, g);
  const float w = wn ? *wn : 0.0f;
  // groups grp .. grp+ng-1 of S shots, each with its own gradient partial
  for (int gi = grp; gi < grp + ng; ++gi) {
    float gacc = 0.0f;
    const int b0 = gi * S, nb = min(S, B - b0);
    for (int s = 0; s < nb; ++s) {
      const int bb = b0 + s;
      const int64_t off = (int64_t)bb * g.plane;
      float acc = 0.0f;
      if (y1) {
        const float* Y = y1 + off;
        const float Lq = lap_taps(z > 0 ? qzm * Y[i - g.nx] : 0.0f, x > 0 ? qxm * Y[i - 1] : 0.0f,
                                  q0 * Y[i], x < g.nx - 1 ? qxp * Y[i + 1] : 0.0f,
                                  z < g.nz - 1 ? qzp * Y[i + g.nx] : 0.0f, k.ih2);
        acc = (k.two_idt2 * (A0 * al0)) * Y[i] + Lq;
        for (int o = 0; o < no; ++o) acc += oc1[o] * Y[on[o]];
      }
      if (y2) {
        const float* Y = y2 + off;
        acc += -(B0 * (A0 * al0)) * Y[i];
        for (int o = 0; o < no; ++o) acc += oc2[o] * Y[on[o]];
      }
      out[off + i] = acc;
      if (in && recon >= 0 && um) {
        const float Lu = lap5(um + off, z, x, i, g, k.ih2);
        gacc += Lu * acc;
        if (recon == 1) {
          float un = uc[off + i];
          if (z == src[2 * bb] && x == src[2 * bb + 1]) un -= w;
          uc[off + i] = ((k.two_idt2 * um[off + i] + (c * c) * Lu) - un / A0) / B0;
        } else if (recon == 2) {
          uc[off + i] = 0.0f;
        }
      }
    }
    if (in && um && gpart) gpart[(int64_t)gi * g.plane + i] += 2.0f * c * A0 * gacc;
  }
}

// The adjoint on the top and bottom bands: the rows outside Deep over Deep's
// columns (x0 >= kW+3 and x0+VB <= nx-kW-3).  There only the top (bottom)
// strip's one-way stencil reaches a cell: from the cells 0, 1 and 2 rows
// outward in its own column and, on the edge row, from strip row kW-1 (the
// roll wrap), and which of those four terms exist depends on the row alone.
// So a lane takes VB consecutive cells of one row, every access is a row
// vector, and each cell sums the same operands in the same order as
// habc_adj_cell (bit-identical).  TOP: the top band, else the bottom band.
template <int S, int VB, bool TOP>
__device__ __forceinline__ void habc_adj_band(
    int z, int x0, int grp, int ng, const float* __restrict__ vp, const float* __restrict__ bw,
    const float* __restrict__ y1, const float* __restrict__ y2, float* __restrict__ out,
    const float* __restrict__ um, float* __restrict__ uc, float* __restrict__ gpart,
    const int32_t* __restrict__ src, const float* __restrict__ wn, int recon, int B,
    const HabcCtx& hc, const seis::TimeConsts& k) {
  const Grid& g = hc.g;
  constexpr int st = TOP ? 0 : 1;  // the strip
  constexpr int er = TOP ? 1 : -1;  // its inward step in rows
  const int dj = TOP ? z : g.nz - 1 - z;
  const int i0 = z * g.nx + x0;
  const float A0 = 1.0f / k.idt2;
  const float B0 = k.idt2;
  // one-way term t reads row trow[t]: t = 0, 1, 2 the cells t rows outward
  // (their stencils reach this row through u, u[+e], u[+2e]), t = 3 the wrap
  bool tv[4];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const int dn = dj - m;
    tv[m] = !(dn < 0 || dn > kW - 1 || (m == 2 && dn == kW - 1));
  }
  tv[3] = dj == 0;
  const int trow[4] = {z, z - er, z - 2 * er, z + (kW - 1) * er};
  const bool hz = z > 0, hzp = z < g.nz - 1;
  auto qw = [&](int zz, int xx) {
    const int j = zz * g.nx + xx;
    const float cn = vp[j];
    return (cn * cn) * A0 * habc_weights(zz, xx, hc, bw[j]).alpha;
  };
  float q0[VB], qzm[VB], qzp[VB], qxm[VB], qxp[VB], al0[VB], c[VB];
  float oc1[4][VB], oc2[4][VB];
  bool on[4][VB];
  bool in[VB];
  bool allin = true, anyin = false;
#pragma unroll
  for (int e = 0; e < VB; ++e) {
    const int x = x0 + e;
    c[e] = vp[i0 + e];
    q0[e] = qw(z, x);
    qzm[e] = hz ? qw(z - 1, x) : 0.0f;
    qzp[e] = hzp ? qw(z + 1, x) : 0.0f;
    qxm[e] = qw(z, x - 1);
    qxp[e] = qw(z, x + 1);
    al0[e] = habc_weights(z, x, hc, bw[i0 + e]).alpha;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      on[t][e] = false;
      oc1[t][e] = 0.0f;
      oc2[t][e] = 0.0f;
      if (!tv[t]) continue;
      const int n = trow[t] * g.nx + x;
      const float be = habc_weights(trow[t], x, hc, bw[n]).beta[st];
      if (be == 0.0f) continue;
      const OwCoef kc = ow_coef(vp[n], hc);
      on[t][e] = true;
      oc1[t][e] = be * (t == 0 ? kc.k0 : t == 1 ? kc.k1 : kc.k2);
      oc2[t][e] = t == 0 ? be * kc.k3 : t == 1 ? be * kc.k4 : 0.0f;
    }
    in[e] = interior(z, x, g);
    allin = allin && in[e];
    anyin = anyin || in[e];
  }
  const float w = wn ? *wn : 0.0f;
  for (int gi = grp; gi < grp + ng; ++gi) {
    float gacc[VB];
    vzero<VB>(gacc);
    const int b0 = gi * S, nb = min(S, B - b0);
    for (int s = 0; s < nb; ++s) {
      const int bb = b0 + s;
      const int64_t off = (int64_t)bb * g.plane;
      const float* Y = y1 + off;
      const float* Y2 = y2 + off;
      float yc[VB], yzm[VB], yzp[VB], acc[VB];
      vload<VB>(Y + i0, yc);
      if (hz) vload<VB>(Y + i0 - g.nx, yzm); else vzero<VB>(yzm);
      if (hzp) vload<VB>(Y + i0 + g.nx, yzp); else vzero<VB>(yzp);
      const float yl = Y[i0 - 1], yr = Y[i0 + VB];
#pragma unroll
      for (int e = 0; e < VB; ++e) {
        const float Lq = lap_taps(hz ? qzm[e] * yzm[e] : 0.0f,
                                  qxm[e] * (e == 0 ? yl : yc[e - 1]), q0[e] * yc[e],
                                  qxp[e] * (e == VB - 1 ? yr : yc[e + 1]),
                                  hzp ? qzp[e] * yzp[e] : 0.0f, k.ih2);
        acc[e] = (k.two_idt2 * (A0 * al0[e])) * yc[e] + Lq;
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        if (!tv[t]) continue;
        float yt[VB];
        if (t == 0) {
#pragma unroll
          for (int e = 0; e < VB; ++e) yt[e] = yc[e];
        } else if (t == 1) {
#pragma unroll
          for (int e = 0; e < VB; ++e) yt[e] = TOP ? yzm[e] : yzp[e];
        } else {
          vload<VB>(Y + trow[t] * g.nx + x0, yt);
        }
#pragma unroll
        for (int e = 0; e < VB; ++e)
          if (on[t][e]) acc[e] += oc1[t][e] * yt[e];
      }
      float y2c[VB];
      vload<VB>(Y2 + i0, y2c);
#pragma unroll
      for (int e = 0; e < VB; ++e) acc[e] += -(B0 * (A0 * al0[e])) * y2c[e];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        if (!tv[t]) continue;
        float yt[VB];
        if (t == 0) {
#pragma unroll
          for (int e = 0; e < VB; ++e) yt[e] = y2c[e];
        } else {
          vload<VB>(Y2 + trow[t] * g.nx + x0, yt);
        }
#pragma unroll
        for (int e = 0; e < VB; ++e)
          if (on[t][e]) acc[e] += oc2[t][e] * yt[e];
      }
      vstore<VB>(out + off + i0, acc);
      if (anyin && recon >= 0) {
        const float* U = um + off;
        float u0[VB], uzm[VB], uzp[VB], un[VB];
        vload<VB>(U + i0, u0);
        if (hz) vload<VB>(U + i0 - g.nx, uzm); else vzero<VB>(uzm);
        if (hzp) vload<VB>(U + i0 + g.nx, uzp); else vzero<VB>(uzp);
        const float ul = U[i0 - 1], ur = U[i0 + VB];
        if (recon == 1) vload<VB>(uc + off + i0, un);
        const int sz = src[2 * bb], sx = src[2 * bb + 1];
#pragma unroll
        for (int e = 0; e < VB; ++e) {
          if (!in[e]) continue;
          const float Lu = lap_taps(hz ? uzm[e] : 0.0f, e == 0 ? ul : u0[e - 1], u0[e],
                                    e == VB - 1 ? ur : u0[e + 1], hzp ? uzp[e] : 0.0f, k.ih2);
          gacc[e] += Lu * acc[e];
          if (recon == 1) {
            float v = un[e];
            if (z == sz && x0 + e == sx) v -= w;
            un[e] = ((k.two_idt2 * u0[e] + (c[e] * c[e]) * Lu) - v / A0) / B0;
          } else if (recon == 2) {
            un[e] = 0.0f;
          }
        }
        if (recon != 0) {
          if (allin) {
            vstore<VB>(uc + off + i0, un);
          } else {
#pragma unroll
            for (int e = 0; e < VB; ++e)
              if (in[e]) uc[off + i0 + e] = un[e];
          }
        }
      }
    }
    if (anyin) {
      float* gp = gpart + (int64_t)gi * g.plane + i0;
#pragma unroll
      for (int e = 0; e < VB; ++e)
        if (in[e]) gp[e] += 2.0f * c[e] * A0 * gacc[e];
    }
  }
}

// The adjoint on the left and right columns outside Deep, over Deep's rows
// (z in [kW+3, nz-kW-3)): only the left (right) strip's one-way stencil
// reaches a cell there, from the cells 0, 1 and 2 columns outward in its own
// row and, on the edge column, from strip column kW-1.  One cell per lane,
// lanes along x, so every term is a row access shifted by up to two cells;
// which terms exist depends on the column alone.  Same operands in the same
// order as habc_adj_cell (bit-identical).  LEFT: the left column, else the
// right one.
template <int S, bool LEFT>
__device__ __forceinline__ void habc_adj_side(
    int z, int x, int grp, int ng, const float* __restrict__ vp, const float* __restrict__ bw,
    const float* __restrict__ y1, const float* __restrict__ y2, float* __restrict__ out,
    const float* __restrict__ um, float* __restrict__ uc, float* __restrict__ gpart,
    const int32_t* __restrict__ src, const float* __restrict__ wn, int recon, int B,
    const HabcCtx& hc, const seis::TimeConsts& k) {
  const Grid& g = hc.g;
  constexpr int st = LEFT ? 2 : 3;  // the strip
  constexpr int ec = LEFT ? 1 : -1;  // its inward step in columns
  const int dj = LEFT ? x : g.nx - 1 - x;
  const int i = z * g.nx + x;
  const float A0 = 1.0f / k.idt2;
  const float B0 = k.idt2;
  bool tv[4];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const int dn = dj - m;
    tv[m] = !(dn < 0 || dn > kW - 1 || (m == 2 && dn == kW - 1));
  }
  tv[3] = dj == 0;
  const int toff[4] = {0, -ec, -2 * ec, (kW - 1) * ec};
  const bool hxm = x > 0, hxp = x < g.nx - 1;
  auto qw = [&](int zz, int xx) {
    const int j = zz * g.nx + xx;
    const float cn = vp[j];
    return (cn * cn) * A0 * habc_weights(zz, xx, hc, bw[j]).alpha;
  };
  const float c = vp[i];
  const float q0 = qw(z, x);
  const float qzm = qw(z - 1, x);
  const float qzp = qw(z + 1, x);
  const float qxm = hxm ? qw(z, x - 1) : 0.0f;
  const float qxp = hxp ? qw(z, x + 1) : 0.0f;
  const float al0 = habc_weights(z, x, hc, bw[i]).alpha;
  float oc1[4], oc2[4];
  bool on[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    on[t] = false;
    oc1[t] = 0.0f;
    oc2[t] = 0.0f;
    if (!tv[t]) continue;
    const int n = i + toff[t];
    const float be = habc_weights(z, x + toff[t], hc, bw[n]).beta[st];
    if (be == 0.0f) continue;
    const OwCoef kc = ow_coef(vp[n], hc);
    on[t] = true;
    oc1[t] = be * (t == 0 ? kc.k0 : t == 1 ? kc.k1 : kc.k2);
    oc2[t] = t == 0 ? be * kc.k3 : t == 1 ? be * kc.k4 : 0.0f;
  }
  const bool in = interior(z, x, g);
  const float w = wn ? *wn : 0.0f;
  for (int gi = grp; gi < grp + ng; ++gi) {
    float gacc = 0.0f;
    const int b0 = gi * S, nb = min(S, B - b0);
    for (int s = 0; s < nb; ++s) {
      const int bb = b0 + s;
      const int64_t off = (int64_t)bb * g.plane;
      const float* Y = y1 + off;
      const float* Y2 = y2 + off;
      const float yc = Y[i];
      const float yl = hxm ? Y[i - 1] : 0.0f;
      const float yr = hxp ? Y[i + 1] : 0.0f;
      const float Lq = lap_taps(qzm * Y[i - g.nx], hxm ? qxm * yl : 0.0f, q0 * yc,
                                hxp ? qxp * yr : 0.0f, qzp * Y[i + g.nx], k.ih2);
      float acc = (k.two_idt2 * (A0 * al0)) * yc + Lq;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        if (!on[t]) continue;
        const float yt = t == 0 ? yc : t == 1 ? (LEFT ? yl : yr) : Y[i + toff[t]];
        acc += oc1[t] * yt;
      }
      const float y2c = Y2[i];
      acc += -(B0 * (A0 * al0)) * y2c;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        if (!on[t]) continue;
        acc += oc2[t] * (t == 0 ? y2c : Y2[i + toff[t]]);
      }
      out[off + i] = acc;
      if (in && recon >= 0) {
        const float Lu = lap5(um + off, z, x, i, g, k.ih2);
        gacc += Lu * acc;
        if (recon == 1) {
          float un = uc[off + i];
          if (z == src[2 * bb] && x == src[2 * bb + 1]) un -= w;
          uc[off + i] = ((k.two_idt2 * um[off + i] + (c * c) * Lu) - un / A0) / B0;
        } else if (recon == 2) {
          uc[off + i] = 0.0f;
        }
      }
    }
    if (in) gpart[(int64_t)gi * g.plane + i] += 2.0f * c * A0 * gacc;
  }
}

template <int S>
__global__ __launch_bounds__(256) void k_habc_adj(
    const float* __restrict__ vp, const float* __restrict__ bw, const float* __restrict__ y1,
    const float* __restrict__ y2, float* __restrict__ out, const float* __restrict__ um,
    float* __restrict__ uc, float* __restrict__ gpart, const int32_t* __restrict__ src,
    const float* __restrict__ wn, int recon, int B, HabcCtx hc, seis::TimeConsts k,
    TileMap tm) {
  int tx, tzi, grp;
  if (!tile_of(tm, tx, tzi, grp)) return;
  const Grid& g = hc.g;
  const int x = tx * kLanes + (int)(threadIdx.x % kLanes);
  const int z = tzi * kRowsOfThreads + (int)(threadIdx.x / kLanes);
  if (x >= g.nx || z >= g.nz) return;
  habc_adj_cell<S>(z, x, grp, vp, bw, y1, y2, out, um, uc, gpart, src, wn, recon, B, hc, k);
}

// Variable-density adjoint (acoustic_rho_habc.py:31-57), exact transpose of
// the step y = (2u - p) + K*M(u) with K = (C*rho)*vp^2 and M the face-averaged
// operator of k_rho_habc_fwd, followed by the strip blend.  M is symmetric
// (the face (i, j) carries the same average in rows i and j; zero outside the
// grid), so with alpha the weight of y in the blended value:
//   out = H_u^T y1 + H_p^T y2,  H_u^T y1 = 2 alpha y1 + M(K alpha y1) + one-way
//   terms,  H_p^T y2 = -alpha y2 + one-way terms.
// One cell per thread for spt shots; the cell's set-up (face averages, K*alpha
// of its neighbours, one-way terms) runs once for them.
__global__ __launch_bounds__(256) void k_rho_adj(
    const float* __restrict__ vp, const float* __restrict__ rho, const float* __restrict__ buoy,
    const float* __restrict__ bw, const float* __restrict__ y1, const float* __restrict__ y2,
    float* __restrict__ out, int B, int spt, HabcCtx hc, float C, TileMap tm) {
  int tx, tzi, grp;
  if (!tile_of(tm, tx, tzi, grp)) return;
  const Grid& g = hc.g;
  const int x = tx * kLanes + (int)(threadIdx.x % kLanes);
  const int z = tzi * kRowsOfThreads + (int)(threadIdx.x / kLanes);
  if (x >= g.nx || z >= g.nz) return;
  const int i = z * g.nx + x;
  const float bc = buoyancy(rho, buoy, z, x, g);
  const float bxp = (bc + buoyancy(rho, buoy, z, x + 1, g)) / 2.0f;
  const float bxn = (buoyancy(rho, buoy, z, x - 1, g) + bc) / 2.0f;
  const float bzp = (bc + buoyancy(rho, buoy, z + 1, x, g)) / 2.0f;
  const float bzn = (buoyancy(rho, buoy, z - 1, x, g) + bc) / 2.0f;
  const bool deep = !near_edge(z, x, g, kW + 3);
  // K*alpha at the cell and its neighbours (0 outside the grid)
  auto ka = [&](int zz, int xx) {
    if (zz < 0 || zz >= g.nz || xx < 0 || xx >= g.nx) return 0.0f;
    const int j = zz * g.nx + xx;
    const float cn = vp[j];
    const float al = deep ? 1.0f : habc_weights(zz, xx, hc, bw[j]).alpha;
    return ((C * rho[j]) * (cn * cn)) * al;
  };
  const float k0 = ka(z, x), kxp = ka(z, x + 1), kxn = ka(z, x - 1), kzp = ka(z + 1, x),
              kzn = ka(z - 1, x);
  const float al0 = deep ? 1.0f : habc_weights(z, x, hc, bw[i]).alpha;
  int on[16];
  float oc1[16], oc2[16];
  const int no = deep ? 0 : oneway_terms(z, x, i, vp, bw, hc, on, oc1, oc2);
  const int b0 = grp * spt, nb = min(spt, B - b0);
  for (int s = 0; s < nb; ++s) {
    const int64_t off = (int64_t)(b0 + s) * g.plane;
    float acc = 0.0f;
    if (y1) {
      const float* Y = y1 + off;
      const float m0 = k0 * Y[i];
      const float mxp = x + 1 < g.nx ? kxp * Y[i + 1] : 0.0f;
      const float mxn = x > 0 ? kxn * Y[i - 1] : 0.0f;
      const float mzp = z + 1 < g.nz ? kzp * Y[i + g.nx] : 0.0f;
      const float mzn = z > 0 ? kzn * Y[i - g.nx] : 0.0f;
      const float Mm = ((bxp * (mxp - m0) + bxn * (mxn - m0)) + bzp * (mzp - m0)) +
                       bzn * (mzn - m0);
      acc = (2.0f * al0) * Y[i] + Mm;
      for (int o = 0; o < no; ++o) acc += oc1[o] * Y[on[o]];
    }
    if (y2) {
      const float* Y = y2 + off;
      acc += -al0 * Y[i];
      for (int o = 0; o < no; ++o) acc += oc2[o] * Y[on[o]];
    }
    out[off + i] = acc;
  }
}

// Adjoint source of the fused variable-density adjoint: lam[rec[t]] += s[t]
// (repeated receivers add).
__global__ __launch_bounds__(256) void k_rho_inject(float* __restrict__ lam,
                                                   const int64_t* __restrict__ rec,
                                                   const float* __restrict__ s, int64_t R) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < R) atomicAdd(lam + rec[t], s[t]);
}


// One reverse step of the variable-density adjoint in one launch.  With lam
// = lambda^k, the cotangent of u^k (k_rho_adj and k_rho_inject done): on the
// undamped interior, where alpha = 1 at the cell and its neighbours, the
// model gradient of the step u^k = F(u^{k-1}, u^{k-2}) and the reverse-time
// rebuild of u^{k-2} over uc (u^k on entry); on the ring, its restore.
//   dJ/dvp  += 2 C rho vp D lam,  D = M(u^{k-1})
//   dJ/drho += C vp^2 D lam - (1/(2 rho^2)) sum_j (u_j - u) (mu - mu_j),
//   mu = K lam  (the face average of 1/rho enters the rows of both its cells)
//   u^{k-2} = 2 u^{k-1} - (u^k - w - K D)   (w at the shot's source cell)
// gpart: [2][G][plane] (vp, then rho partials per group of spt shots).
// recon: 0 none, 1 rebuild + ring from ring_slot, 2 zeros (u^0).
// When out is given, the same thread also does k_rho_adj's work for step
// k-1 at its cell: out = lambda^{k-1} from lam and lam1 = lambda^{k+1}.  Both
// parts read lambda^k at the cell and its four neighbours (complete since
// the previous launch) and write disjoint buffers, so they share the face
// averages of 1/rho and the lambda^k loads; each output keeps the operands
// and order it has in k_rho_adj.
__device__ __forceinline__ void rho_adj_grad_cell(
    int z, int x, int grp, const float* __restrict__ vp, const float* __restrict__ rho,
    const float* __restrict__ buoy, const float* __restrict__ bw, const float* __restrict__ lam,
    const float* __restrict__ lam1, float* __restrict__ out, const float* __restrict__ um,
    float* __restrict__ uc, float* __restrict__ gpart, const int32_t* __restrict__ src,
    const float* __restrict__ wn, const float* __restrict__ ring_slot, int64_t ring_len,
    int recon, int B, int spt, int G, const HabcCtx& hc, float C, int ngrp) {
  const Grid& g = hc.g;
  const int i = z * g.nx + x;
  // shot groups grp .. grp+ngrp-1 (those below G), the set-up below once for all
  const int g_end = min(G, grp + ngrp);
  const bool in = seis::interior(z, x, g);
  if (!in) {  // off the interior: the ring restore of u^{k-2}
    const int64_t j = recon ? seis::ring_index(z, x, g) : -1;
    if (j >= 0)
      for (int bb = grp * spt; bb < min(B, g_end * spt); ++bb)
        uc[(int64_t)bb * g.plane + i] = recon == 1 ? ring_slot[bb * ring_len + j] : 0.0f;
    if (!out) return;
  }
  const bool doA = out != nullptr;
  const float bc = buoyancy(rho, buoy, z, x, g);
  const float bxp = (bc + buoyancy(rho, buoy, z, x + 1, g)) / 2.0f;
  const float bxn = (buoyancy(rho, buoy, z, x - 1, g) + bc) / 2.0f;
  const float bzp = (bc + buoyancy(rho, buoy, z + 1, x, g)) / 2.0f;
  const float bzn = (buoyancy(rho, buoy, z - 1, x, g) + bc) / 2.0f;
  // k_rho_adj's set-up: K*alpha at the cell and its neighbours, one-way terms
  const bool deep = !near_edge(z, x, g, kW + 3);
  auto ka = [&](int zz, int xx) {
    if (zz < 0 || zz >= g.nz || xx < 0 || xx >= g.nx) return 0.0f;
    const int j = zz * g.nx + xx;
    const float cn = vp[j];
    const float al = deep ? 1.0f : habc_weights(zz, xx, hc, bw[j]).alpha;
    return ((C * rho[j]) * (cn * cn)) * al;
  };
  float a0 = 0.0f, axp = 0.0f, axn = 0.0f, azp = 0.0f, azn = 0.0f, al0 = 1.0f;
  int on[16];
  float oc1[16], oc2[16];
  int no = 0;
  if (doA) {
    a0 = ka(z, x); axp = ka(z, x + 1); axn = ka(z, x - 1); azp = ka(z + 1, x);
    azn = ka(z - 1, x);
    al0 = deep ? 1.0f : habc_weights(z, x, hc, bw[i]).alpha;
    no = deep ? 0 : oneway_terms(z, x, i, vp, bw, hc, on, oc1, oc2);
  }
  // the gradient's set-up: K (no blend weight) at the cell and its neighbours
  auto kk = [&](int zz, int xx) {
    if (zz < 0 || zz >= g.nz || xx < 0 || xx >= g.nx) return 0.0f;
    const int j = zz * g.nx + xx;
    const float cn = vp[j];
    return (C * rho[j]) * (cn * cn);
  };
  const float c = vp[i];
  const float r = rho[i];
  float k0 = 0.0f, kxp = 0.0f, kxn = 0.0f, kzp = 0.0f, kzn = 0.0f;
  if (in) {
    k0 = (C * r) * (c * c);
    kxp = kk(z, x + 1); kxn = kk(z, x - 1); kzp = kk(z + 1, x); kzn = kk(z - 1, x);
  }
  const float w = wn ? *wn : 0.0f;
  for (; grp < g_end; ++grp) {
    const int b0 = grp * spt, nb = min(spt, B - b0);
    float gv = 0.0f, gb = 0.0f;
    for (int s = 0; s < nb; ++s) {
      const int bb = b0 + s;
      const int64_t off = (int64_t)bb * g.plane;
      const float* Lm = lam + off;
      const float l0 = Lm[i];
      const float lxp = x + 1 < g.nx ? Lm[i + 1] : 0.0f;
      const float lxn = x > 0 ? Lm[i - 1] : 0.0f;
      const float lzp = z + 1 < g.nz ? Lm[i + g.nx] : 0.0f;
      const float lzn = z > 0 ? Lm[i - g.nx] : 0.0f;
      if (in) {
        const float* U = um + off;
        const float u0 = U[i];
        const float pxp = (x + 1 < g.nx ? U[i + 1] : 0.0f) - u0;
        const float pxn = u0 - (x > 0 ? U[i - 1] : 0.0f);
        const float pzp = (z + 1 < g.nz ? U[i + g.nx] : 0.0f) - u0;
        const float pzn = u0 - (z > 0 ? U[i - g.nx] : 0.0f);
        const float D = ((bxp * pxp - bxn * pxn) + bzp * pzp) - bzn * pzn;
        gv += D * l0;
        const float m0 = k0 * l0;
        const float mxp = x + 1 < g.nx ? kxp * lxp : 0.0f;
        const float mxn = x > 0 ? kxn * lxn : 0.0f;
        const float mzp = z + 1 < g.nz ? kzp * lzp : 0.0f;
        const float mzn = z > 0 ? kzn * lzn : 0.0f;
        gb += ((pxp * (m0 - mxp) - pxn * (m0 - mxn)) + pzp * (m0 - mzp)) - pzn * (m0 - mzn);
        if (recon == 1) {
          float un = uc[off + i];
          if (z == src[2 * bb] && x == src[2 * bb + 1]) un -= w;
          uc[off + i] = 2.0f * u0 - (un - k0 * D);
        } else if (recon == 2) {
          uc[off + i] = 0.0f;
        }
      }
      if (doA) {
        const float* Y = lam + off;
        const float m0 = a0 * l0;
        const float mxp = x + 1 < g.nx ? axp * lxp : 0.0f;
        const float mxn = x > 0 ? axn * lxn : 0.0f;
        const float mzp = z + 1 < g.nz ? azp * lzp : 0.0f;
        const float mzn = z > 0 ? azn * lzn : 0.0f;
        const float Mm = ((bxp * (mxp - m0) + bxn * (mxn - m0)) + bzp * (mzp - m0)) +
                         bzn * (mzn - m0);
        float acc = (2.0f * al0) * l0 + Mm;
        for (int o = 0; o < no; ++o) acc += oc1[o] * Y[on[o]];
        const float* Y2 = lam1 + off;
        acc += -al0 * Y2[i];
        for (int o = 0; o < no; ++o) acc += oc2[o] * Y2[on[o]];
        out[off + i] = acc;
      }
    }
    if (in) {
      gpart[(int64_t)grp * g.plane + i] += ((2.0f * C) * r * c) * gv;
      gpart[((int64_t)G + grp) * g.plane + i] += (C * (c * c)) * gv - (0.5f * bc * bc) * gb;
    }
  }
}

// Deep tile of the variable-density reverse step (cells at least kW+3 from
// every edge: interior, no strip term, blend weight 1 at the cell and its
// neighbours): a lane takes VX consecutive cells of one row, with the face
// averages and K of its cells set up once and every field accessed as row
// vectors; each cell's values use the operands and order of
// rho_adj_grad_cell (bit-identical).
template <int VX>
__device__ __forceinline__ void rho_adj_grad_deep(
    int z, int x0, int grp, const float* __restrict__ vp, const float* __restrict__ rho,
    const float* __restrict__ buoy, const float* __restrict__ lam,
    const float* __restrict__ lam1, float* __restrict__ out, const float* __restrict__ um,
    float* __restrict__ uc, float* __restrict__ gpart, const int32_t* __restrict__ src,
    const float* __restrict__ wn, int recon, int B, int spt, int G, const HabcCtx& hc,
    float C) {
  const Grid& g = hc.g;
  const int i0 = z * g.nx + x0;
  const bool doA = out != nullptr;
  float bcv[VX], bxp[VX], bxn[VX], bzp[VX], bzn[VX], k0[VX], kxp[VX], kxn[VX], kzp[VX],
      kzn[VX], c[VX], r[VX];
  {
    // 1/rho on rows z-1, z, z+1 and at the vector's two ends; K = C rho vp^2
    float bm[VX], b0[VX], bp[VX], vm[VX], v0[VX], vpp[VX], rm[VX], rpp[VX];
    vload<VX>(buoy + i0 - g.nx, bm);
    vload<VX>(buoy + i0, b0);
    vload<VX>(buoy + i0 + g.nx, bp);
    vload<VX>(vp + i0 - g.nx, vm);
    vload<VX>(vp + i0, v0);
    vload<VX>(vp + i0 + g.nx, vpp);
    vload<VX>(rho + i0 - g.nx, rm);
    vload<VX>(rho + i0, r);
    vload<VX>(rho + i0 + g.nx, rpp);
    const float bl = buoy[i0 - 1], br = buoy[i0 + VX];
    const float vl = vp[i0 - 1], vr = vp[i0 + VX];
    const float rl = rho[i0 - 1], rr = rho[i0 + VX];
#pragma unroll
    for (int e = 0; e < VX; ++e) {
      c[e] = v0[e];
      bcv[e] = b0[e];
      const float bxr = e == VX - 1 ? br : b0[e + 1];
      const float bxl = e == 0 ? bl : b0[e - 1];
      bxp[e] = (b0[e] + bxr) / 2.0f;
      bxn[e] = (bxl + b0[e]) / 2.0f;
      bzp[e] = (b0[e] + bp[e]) / 2.0f;
      bzn[e] = (bm[e] + b0[e]) / 2.0f;
      k0[e] = (C * r[e]) * (c[e] * c[e]);
      const float cxr = e == VX - 1 ? vr : v0[e + 1], rxr = e == VX - 1 ? rr : r[e + 1];
      const float cxl = e == 0 ? vl : v0[e - 1], rxl = e == 0 ? rl : r[e - 1];
      kxp[e] = (C * rxr) * (cxr * cxr);
      kxn[e] = (C * rxl) * (cxl * cxl);
      kzp[e] = (C * rpp[e]) * (vpp[e] * vpp[e]);
      kzn[e] = (C * rm[e]) * (vm[e] * vm[e]);
    }
  }
  const float w = wn ? *wn : 0.0f;
  float gv[VX], gb[VX];
  vzero<VX>(gv);
  vzero<VX>(gb);
  const int b0 = grp * spt, nb = min(spt, B - b0);
  for (int s = 0; s < nb; ++s) {
    const int bb = b0 + s;
    const int64_t off = (int64_t)bb * g.plane;
    const float* Lm = lam + off;
    const float* U = um + off;
    float l0[VX], lzm[VX], lzp[VX], u0[VX], uzm[VX], uzp[VX], un[VX];
    vload<VX>(Lm + i0, l0);
    vload<VX>(Lm + i0 - g.nx, lzm);
    vload<VX>(Lm + i0 + g.nx, lzp);
    vload<VX>(U + i0, u0);
    vload<VX>(U + i0 - g.nx, uzm);
    vload<VX>(U + i0 + g.nx, uzp);
    const float ll = Lm[i0 - 1], lr = Lm[i0 + VX];
    const float ul = U[i0 - 1], ur = U[i0 + VX];
    if (recon == 1) vload<VX>(uc + off + i0, un);
    const int sz = src[2 * bb], sx = src[2 * bb + 1];
    float y2[VX], acc[VX];
    if (doA) vload<VX>(lam1 + off + i0, y2);
#pragma unroll
    for (int e = 0; e < VX; ++e) {
      const float lxp = e == VX - 1 ? lr : l0[e + 1];
      const float lxn = e == 0 ? ll : l0[e - 1];
      const float uxp = e == VX - 1 ? ur : u0[e + 1];
      const float uxn = e == 0 ? ul : u0[e - 1];
      const float pxp = uxp - u0[e];
      const float pxn = u0[e] - uxn;
      const float pzp = uzp[e] - u0[e];
      const float pzn = u0[e] - uzm[e];
      const float D = ((bxp[e] * pxp - bxn[e] * pxn) + bzp[e] * pzp) - bzn[e] * pzn;
      gv[e] += D * l0[e];
      const float m0 = k0[e] * l0[e];
      const float mxp = kxp[e] * lxp;
      const float mxn = kxn[e] * lxn;
      const float mzp = kzp[e] * lzp[e];
      const float mzn = kzn[e] * lzm[e];
      gb[e] += ((pxp * (m0 - mxp) - pxn * (m0 - mxn)) + pzp * (m0 - mzp)) - pzn * (m0 - mzn);
      if (recon == 1) {
        float v = un[e];
        if (z == sz && x0 + e == sx) v -= w;
        un[e] = 2.0f * u0[e] - (v - k0[e] * D);
      } else if (recon == 2) {
        un[e] = 0.0f;
      }
      if (doA) {
        // k_rho_adj with blend weight 1: K*1 = K, 2*1, -1
        const float Mm = ((bxp[e] * (mxp - m0) + bxn[e] * (mxn - m0)) + bzp[e] * (mzp - m0)) +
                         bzn[e] * (mzn - m0);
        float a = (2.0f * 1.0f) * l0[e] + Mm;
        a += -1.0f * y2[e];
        acc[e] = a;
      }
    }
    if (recon != 0) vstore<VX>(uc + off + i0, un);
    if (doA) vstore<VX>(out + off + i0, acc);
  }
  float* gp = gpart + (int64_t)grp * g.plane + i0;
  float* gq = gpart + ((int64_t)G + grp) * g.plane + i0;
  float pv[VX], pq[VX];
  vload<VX>(gp, pv);
  vload<VX>(gq, pq);
#pragma unroll
  for (int e = 0; e < VX; ++e) {
    pv[e] += ((2.0f * C) * r[e] * c[e]) * gv[e];
    pq[e] += (C * (c[e] * c[e])) * gv[e] - (0.5f * bcv[e] * bcv[e]) * gb[e];
  }
  vstore<VX>(gp, pv);
  vstore<VX>(gq, pq);
}

// One reverse step of the variable-density adjoint in one launch of two
// block ranges: the frame (the cells outside Deep) one cell per thread for ngf
// shot groups with rho_adj_grad_cell (its per-cell set-up once per ngf
// groups), then Deep's lane vectors with rho_adj_grad_deep.
template <int VX>
__global__ __launch_bounds__(256) void k_rho_adj_grad(
    const float* __restrict__ vp, const float* __restrict__ rho, const float* __restrict__ buoy,
    const float* __restrict__ bw, const float* __restrict__ lam, const float* __restrict__ lam1,
    float* __restrict__ out, const float* __restrict__ um, float* __restrict__ uc,
    float* __restrict__ gpart, const int32_t* __restrict__ src, const float* __restrict__ wn,
    const float* __restrict__ ring_slot, int64_t ring_len, int recon, int B, int spt, int G,
    HabcCtx hc, float C, Deep dp, Frame f, int ngf, TileMap tm) {
  const int nf = f.per * kXcds;
  if ((int)blockIdx.x < nf) {
    int z, x, fg;
    if (frame_cell(f, (int)blockIdx.x, z, x, fg))
      rho_adj_grad_cell(z, x, fg * ngf, vp, rho, buoy, bw, lam, lam1, out, um, uc, gpart, src,
                        wn, ring_slot, ring_len, recon, B, spt, G, hc, C, ngf);
    return;
  }
  int tx, tzi, grp;
  if (!tile_of_w(tm, (int)blockIdx.x - nf, tx, tzi, grp)) return;
  const int x0 = dp.x0 + (tx * kLanes + (int)(threadIdx.x % kLanes)) * VX;
  const int z = dp.z0 + tzi * kRowsOfThreads + (int)(threadIdx.x / kLanes);
  if (x0 >= dp.x1 || z >= dp.z1) return;
  rho_adj_grad_deep<VX>(z, x0, grp, vp, rho, buoy, lam, lam1, out, um, uc, gpart, src, wn,
                        recon, B, spt, G, hc, C);
}

// The fused adjoint step in one launch of two block ranges.  One covers the
// rectangle Deep (z in [z0, z1), x in [x0, x1), at least kW+3
// cells from every edge, x0 and x1 multiples of VX): there no strip term
// reaches, so a lane runs a vectorised copy of the per-cell deep path of
// habc_adj_cell with the same operands in the same order (bit-identical).
// The other runs habc_adj_cell on every other cell, one cell per thread.
// Requires y1, y2, um and gpart (the fused loop's call).
// Block w of the strip ranges of the adjoint (-> false if w is past them):
// [0, f.per*8) the frame's cells (the four corners outside Deep's rows and
// columns), one cell per thread for ngf shot groups (the per-cell set-up once
// per ngf groups); the next tb.per*8 the top and bottom bands in lane vectors
// for ngb groups; the next ts.per*8 the left and right columns over Deep's
// rows, one cell per lane for ngb groups.
template <int S>
__device__ __forceinline__ bool habc_adj_strip_block(
    int w, const float* __restrict__ vp, const float* __restrict__ bw,
    const float* __restrict__ y1, const float* __restrict__ y2, float* __restrict__ out, const float* __restrict__ um,
    float* __restrict__ uc, float* __restrict__ gpart, const int32_t* __restrict__ src,
    const float* __restrict__ wn, int recon, int B, const HabcCtx& hc,
    const seis::TimeConsts& k, const Deep& dp, const Frame& f, int ngf, int G,
    const TileMap& tb, int ngb, const TileMap& ts) {
  const int nf = f.per * kXcds;
  const int nbnd = tb.per * kXcds;
  const int nsd = ts.per * kXcds;
  const Grid& g = hc.g;
  if (w < nf) {
    int zc, xc, gc;
    if (frame_cell(f, w, zc, xc, gc))
      habc_adj_cell<S>(zc, xc, gc * ngf, vp, bw, y1, y2, out, um, uc, gpart, src, wn, recon, B,
                       hc, k, min(ngf, G - gc * ngf));
    return true;
  }
  if (w < nf + nbnd) {
    int tx, tzi, gb;
    if (!tile_of_w(tb, w - nf, tx, tzi, gb)) return true;
    constexpr int VB = kBandVX;
    const int x0 = dp.x0 + (tx * kLanes + (int)(threadIdx.x % kLanes)) * VB;
    const int r = tzi * kRowsOfThreads + (int)(threadIdx.x / kLanes);
    const int ztop = dp.z0;  // rows [0, dp.z0) then [dp.z1, nz)
    if (x0 >= dp.x1 || r >= ztop + (g.nz - dp.z1)) return true;
    const int grp = gb * ngb, ng = min(ngb, G - grp);
    if (r < ztop)
      habc_adj_band<S, VB, true>(r, x0, grp, ng, vp, bw, y1, y2, out, um, uc, gpart, src, wn,
                                 recon, B, hc, k);
    else
      habc_adj_band<S, VB, false>(dp.z1 + (r - ztop), x0, grp, ng, vp, bw, y1, y2, out, um, uc,
                                  gpart, src, wn, recon, B, hc, k);
    return true;
  }
  if (w < nf + nbnd + nsd) {
    // side tiles: tx 0 the left column (x = lane), tx 1 the right one
    // (x = dp.x1 + lane); both are narrower than 64 cells
    int tx, tzi, gb;
    if (!tile_of_w(ts, w - nf - nbnd, tx, tzi, gb)) return true;
    const int lane = (int)(threadIdx.x % kLanes);
    const int z = dp.z0 + tzi * kRowsOfThreads + (int)(threadIdx.x / kLanes);
    const int x = tx == 0 ? lane : dp.x1 + lane;
    if (z >= dp.z1 || (tx == 0 ? x >= dp.x0 : x >= g.nx)) return true;
    const int grp = gb * ngb, ng = min(ngb, G - grp);
    if (tx == 0)
      habc_adj_side<S, true>(z, x, grp, ng, vp, bw, y1, y2, out, um, uc, gpart, src, wn, recon,
                             B, hc, k);
    else
      habc_adj_side<S, false>(z, x, grp, ng, vp, bw, y1, y2, out, um, uc, gpart, src, wn, recon,
                              B, hc, k);
    return true;
  }
  return false;
}

// One tile of Deep (z in [z0, z1), x in [x0, x1), at least kW+3 cells from
// every edge, x0 and x1 multiples of VX): there no strip term reaches, so a
// lane runs a vectorised copy of the per-cell deep path of habc_adj_cell with
// the same operands in the same order (bit-identical), for RZ rows.
template <int S, int VX, int RZ>
__device__ __forceinline__ void habc_adj_deep_tile(
    int tx, int tzi, int grp, const float* __restrict__ vp, const float* __restrict__ bw,
    const float* __restrict__ y1, const float* __restrict__ y2, float* __restrict__ out, const float* __restrict__ um,
    float* __restrict__ uc, float* __restrict__ gpart, const int32_t* __restrict__ src,
    const float* __restrict__ wn, int recon, int B, const HabcCtx& hc,
    const seis::TimeConsts& k, const Deep& dp) {
  const Grid& g = hc.g;
  const int x0 = dp.x0 + (tx * kLanes + (int)(threadIdx.x % kLanes)) * VX;
  const int zb = dp.z0 + (tzi * kRowsOfThreads + (int)(threadIdx.x / kLanes)) * RZ;
  if (x0 >= dp.x1 || zb >= dp.z1) return;
  // every cell of Deep and its neighbours are interior, away from the strips;
  // a thread owns rows zb .. zb+RZ-1 (those below dp.z1) and reads rows
  // zb-1 .. zb+RZ once per shot for all of them
  const int nr = min(RZ, dp.z1 - zb);
  const float A0 = 1.0f / k.idt2;
  const float B0 = k.idt2;
  // q = c^2*A0 on rows zb-1 .. zb+RZ (index p = z - zb + 1) and at the
  // vector's two x-ends of the own rows; c on the own rows
  float qcol[RZ + 2][VX], qL[RZ], qR[RZ], c[RZ][VX];
#pragma unroll
  for (int p = 0; p < RZ + 2; ++p) {
    const int z = min(zb - 1 + p, dp.z1);  // rows past dp.z1 + 1 are never used
    float t[VX];
    vload<VX>(vp + z * g.nx + x0, t);
#pragma unroll
    for (int e = 0; e < VX; ++e) qcol[p][e] = (t[e] * t[e]) * A0 * 1.0f;
  }
#pragma unroll
  for (int r = 0; r < RZ; ++r) {
    const int i = min(zb + r, dp.z1 - 1) * g.nx + x0;
    vload<VX>(vp + i, c[r]);
    const float cl = vp[i - 1], cr = vp[i + VX];
    qL[r] = (cl * cl) * A0 * 1.0f;
    qR[r] = (cr * cr) * A0 * 1.0f;
  }
  const float ca = k.two_idt2 * (A0 * 1.0f);
  const float cb = -(B0 * (A0 * 1.0f));
  const float w = wn ? *wn : 0.0f;
  float gacc[RZ][VX];
#pragma unroll
  for (int r = 0; r < RZ; ++r) vzero<VX>(gacc[r]);
  const int b0 = grp * S, nb = min(S, B - b0);
  for (int s = 0; s < nb; ++s) {
    const int bb = b0 + s;
    const int64_t off = (int64_t)bb * g.plane;
    const float* Y = y1 + off;
    const float* U = um + off;
    const int sz = src[2 * bb], sx = src[2 * bb + 1];
    float ycol[RZ + 2][VX], ucol[RZ + 2][VX];
#pragma unroll
    for (int p = 0; p < RZ + 2; ++p) {
      const int z = min(zb - 1 + p, dp.z1);
      vload<VX>(Y + z * g.nx + x0, ycol[p]);
      if (recon >= 0) vload<VX>(U + z * g.nx + x0, ucol[p]);
    }
#pragma unroll
    for (int r = 0; r < RZ; ++r) {
      if (r >= nr) break;
      const int z = zb + r;
      const int i = z * g.nx + x0;
      const int p = r + 1;
      float y2c[VX], acc[VX];
      vload<VX>(y2 + off + i, y2c);
      const float yl = Y[i - 1], yr = Y[i + VX];
#pragma unroll
      for (int e = 0; e < VX; ++e) {
        const float left = e == 0 ? qL[r] * yl : qcol[p][e - 1] * ycol[p][e - 1];
        const float right = e == VX - 1 ? qR[r] * yr : qcol[p][e + 1] * ycol[p][e + 1];
        const float Lq = lap_taps(qcol[p - 1][e] * ycol[p - 1][e], left, qcol[p][e] * ycol[p][e],
                                  right, qcol[p + 1][e] * ycol[p + 1][e], k.ih2);
        float a = ca * ycol[p][e] + Lq;
        a += cb * y2c[e];
        acc[e] = a;
      }
      vstore<VX>(out + off + i, acc);
      if (recon >= 0) {
        float un[VX];
        const float ul = U[i - 1], ur = U[i + VX];
        if (recon == 1) vload<VX>(uc + off + i, un);
#pragma unroll
        for (int e = 0; e < VX; ++e) {
          const float Lu = lap_taps(ucol[p - 1][e], e == 0 ? ul : ucol[p][e - 1], ucol[p][e],
                                    e == VX - 1 ? ur : ucol[p][e + 1], ucol[p + 1][e], k.ih2);
          gacc[r][e] += Lu * acc[e];
          if (recon == 1) {
            float v = un[e];
            if (z == sz && x0 + e == sx) v -= w;
            un[e] = ((k.two_idt2 * ucol[p][e] + (c[r][e] * c[r][e]) * Lu) - v / A0) / B0;
          } else if (recon == 2) {
            un[e] = 0.0f;
          }
        }
        if (recon != 0) vstore<VX>(uc + off + i, un);
      }
    }
  }
#pragma unroll
  for (int r = 0; r < RZ; ++r) {
    if (r >= nr) break;
    const int i = (zb + r) * g.nx + x0;
    float gv[VX];
    vload<VX>(gpart + (int64_t)grp * g.plane + i, gv);
#pragma unroll
    for (int e = 0; e < VX; ++e) gv[e] += 2.0f * c[r][e] * A0 * gacc[r][e];
    vstore<VX>(gpart + (int64_t)grp * g.plane + i, gv);
  }
}


// The fused adjoint step in one launch: the strip ranges of
// habc_adj_strip_block first (the longer work per cell), then Deep's lane
// vectors (one row per thread: the kernel's register count, and so the strip
// threads' occupancy, follows the larger branch).  Requires y1, y2, um and
// gpart (the fused loop's call).
template <int S, int VX>
__global__ __launch_bounds__(256) void k_habc_adj_deep(
    const float* __restrict__ vp, const float* __restrict__ bw, const float* __restrict__ y1,
    const float* __restrict__ y2, float* __restrict__ out, const float* __restrict__ um,
    float* __restrict__ uc, float* __restrict__ gpart, const int32_t* __restrict__ src,
    const float* __restrict__ wn, int recon, int B, HabcCtx hc, seis::TimeConsts k, Deep dp,
    TileMap tm, Frame f, int ngf, int G, TileMap tb, int ngb, TileMap ts) {
  if (habc_adj_strip_block<S>((int)blockIdx.x, vp, bw, y1, y2, out, um, uc, gpart, src, wn,
                              recon, B, hc, k, dp, f, ngf, G, tb, ngb, ts))
    return;
  const int w = (int)blockIdx.x - (f.per + tb.per + ts.per) * kXcds;
  int tx, tzi, grp;
  if (!tile_of_w(tm, w, tx, tzi, grp)) return;
  habc_adj_deep_tile<S, VX, kHabcRZ>(tx, tzi, grp, vp, bw, y1, y2, out, um, uc, gpart, src, wn,
                                     recon, B, hc, k, dp);
}


// Per-step model gradient (all cells): sum over shots of
// alpha*A0*2c*L(u)*gy + sum_s beta_s*gy*(dl/dc (u1+p0-u0-p1) + dm/dc (2u1-u0-u2)).
__global__ __launch_bounds__(256) void k_habc_vjp_model(const float* __restrict__ vp,
                                                        const float* __restrict__ bw,
                                                        const float* __restrict__ u,
                                                        const float* __restrict__ p,
                                                        const float* __restrict__ gy,
                                                        float* __restrict__ gvp, int B,
                                                        HabcCtx hc, seis::TimeConsts k) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const Grid& g = hc.g;
  if (i >= g.plane) return;
  const int z = (int)(i / g.nx), x = (int)(i - (int64_t)z * g.nx);
  const float c = vp[i];
  const float A0 = 1.0f / k.idt2;
  const HabcW wt = habc_weights(z, x, hc, near_edge(z, x, g, kW) ? bw[i] : 0.0f);
  const float dl = 2.0f * hc.dt / hc.h;
  const float dm = 2.0f * c * (hc.dt * hc.dt) / (hc.h * hc.h);
  float acc = 0.0f;
  for (int b = 0; b < B; ++b) {
    const int64_t off = (int64_t)b * g.plane;
    const float* U = u + off;
    const float* P = p + off;
    const float gv = gy[off + i];
    float t = wt.alpha * A0 * 2.0f * c * lap5(U, z, x, (int)i, g, k.ih2);
    for (int s = 0; s < 4; ++s) {
      if (wt.beta[s] == 0.0f) continue;
      const int e = sstep(s, g);
      const float u2 = sdist(s, z, x, g) == kW - 1 ? U[sedge(s, z, x, g)] : U[i + 2 * e];
      t += wt.beta[s] * (dl * (U[i + e] + P[i] - U[i] - P[i + e]) + dm * (2.0f * U[i + e] - U[i] - u2));
    }
    acc += t * gv;
  }
  gvp[i] = acc;
}

// After k_habc_adj in the fused adjoint: adjoint source s^k at the receivers
// (and its gradient share 2c*A0*L(u^{k-1})*s^k), ring restore of u^{k-2}.
__global__ __launch_bounds__(256) void k_habc_post(
    const float* __restrict__ vp, const float* __restrict__ um, float* __restrict__ uc,
    float* __restrict__ lam, float* __restrict__ gpart, const int64_t* __restrict__ rec,
    const float* __restrict__ rgrad, int64_t R, const float* __restrict__ ring_slot,
    int64_t ring_len, int recon, int B, int S, Grid g, seis::TimeConsts k) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < R) {
    const int64_t f = rec[t];
    const int64_t b = f / g.plane;
    const int i = (int)(f - b * g.plane);
    const int z = i / g.nx, x = i - z * g.nx;
    const float sk = rgrad[t];
    atomicAdd(lam + f, sk);
    if (seis::interior(z, x, g)) {
      const float Lu = lap5(um + b * g.plane, z, x, i, g, k.ih2);
      atomicAdd(gpart + (b / S) * g.plane + i, 2.0f * vp[i] * (1.0f / k.idt2) * Lu * sk);
    }
    return;
  }
  t -= R;
  if (recon != 0 && t < (int64_t)B * ring_len) {
    const int64_t b = t / ring_len;
    const int64_t j = t - b * ring_len;
    uc[b * g.plane + seis::ring_cell(j, g)] = recon == 1 ? ring_slot[t] : 0.0f;
  }
}

__global__ __launch_bounds__(256) void k_habc_record(const float* __restrict__ u,
                                                     const int64_t* __restrict__ rec, int64_t R,
                                                     float* __restrict__ rec_row,
                                                     float* __restrict__ ring_slot,
                                                     int64_t ring_len, int B, Grid g) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < R) {
    rec_row[t] = u[rec[t]];
    return;
  }
  t -= R;
  if (ring_slot && t < (int64_t)B * ring_len) {
    const int64_t b = t / ring_len;
    const int64_t j = t - b * ring_len;
    ring_slot[t] = u[b * g.plane + seis::ring_cell(j, g)];
  }
}

__global__ __launch_bounds__(256) void k_habc_fill(float* __restrict__ p, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0.0f;
}

__global__ __launch_bounds__(256) void k_habc_copy2(const float* __restrict__ a,
                                                    float* __restrict__ ao,
                                                    const float* __restrict__ b,
                                                    float* __restrict__ bo, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) {
    ao[t] = a[t];
    bo[t] = b[t];
  }
}

__global__ __launch_bounds__(256) void k_habc_sum(const float* __restrict__ gpart, int G,
                                                  float* __restrict__ out, int64_t plane) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= plane) return;
  float s = 0.0f;
  for (int q = 0; q < G; ++q) s += gpart[(int64_t)q * plane + t];
  out[t] = s;
}

inline unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

#define SEIS_HABC_LAUNCH(K, S, grid, s, ...)                      \
  do {                                                            \
    switch (S) {                                                  \
      case 1: K<1><<<grid, 256, 0, s>>>(__VA_ARGS__); break;     \
      case 2: K<2><<<grid, 256, 0, s>>>(__VA_ARGS__); break;     \
      case 4: K<4><<<grid, 256, 0, s>>>(__VA_ARGS__); break;     \
      default: K<8><<<grid, 256, 0, s>>>(__VA_ARGS__); break;    \
    }                                                             \
    SEIS_CHECK_LAUNCH();                                          \
  } while (0)

#define SEIS_HABC_VX(K, S_, VX_, grid, s, ...)                       \
  switch (VX_) {                                                      \
    case 1: K<S_, 1><<<grid, 256, 0, s>>>(__VA_ARGS__); break;       \
    case 2: K<S_, 2><<<grid, 256, 0, s>>>(__VA_ARGS__); break;       \
    default: K<S_, 4><<<grid, 256, 0, s>>>(__VA_ARGS__); break;      \
  }

#define SEIS_HABC_LAUNCH_V(K, S, VX, grid, s, ...)                   \
  do {                                                              \
    switch (S) {                                                    \
      case 1: SEIS_HABC_VX(K, 1, VX, grid, s, __VA_ARGS__) break;   \
      case 2: SEIS_HABC_VX(K, 2, VX, grid, s, __VA_ARGS__) break;   \
      case 4: SEIS_HABC_VX(K, 4, VX, grid, s, __VA_ARGS__) break;   \
      default: SEIS_HABC_VX(K, 8, VX, grid, s, __VA_ARGS__) break;  \
    }                                                               \
    SEIS_CHECK_LAUNCH();                                            \
  } while (0)

// Widest lane vector (4, 2, 1) dividing nx with every pointer aligned to it.
int habc_vx(int64_t nx, std::initializer_list<const void*> ptrs) {
  for (int v : {4, 2}) {
    if (nx % v) continue;
    bool ok = true;
    for (const void* p : ptrs) ok = ok && (reinterpret_cast<uintptr_t>(p) % (4 * v)) == 0;
    if (ok) return v;
  }
  return 1;
}

// The frame around dp (the whole grid if dp is empty) as rectangles of tiles;
// with corners_only, just the four corners outside dp's rows and columns (the
// bands over dp's columns go to band_tiles, the columns over dp's rows to
// side_tiles).
Frame make_frame(const Grid& g, const Deep& dp, int groups, bool corners_only = false) {
  Frame f{};
  auto add = [&](int z0, int z1, int x0, int x1) {
    if (z1 <= z0 || x1 <= x0) return;
    const int r = f.n++;
    f.z0[r] = z0; f.z1[r] = z1; f.x0[r] = x0; f.x1[r] = x1;
    f.ntx[r] = (x1 - x0 + kLanes - 1) / kLanes;
    f.ntz[r] = (z1 - z0 + kRowsOfThreads - 1) / kRowsOfThreads;
  };
  if (dp.z1 <= dp.z0 || dp.x1 <= dp.x0) {
    add(0, g.nz, 0, g.nx);
  } else if (corners_only) {
    add(0, dp.z0, 0, dp.x0);
    add(0, dp.z0, dp.x1, g.nx);
    add(dp.z1, g.nz, 0, dp.x0);
    add(dp.z1, g.nz, dp.x1, g.nx);
  } else {
    add(0, dp.z0, 0, g.nx);
    add(dp.z1, g.nz, 0, g.nx);
    add(dp.z0, dp.z1, 0, dp.x0);
    add(dp.z0, dp.z1, dp.x1, g.nx);
  }
  f.groups = groups;
  f.t0[0] = 0;
  for (int r = 0; r < f.n; ++r)
    f.t0[r + 1] = f.t0[r] + (int64_t)f.ntx[r] * f.ntz[r] * groups;
  for (int r = f.n + 1; r < 5; ++r) f.t0[r] = f.t0[f.n];
  f.per = (int)((f.t0[f.n] + kXcds - 1) / kXcds);
  return f;
}

// Tiles of the top and bottom bands of the adjoint (rows [0, dp.z0) and
// [dp.z1, nz) over dp's columns, stacked as one virtual grid of lane vectors
// of vx cells); no tiles if dp is empty.
TileMap band_tiles(const Grid& g, const Deep& dp, int vx, int groups) {
  if (dp.z1 <= dp.z0 || dp.x1 <= dp.x0) return TileMap{0, 0, 0, 0};
  const int rows = dp.z0 + (g.nz - dp.z1);
  const Grid gb{rows, dp.x1 - dp.x0, 0, (int64_t)rows * (dp.x1 - dp.x0)};
  return make_tiles(gb, vx, 1, groups);
}

// Tiles of the left and right columns of the adjoint (x in [0, dp.x0) and
// [dp.x1, nx) over dp's rows): one 64-lane tile column each (both are at
// most kW+3+3 cells wide); no tiles if dp is empty.
TileMap side_tiles(const Grid& g, const Deep& dp, int groups) {
  if (dp.z1 <= dp.z0 || dp.x1 <= dp.x0) return TileMap{0, 0, 0, 0};
  TileMap m;
  m.ntx = 2;
  m.ntz = (dp.z1 - dp.z0 + kRowsOfThreads - 1) / kRowsOfThreads;
  m.groups = groups;
  m.per = (int)(((int64_t)m.ntx * m.ntz * groups + kXcds - 1) / kXcds);
  return m;
}

int habc_check(int64_t B, int64_t nz, int64_t nx, float dt, float h, int S) {
  SEIS_CHECK_ARG(B >= 1 && B <= 65535, "batch size %lld outside [1, 65535]", (long long)B);
  SEIS_CHECK_ARG(nz >= 2 * kW + 2 && nx >= 2 * kW + 2,
                 "HABC needs at least %d cells per axis, got %lld x %lld", 2 * kW + 2,
                 (long long)nz, (long long)nx);
  SEIS_CHECK_ARG(nz * nx < (int64_t)1 << 31, "grid too large");
  SEIS_CHECK_ARG(dt > 0.0f && h > 0.0f, "dt and h must be positive");
  SEIS_CHECK_ARG(S == 1 || S == 2 || S == 4 || S == 8, "shots_per_group %d not in {1,2,4,8}", S);
  return 0;
}

}  // namespace

extern "C" {

int seis_habc_step(const float* vp, const float* b, const float* h1, const float* h2, float* y,
                   int64_t B, int64_t nz, int64_t nx, float dt, float h, int32_t multiple,
                   void* stream) {
  if (int e = habc_check(B, nz, nx, dt, h, 1)) return e;
  SEIS_CHECK_ARG(vp && b && h1 && h2 && y, "null pointer argument");
  const HabcCtx hc{seis::make_grid(nz, nx, 0, false), multiple ? 1 : 0, dt, h};
  const int S = B >= 4 ? 4 : 1;
  const TileMap tm = make_tiles(hc.g, 1, 1, (int)((B + S - 1) / S));
  hipStream_t s = (hipStream_t)stream;
  SEIS_HABC_LAUNCH(k_habc_fwd, S, dim3(tm.per * kXcds), s, vp, b, h1, h2, y, nullptr, nullptr,
                   (int)B, hc, seis::make_consts(dt, h), tm);
  return 0;
}

int seis_habc_step_vjp(const float* vp, const float* b, const float* h1, const float* h2,
                       const float* gy, float* gh1, float* gh2, float* gvp, int64_t B, int64_t nz,
                       int64_t nx, float dt, float h, int32_t multiple, void* stream) {
  if (int e = habc_check(B, nz, nx, dt, h, 1)) return e;
  SEIS_CHECK_ARG(vp && b && h1 && h2 && gy && gh1 && gh2, "null pointer argument");
  const HabcCtx hc{seis::make_grid(nz, nx, 0, false), multiple ? 1 : 0, dt, h};
  const int S = B >= 4 ? 4 : 1;
  const TileMap tm = make_tiles(hc.g, 1, 1, (int)((B + S - 1) / S));
  const seis::TimeConsts k = seis::make_consts(dt, h);
  hipStream_t s = (hipStream_t)stream;
  SEIS_HABC_LAUNCH(k_habc_adj, S, dim3(tm.per * kXcds), s, vp, b, gy, nullptr, gh1, nullptr,
                   nullptr, nullptr, nullptr, nullptr, -1, (int)B, hc, k, tm);
  SEIS_HABC_LAUNCH(k_habc_adj, S, dim3(tm.per * kXcds), s, vp, b, nullptr, gy, gh2, nullptr,
                   nullptr, nullptr, nullptr, nullptr, -1, (int)B, hc, k, tm);
  if (gvp) {
    k_habc_vjp_model<<<nblk(nz * nx), 256, 0, s>>>(vp, b, h1, h2, gy, gvp, (int)B, hc, k);
    SEIS_CHECK_LAUNCH();
  }
  return 0;
}

int seis_habc_forward(const float* vp, const float* b, const float* wavelet, int64_t nt,
                      const int32_t* src_zx, const int64_t* rec_flat, int64_t R, float* rec_out,
                      float* ring, float* snaps, int32_t snap_every, float* illum, float* u_a,
                      float* u_b, float* u_c, int64_t B, int64_t nz, int64_t nx, float dt,
                      float h, int32_t npml, int32_t multiple, int32_t shots_per_group,
                      void* stream) {
  if (int e = habc_check(B, nz, nx, dt, h, shots_per_group)) return e;
  SEIS_CHECK_ARG(vp && b && wavelet && src_zx && u_a && u_b && u_c, "null pointer argument");
  SEIS_CHECK_ARG(!illum || (illum != u_a && illum != u_b && illum != u_c),
                 "illum must not alias a field buffer");
  SEIS_CHECK_ARG(nt >= 1, "nt must be >= 1");
  SEIS_CHECK_ARG(R >= 0 && (R == 0 || (rec_flat && rec_out)), "bad receiver arguments");
  const int64_t ring_len = ring ? seis_ring_len(nz, nx, npml, multiple) : 0;
  SEIS_CHECK_ARG(!ring || ring_len > 0, "grid has no interior inside npml=%d", npml);
  SEIS_CHECK_ARG(!snaps || snap_every > 0, "snapshots need snap_every > 0");
  const HabcCtx hc{seis::make_grid(nz, nx, npml, multiple != 0), multiple ? 1 : 0, dt, h};
  const Grid& g = hc.g;
  const int S = shots_per_group;
  // the deep interior (kW cells from every edge, outside the strips) runs vectorised
  const int VX = habc_vx(nx, {vp, u_a, u_b, u_c});
  Deep dp{kW, (int)nz - kW, (kW + VX - 1) / VX * VX, ((int)nx - kW) / VX * VX};
  if (dp.x1 <= dp.x0 || dp.z1 <= dp.z0) dp = Deep{0, 0, 0, 0};
  const Grid gd{dp.z1 - dp.z0, dp.x1 - dp.x0, 0, (int64_t)(dp.z1 - dp.z0) * (dp.x1 - dp.x0)};
  const int G = (int)((B + S - 1) / S);
  const TileMap tmd = make_tiles(gd, VX, 1, G);
  const int ngf = G < kFrameGroups ? G : kFrameGroups;  // groups per frame thread
  const Frame fr = make_frame(g, dp, (G + ngf - 1) / ngf);
  const seis::TimeConsts k = seis::make_consts(dt, h);
  hipStream_t s = (hipStream_t)stream;
  const int64_t field = B * g.plane;
  // u^m lives in P[m mod 3]; chosen so that u^{nt} ends in u_a, u^{nt-1} in u_b
  float* P[3];
  P[nt % 3] = u_a;
  P[(nt + 2) % 3] = u_b;
  P[(nt + 1) % 3] = u_c;
  for (float* q : {u_a, u_b, u_c}) {
    k_habc_fill<<<nblk(field), 256, 0, s>>>(q, field);
    SEIS_CHECK_LAUNCH();
  }
  if (illum) {
    k_habc_fill<<<nblk(g.plane), 256, 0, s>>>(illum, g.plane);
    SEIS_CHECK_LAUNCH();
  }
  for (int64_t n = 0; n < nt; ++n) {
    float* un = P[n % 3];
    float* up = P[(n + 2) % 3];
    float* out = P[(n + 1) % 3];
    SEIS_PROF(SEIS_KERNEL_HABC_FWD, s);
    SEIS_HABC_LAUNCH_V(k_habc_fwd_deep, S, VX, dim3((tmd.per + fr.per) * kXcds), s, vp, un,
                       up, out, src_zx, wavelet + n, (int)B, hc, k, dp, tmd, b, fr, ngf, G);
    SEIS_PROF(SEIS_KERNEL_HABC_FWD, s);
    if (int e = illuminate(out, (int)B, g.plane, illum, s)) return e;
    float* slot = (ring && n + 1 < nt) ? ring + n * B * ring_len : nullptr;
    const int64_t work = R + (slot ? B * ring_len : 0);
    if (work > 0) {
      k_habc_record<<<nblk(work), 256, 0, s>>>(out, rec_flat, R, R ? rec_out + n * R : nullptr,
                                               slot, ring_len, (int)B, g);
      SEIS_CHECK_LAUNCH();
    }
    if (snaps && n >= snap_every && n % snap_every == 0 && n <= nt - 2) {
      float* sp = snaps + (n / snap_every - 1) * 2 * field;
      k_habc_copy2<<<nblk(field), 256, 0, s>>>(un, sp, out, sp + field, field);
      SEIS_CHECK_LAUNCH();
    }
  }
  return 0;
}

// shots per thread of the per-step variable-density entry point
constexpr int kRhoShots = 16;

int seis_rho_habc_step(const float* vp, const float* rho, const float* b, const float* h1,
                       const float* h2, float* y, int64_t B, int64_t nz, int64_t nx, double dt,
                       double h, int32_t multiple, int32_t blend, void* stream) {
  if (int e = habc_check(B, nz, nx, (float)dt, (float)h, 1)) return e;
  SEIS_CHECK_ARG(vp && rho && h1 && h2 && y && (b || !blend), "null pointer argument");
  const HabcCtx hc{seis::make_grid(nz, nx, 0, false), multiple ? 1 : 0, (float)dt, (float)h};
  const int spt = B < kRhoShots ? (int)B : kRhoShots;
  const TileMap tm = make_tiles(hc.g, 1, 1, (int)((B + spt - 1) / spt));
  const float C = (float)(dt * dt * std::pow(h, -2.0));
  hipStream_t s = (hipStream_t)stream;
  k_rho_habc_fwd<<<dim3(tm.per * kXcds), 256, 0, s>>>(vp, rho, b, nullptr, h1, h2, y, nullptr,
                                                      nullptr, (int)B, spt, hc, C, blend ? 1 : 0,
                                                      tm);
  SEIS_CHECK_LAUNCH();
  return 0;
}

int seis_rho_habc_forward(const float* vp, const float* rho, const float* b,
                          const float* wavelet, int64_t nt, const int32_t* src_zx,
                          const int64_t* rec_flat, int64_t R, float* rec_out, float* ring,
                          float* snaps, int32_t snap_every, float* buoy, float* illum, float* u_a,
                          float* u_b, float* u_c, int64_t B, int64_t nz, int64_t nx, double dt,
                          double h, int32_t npml, int32_t multiple, int32_t shots_per_group,
                          void* stream) {
  if (int e = habc_check(B, nz, nx, (float)dt, (float)h, 1)) return e;
  SEIS_CHECK_ARG(shots_per_group >= 1, "shots_per_group must be >= 1, got %d", shots_per_group);
  SEIS_CHECK_ARG(vp && rho && b && buoy && wavelet && src_zx && u_a && u_b && u_c,
                 "null pointer argument");
  SEIS_CHECK_ARG(!illum || (illum != u_a && illum != u_b && illum != u_c && illum != buoy),
                 "illum must not alias a field buffer");
  SEIS_CHECK_ARG(nt >= 1, "nt must be >= 1");
  SEIS_CHECK_ARG(R >= 0 && (R == 0 || (rec_flat && rec_out)), "bad receiver arguments");
  const int64_t ring_len = ring ? seis_ring_len(nz, nx, npml, multiple) : 0;
  SEIS_CHECK_ARG(!ring || ring_len > 0, "grid has no interior inside npml=%d", npml);
  SEIS_CHECK_ARG(!snaps || snap_every > 0, "snapshots need snap_every > 0");
  const HabcCtx hc{seis::make_grid(nz, nx, npml, multiple != 0), multiple ? 1 : 0, (float)dt,
                   (float)h};
  const Grid& g = hc.g;
  const int S = shots_per_group;
  const TileMap tm = make_tiles(g, 1, 1, (int)((B + S - 1) / S));
  const float C = (float)(dt * dt * std::pow(h, -2.0));
  hipStream_t s = (hipStream_t)stream;
  const int64_t field = B * g.plane;
  // u^m lives in P[m mod 3]; u^{nt} ends in u_a, u^{nt-1} in u_b
  float* P[3];
  P[nt % 3] = u_a;
  P[(nt + 2) % 3] = u_b;
  P[(nt + 1) % 3] = u_c;
  for (float* q : {u_a, u_b, u_c}) {
    k_habc_fill<<<nblk(field), 256, 0, s>>>(q, field);
    SEIS_CHECK_LAUNCH();
  }
  k_reciprocal<<<nblk(g.plane), 256, 0, s>>>(rho, buoy, g.plane);
  SEIS_CHECK_LAUNCH();
  if (illum) {
    k_habc_fill<<<nblk(g.plane), 256, 0, s>>>(illum, g.plane);
    SEIS_CHECK_LAUNCH();
  }
  for (int64_t n = 0; n < nt; ++n) {
    float* un = P[n % 3];
    float* up = P[(n + 2) % 3];
    float* out = P[(n + 1) % 3];
    k_rho_habc_fwd<<<dim3(tm.per * kXcds), 256, 0, s>>>(vp, rho, b, buoy, un, up, out, src_zx,
                                                        wavelet + n, (int)B, S, hc, C, 1, tm);
    SEIS_CHECK_LAUNCH();
    if (int e = illuminate(out, (int)B, g.plane, illum, s)) return e;
    float* slot = (ring && n + 1 < nt) ? ring + n * B * ring_len : nullptr;
    const int64_t work = R + (slot ? B * ring_len : 0);
    if (work > 0) {
      k_habc_record<<<nblk(work), 256, 0, s>>>(out, rec_flat, R, R ? rec_out + n * R : nullptr,
                                               slot, ring_len, (int)B, g);
      SEIS_CHECK_LAUNCH();
    }
    if (snaps && n >= snap_every && n % snap_every == 0 && n <= nt - 2) {
      float* sp = snaps + (n / snap_every - 1) * 2 * field;
      k_habc_copy2<<<nblk(field), 256, 0, s>>>(un, sp, out, sp + field, field);
      SEIS_CHECK_LAUNCH();
    }
  }
  return 0;
}

int seis_rho_habc_backward(const float* vp, const float* rho, const float* b,
                           const float* wavelet, int64_t nt, const int32_t* src_zx,
                           const int64_t* rec_flat, int64_t R, const float* rec_grad,
                           const float* ring, const float* snaps, int32_t snap_every, float* buoy,
                           float* u_a, float* u_b, float* l_a, float* l_b, float* l_c,
                           float* grad_part, float* grad_vp, float* grad_rho, int64_t B,
                           int64_t nz, int64_t nx, double dt, double h, int32_t npml,
                           int32_t multiple, int32_t shots_per_group, void* stream) {
  if (int e = habc_check(B, nz, nx, (float)dt, (float)h, 1)) return e;
  SEIS_CHECK_ARG(shots_per_group >= 1, "shots_per_group must be >= 1, got %d", shots_per_group);
  SEIS_CHECK_ARG(vp && rho && b && buoy && wavelet && src_zx && u_a && u_b && l_a && l_b && l_c &&
                     grad_part && grad_vp && grad_rho,
                 "null pointer argument");
  SEIS_CHECK_ARG(nt >= 1, "nt must be >= 1");
  SEIS_CHECK_ARG(R >= 0 && (R == 0 || (rec_flat && rec_grad)), "bad receiver arguments");
  const int64_t ring_len = seis_ring_len(nz, nx, npml, multiple);
  SEIS_CHECK_ARG(ring_len > 0, "grid has no interior inside npml=%d", npml);
  SEIS_CHECK_ARG(ring || nt <= 2, "ring storage is required for nt > 2");
  SEIS_CHECK_ARG(!snaps || snap_every > 0, "snapshots need snap_every > 0");
  const HabcCtx hc{seis::make_grid(nz, nx, npml, multiple != 0), multiple ? 1 : 0, (float)dt,
                   (float)h};
  const Grid& g = hc.g;
  const int S = shots_per_group;
  const int G = (int)((B + S - 1) / S);
  const TileMap tm = make_tiles(g, 1, 1, G);
  // Deep (kW+3 cells from every edge, column bounds multiples of VX) runs
  // lane vectors of VX cells, the frame around it one cell per thread
  const int VX = habc_vx(nx, {vp, rho, buoy, u_a, u_b, l_a, l_b, l_c, grad_part});
  constexpr int m = kW + 3;
  Deep dp{m, (int)nz - m, (m + VX - 1) / VX * VX, ((int)nx - m) / VX * VX};
  if (dp.x1 <= dp.x0 || dp.z1 <= dp.z0) dp = Deep{0, 0, 0, 0};
  const Grid gd{dp.z1 - dp.z0, dp.x1 - dp.x0, 0, (int64_t)(dp.z1 - dp.z0) * (dp.x1 - dp.x0)};
  TileMap tmd = make_tiles(gd, VX, 1, G);
  if (gd.plane == 0) tmd.per = 0;
  const int ngf = std::min(G, kRhoFrameGroups);
  const Frame fr = make_frame(g, dp, (G + ngf - 1) / ngf);
  const float C = (float)(dt * dt * std::pow(h, -2.0));
  hipStream_t s = (hipStream_t)stream;
  const int64_t field = B * g.plane;
  for (float* q : {l_a, l_b, l_c}) {
    k_habc_fill<<<nblk(field), 256, 0, s>>>(q, field);
    SEIS_CHECK_LAUNCH();
  }
  k_habc_fill<<<nblk(2 * (int64_t)G * g.plane), 256, 0, s>>>(grad_part, 2 * (int64_t)G * g.plane);
  SEIS_CHECK_LAUNCH();
  k_reciprocal<<<nblk(g.plane), 256, 0, s>>>(rho, buoy, g.plane);
  SEIS_CHECK_LAUNCH();
  float* X = u_b;   // u^{k-1}
  float* Y = u_a;   // u^k -> u^{k-2}
  float* L1 = l_a;  // lambda^{k+1}
  float* L2 = l_b;  // lambda^{k+2}
  float* L0 = l_c;  // lambda^k
  // lambda^nt (from zero cotangents of the later states), then per reverse
  // step k one launch: the gradient of step k and lambda^{k-1}, written into
  // the buffer of lambda^{k+2}
  k_rho_adj<<<dim3(tm.per * kXcds), 256, 0, s>>>(vp, rho, buoy, b, L1, L2, L0, (int)B, S, hc, C,
                                                 tm);
  SEIS_CHECK_LAUNCH();
  for (int64_t kk = nt; kk >= 1; --kk) {
    const int recon = kk >= 3 ? 1 : (kk == 2 ? 2 : 0);
    if (R > 0) {
      k_rho_inject<<<nblk(R), 256, 0, s>>>(L0, rec_flat, rec_grad + (kk - 1) * R, R);
      SEIS_CHECK_LAUNCH();
    }
    const float* slot = recon == 1 ? ring + (kk - 3) * B * ring_len : nullptr;
    const dim3 grid((fr.per + tmd.per) * kXcds);
    float* next = kk > 1 ? L2 : nullptr;
    switch (VX) {
      case 4:
        k_rho_adj_grad<4><<<grid, 256, 0, s>>>(vp, rho, buoy, b, L0, L1, next, X, Y, grad_part,
                                                src_zx, wavelet + (kk - 1), slot, ring_len, recon,
                                                (int)B, S, G, hc, C, dp, fr, ngf, tmd);
        break;
      case 2:
        k_rho_adj_grad<2><<<grid, 256, 0, s>>>(vp, rho, buoy, b, L0, L1, next, X, Y, grad_part,
                                                src_zx, wavelet + (kk - 1), slot, ring_len, recon,
                                                (int)B, S, G, hc, C, dp, fr, ngf, tmd);
        break;
      default:
        k_rho_adj_grad<1><<<grid, 256, 0, s>>>(vp, rho, buoy, b, L0, L1, next, X, Y, grad_part,
                                                src_zx, wavelet + (kk - 1), slot, ring_len, recon,
                                                (int)B, S, G, hc, C, dp, fr, ngf, tmd);
    }
    SEIS_CHECK_LAUNCH();
    const int64_t m = kk - 2;
    if (snaps && m >= snap_every && m % snap_every == 0 && m <= nt - 2) {
      const float* sp = snaps + (m / snap_every - 1) * 2 * field;
      k_habc_copy2<<<nblk(field), 256, 0, s>>>(sp, Y, sp + field, X, field);
      SEIS_CHECK_LAUNCH();
    }
    float* t = X; X = Y; Y = t;
    t = L2; L2 = L1; L1 = L0; L0 = t;
  }
  k_habc_sum<<<nblk(g.plane), 256, 0, s>>>(grad_part, G, grad_vp, g.plane);
  SEIS_CHECK_LAUNCH();
  k_habc_sum<<<nblk(g.plane), 256, 0, s>>>(grad_part + (int64_t)G * g.plane, G, grad_rho,
                                           g.plane);
  SEIS_CHECK_LAUNCH();
  return 0;
}

int seis_habc_backward(const float* vp, const float* b, const float* wavelet, int64_t nt,
                       const int32_t* src_zx, const int64_t* rec_flat, int64_t R,
                       const float* rec_grad, const float* ring, const float* snaps,
                       int32_t snap_every, float* u_a, float* u_b, float* l_a, float* l_b,
                       float* l_c, float* grad_part, float* grad, int64_t B, int64_t nz,
                       int64_t nx, float dt, float h, int32_t npml, int32_t multiple,
                       int32_t shots_per_group, void* stream) {
  if (int e = habc_check(B, nz, nx, dt, h, shots_per_group)) return e;
  SEIS_CHECK_ARG(vp && b && wavelet && src_zx && u_a && u_b && l_a && l_b && l_c && grad_part &&
                     grad,
                 "null pointer argument");
  SEIS_CHECK_ARG(nt >= 1, "nt must be >= 1");
  SEIS_CHECK_ARG(R >= 0 && (R == 0 || (rec_flat && rec_grad)), "bad receiver arguments");
  const int64_t ring_len = seis_ring_len(nz, nx, npml, multiple);
  SEIS_CHECK_ARG(ring_len > 0, "grid has no interior inside npml=%d", npml);
  SEIS_CHECK_ARG(ring || nt <= 2, "ring storage is required for nt > 2");
  SEIS_CHECK_ARG(!snaps || snap_every > 0, "snapshots need snap_every > 0");
  const HabcCtx hc{seis::make_grid(nz, nx, npml, multiple != 0), multiple ? 1 : 0, dt, h};
  const Grid& g = hc.g;
  const int S = shots_per_group;
  const int G = (int)((B + S - 1) / S);
  const int VX = habc_vx(nx, {vp, u_a, u_b, l_a, l_b, l_c, grad_part});
  constexpr int m = kW + 3;
  Deep dp{m, (int)nz - m, (m + VX - 1) / VX * VX, ((int)nx - m) / VX * VX};
  if (dp.x1 <= dp.x0 || dp.z1 <= dp.z0) dp = Deep{0, 0, 0, 0};  // no deep cells
  const Grid gd{dp.z1 - dp.z0, dp.x1 - dp.x0, 0, (int64_t)(dp.z1 - dp.z0) * (dp.x1 - dp.x0)};
  const TileMap tmd = make_tiles(gd, VX, kHabcRZ, G);
  const int ngf = G < kFrameGroups ? G : kFrameGroups;  // groups per frame thread
  const Frame fr = make_frame(g, dp, (G + ngf - 1) / ngf, true);
  const int ngb = G < kBandGroups ? G : kBandGroups;  // groups per band thread
  const TileMap tmb = band_tiles(g, dp, kBandVX, (G + ngb - 1) / ngb);
  const TileMap tms = side_tiles(g, dp, (G + ngb - 1) / ngb);
  SEIS_CHECK_ARG(tms.per == 0 || (dp.x0 <= kLanes && (int)nx - dp.x1 <= kLanes),
                 "side columns wider than a wave");
  const seis::TimeConsts k = seis::make_consts(dt, h);
  hipStream_t s = (hipStream_t)stream;
  const int64_t field = B * g.plane;
  for (float* q : {l_a, l_b, l_c}) {
    k_habc_fill<<<nblk(field), 256, 0, s>>>(q, field);
    SEIS_CHECK_LAUNCH();
  }
  k_habc_fill<<<nblk((int64_t)G * g.plane), 256, 0, s>>>(grad_part, (int64_t)G * g.plane);
  SEIS_CHECK_LAUNCH();
  float* X = u_b;                  // u^{k-1}
  float* Y = u_a;                  // u^k -> u^{k-2}
  float* L1 = l_a;                 // lambda^{k+1}
  float* L2 = l_b;                 // lambda^{k+2}
  float* L0 = l_c;                 // lambda^k
  for (int64_t kk = nt; kk >= 1; --kk) {
    const int recon = kk >= 3 ? 1 : (kk == 2 ? 2 : 0);
    SEIS_PROF(SEIS_KERNEL_HABC_BWD, s);
    SEIS_HABC_LAUNCH_V(k_habc_adj_deep, S, VX,
                       dim3((tmd.per + fr.per + tmb.per + tms.per) * kXcds), s, vp, b, L1, L2,
                       L0, X, Y, grad_part, src_zx, wavelet + (kk - 1), recon, (int)B, hc, k,
                       dp, tmd, fr, ngf, G, tmb, ngb, tms);
    SEIS_PROF(SEIS_KERNEL_HABC_BWD, s);
    const int64_t work = R + (recon ? B * ring_len : 0);
    if (work > 0) {
      const float* slot = recon == 1 ? ring + (kk - 3) * B * ring_len : nullptr;
      k_habc_post<<<nblk(work), 256, 0, s>>>(vp, X, Y, L0, grad_part, rec_flat,
                                             R ? rec_grad + (kk - 1) * R : nullptr, R, slot,
                                             ring_len, recon, (int)B, S, g, k);
      SEIS_CHECK_LAUNCH();
    }
    const int64_t m = kk - 2;
    if (snaps && m >= snap_every && m % snap_every == 0 && m <= nt - 2) {
      const float* sp = snaps + (m / snap_every - 1) * 2 * field;
      k_habc_copy2<<<nblk(field), 256, 0, s>>>(sp, Y, sp + field, X, field);
      SEIS_CHECK_LAUNCH();
    }
    float* t = X; X = Y; Y = t;
    t = L2; L2 = L1; L1 = L0; L0 = t;
  }
  k_habc_sum<<<nblk(g.plane), 256, 0, s>>>(grad_part, G, grad, g.plane);
  SEIS_CHECK_LAUNCH();
  return 0;
}

}  // extern "C"
