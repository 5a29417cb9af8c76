// Grid description, XCD-aware tile order and vector access helpers shared
// by the fused time-loop kernels of every equation.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace seis {

// Padded (nz, nx) grid; npml is the ring index of the boundary saving at the
// bottom, left and right edges.  ztop is the ring row at the top: npml, or -1
// for a free surface (seistorch's `multiple` mode: no top boundary, so the
// undamped region reaches row 0 and the top ring row is neither stored nor
// restored, seistorch/equations2d/utils.py:71-78, elastic.py:132-134).
struct Grid {
  int nz, nx;
  int npml;
  int64_t plane;  // nz*nx
  int ztop;
};

__host__ __device__ inline Grid make_grid(int64_t nz, int64_t nx, int npml, bool free_surface) {
  return Grid{(int)nz, (int)nx, npml, nz * nx, free_surface ? -1 : npml};
}

// Rows of the undamped interior: (ztop, nz-npml-1).
__device__ __forceinline__ bool interior_row(int z, const Grid g) {
  return z > g.ztop && z < g.nz - 1 - g.npml;
}

// Undamped interior: rows (ztop, nz-npml-1), columns (npml, nx-npml-1).
__device__ __forceinline__ bool interior(int z, int x, const Grid g) {
  return interior_row(z, g) && x > g.npml && x < g.nx - 1 - g.npml;
}

// Ring sizes: nxi cells in each of the top and bottom rows (no top row with a
// free surface), nzi cells in each of the left and right columns.
__host__ __device__ inline int ring_nxi(const Grid g) { return g.nx - 2 * g.npml - 2; }
__host__ __device__ inline int ring_nzi(const Grid g) { return g.nz - g.npml - 2 - g.ztop; }
__host__ __device__ inline int ring_ntop(const Grid g) { return g.ztop >= 0 ? ring_nxi(g) : 0; }

// j-th cell of the boundary-saving ring: top row (absent with a free
// surface), bottom row, left column, right column, each restricted to the
// 5-point neighbours of the interior.
__device__ __forceinline__ int64_t ring_cell(int64_t j, const Grid g) {
  const int nxi = ring_nxi(g);
  const int nzi = ring_nzi(g);
  const int nt = ring_ntop(g);
  int z, x;
  if (j < nt) {
    z = g.ztop; x = g.npml + 1 + (int)j;
  } else if (j < nt + nxi) {
    z = g.nz - 1 - g.npml; x = g.npml + 1 + (int)(j - nt);
  } else if (j < nt + nxi + nzi) {
    z = g.ztop + 1 + (int)(j - nt - nxi); x = g.npml;
  } else {
    z = g.ztop + 1 + (int)(j - nt - nxi - nzi); x = g.nx - 1 - g.npml;
  }
  return (int64_t)z * g.nx + x;
}

// Ring index j of cell (z, x), the inverse of ring_cell, or -1 off the ring.
__device__ __forceinline__ int64_t ring_index(int z, int x, const Grid g) {
  const int nxi = ring_nxi(g);
  const int nzi = ring_nzi(g);
  const int nt = ring_ntop(g);
  const bool xr = x > g.npml && x < g.nx - 1 - g.npml;
  const bool zr = interior_row(z, g);
  if (nt > 0 && z == g.ztop && xr) return x - g.npml - 1;
  if (z == g.nz - 1 - g.npml && xr) return nt + (x - g.npml - 1);
  if (x == g.npml && zr) return nt + nxi + (z - g.ztop - 1);
  if (x == g.nx - 1 - g.npml && zr) return (int64_t)nt + nxi + nzi + (z - g.ztop - 1);
  return -1;
}

// 5-point Laplacian from its taps: the h^-2-weighted taps accumulated in
// row-major kernel order (z-1, x-1, centre, x+1, z+1) with fused multiply-adds,
// the way a direct 3x3 convolution with weights h^-2*[[0,1,0],[1,-4,1],[0,1,0]]
// accumulates (acoustic.py:60-71).  Every kernel uses this one helper.
__device__ __forceinline__ float lap_taps(float zm, float xm, float c, float xp, float zp,
                                          float ih2) {
  float acc = ih2 * zm;
  acc = __fmaf_rn(ih2, xm, acc);
  acc = __fmaf_rn(-4.0f * ih2, c, acc);
  acc = __fmaf_rn(ih2, xp, acc);
  return __fmaf_rn(ih2, zp, acc);
}

// The damped update of acoustic.py:81-84 rounded op by op in its order:
// A*((2/dt^2*u - B*u_prev) + c^2*L).
__device__ __forceinline__ float acoustic_update(float A, float Bc, float two_idt2, float u,
                                                 float up, float c2, float L) {
  return A * ((two_idt2 * u - Bc * up) + c2 * L);
}

__device__ __forceinline__ float lap5(const float* __restrict__ f, int z, int x, int i,
                                      const Grid g, float ih2) {
  const float c = f[i];
  const float zm = z > 0 ? f[i - g.nx] : 0.0f;
  const float zp = z < g.nz - 1 ? f[i + g.nx] : 0.0f;
  const float xm = x > 0 ? f[i - 1] : 0.0f;
  const float xp = x < g.nx - 1 ? f[i + 1] : 0.0f;
  return lap_taps(zm, xm, c, xp, zp, ih2);
}

namespace tiles {

// Tiles of the fused loops: a 256-thread block is 64 lanes along x by 4
// thread-rows.  Each lane owns VX consecutive x cells (one float, float2 or
// float4 access; the grid width must be a multiple of VX) in RZ consecutive
// rows, for S shots, so a tile is 64*VX x 4*RZ cells.  The RZ+2 rows of the
// stencilled field's column are loaded once into registers (fully unrolled,
// all loads of a thread independent); the x-neighbours come from the lane's
// own vector, plus one scalar load at each end of it (an L1 hit of the
// neighbouring lane's vector).
//
// XCD-aware tile order: workgroups are handed to the 8 XCDs round-robin, so
// workgroup w runs on XCD w % 8.  Tile t = (w % 8) * per + w / 8 gives each
// XCD a contiguous range of tiles in time order.  Tiles are numbered
// x-fastest, so the x- and z-neighbours of a tile (which own its halo cache
// lines) are dispatched by the same XCD within one tile row and find those
// lines in its L2.
constexpr int kBlock = 256;
constexpr int kLanes = 64;
constexpr int kRowsOfThreads = kBlock / kLanes;
constexpr int kXcds = 8;

struct TileMap {
  int ntx, ntz, groups, per;  // tiles in x, tiles in z, shot groups, tiles per XCD
};

inline TileMap make_tiles(const Grid& g, int vx, int rz, int groups) {
  TileMap m;
  m.ntx = (g.nx + kLanes * vx - 1) / (kLanes * vx);
  m.ntz = (g.nz + kRowsOfThreads * rz - 1) / (kRowsOfThreads * rz);
  m.groups = groups;
  const int64_t n = (int64_t)m.ntx * m.ntz * groups;
  m.per = (int)((n + kXcds - 1) / kXcds);
  return m;
}

// -> false if workgroup w has no tile (padding of the XCD split)
__device__ __forceinline__ bool tile_of_w(const TileMap m, int w, int& tx, int& tzi, int& grp) {
  const int64_t t = (int64_t)(w % kXcds) * m.per + w / kXcds;
  if (t >= (int64_t)m.ntx * m.ntz * m.groups) return false;
  tx = (int)(t % m.ntx);
  const int64_t r = t / m.ntx;
  tzi = (int)(r % m.ntz);
  grp = (int)(r / m.ntz);
  return true;
}

// -> false if this workgroup has no tile (padding of the XCD split)
__device__ __forceinline__ bool tile_of(const TileMap m, int& tx, int& tzi, int& grp) {
  const int w = blockIdx.x;
  const int64_t t = (int64_t)(w % kXcds) * m.per + w / kXcds;
  if (t >= (int64_t)m.ntx * m.ntz * m.groups) return false;
  tx = (int)(t % m.ntx);
  const int64_t r = t / m.ntx;
  tzi = (int)(r % m.ntz);
  grp = (int)(r / m.ntz);
  return true;
}

template <int VX>
__device__ __forceinline__ void vload(const float* __restrict__ p, float (&o)[VX]) {
  if constexpr (VX == 1) {
    o[0] = *p;
  } else if constexpr (VX == 2) {
    const float2 t = *reinterpret_cast<const float2*>(p);
    o[0] = t.x; o[1] = t.y;
  } else {
    const float4 t = *reinterpret_cast<const float4*>(p);
    o[0] = t.x; o[1] = t.y; o[2] = t.z; o[3] = t.w;
  }
}

template <int VX>
__device__ __forceinline__ void vstore(float* __restrict__ p, const float (&o)[VX]) {
  if constexpr (VX == 1) {
    *p = o[0];
  } else if constexpr (VX == 2) {
    *reinterpret_cast<float2*>(p) = make_float2(o[0], o[1]);
  } else {
    *reinterpret_cast<float4*>(p) = make_float4(o[0], o[1], o[2], o[3]);
  }
}

// A vector store with the nontemporal hint (global_store ... nt): for fields
// that this launch writes once and only a later launch reads, so that they do
// not displace the halo rows still being read from L2.
template <int VX>
__device__ __forceinline__ void vstore_nt(float* __restrict__ p, const float (&o)[VX]) {
  if constexpr (VX == 4) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    const f4 v = {o[0], o[1], o[2], o[3]};
    __builtin_nontemporal_store(v, reinterpret_cast<f4*>(p));
  } else if constexpr (VX == 2) {
    typedef float f2 __attribute__((ext_vector_type(2)));
    const f2 v = {o[0], o[1]};
    __builtin_nontemporal_store(v, reinterpret_cast<f2*>(p));
  } else {
    __builtin_nontemporal_store(o[0], p);
  }
}

template <int VX>
__device__ __forceinline__ void vzero(float (&o)[VX]) {
#pragma unroll
  for (int e = 0; e < VX; ++e) o[e] = 0.0f;
}

}  // namespace tiles
}  // namespace seis
