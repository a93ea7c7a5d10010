// Device building blocks of charge spreading and force gathering, shared by
// the single-system and batched kernels.
#pragma once

#include "admp_common.hpp"
#include "tiles.hpp"

namespace admp {

__host__ __device__ inline int tile_lo(int K, int nc, int c) {
  return (int)(((long long)K * c + nc - 1) / nc) - 4;
}

// One lane's share of spreading one atom with global atomics:
// Q[(m0 - 3 + a, m0 - 3 + b, m0 - 3 + c) mod K] += q wx[a] wy[b] wz[c] for the
// lane's x offset a < 6 (36 adds).
template <typename T>
__device__ __forceinline__ void spread_lane(const Geom& g, double x, double y, double z, double q,
                                            int a, T* __restrict__ mesh) {
  int m0[3];
  double f[3];
  mesh_ref(g, x, y, z, m0, f);
  T wx[6], wy[6], wz[6], dd[6];
  bspline6_rounded<T>(f[0], wx, dd);
  bspline6_rounded<T>(f[1], wy, dd);
  bspline6_rounded<T>(f[2], wz, dd);
  const int K0 = g.k[0], K1 = g.k[1], K2 = g.k[2];
  const T qa = (T)q * wx[a];
  const int ix = wrapk(pmod32(m0[0] - 3, K0) + a, K0);
  const int by = pmod32(m0[1] - 3, K1), bz = pmod32(m0[2] - 3, K2);
  int iz[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) iz[c] = wrapk(bz + c, K2);
#pragma unroll
  for (int b = 0; b < 6; ++b) {
    const int iy = wrapk(by + b, K1);
    T* row = mesh + ((int64_t)ix * K1 + iy) * K2;
    const T qab = qa * wy[b];
#pragma unroll
    for (int c = 0; c < 6; ++c) unsafeAtomicAdd(row + iz[c], qab * wz[c]);
  }
}

// Zero a float64 tile of n values in LDS with 16-byte stores; ends with a
// barrier.
__device__ __forceinline__ void zero_tile(double* tile, int n) {
  typedef double d2 __attribute__((ext_vector_type(2)));
  d2* t2 = reinterpret_cast<d2*>(tile);
  const d2 z2 = {0.0, 0.0};
  for (int k = threadIdx.x; k < (n >> 1); k += blockDim.x) t2[k] = z2;
  if ((n & 1) && threadIdx.x == 0) tile[n - 1] = 0.0;
  __syncthreads();
}

// Mesh index (before wrapping) of tile point 0 of cell lc.
__device__ __forceinline__ void tile_origin(const TileGeom& tg, int lc, int* lo0, int* lo1,
                                            int* lo2) {
  const int cz = lc % tg.nc[2];
  const int cy = (lc / tg.nc[2]) % tg.nc[1];
  const int cx = lc / (tg.nc[1] * tg.nc[2]);
  *lo0 = tile_lo(tg.K[0], tg.nc[0], cx);
  *lo1 = tile_lo(tg.K[1], tg.nc[1], cy);
  *lo2 = tile_lo(tg.K[2], tg.nc[2], cz);
}

// Add the nonzero points of a finished tile to the mesh with global atomics.
// Tile offsets are mesh indices from lo, modulo the mesh; consecutive lanes
// take consecutive z of one tile row (rows_per_pass rows at a time), so a
// wavefront's atomics fall on few mesh lines.
template <typename T>
__device__ __forceinline__ void flush_tile(const TileGeom& tg, const double* tile, int lo0,
                                           int lo1, int lo2, T* __restrict__ mesh) {
  const int W1 = tg.W[1], W2 = tg.W[2];
  const int b0 = pmod32(lo0, tg.K[0]), b1 = pmod32(lo1, tg.K[1]), b2 = pmod32(lo2, tg.K[2]);
  const int rows_per_pass = blockDim.x / W2;
  const int lz = threadIdx.x % W2, r0 = threadIdx.x / W2;
  if (r0 < rows_per_pass) {
    const int iz = wrapk(b2 + lz, tg.K[2]);
    for (int row = r0; row < tg.W[0] * W1; row += rows_per_pass) {
      const double v = tile[row * W2 + lz];
      if (v == 0.0) continue;
      const int lx = row / W1, ly = row - lx * W1;
      const int ix = wrapk(b0 + lx, tg.K[0]), iy = wrapk(b1 + ly, tg.K[1]);
      unsafeAtomicAdd(mesh + ((int64_t)ix * tg.K[1] + iy) * tg.K[2] + iz, (T)v);
    }
  }
}

// Per-atom stencil record, formed once per evaluation and read by spread and
// gather: the reference point m0 (mesh_ref) and the B-spline weights and
// derivatives of the three axes, evaluated in float64 and rounded once to T:
//   w[0..5] x, w[6..11] y, w[12..17] z, w[18..23] dx, w[24..29] dy, w[30..35] dz.
constexpr int kStencilVals = 36;

template <typename T>
__device__ __forceinline__ void store_record(const int m0[3], const T wx[6], const T wy[6],
                                             const T wz[6], const T dx[6], const T dy[6],
                                             const T dz[6], int4* __restrict__ m,
                                             T* __restrict__ w) {
  const T* part[6] = {wx, wy, wz, dx, dy, dz};
  T wv[kStencilVals];
#pragma unroll
  for (int p = 0; p < 6; ++p)
#pragma unroll
    for (int j = 0; j < 6; ++j) wv[6 * p + j] = part[p][j];
  *m = make_int4(m0[0], m0[1], m0[2], 0);
  typedef T t4 __attribute__((ext_vector_type(4)));
  t4* w4 = reinterpret_cast<t4*>(w);
#pragma unroll
  for (int k = 0; k < kStencilVals / 4; ++k) {
    t4 v;
    v.x = wv[4 * k];
    v.y = wv[4 * k + 1];
    v.z = wv[4 * k + 2];
    v.w = wv[4 * k + 3];
    w4[k] = v;
  }
}

template <typename T>
__device__ __forceinline__ void stencil_record(const Geom& g, double x, double y, double z,
                                               int4* __restrict__ m, T* __restrict__ w) {
  int m0[3];
  double f[3];
  mesh_ref(g, x, y, z, m0, f);
  T wx[6], wy[6], wz[6], dx[6], dy[6], dz[6];
  bspline6_rounded<T>(f[0], wx, dx);
  bspline6_rounded<T>(f[1], wy, dy);
  bspline6_rounded<T>(f[2], wz, dz);
  store_record<T>(m0, wx, wy, wz, dx, dy, dz, m, w);
}

// One work-group spreads the atoms of cell lc (sorted slots [b, e)) into a
// float64 LDS tile (dynamic LDS, tg.ntile doubles), then adds the tile's
// nonzero points to the mesh with global atomics (one per point and cell
// instead of one per point and atom).  Atoms whose stencil leaves the tile add
// to the mesh directly.  Eight lanes per atom: lanes 0-5 own x offsets; with
// stm non-null, lane 6 writes the atom's stencil record for the gather.
template <typename T>
__device__ __forceinline__ void spread_cell_tile(const Geom& g, const TileGeom& tg, int lc, int b,
                                                 int e, const int* __restrict__ sorted,
                                                 const double* __restrict__ pos,
                                                 const double* __restrict__ q,
                                                 T* __restrict__ mesh,
                                                 int4* __restrict__ stm = nullptr,
                                                 T* __restrict__ stw = nullptr) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* tile = reinterpret_cast<double*>(smem_raw);  // float64 LDS atomics in both precisions
  const int W0 = tg.W[0], W1 = tg.W[1], W2 = tg.W[2];
  zero_tile(tile, tg.ntile);
  int lo0, lo1, lo2;
  tile_origin(tg, lc, &lo0, &lo1, &lo2);
  const int a = threadIdx.x & 7;
  for (int slot = b + (threadIdx.x >> 3); slot < e; slot += blockDim.x >> 3) {
    if (a == 7 || (a == 6 && !stm)) continue;
    const int i = sorted[slot];
    int m0[3];
    double f[3];
    mesh_ref(g, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], m0, f);
    T wx[6], wy[6], wz[6], dx[6], dy[6], dz[6];
    bspline6_rounded<T>(f[0], wx, dx);
    bspline6_rounded<T>(f[1], wy, dy);
    bspline6_rounded<T>(f[2], wz, dz);
    if (a == 6) {  // the same weights, stored once for the gather
      store_record<T>(m0, wx, wy, wz, dx, dy, dz, stm + i, stw + (int64_t)kStencilVals * i);
      continue;
    }
    const T qa = (T)q[i] * wx[a];
    const int lx = wrapk(pmod32(m0[0] - 3 - lo0, tg.K[0]) + a, tg.K[0]);
    const int by = pmod32(m0[1] - 3 - lo1, tg.K[1]), bz = pmod32(m0[2] - 3 - lo2, tg.K[2]);
    int ly[6], lz[6];
    bool inside = lx < W0;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      ly[j] = wrapk(by + j, tg.K[1]);
      lz[j] = wrapk(bz + j, tg.K[2]);
      inside = inside && ly[j] < W1 && lz[j] < W2;
    }
    if (inside) {
#pragma unroll
      for (int bb = 0; bb < 6; ++bb) {
        double* row = tile + (lx * W1 + ly[bb]) * W2;
        const T qab = qa * wy[bb];
#pragma unroll
        for (int cc = 0; cc < 6; ++cc) atomicAdd(row + lz[cc], (double)(qab * wz[cc]));
      }
    } else {
      const int ix = wrapk(pmod32(m0[0] - 3, tg.K[0]) + a, tg.K[0]);
      const int gy = pmod32(m0[1] - 3, tg.K[1]), gz = pmod32(m0[2] - 3, tg.K[2]);
#pragma unroll
      for (int bb = 0; bb < 6; ++bb) {
        const int iy = wrapk(gy + bb, tg.K[1]);
        T* row = mesh + ((int64_t)ix * tg.K[1] + iy) * tg.K[2];
        const T qab = qa * wy[bb];
#pragma unroll
        for (int cc = 0; cc < 6; ++cc) unsafeAtomicAdd(row + wrapk(gz + cc, tg.K[2]), qab * wz[cc]);
      }
    }
  }
  __syncthreads();
  flush_tile(tg, tile, lo0, lo1, lo2, mesh);
}

// Six consecutive mesh values (element-aligned only) in two or three vector loads.
__device__ __forceinline__ void load6(const float* r, float p[6]) {
  typedef float f4 __attribute__((ext_vector_type(4), aligned(4)));
  typedef float f2 __attribute__((ext_vector_type(2), aligned(4)));
  const f4 u = *reinterpret_cast<const f4*>(r);
  const f2 v = *reinterpret_cast<const f2*>(r + 4);
  p[0] = u.x;
  p[1] = u.y;
  p[2] = u.z;
  p[3] = u.w;
  p[4] = v.x;
  p[5] = v.y;
}
__device__ __forceinline__ void load6(const double* r, double p[6]) {
  typedef double d2 __attribute__((ext_vector_type(2), aligned(8)));
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const d2 u = *reinterpret_cast<const d2*>(r + 2 * k);
    p[2 * k] = u.x;
    p[2 * k + 1] = u.y;
  }
}

// One lane's share of the gather for one atom (lane a < 6 owns x offset a);
// the caller sums the outputs over the atom's 8 lanes.  With phi = dE/dQ:
//   sx, sy, sz = sum_g phi(g) dtheta/du_d (u = f + j, du/dR = -1)
//   sp        = sum_g phi(g) theta(g)
template <typename T>
__device__ __forceinline__ void gather_lane(const Geom& g, double x, double y, double z, int a,
                                            const T* __restrict__ phi, T* sx, T* sy, T* sz, T* sp) {
  int m0[3];
  double f[3];
  mesh_ref(g, x, y, z, m0, f);
  T wx[6], wy[6], wz[6], dx[6], dy[6], dz[6];
  bspline6_rounded<T>(f[0], wx, dx);
  bspline6_rounded<T>(f[1], wy, dy);
  bspline6_rounded<T>(f[2], wz, dz);
  T ax = 0, ay = 0, az = 0, ap = 0;
  if (a < kOrder) {
    const int K0 = g.k[0], K1 = g.k[1], K2 = g.k[2];
    const int ix = wrapk(pmod32(m0[0] - 3, K0) + a, K0);
    const int by = pmod32(m0[1] - 3, K1), bz = pmod32(m0[2] - 3, K2);
    int iz[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) iz[c] = wrapk(bz + c, K2);
    const bool run = bz + 5 < K2;  // the 6 z points are consecutive (no wrap)
#pragma unroll
    for (int b = 0; b < 6; ++b) {
      const int iy = wrapk(by + b, K1);
      const T* row = phi + ((int64_t)ix * K1 + iy) * K2;
      T p[6];
      if (run) load6(row + bz, p);
      else {
#pragma unroll
        for (int c = 0; c < 6; ++c) p[c] = row[iz[c]];
      }
      T rz = 0, rdz = 0;
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        rz += p[c] * wz[c];
        rdz += p[c] * dz[c];
      }
      ap += wy[b] * rz;
      ay += dy[b] * rz;
      az += wy[b] * rdz;
    }
    ax = dx[a] * ap;
    ay *= wx[a];
    az *= wx[a];
    ap *= wx[a];
  }
  *sx = ax;
  *sy = ay;
  *sz = az;
  *sp = ap;
}

// gather_lane from a stencil record (reference point m and weights w of one
// atom); same sums in the same order as gather_lane.
template <typename T>
__device__ __forceinline__ void gather_lane_rec(int K0, int K1, int K2, int4 m,
                                                const T* __restrict__ wrec, int a,
                                                const T* __restrict__ phi, T* sx, T* sy, T* sz,
                                                T* sp) {
  const T* w = static_cast<const T*>(__builtin_assume_aligned(wrec, 16));
  T ax = 0, ay = 0, az = 0, ap = 0;
  if (a < kOrder) {
    T wy[6], wz[6], dy[6], dz[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      wy[j] = w[6 + j];
      wz[j] = w[12 + j];
      dy[j] = w[24 + j];
      dz[j] = w[30 + j];
    }
    const int ix = wrapk(pmod32(m.x - 3, K0) + a, K0);
    const int by = pmod32(m.y - 3, K1), bz = pmod32(m.z - 3, K2);
    int iz[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) iz[c] = wrapk(bz + c, K2);
    const bool run = bz + 5 < K2;  // the 6 z points are consecutive (no wrap)
#pragma unroll
    for (int b = 0; b < 6; ++b) {
      const int iy = wrapk(by + b, K1);
      const T* row = phi + ((int64_t)ix * K1 + iy) * K2;
      T p[6];
      if (run) load6(row + bz, p);
      else {
#pragma unroll
        for (int c = 0; c < 6; ++c) p[c] = row[iz[c]];
      }
      T rz = 0, rdz = 0;
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        rz += p[c] * wz[c];
        rdz += p[c] * dz[c];
      }
      ap += wy[b] * rz;
      ay += dy[b] * rz;
      az += wy[b] * rdz;
    }
    const T wxa = w[a], dxa = w[18 + a];
    ax = dxa * ap;
    ay *= wxa;
    az *= wxa;
    ap *= wxa;
  }
  *sx = ax;
  *sy = ay;
  *sz = az;
  *sp = ap;
}

}  // namespace admp
