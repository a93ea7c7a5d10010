// Cell-tiled charge spreading for gfx950.
//
// The atoms are already binned into the real-space cells (fractional-space
// grid nc[0] x nc[1] x nc[2]).  Cell c along dimension d owns the mesh points
// with ceil(K c / nc) <= m0 <= ceil(K (c + 1) / nc); with the 6-point stencil
// and one point of margin on each side its tile is
//   lo(c) = ceil(K c / nc) - 4,   W = ceil(K / nc) + 8
// points wide.  One work-group per cell accumulates q * theta of its atoms in
// an LDS tile (LDS atomics), then writes the whole tile to a per-cell buffer
// without global atomics.  A second kernel forms every mesh value as the sum
// of the (at most a few) tiles that cover it, in a fixed order.  Atoms whose
// stencil leaves the tile (possible only through rounding at a cell face) add
// their weights to the mesh with global atomics instead, so the result is
// exact in every case.
//
// Same values as the reference spread (src/torch_admp/recip.py:216-252).
#include "nblist_kernels.hpp"
#include "pme_kernels.hpp"
#include "tiles.hpp"

namespace admp {

__host__ __device__ inline int tile_lo(int K, int nc, int c) {
  return (int)(((long long)K * c + nc - 1) / nc) - 4;
}

TileGeom make_tile_geom(const CellGrid& cg, const int32_t* kmesh, size_t elem_bytes) {
  (void)elem_bytes;  // the LDS tile is float64 in both precisions
  TileGeom t;
  t.ok = 1;
  size_t ntile = 1;
  for (int d = 0; d < 3; ++d) {
    t.nc[d] = cg.nc[d];
    t.K[d] = kmesh[d];
    t.W[d] = (kmesh[d] + cg.nc[d] - 1) / cg.nc[d] + 8;
    ntile *= t.W[d];
    // largest number of (tile, offset) pairs that cover one mesh index
    int maxc = 0;
    for (int gx = 0; gx < kmesh[d]; ++gx) {
      int cnt = 0;
      for (int c = 0; c < cg.nc[d]; ++c)
        for (int o = pmod((long long)gx - tile_lo(kmesh[d], cg.nc[d], c), kmesh[d]); o < t.W[d];
             o += kmesh[d])
          ++cnt;
      maxc = std::max(maxc, cnt);
    }
    if (maxc > kMaxCover) t.ok = 0;
  }
  t.ntile = (int)std::min<size_t>(ntile, 1u << 30);
  if (ntile * sizeof(double) > kMaxTileBytes) t.ok = 0;
  if (cg.obc) t.ok = 0;
  return t;
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_spread_tiles(const double* __restrict__ pos,
                                                         const double* __restrict__ q, Geom g,
                                                         TileGeom tg, const int* __restrict__ sorted,
                                                         const int* __restrict__ start,
                                                         T* __restrict__ tiles,
                                                         T* __restrict__ mesh) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* tile = reinterpret_cast<double*>(smem_raw);  // float64 LDS atomics in both precisions
  const int c = blockIdx.x;
  const int W0 = tg.W[0], W1 = tg.W[1], W2 = tg.W[2];
  for (int k = threadIdx.x; k < tg.ntile; k += blockDim.x) tile[k] = 0.0;
  __syncthreads();
  const int cz = c % tg.nc[2];
  const int cy = (c / tg.nc[2]) % tg.nc[1];
  const int cx = c / (tg.nc[1] * tg.nc[2]);
  const int lo0 = tile_lo(tg.K[0], tg.nc[0], cx);
  const int lo1 = tile_lo(tg.K[1], tg.nc[1], cy);
  const int lo2 = tile_lo(tg.K[2], tg.nc[2], cz);
  const int b = start[c * kSub], e = start[(c + 1) * kSub];
  const int a = threadIdx.x & 7;
  for (int slot = b + (threadIdx.x >> 3); slot < e; slot += blockDim.x >> 3) {
    if (a >= kOrder) continue;
    const int i = sorted[slot];
    int m0[3];
    double f[3];
    mesh_ref(g, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], m0, f);
    T wx[6], wy[6], wz[6], dd[6];
    bspline6<T>((T)f[0], wx, dd);
    bspline6<T>((T)f[1], wy, dd);
    bspline6<T>((T)f[2], wz, dd);
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
  T* out = tiles + (int64_t)c * tg.ntile;
  for (int k = threadIdx.x; k < tg.ntile; k += blockDim.x) out[k] = (T)tile[k];
}

// (tile, offset) pairs covering mesh index x along one dimension
__device__ __forceinline__ int cover(int x, int K, int nc, int W, int* t, int* o) {
  int n = 0;
  for (int c = 0; c < nc; ++c) {
    // x - lo(c) lies in (-K, K + 4]
    int off = x - tile_lo(K, nc, c);
    if (off < 0) off += K;
    if (off >= K) off -= K;
    for (; off < W && n < kMaxCover; off += K) {
      t[n] = c;
      o[n] = off;
      ++n;
    }
  }
  return n;
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_spread_reduce(const T* __restrict__ tiles, TileGeom tg,
                                                          T* __restrict__ mesh) {
  const int64_t nm = (int64_t)tg.K[0] * tg.K[1] * tg.K[2];
  for (int64_t gidx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; gidx < nm;
       gidx += (int64_t)gridDim.x * blockDim.x) {
    const int gz = (int)(gidx % tg.K[2]);
    const int gy = (int)((gidx / tg.K[2]) % tg.K[1]);
    const int gx = (int)(gidx / ((int64_t)tg.K[1] * tg.K[2]));
    int tx[kMaxCover], ox[kMaxCover], ty[kMaxCover], oy[kMaxCover], tz[kMaxCover], oz[kMaxCover];
    const int nx = cover(gx, tg.K[0], tg.nc[0], tg.W[0], tx, ox);
    const int ny = cover(gy, tg.K[1], tg.nc[1], tg.W[1], ty, oy);
    const int nz = cover(gz, tg.K[2], tg.nc[2], tg.W[2], tz, oz);
    T s = T(0);
    for (int a = 0; a < nx; ++a)
      for (int b = 0; b < ny; ++b) {
        const int64_t cab = ((int64_t)tx[a] * tg.nc[1] + ty[b]) * tg.nc[2];
        const int oab = (ox[a] * tg.W[1] + oy[b]) * tg.W[2];
        for (int c = 0; c < nz; ++c) s += tiles[(cab + tz[c]) * tg.ntile + oab + oz[c]];
      }
    mesh[gidx] += s;
  }
}

template <typename T>
hipError_t launch_spread_tiled(const double* pos, const double* q, int64_t n, const Geom& g,
                               const TileGeom& tg, int ncell, const int* sorted, const int* start,
                               T* tiles, T* mesh, hipStream_t s) {
  const int64_t nm = (int64_t)g.k[0] * g.k[1] * g.k[2];
  k_fill<T><<<(int)std::min<int64_t>((nm + kBlock - 1) / kBlock, 4096), kBlock, 0, s>>>(mesh, nm,
                                                                                     T(0));
  if (n > 0) {
    const size_t sh = (size_t)tg.ntile * sizeof(double);
    k_spread_tiles<T><<<ncell, kBlock, sh, s>>>(pos, q, g, tg, sorted, start, tiles, mesh);
    k_spread_reduce<T><<<(int)std::min<int64_t>((nm + kBlock - 1) / kBlock, 8192), kBlock, 0, s>>>(
        tiles, tg, mesh);
  }
  return hipGetLastError();
}

template hipError_t launch_spread_tiled<float>(const double*, const double*, int64_t, const Geom&,
                                               const TileGeom&, int, const int*, const int*,
                                               float*, float*, hipStream_t);
template hipError_t launch_spread_tiled<double>(const double*, const double*, int64_t,
                                                const Geom&, const TileGeom&, int, const int*,
                                                const int*, double*, double*, hipStream_t);

}  // namespace admp
