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
#include "mesh_device.hpp"

namespace admp {

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
  const int c = blockIdx.x;
  spread_cell_tile<T>(g, tg, c, start[c * kSub], start[(c + 1) * kSub], sorted, pos, q, mesh,
                      tiles + (int64_t)c * tg.ntile);
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_spread_reduce(const T* __restrict__ tiles, TileGeom tg,
                                                          T* __restrict__ mesh) {
  const int64_t nm = (int64_t)tg.K[0] * tg.K[1] * tg.K[2];
  for (int64_t gidx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; gidx < nm;
       gidx += (int64_t)gridDim.x * blockDim.x)
    mesh[gidx] += tile_sum<T>(tg, gidx, tiles);
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
