// Cell-tiled charge spreading for gfx950.
//
// The atoms are already binned into the real-space cells (fractional-space
// grid nc[0] x nc[1] x nc[2]).  Cell c along dimension d owns the mesh points
// with ceil(K c / nc) <= m0 <= ceil(K (c + 1) / nc); with the 6-point stencil
// and one point of margin on each side its tile is
//   lo(c) = ceil(K c / nc) - 4,   W = ceil(K / nc) + 8
// points wide.  One work-group per cell accumulates q * theta of its atoms in
// an LDS tile (LDS atomics), then adds the tile's nonzero points to the mesh
// with global atomics: one atomic per point and cell instead of one per point
// and atom.  Atoms whose stencil leaves the tile (possible only through
// rounding at a cell face) add their weights to the mesh directly.
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
  }
  t.ntile = (int)std::min<size_t>(ntile, 1u << 30);
  if (ntile * sizeof(double) > kMaxTileBytes) t.ok = 0;
  if (t.W[2] > kBlock) t.ok = 0;  // the tile flush puts one tile row's z points on lanes
  if (cg.obc) t.ok = 0;
  return t;
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_spread_tiles(const double* __restrict__ pos,
                                                         const double* __restrict__ q, Geom g,
                                                         TileGeom tg, const int* __restrict__ sorted,
                                                         const int* __restrict__ start,
                                                         T* __restrict__ mesh) {
  const int c = blockIdx.x;
  spread_cell_tile<T>(g, tg, c, start[c * kSub], start[(c + 1) * kSub], sorted, pos, q, mesh);
}

template <typename T>
hipError_t launch_spread_tiled(const double* pos, const double* q, int64_t n, const Geom& g,
                               const TileGeom& tg, int ncell, const int* sorted, const int* start,
                               T* mesh, hipStream_t s) {
  const int64_t nm = (int64_t)g.k[0] * g.k[1] * g.k[2];
  k_fill<T><<<(int)std::min<int64_t>((nm + kBlock - 1) / kBlock, 4096), kBlock, 0, s>>>(mesh, nm,
                                                                                     T(0));
  if (n > 0) {
    const size_t sh = (size_t)tg.ntile * sizeof(double);
    k_spread_tiles<T><<<ncell, kBlock, sh, s>>>(pos, q, g, tg, sorted, start, mesh);
  }
  return hipGetLastError();
}

template hipError_t launch_spread_tiled<float>(const double*, const double*, int64_t, const Geom&,
                                               const TileGeom&, int, const int*, const int*,
                                               float*, hipStream_t);
template hipError_t launch_spread_tiled<double>(const double*, const double*, int64_t,
                                                const Geom&, const TileGeom&, int, const int*,
                                                const int*, double*, hipStream_t);

}  // namespace admp
