// Cell-tiled spreading (see tiles.hip).
#pragma once

#include "admp_common.hpp"
#include "nblist_kernels.hpp"

namespace admp {

constexpr size_t kMaxTileBytes = 65536;  // LDS budget of one tile

struct TileGeom {
  int nc[3];
  int K[3];
  int W[3];
  int ntile;  // W0 * W1 * W2
  int ok;     // the tile fits LDS
};

// Tile geometry for the cell grid; ok == 0 when the tiled path does not apply.
TileGeom make_tile_geom(const CellGrid& cg, const int32_t* kmesh, size_t elem_bytes);

template <typename T>
hipError_t launch_spread_tiled(const double* pos, const double* q, int64_t n, const Geom& g,
                               const TileGeom& tg, int ncell, const int* sorted, const int* start,
                               T* mesh, hipStream_t s);

}  // namespace admp
