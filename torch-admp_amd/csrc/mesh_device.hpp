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

// ---- owner-computes spread over mesh bricks
// A brick is a block of kBrickX x kBrickY x kBrickZ mesh points (smaller at the
// upper mesh faces), owned by one work-group of kBrickWaves wavefronts; lane
// (x, y) of every wavefront holds a column of kBrickZ points in float64
// registers.  The work-group collects every atom whose 6-point stencil
// reaches the brick from the cells of the binned atoms (packed positions in
// cell order), lays each collected atom's weights onto the brick's x, y and z
// lines, and each wavefront adds q wx wy wz of its share of them to its
// columns.  The wavefronts' sums are added in LDS and every point is stored
// once: no atomics, and the mesh needs no zeroing.  The brick holding an
// atom's reference point m0 (mod K) writes its stencil record for the gather.
constexpr int kBrickX = 8, kBrickY = 8, kBrickZ = 8;
constexpr int kBrickWaves = 4;
constexpr int kBrickThreads = 64 * kBrickWaves;
constexpr int kBrickPer = 4;      // candidates tested per thread and segment
constexpr int kBrickSeg = kBrickPer * kBrickThreads;  // candidates per segment
constexpr int kBrickList = kBrickSeg + kBrickThreads;  // collected atoms held (a pass runs
                                                       // once kBrickThreads are pending)
constexpr int kBrickCells = 64;   // cells whose candidates are scanned together
// Smallest mesh dimension of the brick path: with K >= kBrickZ + 6 a stencil
// meets a brick in one run of points (no second meeting through the wrap).
constexpr int kBrickMinK = 14;
// Mesh points per atom from which a system's bricks are kBrickWaves times
// deeper along z (the serving stream's rule meshes have ~300, the 24k-atom
// water box at 64^3 has 11).
constexpr int kDeepBrickPts = 64;
static_assert(kBrickX == 8 && kBrickY == 8 && kBrickZ == 8,
              "lane (x, y) of a wavefront owns one column; lines of 8 points");

__host__ __device__ inline int floordiv(long long a, long long b) {
  return (int)(a >= 0 ? a / b : -((-a + b - 1) / b));
}

// Inclusive prefix sum over the 64 lanes (row shifts, then row broadcasts).
__device__ __forceinline__ int wave_incl_scan(int v) {
  v += dpp_move<0x111, 0xf>(v);  // row_shr:1
  v += dpp_move<0x112, 0xf>(v);  // row_shr:2
  v += dpp_move<0x114, 0xf>(v);  // row_shr:4
  v += dpp_move<0x118, 0xf>(v);  // row_shr:8
  v += dpp_move<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_move<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
  return v;
}

// Collected atoms whose weights are laid out at once: one per thread, half
// the threads in float64 (the LDS of the lines).  With the brick deepened
// along z (ZW > 1, dense meshes: a few atoms per brick) 64 in both precisions.
template <typename T, int ZW = 1>
constexpr int brick_chunk() {
  return ZW > 1 ? 64 : (sizeof(T) == 4 ? kBrickThreads : kBrickThreads / 2);
}

// ZW = 1: a brick of kBrickX x kBrickY x kBrickZ points, the work-group's
// wavefronts split the collected atoms and their column sums are added in
// LDS.  ZW = kBrickWaves: a brick ZW times deeper along z, wavefront w owns
// the points [w kBrickZ, (w + 1) kBrickZ) of every column and adds every
// collected atom (no cross-wavefront sums); for meshes with many points per
// atom this takes ZW times fewer work-groups through the same scan.
template <typename T, int ZW = 1>
struct BrickStage {
  int slot[kBrickList];             // sorted slot of each collected atom
  int rel[kBrickList];              // stencil start - brick origin + 8, one byte per axis
  int cpre[kBrickCells + 1];        // prefix of the scanned cells' atom counts
  int cbeg[kBrickCells];            // first sorted slot of each scanned cell
  int wcnt[kBrickWaves];
  union {
    struct {
      // weights laid onto the brick's lines (zero where the stencil misses)
      __align__(16) T qx[brick_chunk<T, ZW>()][kBrickX];  // q wx
      __align__(16) T wy[brick_chunk<T, ZW>()][kBrickY];
      __align__(16) T wz[brick_chunk<T, ZW>()][kBrickZ * ZW];
    } w;
    double red[kBrickWaves - 1][kBrickZ * 64];  // the other wavefronts' column sums
  } u;
};

// Wrapped positions give 0 <= m0 <= K, so m0 - 3 - lo lies in (-2K, K).
__device__ __forceinline__ int brick_rel(int m0, int lo, int K) {
  int r = m0 - 3 - lo;
  r += r < 0 ? K : 0;
  r += r < 0 ? K : 0;
  return r;
}

// w (6 weights from offset s of an 8-point line) laid onto the line, times scale.
template <typename T>
__device__ __forceinline__ void lay_line(const T w[6], int s, T scale, T* __restrict__ line) {
  typedef T tv4 __attribute__((ext_vector_type(4)));
  T v[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    T t = T(0);
#pragma unroll
    for (int c = 0; c < 6; ++c) t = (u - s == c) ? scale * w[c] : t;
    v[u] = t;
  }
  tv4* l4 = reinterpret_cast<tv4*>(line);
  l4[0] = tv4{v[0], v[1], v[2], v[3]};
  l4[1] = tv4{v[4], v[5], v[6], v[7]};
}

// The same onto a line of N points (a multiple of 4): zeros, then the six
// weights that fall inside the line (one thread writes the whole line).
template <typename T, int N>
__device__ __forceinline__ void lay_line_n(const T w[6], int s, T scale, T* __restrict__ line) {
  typedef T tv4 __attribute__((ext_vector_type(4)));
  tv4* l4 = reinterpret_cast<tv4*>(line);
#pragma unroll
  for (int u = 0; u < N / 4; ++u) l4[u] = tv4{T(0), T(0), T(0), T(0)};
#pragma unroll
  for (int c = 0; c < 6; ++c)
    if (s + c >= 0 && s + c < N) line[s + c] = scale * w[c];
}

// The collected atoms [0, cnt), brick_chunk<T>() at a time: weights (float64
// B-splines rounded once to T, as every other spread path) laid onto the
// brick's lines, stencil records of the atoms this brick owns, then each
// wavefront's contiguous share of the column sums.  Ends with a barrier.
template <typename T, int ZW>
__device__ __forceinline__ void brick_pass(const Geom& g, const int lo[3], const int W[3], int cnt,
                                           BrickStage<T, ZW>& S, const P4<double>* __restrict__ ps,
                                           const int* __restrict__ sorted,
                                           const double* __restrict__ qv, int4* __restrict__ stm,
                                           T* __restrict__ stw, double (&acc)[kBrickZ]) {
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int x = lane >> 3, y = lane & 7;
  typedef T tv4 __attribute__((ext_vector_type(4)));
  constexpr int kChunk = brick_chunk<T, ZW>();
  for (int c0 = 0; c0 < cnt; c0 += kChunk) {
    const int k = c0 + tid;
    if (tid < kChunk && k < cnt) {
      const int slot = S.slot[k], rel = S.rel[k];
      const P4<double> p = ps[slot];
      const int iv = sorted[slot];
      int m0[3];
      double f[3];
      mesh_ref(g, p.x, p.y, p.z, m0, f);
      bool own = true;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const int o = (m0[d] == g.k[d] ? 0 : m0[d]) - lo[d];
        own = own && o >= 0 && o < W[d];
      }
      T* rec = stw + (int64_t)kStencilVals * iv;
      // one axis at a time (register pressure); the owner stores the record
      // piecewise: weights at 6 d, derivatives at 18 + 6 d
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        T w[6], dw[6];
        bspline6_rounded<T>(f[d], w, dw);
        const int s = ((rel >> (8 * d)) & 255) - 8;
        if (d == 0) lay_line<T>(w, s, (T)(qv ? qv[iv] : p.q), S.u.w.qx[tid]);
        else if (d == 1) lay_line<T>(w, s, T(1), S.u.w.wy[tid]);
        else if (ZW == 1) lay_line<T>(w, s, T(1), S.u.w.wz[tid]);
        else lay_line_n<T, kBrickZ * ZW>(w, s, T(1), S.u.w.wz[tid]);
        if (own) {
          typedef T tv2 __attribute__((ext_vector_type(2)));
          tv2* pw = reinterpret_cast<tv2*>(rec + 6 * d);
          tv2* pd = reinterpret_cast<tv2*>(rec + 18 + 6 * d);
#pragma unroll
          for (int j = 0; j < 3; ++j) {
            pw[j] = tv2{w[2 * j], w[2 * j + 1]};
            pd[j] = tv2{dw[2 * j], dw[2 * j + 1]};
          }
        }
      }
      if (own) stm[iv] = make_int4(m0[0], m0[1], m0[2], 0);
    }
    __syncthreads();
    const int nk = min(kChunk, cnt - c0);
    // ZW = 1: wavefront w takes its share of the atoms; ZW > 1: every atom,
    // at the z points [w kBrickZ, (w + 1) kBrickZ) of the lines
    const int qn = ZW > 1 ? nk : (nk + kBrickWaves - 1) / kBrickWaves;
    const int e1 = ZW > 1 ? nk : min((wave + 1) * qn, nk);
    int e = ZW > 1 ? 0 : wave * qn;
    const int zo = ZW > 1 ? wave * kBrickZ : 0;
    // four collected atoms a step, their LDS reads issued together
    for (; e + 4 <= e1; e += 4) {
      T a[4], b[4];
      tv4 z0[4], z1[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        a[u] = S.u.w.qx[e + u][x];
        b[u] = S.u.w.wy[e + u][y];
        z0[u] = reinterpret_cast<const tv4*>(S.u.w.wz[e + u] + zo)[0];
        z1[u] = reinterpret_cast<const tv4*>(S.u.w.wz[e + u] + zo)[1];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const T v = a[u] * b[u];
        acc[0] += (double)(v * z0[u].x);
        acc[1] += (double)(v * z0[u].y);
        acc[2] += (double)(v * z0[u].z);
        acc[3] += (double)(v * z0[u].w);
        acc[4] += (double)(v * z1[u].x);
        acc[5] += (double)(v * z1[u].y);
        acc[6] += (double)(v * z1[u].z);
        acc[7] += (double)(v * z1[u].w);
      }
    }
    for (; e < e1; ++e) {
      const T v = S.u.w.qx[e][x] * S.u.w.wy[e][y];
      const tv4 z0 = reinterpret_cast<const tv4*>(S.u.w.wz[e] + zo)[0];
      const tv4 z1 = reinterpret_cast<const tv4*>(S.u.w.wz[e] + zo)[1];
      acc[0] += (double)(v * z0.x);
      acc[1] += (double)(v * z0.y);
      acc[2] += (double)(v * z0.z);
      acc[3] += (double)(v * z0.w);
      acc[4] += (double)(v * z1.x);
      acc[5] += (double)(v * z1.y);
      acc[6] += (double)(v * z1.z);
      acc[7] += (double)(v * z1.w);
    }
    __syncthreads();
  }
}

// Brick lb (x-major over the brick grid nbr) of one system's mesh; launched
// with kBrickThreads threads.  Atoms come binned: cell c (global index cell0 +
// local, nc cells per axis of the fractional grid) holds the sorted slots
// [start[c * kSub], start[(c + 1) * kSub]) with packed wrapped positions
// ps[slot] and their mesh reference points mref[slot] (x | y << 16, z); the
// charges are ps[slot].q, or qv[atom] when qv is given.  Cell c holds atoms with ceil(K c / nc) - 1 <= m0 <= ceil(K (c +
// 1) / nc) up to the rounding of the binning, and a brick [lo, lo + W)
// receives from atoms with lo - 2 <= m0 <= lo + W + 2: the cells scanned cover
// that range with one more point on each side.
template <typename T, int ZW>
__device__ __forceinline__ void spread_brick(const Geom& g, const int nc[3], int64_t cell0,
                                             const int nbr[3], int64_t lb,
                                             const P4<double>* __restrict__ ps,
                                             const int2* __restrict__ mref,
                                             const int* __restrict__ sorted,
                                             const int* __restrict__ start,
                                             const double* __restrict__ qv, int4* __restrict__ stm,
                                             T* __restrict__ stw, T* __restrict__ mesh,
                                             BrickStage<T, ZW>& S) {
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int K[3] = {g.k[0], g.k[1], g.k[2]};
  const int bz = (int)(lb % nbr[2]), by = (int)((lb / nbr[2]) % nbr[1]),
            bx = (int)(lb / ((int64_t)nbr[1] * nbr[2]));
  const int lo[3] = {bx * kBrickX, by * kBrickY, bz * kBrickZ * ZW};
  const int W[3] = {min(kBrickX, K[0] - lo[0]), min(kBrickY, K[1] - lo[1]),
                    min(kBrickZ * ZW, K[2] - lo[2])};
  int cmin[3], clen[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int a = floordiv((long long)(lo[d] - 4) * nc[d], K[d]);
    const int e = floordiv((long long)(lo[d] + W[d] + 3) * nc[d], K[d]);
    clen[d] = min(e - a + 1, nc[d]);
    cmin[d] = clen[d] == nc[d] ? 0 : a;
  }
  double acc[kBrickZ];
#pragma unroll
  for (int z = 0; z < kBrickZ; ++z) acc[z] = 0.0;
  int cnt = 0;  // collected atoms not yet summed (the same in every thread)
  const int ncl = clen[0] * clen[1] * clen[2];
  for (int g0 = 0; g0 < ncl; g0 += kBrickCells) {
    const int m = min(kBrickCells, ncl - g0);
    if (wave == 0) {
      // cells g0 .. g0 + m - 1 and the inclusive prefix of their counts
      int n = 0;
      if (lane < m) {
        const int cc = g0 + lane;
        const int o2 = cc % clen[2], o1 = (cc / clen[2]) % clen[1], o0 = cc / (clen[1] * clen[2]);
        const int c0 = pmod32(cmin[0] + o0, nc[0]), c1 = pmod32(cmin[1] + o1, nc[1]),
                  c2 = pmod32(cmin[2] + o2, nc[2]);
        const int64_t cid = cell0 + ((int64_t)c0 * nc[1] + c1) * nc[2] + c2;
        const int b = start[cid * kSub];
        S.cbeg[lane] = b;
        n = start[(cid + 1) * kSub] - b;
      }
      n = wave_incl_scan(n);
      if (lane < m) S.cpre[lane + 1] = n;
      if (lane == 0) S.cpre[0] = 0;
    }
    __syncthreads();
    const int tot = S.cpre[m];
    // segments of kBrickSeg candidates; thread tid tests candidates
    // s0 + kBrickPer tid + r with their loads in flight together
    for (int s0 = 0; s0 < tot; s0 += kBrickSeg) {
      const int ns = min(kBrickSeg, tot - s0);
      const int j0 = kBrickPer * tid;
      int slot4[kBrickPer], rel4[kBrickPer];
      int nh = 0;
#pragma unroll
      for (int r = 0; r < kBrickPer; ++r) rel4[r] = -1;
      if (j0 < ns) {
        const int t0 = s0 + j0;
        int a = 0, b = m;  // largest e with cpre[e] <= t0
        while (b - a > 1) {
          const int mid = (a + b) >> 1;
          if (S.cpre[mid] <= t0) a = mid;
          else b = mid;
        }
        int2 mr[kBrickPer];
#pragma unroll
        for (int r = 0; r < kBrickPer; ++r) {
          const int t = t0 + r;
          while (a + 1 < m && S.cpre[a + 1] <= t) ++a;
          slot4[r] = S.cbeg[a] + (t - S.cpre[a]);
          if (j0 + r < ns) mr[r] = mref[slot4[r]];
        }
#pragma unroll
        for (int r = 0; r < kBrickPer; ++r) {
          if (j0 + r >= ns) continue;
          const int m0[3] = {mr[r].x & 0xffff, (int)((unsigned)mr[r].x >> 16), mr[r].y};
          bool hit = true;
          int rel = 0;
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const int q = brick_rel(m0[d], lo[d], K[d]);
            hit = hit && (q < W[d] || q + 5 >= K[d]);
            rel |= ((q < W[d] ? q : q - K[d]) + 8) << (8 * d);
          }
          rel4[r] = hit ? rel : -1;
          nh += hit ? 1 : 0;
        }
      }
      // collected atoms appended in candidate order
      const int inc = wave_incl_scan(nh);
      if (lane == 63) S.wcnt[wave] = inc;
      __syncthreads();
      int at = cnt + inc - nh, add = 0;
#pragma unroll
      for (int w = 0; w < kBrickWaves; ++w) {
        at += w < wave ? S.wcnt[w] : 0;
        add += S.wcnt[w];
      }
#pragma unroll
      for (int r = 0; r < kBrickPer; ++r)
        if (rel4[r] >= 0) {
          S.slot[at] = slot4[r];
          S.rel[at] = rel4[r];
          ++at;
        }
      cnt += add;
      __syncthreads();
      if (cnt >= kBrickThreads) {
        brick_pass<T, ZW>(g, lo, W, cnt, S, ps, sorted, qv, stm, stw, acc);
        cnt = 0;
      }
    }
    __syncthreads();  // every thread has read cpre before the next group rewrites it
  }
  if (cnt > 0) brick_pass<T, ZW>(g, lo, W, cnt, S, ps, sorted, qv, stm, stw, acc);
  if (ZW > 1) {
    // every wavefront stores its own z points of the columns
    const int x = lane >> 3, y = lane & 7, z0 = wave * kBrickZ;
    if (x < W[0] && y < W[1] && z0 < W[2]) {
      T* col = mesh + ((int64_t)(lo[0] + x) * K[1] + lo[1] + y) * K[2] + lo[2] + z0;
#pragma unroll
      for (int z = 0; z < kBrickZ; ++z)
        if (z0 + z < W[2]) col[z] = (T)acc[z];
    }
    return;
  }
  // the wavefronts' column sums; wavefront 0 stores the brick
  if (wave > 0) {
#pragma unroll
    for (int z = 0; z < kBrickZ; ++z) S.u.red[wave - 1][z * 64 + lane] = acc[z];
  }
  __syncthreads();
  if (wave == 0) {
    const int x = lane >> 3, y = lane & 7;
#pragma unroll
    for (int z = 0; z < kBrickZ; ++z)
#pragma unroll
      for (int w = 0; w < kBrickWaves - 1; ++w) acc[z] += S.u.red[w][z * 64 + lane];
    if (x < W[0] && y < W[1]) {
      T* col = mesh + ((int64_t)(lo[0] + x) * K[1] + lo[1] + y) * K[2] + lo[2];
#pragma unroll
      for (int z = 0; z < kBrickZ; ++z)
        if (z < W[2]) col[z] = (T)acc[z];
    }
  }
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
