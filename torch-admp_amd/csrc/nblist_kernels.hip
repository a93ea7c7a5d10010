// Cell-list neighbour search, pair distances, pair energies and the fused
// cell-list real-space Ewald sum for gfx950.
//
// Replaces TorchNeighborList (src/torch_admp/nblist.py:304-551,
// src/torch_admp/spatial.py:15-36) and the P-sized torch chains of
// src/torch_admp/pme.py:312-349, 529-561 and src/torch_admp/qeq.py:106-121.
#include "nblist_kernels.hpp"
#include "real_device.hpp"

namespace admp {

// ---------------------------------------------------------------------------
// Cell grid in fractional space.  Along dimension d the grid has nc[d] cells;
// with nc >= 3 the stencil is the 3 neighbouring cells (periodic wrap with an
// image shift), otherwise the whole box is one cell and the stencil runs over
// the image shifts -nimg..nimg.  Cells are at least rc + 1e-6 h wide, so no
// pair within rc is missed.
CellGrid make_grid(const double* box, double cutoff, bool obc) {
  CellGrid cg;
  for (int i = 0; i < 9; ++i) cg.h[i] = box ? box[i] : (i % 4 == 0 ? 1.0 : 0.0);
  cg.obc = obc ? 1 : 0;
  cg.rc2 = cutoff * cutoff;
  if (obc) {
    for (int d = 0; d < 3; ++d) {
      cg.nc[d] = 1;
      cg.nimg[d] = 0;
    }
    for (int i = 0; i < 9; ++i) cg.a[i] = (i % 4 == 0 ? 1.0 : 0.0);
    cg.ncell = 1;
    return cg;
  }
  const double* h = box;
  const double c0[3] = {h[4] * h[8] - h[5] * h[7], h[5] * h[6] - h[3] * h[8], h[3] * h[7] - h[4] * h[6]};
  const double det = h[0] * c0[0] + h[1] * c0[1] + h[2] * c0[2];
  // inverse (A = h^-1) by cofactors
  cg.a[0] = c0[0] / det;
  cg.a[3] = c0[1] / det;
  cg.a[6] = c0[2] / det;
  cg.a[1] = (h[2] * h[7] - h[1] * h[8]) / det;
  cg.a[4] = (h[0] * h[8] - h[2] * h[6]) / det;
  cg.a[7] = (h[1] * h[6] - h[0] * h[7]) / det;
  cg.a[2] = (h[1] * h[5] - h[2] * h[4]) / det;
  cg.a[5] = (h[2] * h[3] - h[0] * h[5]) / det;
  cg.a[8] = (h[0] * h[4] - h[1] * h[3]) / det;
  const double vol = fabs(det);
  int64_t ncell = 1;
  for (int d = 0; d < 3; ++d) {
    const int e = (d + 1) % 3, f = (d + 2) % 3;
    const double cx = h[e * 3 + 1] * h[f * 3 + 2] - h[e * 3 + 2] * h[f * 3 + 1];
    const double cy = h[e * 3 + 2] * h[f * 3 + 0] - h[e * 3 + 0] * h[f * 3 + 2];
    const double cz = h[e * 3 + 0] * h[f * 3 + 1] - h[e * 3 + 1] * h[f * 3 + 0];
    const double height = vol / sqrt(cx * cx + cy * cy + cz * cz);
    int nc = (int)floor(height / (cutoff + 1e-6 * height));
    if (nc > 1024) nc = 1024;
    if (nc < 3) {
      cg.nc[d] = 1;
      cg.nimg[d] = (int)ceil(cutoff / height);
    } else {
      cg.nc[d] = nc;
      cg.nimg[d] = 1;
    }
    ncell *= cg.nc[d];
  }
  // bound the cell count (cells wider than rc stay correct, only slower)
  while (ncell > kMaxCells) {
    int d = 0;
    for (int k = 1; k < 3; ++k)
      if (cg.nc[k] > cg.nc[d]) d = k;
    ncell /= cg.nc[d];
    cg.nc[d] = std::max(3, cg.nc[d] / 2);
    ncell *= cg.nc[d];
  }
  cg.ncell = (int)ncell;
  return cg;
}

// Bin: wrapped fractional s = (r A + 1e-7) mod 1 - 1e-7 and wrapped
// position w = s h (src/torch_admp/nblist.py:447-470); cell of s.
__global__ void __launch_bounds__(kBlock) k_nb_bin(const double* __restrict__ pos, int64_t n,
                                                   CellGrid cg, double* __restrict__ w,
                                                   int* __restrict__ cell_of,
                                                   int* __restrict__ count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double r[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
  int cid = 0;
  if (cg.obc) {
    for (int d = 0; d < 3; ++d) w[3 * i + d] = r[d];
  } else {
    double s[3];
    int c[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double v = r[0] * cg.a[d] + r[1] * cg.a[3 + d] + r[2] * cg.a[6 + d];
      v = v + 1e-7;
      v = v - floor(v);
      s[d] = v - 1e-7;
      int cc = (int)floor(s[d] * cg.nc[d]);
      c[d] = cc < 0 ? 0 : (cc >= cg.nc[d] ? cg.nc[d] - 1 : cc);
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) w[3 * i + j] = s[0] * cg.h[j] + s[1] * cg.h[3 + j] + s[2] * cg.h[6 + j];
    cid = (c[0] * cg.nc[1] + c[1]) * cg.nc[2] + c[2];
  }
  cell_of[i] = cid;
  atomicAdd(&count[cid * kSub + (int)(i & (kSub - 1))], 1);
}

__global__ void __launch_bounds__(kBlock) k_nb_scatter(int64_t n, const int* __restrict__ cell_of,
                                                       const int* __restrict__ start,
                                                       int* __restrict__ cursor,
                                                       int* __restrict__ sorted) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int key = cell_of[i] * kSub + (int)(i & (kSub - 1));
  sorted[start[key] + atomicAdd(&cursor[key], 1)] = (int)i;
}

// Deterministic in-cell order: the rank of atom i among the atoms of its cell
// is the number of cell members with a smaller index (one thread per atom).
__global__ void __launch_bounds__(kBlock) k_nb_rank(int64_t n, const int* __restrict__ cell_of,
                                                    const int* __restrict__ start,
                                                    const int* __restrict__ unsorted,
                                                    int* __restrict__ sorted) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  const int i = unsorted[x];
  const int c = cell_of[i];
  const int b = start[c * kSub], e = start[(c + 1) * kSub];
  int rank = 0;
  for (int y = b; y < e; ++y) rank += unsorted[y] < i;
  sorted[b + rank] = i;
}

// Count (MODE 0) or write (MODE 1) the ordered pairs of the home atoms of one
// cell: PBC keeps 0.01 < |w_j + n h - w_i| < rc (src/torch_admp/nblist.py:433-445),
// OBC keeps |r_j - r_i| < rc, j != i (nblist.py:472-490).  Work-group
// (cell, split) owns the split-th contiguous share of the cell's sorted home
// slots.  Candidates of the full stencil (both orderings are listed) are
// staged in LDS in float64, kNbCand at a time; one wavefront takes one home
// atom, its lanes test 64 candidates per step and the survivors of a step are
// compacted by ballot.  The distance test compares squared distances and takes
// the square root only within 1e-12 (relative) of rc^2 or 1e-4, where it
// decides exactly as sqrt(d2) < rc and d > 0.01 do.  Both passes enumerate the
// pairs in the same order (stencil entry, then the candidate's cell slot), so
// pass 1 writes each home atom's pairs at the offset pass 0 counted: one row
// per atom in cell order (row h = the atom of sorted slot h), each row in
// ascending (stencil entry, slot).  With `keys`, pass 1 also stores that order
// as keys[o] = entry << 24 | slot of the second atom, which locates the
// reverse pair (k_nb_reverse).
constexpr int kNbThreads = 512;
constexpr int kNbCand = 2048;     // staged candidates per chunk
constexpr int kNbHomeBlk = 256;   // home atoms whose counters sit in LDS at once

__device__ __forceinline__ bool pair_kept(double d2, double rc, double rc2_lo, double rc2_hi,
                                          bool obc, bool same) {
  constexpr double kMin2 = 0.01 * 0.01;
  if (obc) {
    if (same || d2 >= rc2_hi) return false;
    if (d2 < rc2_lo) return true;
    return sqrt(d2) < rc;
  }
  if (d2 >= rc2_hi) return false;
  if (d2 < rc2_lo && d2 > kMin2 * (1.0 + 1e-12)) return true;
  const double d = sqrt(d2);
  return d < rc && d * d > kMin2 && d > 0.01;
}

template <int MODE>
__global__ void __launch_bounds__(kNbThreads) k_nb_cell_pairs(
    CellGrid cg, int S, const double* __restrict__ w, const int* __restrict__ start,
    const int* __restrict__ sorted, int64_t* __restrict__ pcount,
    const int64_t* __restrict__ offsets, int64_t* __restrict__ pairs, int64_t capacity,
    int* __restrict__ keys, int64_t* __restrict__ scales) {
  __shared__ double cx[kNbCand], cy[kNbCand], cz[kNbCand];
  __shared__ int cj[kNbCand];
  __shared__ int ckey[MODE == 1 ? kNbCand : 1];
  __shared__ int64_t hcnt[kNbHomeBlk];
  __shared__ int st_pre[kMaxStencil + 1];
  __shared__ int st_beg[kMaxStencil];
  __shared__ double st_sh[kMaxStencil * 3];
  const int c = blockIdx.x, split = blockIdx.y;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int nw = kNbThreads / 64;
  const int ci[3] = {c / (cg.nc[1] * cg.nc[2]), (c / cg.nc[2]) % cg.nc[1], c % cg.nc[2]};
  const int cb = start[c * kSub], ce = start[(c + 1) * kSub];
  const int share = (ce - cb + S - 1) / S;
  const int hb = cb + split * share, he = min(ce, hb + share);
  if (hb >= he) return;
  const double rc = sqrt(cg.rc2);
  const double rc2_lo = cg.rc2 * (1.0 - 1e-12), rc2_hi = cg.rc2 * (1.0 + 1e-12);
  const bool obc = cg.obc != 0;
  // rectangular cells: candidates farther than rc from the home cell's box
  // pair with no home atom (home atoms lie in the box up to the binning's
  // rounding; 1e-4 A of slack)
  const bool prune = !obc && cg.h[1] == 0.0 && cg.h[2] == 0.0 && cg.h[3] == 0.0 &&
                     cg.h[5] == 0.0 && cg.h[6] == 0.0 && cg.h[7] == 0.0;
  double blo[3], bhi[3];
  blo[0] = cg.h[0] * ci[0] / cg.nc[0];
  blo[1] = cg.h[4] * ci[1] / cg.nc[1];
  blo[2] = cg.h[8] * ci[2] / cg.nc[2];
  bhi[0] = cg.h[0] * (ci[0] + 1) / cg.nc[0];
  bhi[1] = cg.h[4] * (ci[1] + 1) / cg.nc[1];
  bhi[2] = cg.h[8] * (ci[2] + 1) / cg.nc[2];
  const double prune_r = rc + 1e-4, prune_r2 = prune_r * prune_r;
  const int nst = stencil_len(cg, 0) * stencil_len(cg, 1) * stencil_len(cg, 2);

  for (int h0 = hb; h0 < he; h0 += kNbHomeBlk) {
    const int nh = min(kNbHomeBlk, he - h0);
    for (int t = threadIdx.x; t < nh; t += kNbThreads) hcnt[t] = MODE == 1 ? offsets[h0 + t] : 0;
    for (int st0 = 0; st0 < nst; st0 += kMaxStencil) {
      __syncthreads();
      const int m = load_stencil_chunk(cg, 0, ci, st0, nst, start, st_pre, st_beg, st_sh);
      const int ctot = st_pre[m];
      for (int k0 = 0; k0 < ctot; k0 += kNbCand) {
        // stage this chunk's candidates that can reach the home cell, compacted
        // in candidate order (one ballot round of kNbThreads per pass)
        int cc = 0;
        for (int r0 = k0; r0 < min(ctot, k0 + kNbCand); r0 += kNbThreads) {
          const int kg = r0 + threadIdx.x;
          bool keep = false;
          double px = 0.0, py = 0.0, pz = 0.0;
          int j = 0, key = 0;
          if (kg < min(ctot, k0 + kNbCand)) {
            int lo = 0, hi = m;
            while (hi - lo > 1) {
              const int mid = (lo + hi) >> 1;
              if (st_pre[mid] <= kg) lo = mid;
              else hi = mid;
            }
            const int slot = st_beg[lo] + (kg - st_pre[lo]);
            j = sorted[slot];
            key = ((st0 + lo) << 24) | slot;
            px = w[3 * j] + st_sh[lo * 3];
            py = w[3 * j + 1] + st_sh[lo * 3 + 1];
            pz = w[3 * j + 2] + st_sh[lo * 3 + 2];
            keep = true;
            if (prune) {
              const double dx = fmax(0.0, fmax(blo[0] - px, px - bhi[0]));
              const double dy = fmax(0.0, fmax(blo[1] - py, py - bhi[1]));
              const double dz = fmax(0.0, fmax(blo[2] - pz, pz - bhi[2]));
              keep = dx * dx + dy * dy + dz * dz < prune_r2;
            }
          }
          __shared__ int wcnt[nw];
          const unsigned long long msk = __ballot(keep);
          if (lane == 0) wcnt[wave] = __popcll(msk);
          __syncthreads();
          int at = cc + __popcll(msk & ((1ull << lane) - 1ull)), tot = 0;
#pragma unroll
          for (int v = 0; v < nw; ++v) {
            at += v < wave ? wcnt[v] : 0;
            tot += wcnt[v];
          }
          if (keep) {
            cx[at] = px;
            cy[at] = py;
            cz[at] = pz;
            cj[at] = j;
            if (MODE == 1) ckey[at] = key;
          }
          cc += tot;
          __syncthreads();
        }
        // one wavefront per home atom, lanes over the staged candidates
        for (int t = wave; t < nh; t += nw) {
          const int i = sorted[h0 + t];
          const double wx = w[3 * i], wy = w[3 * i + 1], wz = w[3 * i + 2];
          int64_t cnt = hcnt[t];
          for (int k = lane; k - lane < cc; k += 64) {
            bool keep = false;
            int j = 0;
            if (k < cc) {
              j = cj[k];
              const double dx = cx[k] - wx, dy = cy[k] - wy, dz = cz[k] - wz;
              keep = pair_kept(dx * dx + dy * dy + dz * dz, rc, rc2_lo, rc2_hi, obc, j == i);
            }
            const unsigned long long msk = __ballot(keep);
            if (MODE == 1 && keep) {
              const int64_t o = cnt + __popcll(msk & ((1ull << lane) - 1ull));
              if (o < capacity) {
                pairs[2 * o] = i;
                pairs[2 * o + 1] = j;
                if (keys) keys[o] = ckey[k];
                if (scales) scales[o] = i < j ? 1 : 0;
              }
            }
            cnt += __popcll(msk);
          }
          if (lane == 0) hcnt[t] = cnt;
        }
        __syncthreads();
      }
    }
    __syncthreads();
    if (MODE == 0)
      for (int t = threadIdx.x; t < nh; t += kNbThreads) pcount[h0 + t] = hcnt[t];
  }
}

// ---------------------------------------------------------------------------
// Exclusive scan (3 kernels): block sums of 2048 items, single-block scan of
// the block sums, then the per-block scan with its offset.
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

template <typename I>
__device__ I block_exclusive_scan(I v, I* lds, I* total) {
  lds[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < kBlock; off <<= 1) {
    const I add = threadIdx.x >= off ? lds[threadIdx.x - off] : I(0);
    __syncthreads();
    lds[threadIdx.x] += add;
    __syncthreads();
  }
  const I incl = lds[threadIdx.x];
  if (total) *total = lds[kBlock - 1];
  __syncthreads();
  return incl - v;
}

template <typename I>
__global__ void __launch_bounds__(kBlock) k_scan_sums(const I* __restrict__ in, int64_t n,
                                                      I* __restrict__ sums) {
  __shared__ I lds[kBlock];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  I s = 0;
  for (int k = 0; k < kScanItems; ++k)
    if (base + k < n) s += in[base + k];
  I tot;
  block_exclusive_scan<I>(s, lds, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

template <typename I>
__global__ void __launch_bounds__(kBlock) k_scan_top(I* __restrict__ sums, int nb,
                                                     I* __restrict__ total) {
  __shared__ I lds[kBlock];
  I carry = 0;
  for (int b0 = 0; b0 < nb; b0 += kBlock) {
    const int i = b0 + threadIdx.x;
    const I v = i < nb ? sums[i] : I(0);
    I tot;
    const I ex = block_exclusive_scan<I>(v, lds, &tot);
    if (i < nb) sums[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

template <typename I>
__global__ void __launch_bounds__(kBlock) k_scan_apply(const I* __restrict__ in, int64_t n,
                                                       const I* __restrict__ sums,
                                                       I* __restrict__ out) {
  __shared__ I lds[kBlock];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  I vals[kScanItems];
  I s = 0;
  for (int k = 0; k < kScanItems; ++k) {
    vals[k] = base + k < n ? in[base + k] : I(0);
    s += vals[k];
  }
  I run = block_exclusive_scan<I>(s, lds, (I*)nullptr) + sums[blockIdx.x];
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) out[base + k] = run;
    run += vals[k];
  }
}

// Single-block exclusive scan for short arrays (n <= kScanSingleMax); writes
// out[0..n] with out[n] = total.  Wave w owns one contiguous segment and reads
// it 64 consecutive items at a time (coalesced), kScanBatch loads in flight:
// segment sums first, then each 64-item step is a wave inclusive scan plus the
// running carry of the segments before it.
constexpr int kScanSingleThreads = 1024;
constexpr int kScanSingleItems = 32;
constexpr int64_t kScanSingleMax = (int64_t)kScanSingleThreads * kScanSingleItems;
constexpr int kScanBatch = 8;

template <typename I>
__device__ __forceinline__ I wave_inclusive_scan(I v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const I u = __shfl_up(v, off, 64);
    if (lane >= off) v += u;
  }
  return v;
}

template <typename I>
__global__ void __launch_bounds__(kScanSingleThreads) k_scan_single(const I* __restrict__ in,
                                                                    int64_t n,
                                                                    I* __restrict__ out) {
  constexpr int kWaves = kScanSingleThreads / 64;
  __shared__ I wsum[kWaves];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t seg = ((n + kWaves - 1) / kWaves + 63) & ~(int64_t)63;
  const int64_t lo = (int64_t)w * seg < n ? (int64_t)w * seg : n;
  const int64_t hi = lo + seg < n ? lo + seg : n;
  I s = 0;
  for (int64_t c0 = lo; c0 < hi; c0 += 64 * kScanBatch) {
    I v[kScanBatch];
#pragma unroll
    for (int b = 0; b < kScanBatch; ++b) {
      const int64_t i = c0 + b * 64 + lane;
      v[b] = i < hi ? in[i] : I(0);
    }
#pragma unroll
    for (int b = 0; b < kScanBatch; ++b) s += v[b];
  }
  s = __shfl(wave_inclusive_scan(s), 63, 64);
  if (lane == 0) wsum[w] = s;
  __syncthreads();
  I carry = 0;
  for (int k = 0; k < w; ++k) carry += wsum[k];
  for (int64_t c0 = lo; c0 < hi; c0 += 64 * kScanBatch) {
    I v[kScanBatch];
#pragma unroll
    for (int b = 0; b < kScanBatch; ++b) {
      const int64_t i = c0 + b * 64 + lane;
      v[b] = i < hi ? in[i] : I(0);
    }
#pragma unroll
    for (int b = 0; b < kScanBatch; ++b) {
      const int64_t i = c0 + b * 64 + lane;
      const I incl = wave_inclusive_scan(v[b]);
      if (i < hi) out[i] = carry + incl - v[b];
      carry += __shfl(incl, 63, 64);
    }
  }
  if (threadIdx.x == 0) {
    I t = 0;
    for (int k = 0; k < kWaves; ++k) t += wsum[k];
    out[n] = t;
  }
}

template <typename I>
size_t scan_scratch_count(int64_t n) {
  return (size_t)((n + kScanTile - 1) / kScanTile) + 1;
}

// out[0..n] = exclusive scan of in[0..n-1] with out[n] = total (n + 1 entries).
template <typename I>
void launch_scan(const I* in, int64_t n, I* out, I* scratch, hipStream_t s) {
  const int nb = (int)((n + kScanTile - 1) / kScanTile);
  if (nb == 0) {
    k_fill<I><<<1, 64, 0, s>>>(out, 1, I(0));
    return;
  }
  if (n <= kScanSingleMax) {
    k_scan_single<I><<<1, kScanSingleThreads, 0, s>>>(in, n, out);
    return;
  }
  k_scan_sums<I><<<nb, kBlock, 0, s>>>(in, n, scratch);
  k_scan_top<I><<<1, kBlock, 0, s>>>(scratch, nb, out + n);
  k_scan_apply<I><<<nb, kBlock, 0, s>>>(in, n, scratch, out);
}

template void launch_scan<int>(const int*, int64_t, int*, int*, hipStream_t);
template void launch_scan<int64_t>(const int64_t*, int64_t, int64_t*, int64_t*, hipStream_t);

// ---------------------------------------------------------------------------
// Pair distances: dr = pbc_shift(r_i - r_j) (src/torch_admp/spatial.py:15-36)
__device__ __forceinline__ void pair_disp(const double* __restrict__ pos, int64_t i, int64_t j,
                                          const CellGrid& cg, double v[3], double nimg[3]) {
  if (cg.obc) {
    for (int d = 0; d < 3; ++d) {
      v[d] = pos[3 * j + d] - pos[3 * i + d];
      nimg[d] = 0.0;
    }
    return;
  }
  double x[3];
  for (int d = 0; d < 3; ++d) x[d] = pos[3 * i + d] - pos[3 * j + d];
  double s[3];
  for (int d = 0; d < 3; ++d) {
    s[d] = x[0] * cg.a[d] + x[1] * cg.a[3 + d] + x[2] * cg.a[6 + d];
    nimg[d] = floor(s[d] + 0.5);
    s[d] -= nimg[d];
  }
  for (int d = 0; d < 3; ++d) v[d] = s[0] * cg.h[d] + s[1] * cg.h[3 + d] + s[2] * cg.h[6 + d];
}

__global__ void __launch_bounds__(kBlock) k_pair_ds(const double* __restrict__ pos,
                                                    const int64_t* __restrict__ pairs,
                                                    int64_t np, CellGrid cg,
                                                    double* __restrict__ ds,
                                                    double* __restrict__ dr) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  double v[3], nimg[3];
  pair_disp(pos, pairs[2 * p], pairs[2 * p + 1], cg, v, nimg);
  ds[p] = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (dr)
    for (int d = 0; d < 3; ++d) dr[3 * p + d] = v[d];
}

// dL/dr and dL/dbox of L(ds).  The displacement is recomputed from the
// positions (or read from dr when given).  Pair lists grouped by their first
// atom put runs of equal i in consecutive lanes: a segmented inclusive scan
// over DPP row shifts sums each run piece within a row of 16 lanes and its
// last lane adds once; the second atom adds per pair.  The box term (image
// pairs only) is summed per work-group into partial[block][9], reduced in a
// fixed order by k_reduce_box.
constexpr int kDsBwdBlocks = 1024;
// the row form: up to one row per wavefront for 32k rows (a row's pairs are
// chains of dependent loads; more wavefronts in flight hide them: 24k rows
// 140 -> 120 us)
constexpr int kDsBwdRowBlocks = 8192;

__global__ void __launch_bounds__(kBlock) k_pair_ds_bwd(const double* __restrict__ gds,
                                                        const int64_t* __restrict__ pairs,
                                                        int64_t np, const double* __restrict__ pos,
                                                        CellGrid cg, const double* __restrict__ dr,
                                                        const double* __restrict__ ds,
                                                        double* __restrict__ gpos,
                                                        double* __restrict__ partial) {
  __shared__ double red[9 * (kBlock / 64)];
  double gb[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const int lane = threadIdx.x & 63;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  // every lane runs the same number of iterations (DPP needs the whole row)
  const int64_t iters = (np + stride - 1) / stride;
  for (int64_t it = 0; it < iters; ++it) {
    const int64_t p = it * stride + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    int key = -1 - lane;  // lanes past the end form runs of their own
    int64_t j = 0;
    bool act = false;
    if (p < np) {
      const int64_t i = pairs[2 * p];
      key = (int)i;  // rows with a zero gradient keep their run contiguous
      const double d = ds[p];
      const double g = gds[p];
      if (d > 0.0 && g != 0.0) {
        j = pairs[2 * p + 1];
        double v[3], nimg[3];
        if (dr) {
          for (int k = 0; k < 3; ++k) v[k] = dr[3 * p + k];
          if (!cg.obc) {
            double x[3];
            for (int k = 0; k < 3; ++k) x[k] = pos[3 * i + k] - pos[3 * j + k];
            for (int k = 0; k < 3; ++k)
              nimg[k] = floor(x[0] * cg.a[k] + x[1] * cg.a[3 + k] + x[2] * cg.a[6 + k] + 0.5);
          }
        } else {
          pair_disp(pos, i, j, cg, v, nimg);
        }
        // PBC: v = r_i - r_j - n h; OBC: v = r_j - r_i
        const double si = cg.obc ? -1.0 : 1.0;
        const double f = si * g / d;
        gx = f * v[0];
        gy = f * v[1];
        gz = f * v[2];
        act = true;
        if (!cg.obc)
          for (int c = 0; c < 3; ++c)
            if (nimg[c] != 0.0)
              for (int b = 0; b < 3; ++b) gb[c * 3 + b] -= g * nimg[c] * v[b] / d;
      }
    }
    if (act) {
      unsafeAtomicAdd(&gpos[3 * j], -gx);
      unsafeAtomicAdd(&gpos[3 * j + 1], -gy);
      unsafeAtomicAdd(&gpos[3 * j + 2], -gz);
    }
    // segmented inclusive sum of the i terms within each row of 16 lanes: rs is
    // the row lane where this lane's run of equal keys starts (a max-scan of
    // the run heads), and a lane adds the partial sum K lanes back only when
    // that lane lies in the same contiguous run.  Keys that recur further on
    // in the row start runs of their own, so any pair order is summed exactly.
    const int rl = lane & 15;
    const int prev = row_shr<1>(~key, key);
    int rs = (rl == 0 || prev != key) ? rl : 0;
    rs = max(rs, row_shr<1>(0, rs));
    rs = max(rs, row_shr<2>(0, rs));
    rs = max(rs, row_shr<4>(0, rs));
    rs = max(rs, row_shr<8>(0, rs));
#define ADMP_SEG_STEP(K)                                                            \
  {                                                                                 \
    const double tx = row_shr<K>(gx), ty = row_shr<K>(gy), tz = row_shr<K>(gz);     \
    const bool same = rl - K >= rs;                                                 \
    gx += same ? tx : 0.0;                                                          \
    gy += same ? ty : 0.0;                                                          \
    gz += same ? tz : 0.0;                                                          \
  }
    ADMP_SEG_STEP(1)
    ADMP_SEG_STEP(2)
    ADMP_SEG_STEP(4)
    ADMP_SEG_STEP(8)
#undef ADMP_SEG_STEP
    const int next = row_shl1(~key, key);
    if (key >= 0 && (rl == 15 || next != key) && (gx != 0.0 || gy != 0.0 || gz != 0.0)) {
      unsafeAtomicAdd(&gpos[3 * (int64_t)key], gx);
      unsafeAtomicAdd(&gpos[3 * (int64_t)key + 1], gy);
      unsafeAtomicAdd(&gpos[3 * (int64_t)key + 2], gz);
    }
  }
  if (partial) {
    block_sum<9>(gb, red);
    if (threadIdx.x == 0)
      for (int k = 0; k < 9; ++k) partial[(int64_t)blockIdx.x * 9 + k] = gb[k];
  }
}

// Stencil-entry starts of every row: es[i * (nst + 1) + e] = offset within
// row i of its first pair whose entry is >= e (keys[o] >> 24), es[.. + nst] =
// the row length.  One wavefront per row; lane k writes the starts of the
// entries between the previous pair's entry and its own.
__global__ void __launch_bounds__(kBlock) k_nb_entry_starts(int64_t n, int nst,
                                                            const int64_t* __restrict__ off,
                                                            const int* __restrict__ keys,
                                                            int* __restrict__ es) {
  const int lane = threadIdx.x & 63;
  const int64_t i = xcd_block(blockIdx.x, gridDim.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (i >= n) return;  // whole wavefronts
  const int64_t b = off[i], e = off[i + 1];
  int* row = es + i * (int64_t)(nst + 1);
  int prev = -1;  // entry of the pair before this chunk
  for (int64_t o0 = b; o0 < e; o0 += 64) {
    const int64_t o = o0 + lane;
    const int ent = o < e ? keys[o] >> 24 : nst;
    int pe = __shfl_up(ent, 1, 64);
    if (lane == 0) pe = prev;
    if (o < e)
      for (int x = pe + 1; x <= ent; ++x) row[x] = (int)(o - b);
    prev = __shfl(ent, (int)min<int64_t>(63, e - 1 - o0), 64);
  }
  for (int x = prev + 1 + lane; x <= nst; x += 64) row[x] = (int)(e - b);
}

// Reverse index of an ordered pair list built by k_nb_cell_pairs: row h (the
// atom of cell slot h) = [off[h], off[h + 1]), each row in ascending key
// order, keys[o] = stencil entry << 24 | slot of the second atom.  The pair
// (j, i) of o = (i, j) lies in the row of j's slot among the pairs of the
// mirrored entry nst - 1 - e (the stencil is symmetric about its centre; es
// gives where that entry's pairs start), with the slot of i: a binary search
// over that group of a few pairs.  rev[o] = its index, -1 if absent.  One
// wavefront per row; rows of one cell region stay on one XCD, so the rows
// searched (nearby cells) are mostly in its L2.
__global__ void __launch_bounds__(kBlock) k_nb_reverse(int64_t n, int nst,
                                                       const int64_t* __restrict__ off,
                                                       const int* __restrict__ keys,
                                                       const int* __restrict__ es,
                                                       int* __restrict__ rev) {
  const int lane = threadIdx.x & 63;
  const int64_t h = xcd_block(blockIdx.x, gridDim.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (h >= n) return;  // whole wavefronts
  const int64_t b = off[h], e = off[h + 1];
  for (int64_t o = b + lane; o < e; o += 64) {
    const int k = keys[o];
    const int sj = k & 0xffffff;
    const int em = nst - 1 - (k >> 24);
    const int target = (em << 24) | (int)h;
    const int* ej = es + (int64_t)sj * (nst + 1);
    const int64_t rj = off[sj];
    int64_t lo = rj + ej[em], hi = rj + ej[em + 1];
    int found = -1;
    if (hi > lo) {
      while (hi - lo > 1) {
        const int64_t mid = (lo + hi) >> 1;
        if (keys[mid] <= target) lo = mid;
        else hi = mid;
      }
      if (keys[lo] == target) found = (int)lo;
    }
    rev[o] = found;
  }
}

// dL/dr and dL/dbox of L(ds) for a list in row form: row r = pairs
// [row_ptr[r], row_ptr[r + 1]) share their first atom i (nrow rows, some
// empty) and rev[o] is the index of the reverse pair (j, i) of o = (i, j), or
// -1.  The second-atom term of pair rev(o) acts on atom i along -u_rev = u_o,
// so row r sums (g_o + g_rev(o)) d ds_o / d r_i over its own pairs: one
// wavefront per row, one float64 atomic per row component (grad_pos zeroed by
// the caller), no per-pair atomics.  A pair without a reverse adds its
// second-atom term by atomics.  Each wavefront takes a contiguous range of
// rows and each XCD a contiguous range of wavefronts, so the reverse pairs
// read (rows of nearby atoms in a cell-ordered list) stay in its L2.  The box
// term (image pairs, each pair's own g) goes to partial[block][9] as in
// k_pair_ds_bwd.
__global__ void __launch_bounds__(kBlock) k_pair_ds_bwd_rows(
    const double* __restrict__ gds, const int64_t* __restrict__ pairs, int64_t n,
    const int64_t* __restrict__ row_ptr, const int* __restrict__ rev,
    const double* __restrict__ pos, CellGrid cg, const double* __restrict__ dr,
    const double* __restrict__ ds, double* __restrict__ gpos, double* __restrict__ partial) {
  __shared__ double red[9 * (kBlock / 64)];
  double gb[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const int lane = threadIdx.x & 63;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const int64_t w = xcd_block(blockIdx.x, gridDim.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int64_t per = (n + nw - 1) / nw;
  const int64_t r0 = w * per, r1 = min(n, r0 + per);
  const double si = cg.obc ? -1.0 : 1.0;  // PBC: v = r_i - r_j - n h; OBC: v = r_j - r_i
  for (int64_t r = r0; r < r1; ++r) {
    const int64_t b = row_ptr[r], e = row_ptr[r + 1];
    if (e <= b) continue;  // wave-uniform
    const int64_t i = pairs[2 * b];
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int64_t o = b + lane; o < e; o += 64) {
      const double d = ds[o];
      const double g = gds[o];
      const int r = rev[o];
      const double gr = r >= 0 ? gds[r] : 0.0;
      if (d > 0.0 && (g != 0.0 || gr != 0.0)) {
        const int64_t j = pairs[2 * o + 1];
        double v[3], nimg[3];
        if (dr) {
          for (int k = 0; k < 3; ++k) v[k] = dr[3 * o + k];
          if (!cg.obc) {
            double x[3];
            for (int k = 0; k < 3; ++k) x[k] = pos[3 * i + k] - pos[3 * j + k];
            for (int k = 0; k < 3; ++k)
              nimg[k] = floor(x[0] * cg.a[k] + x[1] * cg.a[3 + k] + x[2] * cg.a[6 + k] + 0.5);
          }
        } else {
          pair_disp(pos, i, j, cg, v, nimg);
        }
        const double f = si * (g + gr) / d;
        sx += f * v[0];
        sy += f * v[1];
        sz += f * v[2];
        if (r < 0 && g != 0.0) {
          const double fj = si * g / d;
          unsafeAtomicAdd(&gpos[3 * j], -fj * v[0]);
          unsafeAtomicAdd(&gpos[3 * j + 1], -fj * v[1]);
          unsafeAtomicAdd(&gpos[3 * j + 2], -fj * v[2]);
        }
        if (!cg.obc && g != 0.0)
          for (int c = 0; c < 3; ++c)
            if (nimg[c] != 0.0)
              for (int k = 0; k < 3; ++k) gb[c * 3 + k] -= g * nimg[c] * v[k] / d;
      }
    }
    {
      sx = wave_sum(sx);
      sy = wave_sum(sy);
      sz = wave_sum(sz);
      if (lane == 0 && (sx != 0.0 || sy != 0.0 || sz != 0.0)) {
        unsafeAtomicAdd(&gpos[3 * i], sx);
        unsafeAtomicAdd(&gpos[3 * i + 1], sy);
        unsafeAtomicAdd(&gpos[3 * i + 2], sz);
      }
    }
  }
  if (partial) {
    block_sum<9>(gb, red);
    if (threadIdx.x == 0)
      for (int k = 0; k < 9; ++k) partial[(int64_t)blockIdx.x * 9 + k] = gb[k];
  }
}

// gbox[k] += sum over blocks of partial[b][k], in block order
__global__ void __launch_bounds__(kBlock) k_reduce_box(const double* __restrict__ partial,
                                                       int nblocks, double* __restrict__ gbox) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int k = wv; k < 9; k += kBlock / 64) {
    double s = 0.0;
    for (int b = lane; b < nblocks; b += 64) s += partial[(int64_t)b * 9 + k];
    s = wave_sum(s);
    if (lane == 0) gbox[k] += s;
  }
}

// ---------------------------------------------------------------------------
// Pair energies over a caller-supplied list; grid (blocks per frame, nf).
__global__ void __launch_bounds__(kBlock) k_pair_energy(int mode, int64_t n, int64_t np,
                                                        const double* __restrict__ q,
                                                        const double* __restrict__ eta,
                                                        const int64_t* __restrict__ pairs,
                                                        const double* __restrict__ ds,
                                                        const double* __restrict__ bs,
                                                        const int64_t* __restrict__ bsl,
                                                        double kappa, double eps,
                                                        double* __restrict__ partials,
                                                        double* __restrict__ gds,
                                                        double* __restrict__ gq,
                                                        double* __restrict__ geta) {
  __shared__ double red[kBlock / 64];
  const int64_t f = blockIdx.y;
  const double* qf = q + f * n;
  const double* etaf = eta ? eta + f * n : nullptr;
  const int64_t* pf = pairs + f * np * 2;
  double e_acc = 0.0;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < np;
       p += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = pf[2 * p], j = pf[2 * p + 1];
    const double d = ds[f * np + p];
    const double b = bsl ? (double)bsl[f * np + p] : bs[f * np + p];
    const double qi = qf[i], qj = qf[j];
    const double thr = mode == 1 ? 1e-8 : 1e-4;
    const bool ok = fabs(d) > thr;
    const double inv = ok ? 1.0 / d : 0.0;
    const double dinv = ok ? -inv * inv : 0.0;
    double s, ds_s;  // screening function and its derivative in d
    if (mode == 0) {
      s = erfc(kappa * d);
      ds_s = -2.0 / kSqrtPi * kappa * exp(-kappa * kappa * d * d);
    } else if (mode == 1) {
      s = 1.0;
      ds_s = 0.0;
    } else {
      double eij = sqrt((etaf[i] * etaf[i] + etaf[j] * etaf[j]) * 0.5);
      const bool guard = eij == 0.0;
      if (guard) eij = 1e-10;  // src/torch_admp/qeq.py:111
      const double x = d / (2.0 * eij);
      const double gx = exp(-x * x);
      s = -erfc(x);
      ds_s = 2.0 / kSqrtPi * gx / (2.0 * eij);
      if (geta && !guard && b != 0.0) {
        // d s / d eij = -(2 / sqrt(pi)) exp(-x^2) d / (2 eij^2); d eij / d eta_i = eta_i / (2 eij)
        const double de = eps * b * qi * qj * inv * (-gx * d / (kSqrtPi * eij * eij));
        unsafeAtomicAdd(&geta[f * n + i], de * etaf[i] / (2.0 * eij));
        unsafeAtomicAdd(&geta[f * n + j], de * etaf[j] / (2.0 * eij));
      }
    }
    const double pre = eps * b;
    const double e = pre * s * qi * qj * inv;
    e_acc += e;
    if (gds) gds[f * np + p] = pre * qi * qj * (ds_s * inv + s * dinv);
    if (gq && pre != 0.0) {
      unsafeAtomicAdd(&gq[f * n + i], pre * s * qj * inv);
      unsafeAtomicAdd(&gq[f * n + j], pre * s * qi * inv);
    }
  }
  double arr[1] = {e_acc};
  block_sum<1>(arr, red);
  if (threadIdx.x == 0) partials[f * gridDim.x + blockIdx.x] = arr[0];
}

// Buffer scales of an ordered pair list (src/torch_admp/nblist.py:492-515):
// 1 where i < j, else 0; with ds, also 0 unless lo < ds < hi.
__global__ void __launch_bounds__(kBlock) k_pair_scales(const int64_t* __restrict__ pairs,
                                                        int64_t np,
                                                        const double* __restrict__ ds, double lo,
                                                        double hi, int64_t* __restrict__ out) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < np;
       p += (int64_t)gridDim.x * blockDim.x) {
    bool keep = pairs[2 * p] < pairs[2 * p + 1];
    if (ds) {
      const double d = ds[p];
      keep = keep && d > lo && d < hi;
    }
    out[p] = keep ? 1 : 0;
  }
}

hipError_t launch_pair_scales(const int64_t* pairs, int64_t np, const double* ds, double lo,
                              double hi, int64_t* out, hipStream_t s) {
  if (np > 0) {
    const unsigned nb = (unsigned)std::min<int64_t>((np + kBlock - 1) / kBlock, 8192);
    k_pair_scales<<<nb, kBlock, 0, s>>>(pairs, np, ds, lo, hi, out);
  }
  return hipGetLastError();
}

__global__ void __launch_bounds__(kBlock) k_reduce_frames(const double* __restrict__ p, int nb,
                                                          double* __restrict__ out) {
  __shared__ double red[kBlock / 64];
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) s += p[(int64_t)blockIdx.x * nb + b];
  double arr[1] = {s};
  block_sum<1>(arr, red);
  if (threadIdx.x == 0) out[blockIdx.x] = arr[0];
}

// ---------------------------------------------------------------------------
// Cell-sorted copy of the wrapped positions and charges.  Positions stay in
// float64 so that pair displacements are formed without cancellation; the
// pair arithmetic itself runs in the bulk type.
__global__ void __launch_bounds__(kBlock) k_pack_sorted(int64_t n, const int* __restrict__ sorted,
                                                        const double* __restrict__ w,
                                                        const double* __restrict__ q,
                                                        P4<double>* __restrict__ ps) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = sorted[t];
  P4<double> v;
  v.x = w[3 * i];
  v.y = w[3 * i + 1];
  v.z = w[3 * i + 2];
  v.q = q[i];
  ps[t] = v;
}



// ---------------------------------------------------------------------------
// Largest displacement |r_i - ref_i| (minimum image under PBC) for the Verlet
// skin test.  Distances are non-negative, so their float64 bit patterns order
// like unsigned integers: one 64-bit atomic max per wavefront into out (zeroed
// by the launcher).
__global__ void __launch_bounds__(kBlock) k_max_disp(const double* __restrict__ pos,
                                                     const double* __restrict__ ref, int64_t n,
                                                     CellGrid cg,
                                                     unsigned long long* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double d2 = 0.0;
  if (i < n) {
    double v[3], nimg[3];
    double x[3];  // x = pos_i - ref_i, then its minimum image
    for (int k = 0; k < 3; ++k) x[k] = pos[3 * i + k] - ref[3 * i + k];
    if (cg.obc) {
      for (int k = 0; k < 3; ++k) v[k] = x[k];
    } else {
      double s[3];
      for (int k = 0; k < 3; ++k) {
        s[k] = x[0] * cg.a[k] + x[1] * cg.a[3 + k] + x[2] * cg.a[6 + k];
        nimg[k] = floor(s[k] + 0.5);
        s[k] -= nimg[k];
      }
      for (int k = 0; k < 3; ++k) v[k] = s[0] * cg.h[k] + s[1] * cg.h[3 + k] + s[2] * cg.h[6 + k];
    }
    d2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  }
  double m = sqrt(d2);
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

hipError_t launch_max_disp(const double* pos, const double* ref, int64_t n, const CellGrid& cg,
                           double* out, hipStream_t s) {
  k_fill<unsigned long long><<<1, 64, 0, s>>>((unsigned long long*)out, 1, 0ull);
  if (n > 0)
    k_max_disp<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(
        pos, ref, n, cg, (unsigned long long*)out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// host launchers

void nb_layout(int64_t n, const CellGrid& cg, NbLayout* L, bool with_entries) {
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += (bytes + 255) & ~size_t(255);
    return o;
  };
  L->w = take(sizeof(double) * 3 * (size_t)std::max<int64_t>(n, 1));
  L->cell_of = take(sizeof(int) * (size_t)std::max<int64_t>(n, 1));
  const size_t nkeys = (size_t)cg.ncell * kSub + 1;
  L->count = take(sizeof(int) * 2 * nkeys);  // counts, then the scatter cursors
  L->cursor = L->count + sizeof(int) * nkeys;
  L->start = take(sizeof(int) * nkeys);
  L->sorted = take(sizeof(int) * (size_t)std::max<int64_t>(n, 1));
  L->unsorted = take(sizeof(int) * (size_t)std::max<int64_t>(n, 1));
  L->pcount = take(sizeof(int64_t) * (size_t)(n + 1));
  L->offsets = take(sizeof(int64_t) * (size_t)(n + 1));
  const size_t sc = std::max(scan_scratch_count<int64_t>(n + 1), scan_scratch_count<int>(nkeys));
  L->scan = take(sizeof(int64_t) * sc);
  L->entry = with_entries
                 ? take(sizeof(int) * (size_t)std::max<int64_t>(n, 1) *
                        (size_t)(nb_stencil_entries(cg) + 1))
                 : 0;
  L->bytes = off;
}

hipError_t launch_bin(const double* pos, int64_t n, const CellGrid& cg, char* ws,
                      const NbLayout& L, hipStream_t s) {
  int* count = (int*)(ws + L.count);
  const int nfill = 2 * (cg.ncell * kSub + 1);
  k_fill<int><<<std::min((nfill + kBlock - 1) / kBlock, 4096), kBlock, 0, s>>>(count, nfill, 0);
  const unsigned nb = (unsigned)((n + kBlock - 1) / kBlock);
  if (n > 0)
    k_nb_bin<<<nb, kBlock, 0, s>>>(pos, n, cg, (double*)(ws + L.w), (int*)(ws + L.cell_of), count);
  return launch_cell_sort(n, cg.ncell, ws, L, s);
}

hipError_t launch_cell_scatter(int64_t n, int64_t ncell, char* ws, const NbLayout& L,
                               hipStream_t s) {
  int* count = (int*)(ws + L.count);
  int* start = (int*)(ws + L.start);
  int* cursor = (int*)(ws + L.cursor);
  launch_scan<int>(count, ncell * kSub, start, (int*)(ws + L.scan), s);
  if (n > 0)
    k_nb_scatter<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(
        n, (int*)(ws + L.cell_of), start, cursor, (int*)(ws + L.unsorted));
  return hipGetLastError();
}

hipError_t launch_cell_sort(int64_t n, int64_t ncell, char* ws, const NbLayout& L, hipStream_t s) {
  hipError_t e = launch_cell_scatter(n, ncell, ws, L, s);
  if (e == hipSuccess && n > 0) {
    const int* start = (const int*)(ws + L.start);
    k_nb_rank<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(
        n, (int*)(ws + L.cell_of), start, (int*)(ws + L.unsorted), (int*)(ws + L.sorted));
  }
  return hipGetLastError();
}

hipError_t launch_pairs(int mode, int64_t n, const CellGrid& cg, char* ws, const NbLayout& L,
                        int64_t* pairs, int64_t capacity, int* keys, int64_t* scales,
                        hipStream_t s) {
  const unsigned nb = (unsigned)((n + kBlock - 1) / kBlock);
  if (n == 0) return hipGetLastError();
  const double* w = (const double*)(ws + L.w);
  const int* cell_of = (const int*)(ws + L.cell_of);
  const int* start = (const int*)(ws + L.start);
  const int* sorted = (const int*)(ws + L.sorted);
  int64_t* pc = (int64_t*)(ws + L.pcount);
  int64_t* off = (int64_t*)(ws + L.offsets);
  (void)nb;
  (void)cell_of;
  // home splits per cell: at least ~512 work-groups, at most one per 64 home atoms
  const int64_t ncell = cg.ncell;
  const int max_occ = (int)((n + ncell - 1) / ncell);
  int S = (int)std::max<int64_t>(1, (512 + ncell - 1) / ncell);
  S = std::max(1, std::min(S, (max_occ + 63) / 64));
  const dim3 grid((unsigned)ncell, (unsigned)S);
  if (mode == 0) {
    k_nb_cell_pairs<0><<<grid, kNbThreads, 0, s>>>(cg, S, w, start, sorted, pc, nullptr, nullptr,
                                                   0, nullptr, nullptr);
    launch_scan<int64_t>(pc, n, off, (int64_t*)(ws + L.scan), s);
  } else {
    k_nb_cell_pairs<1><<<grid, kNbThreads, 0, s>>>(cg, S, w, start, sorted, nullptr, off, pairs,
                                                   capacity, keys, scales);
  }
  return hipGetLastError();
}

int nb_stencil_entries(const CellGrid& cg) {
  int m = 1;
  for (int d = 0; d < 3; ++d) m *= cg.nc[d] >= 3 ? 3 : 2 * cg.nimg[d] + 1;
  return m;
}

bool nb_reverse_supported(int64_t n, const CellGrid& cg, int64_t np) {
  return n < ((int64_t)1 << 24) && nb_stencil_entries(cg) <= 127 && np < ((int64_t)1 << 31);
}

hipError_t launch_nb_reverse(int64_t n, const CellGrid& cg, char* ws, const NbLayout& L,
                             const int* keys, int64_t np, int64_t* row_ptr, int* rev,
                             hipStream_t s) {
  const int64_t* off = (const int64_t*)(ws + L.offsets);
  if (row_ptr)
    if (hipMemcpyAsync(row_ptr, off, sizeof(int64_t) * (size_t)(n + 1), hipMemcpyDeviceToDevice,
                       s) != hipSuccess)
      return hipGetLastError();
  if (rev && np > 0) {
    if (!keys || !L.entry || !nb_reverse_supported(n, cg, np)) {
      k_fill<int><<<(unsigned)std::min<int64_t>((np + kBlock - 1) / kBlock, 4096), kBlock, 0, s>>>(
          rev, np, -1);
    } else {
      const unsigned nb = (unsigned)((n * 64 + kBlock - 1) / kBlock);
      int* es = (int*)(ws + L.entry);
      const int nst = nb_stencil_entries(cg);
      k_nb_entry_starts<<<nb, kBlock, 0, s>>>(n, nst, off, keys, es);
      k_nb_reverse<<<nb, kBlock, 0, s>>>(n, nst, off, keys, es, rev);
    }
  }
  return hipGetLastError();
}

hipError_t launch_pair_ds(const double* pos, const int64_t* pairs, int64_t np, const CellGrid& cg,
                          double* ds, double* dr, hipStream_t s) {
  if (np > 0)
    k_pair_ds<<<(unsigned)((np + kBlock - 1) / kBlock), kBlock, 0, s>>>(pos, pairs, np, cg, ds, dr);
  return hipGetLastError();
}

size_t pair_ds_bwd_scratch_bytes(int64_t np) {
  (void)np;
  return sizeof(double) * 9 * std::max(kDsBwdBlocks, kDsBwdRowBlocks);
}

hipError_t launch_pair_ds_bwd(const double* gds, const int64_t* pairs, int64_t np,
                              const double* pos, const CellGrid& cg, const double* dr,
                              const double* ds, double* gpos, double* gbox, double* partial,
                              const int64_t* row_ptr, const int* rev, int64_t n,
                              hipStream_t s) {
  if (np <= 0) return hipGetLastError();
  const bool box_term = gbox && !cg.obc;
  int nb;
  if (row_ptr && rev) {
    nb = (int)std::min<int64_t>((n * 64 + kBlock - 1) / kBlock, kDsBwdRowBlocks);
    if (nb > 0)
      k_pair_ds_bwd_rows<<<nb, kBlock, 0, s>>>(gds, pairs, n, row_ptr, rev, pos, cg, dr, ds, gpos,
                                               box_term ? partial : nullptr);
  } else {
    nb = (int)std::min<int64_t>((np + kBlock - 1) / kBlock, kDsBwdBlocks);
    k_pair_ds_bwd<<<nb, kBlock, 0, s>>>(gds, pairs, np, pos, cg, dr, ds, gpos,
                                        box_term ? partial : nullptr);
  }
  if (nb <= 0) return hipGetLastError();
  if (box_term) k_reduce_box<<<1, kBlock, 0, s>>>(partial, nb, gbox);
  return hipGetLastError();
}

hipError_t launch_pair_energy(int mode, int64_t nf, int64_t n, int64_t np, const double* q,
                              const double* eta, const int64_t* pairs, const double* ds,
                              const double* bs, const int64_t* bsl, double kappa, double eps,
                              double* energy, double* gds, double* gq, double* geta,
                              double* partials, hipStream_t s) {
  // one pair per thread up to kPairBlocksPerFrame work-groups a frame (the
  // partial sums are then reduced in that fixed order)
  const int bpf = (int)std::max<int64_t>(
      1, std::min<int64_t>((np + kBlock - 1) / kBlock, kPairBlocksPerFrame));
  const int64_t m = nf * n;
  if (m > 0) {
    const unsigned nb = (unsigned)std::min<int64_t>((m + kBlock - 1) / kBlock, 4096);
    if (gq) k_fill<double><<<nb, kBlock, 0, s>>>(gq, m, 0.0);
    if (geta) k_fill<double><<<nb, kBlock, 0, s>>>(geta, m, 0.0);
  }
  dim3 grid(bpf, (unsigned)nf);
  k_pair_energy<<<grid, kBlock, 0, s>>>(mode, n, np, q, eta, pairs, ds, bs, bsl, kappa, eps,
                                        partials, gds, gq, geta);
  k_reduce_frames<<<(unsigned)nf, kBlock, 0, s>>>(partials, bpf, energy);
  return hipGetLastError();
}

hipError_t launch_pack(const double* q, int64_t n, char* ws, const NbLayout& L, P4<double>* ps,
                       hipStream_t s) {
  if (n > 0)
    k_pack_sorted<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(
        n, (const int*)(ws + L.sorted), (const double*)(ws + L.w), q, ps);
  return hipGetLastError();
}

}  // namespace admp
