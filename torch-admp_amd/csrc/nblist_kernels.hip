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

// Count (MODE 0) or write (MODE 1) the ordered pairs of atom i:
// PBC keeps 0.01 < |w_j + n h - w_i| < rc (src/torch_admp/nblist.py:433-445);
// OBC keeps |r_j - r_i| < rc, j != i (nblist.py:472-490).
template <int MODE>
__global__ void __launch_bounds__(kBlock) k_nb_pairs(int64_t n, CellGrid cg,
                                                     const double* __restrict__ w,
                                                     const int* __restrict__ cell_of,
                                                     const int* __restrict__ start,
                                                     const int* __restrict__ sorted,
                                                     int64_t* __restrict__ counts,
                                                     const int64_t* __restrict__ offsets,
                                                     int64_t* __restrict__ pairs,
                                                     int64_t capacity) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = sorted[t];
  const double wi[3] = {w[3 * i], w[3 * i + 1], w[3 * i + 2]};
  const int cid = cell_of[i];
  const int ci[3] = {cid / (cg.nc[1] * cg.nc[2]), (cid / cg.nc[2]) % cg.nc[1], cid % cg.nc[2]};
  const double rmin2 = 0.01 * 0.01;
  const double rc = sqrt(cg.rc2);
  int64_t cnt = 0;
  int64_t out = MODE == 1 ? offsets[i] : 0;
  const int l0 = stencil_len(cg, 0), l1 = stencil_len(cg, 1), l2 = stencil_len(cg, 2);
  for (int o0 = 0; o0 < l0; ++o0) {
    int c0, s0;
    stencil_at(cg, 0, ci[0], o0, &c0, &s0);
    for (int o1 = 0; o1 < l1; ++o1) {
      int c1, s1;
      stencil_at(cg, 1, ci[1], o1, &c1, &s1);
      for (int o2 = 0; o2 < l2; ++o2) {
        int c2, s2;
        stencil_at(cg, 2, ci[2], o2, &c2, &s2);
        double sh[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) sh[j] = s0 * cg.h[j] + s1 * cg.h[3 + j] + s2 * cg.h[6 + j];
        const int cc = (c0 * cg.nc[1] + c1) * cg.nc[2] + c2;
        const int xe = start[(cc + 1) * kSub];
        for (int x = start[cc * kSub]; x < xe; ++x) {
          const int j = sorted[x];
          const double dx = (w[3 * j] + sh[0]) - wi[0];
          const double dy = (w[3 * j + 1] + sh[1]) - wi[1];
          const double dz = (w[3 * j + 2] + sh[2]) - wi[2];
          const double d = sqrt(dx * dx + dy * dy + dz * dz);
          bool keep;
          if (cg.obc) keep = (j != i) && d < rc;
          else keep = d < rc && d * d > rmin2 && d > 0.01;
          if (keep) {
            if (MODE == 1 && out + cnt < capacity) {
              pairs[2 * (out + cnt)] = i;
              pairs[2 * (out + cnt) + 1] = j;
            }
            ++cnt;
          }
        }
      }
    }
  }
  if (MODE == 0) counts[i] = cnt;
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
// out[0..n] with out[n] = total.
constexpr int kScanSingleThreads = 1024;
constexpr int kScanSingleItems = 16;
constexpr int64_t kScanSingleMax = (int64_t)kScanSingleThreads * kScanSingleItems;

template <typename I>
__global__ void __launch_bounds__(kScanSingleThreads) k_scan_single(const I* __restrict__ in,
                                                                    int64_t n,
                                                                    I* __restrict__ out) {
  __shared__ I lds[kScanSingleThreads];
  const int per = (int)((n + kScanSingleThreads - 1) / kScanSingleThreads);
  const int64_t base = (int64_t)threadIdx.x * per;
  I s = 0;
  for (int k = 0; k < per; ++k)
    if (base + k < n) s += in[base + k];
  lds[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < kScanSingleThreads; off <<= 1) {
    const I add = threadIdx.x >= off ? lds[threadIdx.x - off] : I(0);
    __syncthreads();
    lds[threadIdx.x] += add;
    __syncthreads();
  }
  I run = lds[threadIdx.x] - s;
  for (int k = 0; k < per; ++k)
    if (base + k < n) {
      const I v = in[base + k];
      out[base + k] = run;
      run += v;
    }
  if (threadIdx.x == kScanSingleThreads - 1) out[n] = lds[threadIdx.x];
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
__global__ void __launch_bounds__(kBlock) k_pair_ds(const double* __restrict__ pos,
                                                    const int64_t* __restrict__ pairs,
                                                    int64_t np, CellGrid cg,
                                                    double* __restrict__ ds,
                                                    double* __restrict__ dr) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const int64_t i = pairs[2 * p], j = pairs[2 * p + 1];
  double v[3];
  if (cg.obc) {
    for (int d = 0; d < 3; ++d) v[d] = pos[3 * j + d] - pos[3 * i + d];
  } else {
    double x[3];
    for (int d = 0; d < 3; ++d) x[d] = pos[3 * i + d] - pos[3 * j + d];
    double s[3];
    for (int d = 0; d < 3; ++d) {
      s[d] = x[0] * cg.a[d] + x[1] * cg.a[3 + d] + x[2] * cg.a[6 + d];
      s[d] -= floor(s[d] + 0.5);
    }
    for (int d = 0; d < 3; ++d) v[d] = s[0] * cg.h[d] + s[1] * cg.h[3 + d] + s[2] * cg.h[6 + d];
  }
  ds[p] = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (dr)
    for (int d = 0; d < 3; ++d) dr[3 * p + d] = v[d];
}

__global__ void __launch_bounds__(kBlock) k_pair_ds_bwd(const double* __restrict__ gds,
                                                        const int64_t* __restrict__ pairs,
                                                        int64_t np, const double* __restrict__ pos,
                                                        CellGrid cg, const double* __restrict__ dr,
                                                        const double* __restrict__ ds,
                                                        double* __restrict__ gpos,
                                                        double* __restrict__ gbox) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const double d = ds[p];
  const double g = gds[p];
  if (d <= 0.0 || g == 0.0) return;
  const int64_t i = pairs[2 * p], j = pairs[2 * p + 1];
  const double u[3] = {dr[3 * p] / d, dr[3 * p + 1] / d, dr[3 * p + 2] / d};
  // PBC: dr = r_i - r_j - n h; OBC: dr = r_j - r_i
  const double si = cg.obc ? -1.0 : 1.0;
  for (int k = 0; k < 3; ++k) {
    unsafeAtomicAdd(&gpos[3 * i + k], si * g * u[k]);
    unsafeAtomicAdd(&gpos[3 * j + k], -si * g * u[k]);
  }
  if (gbox && !cg.obc) {
    double x[3], nimg[3];
    for (int k = 0; k < 3; ++k) x[k] = pos[3 * i + k] - pos[3 * j + k];
    for (int k = 0; k < 3; ++k) {
      const double s = x[0] * cg.a[k] + x[1] * cg.a[3 + k] + x[2] * cg.a[6 + k];
      nimg[k] = floor(s + 0.5);
    }
    for (int c = 0; c < 3; ++c)
      if (nimg[c] != 0.0)
        for (int b = 0; b < 3; ++b) unsafeAtomicAdd(&gbox[c * 3 + b], -g * nimg[c] * u[b]);
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
    const double b = bs[f * np + p];
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
// host launchers

void nb_layout(int64_t n, const CellGrid& cg, NbLayout* L) {
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
                        int64_t* pairs, int64_t capacity, hipStream_t s) {
  const unsigned nb = (unsigned)((n + kBlock - 1) / kBlock);
  if (n == 0) return hipGetLastError();
  const double* w = (const double*)(ws + L.w);
  const int* cell_of = (const int*)(ws + L.cell_of);
  const int* start = (const int*)(ws + L.start);
  const int* sorted = (const int*)(ws + L.sorted);
  int64_t* pc = (int64_t*)(ws + L.pcount);
  int64_t* off = (int64_t*)(ws + L.offsets);
  if (mode == 0) {
    k_nb_pairs<0><<<nb, kBlock, 0, s>>>(n, cg, w, cell_of, start, sorted, pc, nullptr, nullptr,
                                        0);
    launch_scan<int64_t>(pc, n, off, (int64_t*)(ws + L.scan), s);
  } else {
    k_nb_pairs<1><<<nb, kBlock, 0, s>>>(n, cg, w, cell_of, start, sorted, nullptr, off, pairs,
                                        capacity);
  }
  return hipGetLastError();
}

hipError_t launch_pair_ds(const double* pos, const int64_t* pairs, int64_t np, const CellGrid& cg,
                          double* ds, double* dr, hipStream_t s) {
  if (np > 0)
    k_pair_ds<<<(unsigned)((np + kBlock - 1) / kBlock), kBlock, 0, s>>>(pos, pairs, np, cg, ds, dr);
  return hipGetLastError();
}

hipError_t launch_pair_ds_bwd(const double* gds, const int64_t* pairs, int64_t np,
                              const double* pos, const CellGrid& cg, const double* dr,
                              const double* ds, double* gpos, double* gbox, hipStream_t s) {
  if (np > 0)
    k_pair_ds_bwd<<<(unsigned)((np + kBlock - 1) / kBlock), kBlock, 0, s>>>(gds, pairs, np, pos, cg,
                                                                          dr, ds, gpos, gbox);
  return hipGetLastError();
}

hipError_t launch_pair_energy(int mode, int64_t nf, int64_t n, int64_t np, const double* q,
                              const double* eta, const int64_t* pairs, const double* ds,
                              const double* bs, double kappa, double eps, double* energy,
                              double* gds, double* gq, double* geta, double* partials,
                              hipStream_t s) {
  const int bpf = kPairBlocksPerFrame;
  const int64_t m = nf * n;
  if (m > 0) {
    const unsigned nb = (unsigned)std::min<int64_t>((m + kBlock - 1) / kBlock, 4096);
    if (gq) k_fill<double><<<nb, kBlock, 0, s>>>(gq, m, 0.0);
    if (geta) k_fill<double><<<nb, kBlock, 0, s>>>(geta, m, 0.0);
  }
  dim3 grid(bpf, (unsigned)nf);
  k_pair_energy<<<grid, kBlock, 0, s>>>(mode, n, np, q, eta, pairs, ds, bs, kappa, eps, partials,
                                        gds, gq, geta);
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
