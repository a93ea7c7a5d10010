// Batched PME energy + forces + dE/dbox for many independent periodic
// systems of different atom counts, boxes and meshes in one launch sequence.
//
// Every kernel reads its system's parameters from a device table (SysDesc)
// and finds the system of a global index (atom, cell, mesh point, FFT block,
// gather block) by binary search over prefix offsets, so a batch of any mix
// costs the same 14 launches as a single system; admp_pme_fused is a batch of
// one whose table is written by the first kernel from its arguments (no copy,
// so the sequence can be captured in a HIP graph).  Reductions are per system
// and in a fixed order.
//
// Per system the result equals CoulombForceModule.forward on the
// TorchNeighborList pairs followed by autograd (src/torch_admp/pme.py:261-310,
// src/torch_admp/utils.py:84-112) when every box height is at least 2 rcut;
// below that every periodic image inside the cutoff is summed.
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "batch.hpp"
#include "fft_device.hpp"
#include "mesh_device.hpp"
#include "real_device.hpp"

namespace admp {

__device__ __forceinline__ int find_seg(const int64_t* __restrict__ off, int nseg, int64_t x) {
  int lo = 0, hi = nseg;  // largest s with off[s] <= x
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= x) lo = mid;
    else hi = mid;
  }
  return lo;
}

// ------------------------------------------------------------------ prologue
// Zero `bytes` (a multiple of 4) at p (16-byte aligned) with 16-byte stores.
__device__ __forceinline__ void zero_words(void* p, int64_t bytes, int64_t t0, int64_t step) {
  int4* p4 = reinterpret_cast<int4*>(p);
  const int64_t n4 = bytes >> 4;
  for (int64_t i = t0; i < n4; i += step) p4[i] = make_int4(0, 0, 0, 0);
  int* p1 = reinterpret_cast<int*>(p);
  for (int64_t i = (n4 << 2) + t0; i < (bytes >> 2); i += step) p1[i] = 0;
}

// Zero the bin counters and scatter cursors, the meshes, the real-space
// gradient accumulators and the per-system reduction tickets; for a batch of one also write the system table and the
// segment offsets from the kernel arguments.
struct OneSystem {
  SysDesc d;
  int64_t tot[kSegCount];  // segment sizes (offsets are [0, tot])
};

template <typename T>
__global__ void __launch_bounds__(kBlock) k_b_prologue(int inline_table, OneSystem one,
                                                       SysDesc* __restrict__ tab,
                                                       int64_t* __restrict__ offs,
                                                       int* __restrict__ count, int64_t ncount,
                                                       T* __restrict__ mesh, int64_t nmesh,
                                                       double* __restrict__ grad, int64_t ngrad,
                                                       int* __restrict__ tickets, int ntickets) {
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  if (inline_table) {
    // the table word by word from the kernel arguments
    static_assert(sizeof(SysDesc) % 4 == 0, "SysDesc is a whole number of words");
    const int* src = reinterpret_cast<const int*>(&one.d);
    int* dst = reinterpret_cast<int*>(tab);
    for (int64_t i = t0; i < (int64_t)(sizeof(SysDesc) / 4); i += step) dst[i] = src[i];
    if (t0 < kSegCount) {
      offs[2 * t0] = 0;
      offs[2 * t0 + 1] = one.tot[t0];
    }
  }
  zero_words(count, (int64_t)sizeof(int) * ncount, t0, step);
  zero_words(mesh, (int64_t)sizeof(T) * nmesh, t0, step);
  zero_words(grad, (int64_t)sizeof(double) * ngrad, t0, step);
  zero_words(tickets, (int64_t)sizeof(int) * ntickets, t0, step);
}

// ------------------------------------------------------------------ binning
__global__ void __launch_bounds__(kBlock) k_b_bin(int64_t ntot, int nsys,
                                                  const int64_t* __restrict__ atom_off,
                                                  const SysDesc* __restrict__ tab,
                                                  const double* __restrict__ pos,
                                                  double* __restrict__ w, int* __restrict__ cell_of,
                                                  int* __restrict__ sys_of,
                                                  int* __restrict__ count,
                                                  int* __restrict__ rank) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ntot) return;
  const int s = find_seg(atom_off, nsys, i);
  const CellGrid& cg = tab[s].cg;
  const double r[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
  double sf[3];
  int c[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double v = r[0] * cg.a[d] + r[1] * cg.a[3 + d] + r[2] * cg.a[6 + d];
    v = v + 1e-7;
    v = v - floor(v);
    sf[d] = v - 1e-7;
    const int nc = cg.nc[d];
    int cc = (int)floor(sf[d] * nc);
    c[d] = cc < 0 ? 0 : (cc >= nc ? nc - 1 : cc);
  }
#pragma unroll
  for (int j = 0; j < 3; ++j)
    w[3 * i + j] = sf[0] * cg.h[j] + sf[1] * cg.h[3 + j] + sf[2] * cg.h[6 + j];
  const int cid = (int)tab[s].cell0 + (c[0] * cg.nc[1] + c[1]) * cg.nc[2] + c[2];
  cell_of[i] = cid;
  sys_of[i] = s;
  // the atom's place among its counter's atoms (the cell-order pass adds the
  // counter's start)
  rank[i] = atomicAdd(&count[cid * kSub + (int)(i & (kSub - 1))], 1);
}

// Cell order and packing, one thread per atom: slot = start[key] + the
// atom's rank from the binning's counter atomic (key = cell * kSub + i mod
// kSub), then sorted[slot] = i,
// ps[slot] = (wrapped position, charge) and, for systems spread by bricks,
// mref[slot] = the mesh reference point m0 of the wrapped position (x | y << 16, z).  The order within a cell is that of
// the atomics; every consumer (centre-cell pair test, spread, gather) works
// with any order.  With start == nullptr each work-group forms the exclusive
// scan of the ncount counters in LDS itself (ncount <= kScatterScanMax) and
// work-group 0 stores it as start; otherwise start was scanned beforehand.
constexpr int kScatterScanMax = 8192;

__global__ void __launch_bounds__(kBlock) k_b_scatter_pack(int64_t ntot, int ncount,
                                                           const int* __restrict__ count,
                                                           int* __restrict__ start_out,
                                                           const int* __restrict__ start,
                                                           const int* __restrict__ rank,
                                                           const int* __restrict__ cell_of,
                                                           const double* __restrict__ w,
                                                           const double* __restrict__ q,
                                                           int* __restrict__ sorted,
                                                           P4<double>* __restrict__ ps,
                                                           const int* __restrict__ sys_of,
                                                           const SysDesc* __restrict__ tab,
                                                           int2* __restrict__ mref) {
  __shared__ int st[kScatterScanMax + 1];
  __shared__ int part[kBlock];
  if (!start) {
    // thread t sums counts [t * per, (t + 1) * per); scan of the sums; write back
    const int per = (ncount + kBlock - 1) / kBlock;
    const int base = threadIdx.x * per;
    int sum = 0;
    for (int k = 0; k < per; ++k)
      if (base + k < ncount) sum += count[base + k];
    // inclusive scan of the per-thread sums: DPP within each wavefront, then
    // the wavefront totals
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int inc = sum;
    inc += dpp_move<0x111, 0xf>(inc);  // row_shr:1
    inc += dpp_move<0x112, 0xf>(inc);  // row_shr:2
    inc += dpp_move<0x114, 0xf>(inc);  // row_shr:4
    inc += dpp_move<0x118, 0xf>(inc);  // row_shr:8
    inc += dpp_move<0x142, 0xa>(inc);  // row_bcast:15 into rows 1 and 3
    inc += dpp_move<0x143, 0xc>(inc);  // row_bcast:31 into rows 2 and 3
    if (lane == 63) part[wid] = inc;
    __syncthreads();
    for (int w = 0; w < wid; ++w) inc += part[w];
    if (threadIdx.x == kBlock - 1) part[kBlock / 64] = inc;
    int run = inc - sum;
    for (int k = 0; k < per; ++k)
      if (base + k < ncount) {
        st[base + k] = run;
        run += count[base + k];
      }
    __syncthreads();
    if (threadIdx.x == 0) st[ncount] = part[kBlock / 64];
    __syncthreads();
    if (blockIdx.x == 0)
      for (int k = threadIdx.x; k <= ncount; k += blockDim.x) start_out[k] = st[k];
  }
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ntot) return;
  const int key = cell_of[i] * kSub + (int)(i & (kSub - 1));
  const int slot = (start ? start[key] : st[key]) + rank[i];
  sorted[slot] = (int)i;
  P4<double> v;
  v.x = w[3 * i];
  v.y = w[3 * i + 1];
  v.z = w[3 * i + 2];
  v.q = q[i];
  ps[slot] = v;
  // the mesh reference point of the wrapped position, for the brick spread's scan
  const SysDesc* D = tab + sys_of[i];
  if (D->brick) {
    int m0[3];
    double f[3];
    mesh_ref(D->g, v.x, v.y, v.z, m0, f);
    mref[slot] = make_int2(m0[0] | (m0[1] << 16), m0[2]);
  }
}

// ------------------------------------------------------------------ real space
// One work-group per (global cell, candidate split).  Partials are stored
// value-major, rows cell-major (row = cell * splits + split), so that a
// system's rows are contiguous and the reduction reads them coalesced.
// The register cap keeps room on each SIMD for the r2c and FFT wavefronts
// that run beside this kernel on the other stream: at 117 registers r2c ran
// 28 us beside it instead of 14.
template <typename T>
__global__ void __launch_bounds__(kRealBlock) __attribute__((amdgpu_num_vgpr(104))) k_b_real(int nsys,
                                                   const int64_t* __restrict__ real_off,
                                                   const SysDesc* __restrict__ tab,
                                                   const P4<double>* __restrict__ ps,
                                                   const int* __restrict__ sorted,
                                                   const int* __restrict__ start,
                                                   double* __restrict__ grad,
                                                   double* __restrict__ partials,
                                                   int64_t rows) {
  // work-group = (cell, split) of a system with its own split count; the
  // items of a system are numbered cell-major from real_off[sid]
  const int64_t item = blockIdx.x;
  const int sid = find_seg(real_off, nsys, item);
  const SysDesc* D = tab + sid;
  const int S = D->nsplit;
  const int64_t local = item - real_off[sid];
  real_cell_body<T>(D->cg, D->cell0, (int)(local / S), (int)(local % S), S, D->kappa, (T)D->eps,
                    ps, sorted, start, grad, partials + item, rows);
}

// ------------------------------------------------------------------ spread
// systems whose tiles fit LDS: one work-group per cell (global cell index),
// kSpreadThreads threads (64 atoms a pass: at ~70 atoms a cell, 512 measured
// faster than 256 and 768)
constexpr int kSpreadThreads = 512;

template <typename T>
__global__ void __launch_bounds__(kSpreadThreads) k_b_spread_tiles(int nsys,
                                                           const int64_t* __restrict__ cell_off,
                                                           const SysDesc* __restrict__ tab,
                                                           const double* __restrict__ pos,
                                                           const double* __restrict__ qv,
                                                           const int* __restrict__ sorted,
                                                           const int* __restrict__ start,
                                                           int4* __restrict__ stm,
                                                           T* __restrict__ stw,
                                                           T* __restrict__ mesh) {
  const int64_t c = blockIdx.x;
  const SysDesc* D = tab + find_seg(cell_off, nsys, c);
  if (!D->tg.ok || D->brick) return;
  const int lc = (int)(c - D->cell0);
  spread_cell_tile<T>(D->g, D->tg, lc, start[c * kSub], start[(c + 1) * kSub], sorted, pos, qv,
                      mesh + D->mesh0, stm, stw);
}

// systems with every mesh dimension >= kBrickMinK: one work-group per mesh
// brick (global brick index, XCD-local), see spread_brick; ZW = the system
// table's brick value (the launch covers only the systems of that kind)
template <typename T, int ZW>
__global__ void __launch_bounds__(kBrickThreads) k_b_spread_bricks(int nsys,
                                                        const int64_t* __restrict__ brick_off,
                                                        const SysDesc* __restrict__ tab,
                                                        const P4<double>* __restrict__ ps,
                                                        const int2* __restrict__ mref,
                                                        const int* __restrict__ sorted,
                                                        const int* __restrict__ start,
                                                        const double* __restrict__ qv,
                                                        int4* __restrict__ stm,
                                                        T* __restrict__ stw,
                                                        T* __restrict__ mesh) {
  __shared__ BrickStage<T, ZW> S;
  const int64_t b = xcd_block(blockIdx.x, gridDim.x);
  const int sid = find_seg(brick_off, nsys, b);
  const SysDesc* D = tab + sid;
  if (D->brick != ZW) return;
  const Geom g = D->g;
  const int nc[3] = {D->cg.nc[0], D->cg.nc[1], D->cg.nc[2]};
  const int nbr[3] = {D->nbr[0], D->nbr[1], D->nbr[2]};
  spread_brick<T, ZW>(g, nc, D->cell0, nbr, b - brick_off[sid], ps, mref, sorted, start, qv,
                      stm, stw, mesh + D->mesh0, S);
}

// atoms of systems whose tiles do not fit: global atomics into their mesh
template <typename T>
__global__ void __launch_bounds__(kBlock) k_b_spread_atoms(int64_t ntot,
                                                           const int* __restrict__ sys_of,
                                                           const SysDesc* __restrict__ tab,
                                                           const double* __restrict__ pos,
                                                           const double* __restrict__ qv,
                                                           int4* __restrict__ stm,
                                                           T* __restrict__ stw,
                                                           T* __restrict__ mesh) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = t >> 3;
  const int a = (int)(t & 7);
  if (i >= ntot || a == 7) return;
  const SysDesc* D = tab + sys_of[i];
  if (D->tg.ok || D->brick) return;
  if (a == 6)  // the stencil record for the gather
    stencil_record<T>(D->g, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], stm + i,
                      stw + (int64_t)kStencilVals * i);
  else
    spread_lane<T>(D->g, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], qv[i], a, mesh + D->mesh0);
}

// ------------------------------------------------------------------ FFT passes
template <typename T>
__global__ void __launch_bounds__(kBlock) k_b_r2c(int nsys, const int64_t* __restrict__ off,
                                                  const SysDesc* __restrict__ tab,
                                                  const T* __restrict__ meshes,
                                                  cplx<T>* __restrict__ specs) {
  const int sid = find_seg(off, nsys, blockIdx.x);
  const SysDesc* D = tab + sid;
  const int C = D->fchunk[0];
  const int L = D->g.k[2];
  const T* mesh = meshes + D->mesh0;
  const int64_t nrows = (int64_t)D->g.k[0] * D->g.k[1];
  r2c_rows<T>([=](int64_t row, int z) { return mesh[row * L + z]; }, specs + D->spec0, nrows,
              (blockIdx.x - off[sid]) * 2 * C, L, C, D->rad[2], D->plane != 0, D->kzp);
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_b_c2r(int nsys, const int64_t* __restrict__ off,
                                                  const SysDesc* __restrict__ tab,
                                                  const cplx<T>* __restrict__ specs,
                                                  T* __restrict__ meshes) {
  const int sid = find_seg(off, nsys, blockIdx.x);
  const SysDesc* D = tab + sid;
  const int C = D->fchunk[0];
  c2r_rows<T>(specs + D->spec0, meshes + D->mesh0, (int64_t)D->g.k[0] * D->g.k[1],
              (blockIdx.x - off[sid]) * 2 * C, D->g.k[2], C, D->rad[2], D->plane != 0, D->kzp);
}

// lines along y: element (x, y, kz) at (x * K1 + y) * kzp + kz
template <typename T, int SIGN>
__global__ void __launch_bounds__(kBlock) k_b_cols_y(int nsys, const int64_t* __restrict__ off,
                                                     const SysDesc* __restrict__ tab,
                                                     cplx<T>* __restrict__ specs) {
  const int sid = find_seg(off, nsys, blockIdx.x);
  const SysDesc* D = tab + sid;
  const int K1 = D->g.k[1], Kzh = D->g.k[2] / 2 + 1, C = D->fchunk[1];
  const int nchunk = (Kzh + C - 1) / C;
  const int64_t lb = blockIdx.x - off[sid];
  const int x = (int)(lb / nchunk), chunk = (int)(lb % nchunk);
  strided_lines<T, SIGN>(specs + D->spec0 + (int64_t)x * K1 * D->kzp + (int64_t)chunk * C, K1,
                         D->kzp,
                         min(C, Kzh - chunk * C), C, D->rad[1]);
}

// lines along x with the influence function; per-block partials (10 values
// times eps, value-major with nrows rows)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_b_cols_x(int nsys, const int64_t* __restrict__ off,
                                                     const int64_t* __restrict__ xoff,
                                                     const SysDesc* __restrict__ tab,
                                                     cplx<T>* __restrict__ specs,
                                                     double* __restrict__ partials, int64_t nrows) {
  const int sid = find_seg(off, nsys, blockIdx.x);
  const SysDesc* D = tab + sid;
  if (D->plane) return;
  const int64_t row = xoff[sid] + (blockIdx.x - off[sid]);  // the system's partial row
  InflParams P;
  P.g = D->g;
  P.kappa = D->kappa;
  P.phi_scale = 1.0;
  P.eps = D->eps;
  P.conv = D->conv;
  P.ortho = D->ortho;
  const int Kzh = D->g.k[2] / 2 + 1, C = D->fchunk[2];
  const int nchunk = (Kzh + C - 1) / C;
  const int64_t lb = blockIdx.x - off[sid];
  __shared__ double vals[10];
  x_convolve_lines<T>(specs + D->spec0, (int)(lb / nchunk), (int)(lb % nchunk) * C, C, D->rad[0],
                      P, vals, D->kzp);
  __syncthreads();
  if (threadIdx.x < 10) partials[(int64_t)threadIdx.x * nrows + row] = vals[threadIdx.x] * P.eps;
}

// one kz plane per work-group (systems with plane set; their partial rows are
// the system's x-pass rows xoff[sid] + kz)
template <typename T>
__global__ void __launch_bounds__(kPlaneThreads) k_b_plane(int nsys,
                                                         const int64_t* __restrict__ off,
                                                         const int64_t* __restrict__ xoff,
                                                         const SysDesc* __restrict__ tab,
                                                         cplx<T>* __restrict__ specs,
                                                         double* __restrict__ partials,
                                                         int64_t nrows) {
  const int sid = find_seg(off, nsys, blockIdx.x);
  const SysDesc* D = tab + sid;
  if (!D->plane) return;
  const int64_t row = xoff[sid] + (blockIdx.x - off[sid]);
  InflParams P;
  P.g = D->g;
  P.kappa = D->kappa;
  P.phi_scale = 1.0;
  P.eps = D->eps;
  P.conv = D->conv;
  P.ortho = D->ortho;
  __shared__ double vals[10];
  plane_convolve<T>(specs + D->spec0, (int)(blockIdx.x - off[sid]), D->rad[0], D->rad[1], P,
                    vals);
  __syncthreads();
  if (threadIdx.x < 10) partials[(int64_t)threadIdx.x * nrows + row] = vals[threadIdx.x] * P.eps;
}

// ------------------------------------------------------------------ gather
// kGatherAtomsPerBlock atoms of one system per work-group, 8 lanes per atom.
// Forces are final here: -(real-space gradient rows + reciprocal gradient).  The block's dE/dA partials (9 values) are stored value-major.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_b_gather(int nsys, const int64_t* __restrict__ off,
                                                     const SysDesc* __restrict__ tab,
                                                     const double* __restrict__ pos,
                                                     const double* __restrict__ qv,
                                                     const T* __restrict__ meshes,
                                                     const double* __restrict__ grad,
                                                     const int4* __restrict__ stm,
                                                     const T* __restrict__ stw,
                                                     double* __restrict__ forces,
                                                     double* __restrict__ partials, int64_t nrows) {
  __shared__ double red[kGatherNV * (kBlock / 64)];
  const int sid = find_seg(off, nsys, blockIdx.x);
  const SysDesc* D = tab + sid;
  const Geom& g = D->g;
  const int a = threadIdx.x & 7;
  const int64_t atom =
      D->atom0 + (blockIdx.x - off[sid]) * kGatherAtomsPerBlock + (threadIdx.x >> 3);
  const bool valid = atom < D->atom0 + D->n;
  double acc[kGatherNV];
#pragma unroll
  for (int i = 0; i < kGatherNV; ++i) acc[i] = 0.0;
  T sx = 0, sy = 0, sz = 0, sp = 0;
  double x = 0.0, y = 0.0, z = 0.0, qi = 0.0, g_own = 0.0;
  if (valid) {
    // loads that do not depend on the mesh sums, issued first
    qi = qv[atom];
    if (a < 3) g_own = grad[3 * atom + a];
    x = pos[3 * atom];
    y = pos[3 * atom + 1];
    z = pos[3 * atom + 2];
    gather_lane_rec<T>(g.k[0], g.k[1], g.k[2], stm[atom], stw + (int64_t)kStencilVals * atom, a,
                       meshes + D->mesh0, &sx, &sy, &sz, &sp);
  }
  sx = group8_sum(sx);
  sy = group8_sum(sy);
  sz = group8_sum(sz);
  // the atom's real-space gradient row in lane 0 (lanes 0-2 hold one
  // component each; quad_perm broadcasts within the quad)
  const double gr0 = dpp_move<0x00, 0xf>(g_own), gr1 = dpp_move<0x55, 0xf>(g_own),
               gr2 = dpp_move<0xaa, 0xf>(g_own);
  if (valid && a < 3) {
    // lane a (< 3) owns force component a
    const double dR[3] = {-qi * (double)sx, -qi * (double)sy, -qi * (double)sz};
    double gj = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) gj += g.k[d] * g.a[a * 3 + d] * dR[d];
    gj += g_own;
    forces[3 * atom + a] = -gj;
    if (a == 0) {
      // dE/dA from the reciprocal sum, r (x) (k dR), and from the real-space
      // gradient row g, r (x) (h g): the finishing step applies A = h^-1 on
      // both sides, so h g enters as -A (r (x) g) does in the box derivative
      const double r[3] = {x, y, z};
      const double gv[3] = {gr0, gr1, gr2};
      double u[3];
#pragma unroll
      for (int d = 0; d < 3; ++d)
        u[d] = dR[d] * g.k[d] + (g.h[d * 3] * gv[0] + g.h[d * 3 + 1] * gv[1] + g.h[d * 3 + 2] * gv[2]);
#pragma unroll
      for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int d = 0; d < 3; ++d) acc[j * 3 + d] = r[j] * u[d];
    }
  }
  block_sum<kGatherNV>(acc, red);  // the block total is in thread 0
  if (threadIdx.x == 0)
    for (int i = 0; i < kGatherNV; ++i) partials[(int64_t)i * nrows + blockIdx.x] = acc[i];
}

// ------------------------------------------------------------------ reductions
// Work-group (s, v) reduces value v of system s over its rows (a fixed
// order: thread t sums rows t, t + 256, ..., then a block tree) into
// red[s][kSysValues]; the last work-group of a system to finish (a ticket
// counter) turns the values into energies and dE/dbox.
// Each system's ticket counter on a 128-byte line of its own.
constexpr int kTicketStride = 32;

struct PartialSets {
  const double* real;  // kRealNV values x rows_real (rows = (cell, split) items)
  const double* recip; // 10 values x rows_recip (x-pass blocks)
  const double* gath;  // kGatherNV values x rows_gath (gather blocks)
  int64_t rows_real, rows_recip, rows_gath;
  int nsplit;
};

__global__ void __launch_bounds__(kBlock) k_b_reduce_finalize(
    int nsys, const int64_t* __restrict__ offs, const SysDesc* __restrict__ tab, PartialSets ps,
    double* __restrict__ red, int* __restrict__ tickets, double* __restrict__ energy,
    double* __restrict__ dEdbox) {
  __shared__ double wsum[kBlock / 64];
  const int s = blockIdx.x;
  const int v = blockIdx.y;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  {
    const double* p;
    int64_t r0, r1;
    const int64_t* cell = offs + (int64_t)(nsys + 1) * kSegReal;  // real-space rows
    const int64_t* xo = offs + (int64_t)(nsys + 1) * kSegFftX;
    const int64_t* go = offs + (int64_t)(nsys + 1) * kSegGather;
    if (v < kRealNV) {
      p = ps.real + (int64_t)v * ps.rows_real;
      r0 = cell[s];
      r1 = cell[s + 1];
    } else if (v < kRealNV + 10) {
      p = ps.recip + (int64_t)(v - kRealNV) * ps.rows_recip;
      r0 = xo[s];
      r1 = xo[s + 1];
    } else {
      p = ps.gath + (int64_t)(v - kRealNV - 10) * ps.rows_gath;
      r0 = go[s];
      r1 = go[s + 1];
    }
    const int nt = blockDim.x;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int64_t r = r0 + threadIdx.x;
    for (; r + 3 * nt < r1; r += 4 * nt) {
      a0 += p[r];
      a1 += p[r + nt];
      a2 += p[r + 2 * nt];
      a3 += p[r + 3 * nt];
    }
    for (; r < r1; r += nt) a0 += p[r];
    const double acc = wave_sum((a0 + a1) + (a2 + a3));
    if (lane == 0) wsum[wid] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += wsum[w];
      red[(int64_t)s * kSysValues + v] = t;
    }
  }
  __threadfence();
  __syncthreads();
  __shared__ int last;
  if (threadIdx.x == 0) last = atomicAdd(&tickets[s * kTicketStride], 1) == (int)gridDim.y - 1;
  __syncthreads();
  if (last && threadIdx.x == 0) {
    __threadfence();
    const SysDesc* D = tab + s;
    const double* rr = red + (int64_t)s * kSysValues;
    double vals[kSysValues];
    for (int k = 0; k < kSysValues; ++k)
      vals[k] = __hip_atomic_load(rr + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // real-space values in finalize_system's order: energy and pair virial,
    // sum r (x) dE/dr (zero here: the gather folded it into its dE/dA
    // terms), sum q^2, sum q
    double real[18];
    for (int k = 0; k < 7; ++k) real[k] = vals[k];
    for (int k = 7; k < 16; ++k) real[k] = 0.0;
    real[16] = vals[7];
    real[17] = vals[8];
    finalize_system(real, vals + kRealNV, vals + kRealNV + 10, D->g, D->kappa, D->eps,
                    energy + 5 * s, dEdbox ? dEdbox + 9 * s : nullptr);
  }
}

// The same reductions and finalisation with one work-group per system (no
// tickets, no fences): wavefront w sums values w, w + 4, ... over their rows,
// then thread 0 finalises.  For batches of many systems, where the per-value
// work-groups' ticket atomics and fences cost ~72 ns each, serialised (a batch
// of 256 systems: 6656 work-groups, 0.48 ms).
__global__ void __launch_bounds__(kBlock) k_b_reduce_finalize_sys(
    int nsys, const int64_t* __restrict__ offs, const SysDesc* __restrict__ tab, PartialSets ps,
    double* __restrict__ energy, double* __restrict__ dEdbox) {
  __shared__ double vals[kSysValues];
  const int s = blockIdx.x;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int64_t* cell = offs + (int64_t)(nsys + 1) * kSegReal;
  const int64_t* xo = offs + (int64_t)(nsys + 1) * kSegFftX;
  const int64_t* go = offs + (int64_t)(nsys + 1) * kSegGather;
  for (int v = wid; v < kSysValues; v += kBlock / 64) {
    const double* p;
    int64_t r0, r1;
    if (v < kRealNV) {
      p = ps.real + (int64_t)v * ps.rows_real;
      r0 = cell[s];
      r1 = cell[s + 1];
    } else if (v < kRealNV + 10) {
      p = ps.recip + (int64_t)(v - kRealNV) * ps.rows_recip;
      r0 = xo[s];
      r1 = xo[s + 1];
    } else {
      p = ps.gath + (int64_t)(v - kRealNV - 10) * ps.rows_gath;
      r0 = go[s];
      r1 = go[s + 1];
    }
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int64_t r = r0 + lane;
    for (; r + 3 * 64 < r1; r += 4 * 64) {
      a0 += p[r];
      a1 += p[r + 64];
      a2 += p[r + 128];
      a3 += p[r + 192];
    }
    for (; r < r1; r += 64) a0 += p[r];
    const double t = wave_sum((a0 + a1) + (a2 + a3));
    if (lane == 0) vals[v] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const SysDesc* D = tab + s;
    double real[18];
    for (int k = 0; k < 7; ++k) real[k] = vals[k];
    for (int k = 7; k < 16; ++k) real[k] = 0.0;
    real[16] = vals[7];
    real[17] = vals[8];
    finalize_system(real, vals + kRealNV, vals + kRealNV + 10, D->g, D->kappa, D->eps,
                    energy + 5 * s, dEdbox ? dEdbox + 9 * s : nullptr);
  }
}

// Batches of at least this many systems finalise with one work-group per
// system; smaller ones (a single system: 26 work-groups in parallel) keep the
// per-value work-groups.
constexpr int kFinalizePerSystem = 16;

// ------------------------------------------------------------------ host
// Lines per FFT work-group: enough for about one point per thread, within the
// LDS budget; a function of the system alone, so that a system's result does
// not depend on the rest of the batch.
static int pick_lines(int L, size_t elem, int target) {
  int C = 1;
  while (C < 16 && L * C < target) C <<= 1;
  while (C > 1 && (size_t)(2 * L * C + L) * elem > 65536) C >>= 1;
  return C;
}

static size_t al(size_t b) { return (b + 255) & ~size_t(255); }

constexpr int64_t kWideLinesElems = int64_t(1) << 22;

// Candidate splits per home cell: enough work-groups to fill the chip for
// small batches, one per cell once the batch has that many cells.
static int real_splits(int64_t ctot) {
  const int64_t c = std::max<int64_t>(ctot, 1);
  const int64_t s = (kRealTargetBlocks + c - 1) / c;
  return (int)std::min<int64_t>(std::max<int64_t>(s, 1), kRealSplit);
}

// The y and x passes of a K0 x K1 plane fit one work-group: the plane within
// 64 KiB of LDS and kPlaneLoad elements per thread, radices the in-place
// passes have, and every pass within kPlaneB butterflies per thread.
static bool plane_fits(int K0, int K1, const Radix& rx, const Radix& ry, size_t lds) {
  if (lds > 65536 || !radix_ok(rx) || !radix_ok(ry)) return false;
  if ((int64_t)K0 * K1 > (int64_t)kPlaneLoad * kPlaneThreads || 2 * (K0 + K1) > kPlaneThreads)
    return false;
  auto ok = [](const Radix& rd, int L, int C) {
    for (int k = 0; k < rd.n; ++k) {
      const int r = rd.r[k];
      if (r != 2 && r != 3 && r != 4 && r != 5 && r != 8) return false;
      if ((int64_t)(L / r) * C > (int64_t)kPlaneB * kPlaneThreads) return false;
    }
    return true;
  };
  return ok(rx, K0, K1) && ok(ry, K1, K0);
}

BatchPlan plan_batch(int nsys, const int64_t* n_atoms, const double* boxes, const double* cutoffs,
                     const double* kappas, const int32_t* kmesh, int precision, int convention,
                     double eps) {
  BatchPlan P;
  P.nsys = nsys;
  P.precision = precision;
  P.status = ADMP_OK;
  P.sys.resize(nsys);
  for (int k = 0; k < kSegCount; ++k) P.off[k].assign(nsys + 1, 0);
  const size_t es = precision == ADMP_F32 ? 4 : 8;
  const size_t ces = 2 * es;
  P.lds_fft[0] = P.lds_fft[1] = P.lds_fft[2] = 0;
  P.lds_tile = 0;
  P.lds_plane = 0;
  P.any_plane = P.any_cols = 0;
  P.any_untiled = 0;
  P.any_brick = P.any_nonbrick = 0;
  // Elements per work-group of the line passes: kBlock for small batches
  // (a single 64^3 system needs the work-groups to hide latency), twice that
  // once the half spectra reach kWideLinesElems elements, where more lines per
  // work-group give each Stockham pass more butterflies between barriers
  // (256 mixed serving requests: 14.8 -> 12.2 ms float64, 11.9 -> 8.8 ms
  // float32 per batch).
  int64_t spec_elems = 0;
  for (int s = 0; s < nsys; ++s) {
    const int32_t* km = kmesh + 3 * s;
    spec_elems += (int64_t)km[0] * km[1] * (km[2] / 2 + 1);
  }
  const int line_target = spec_elems >= kWideLinesElems ? 2 * kBlock : kBlock;
  int64_t spec = 0;
  for (int s = 0; s < nsys; ++s) {
    SysDesc& D = P.sys[s];
    std::memset(&D, 0, sizeof(D));
    const double* box = boxes + 9 * s;
    const int32_t* km = kmesh + 3 * s;
    D.cg = make_grid(box, cutoffs[s], false);
    for (int i = 0; i < 9; ++i) {
      D.g.h[i] = D.cg.h[i];
      D.g.a[i] = D.cg.a[i];
    }
    D.g.vol = box[0] * (box[4] * box[8] - box[5] * box[7]) - box[1] * (box[3] * box[8] - box[5] * box[6]) +
              box[2] * (box[3] * box[7] - box[4] * box[6]);
    for (int d = 0; d < 3; ++d) {
      D.g.k[d] = km[d];
      D.rad[d] = factorize(km[d]);
    }
    D.tg = make_tile_geom(D.cg, km, es);
    D.brick = km[0] >= kBrickMinK && km[1] >= kBrickMinK && km[2] >= kBrickMinK &&
              km[0] < 32768 && km[1] < 32768 ? 1 : 0;
    // deep bricks where a brick of kBrickX kBrickY kBrickZ points would see
    // few atoms (>= kDeepBrickPts mesh points per atom) and a deep brick
    // plus a stencil fits along z
    const double pts = (double)km[0] * km[1] * km[2];
    if (D.brick && pts >= kDeepBrickPts * (double)std::max<int64_t>(n_atoms[s], 1) &&
        km[2] >= kBrickZ * kBrickWaves + 6)
      D.brick = kBrickWaves;
    const int zb = kBrickZ * D.brick;
    D.nbr[0] = D.brick ? (km[0] + kBrickX - 1) / kBrickX : 0;
    D.nbr[1] = D.brick ? (km[1] + kBrickY - 1) / kBrickY : 0;
    D.nbr[2] = D.brick ? (km[2] + zb - 1) / zb : 0;
    D.kappa = kappas ? kappas[s] : 0.0;
    D.eps = eps;
    D.conv = convention;
    D.ortho = (box[1] == 0.0 && box[2] == 0.0 && box[3] == 0.0 && box[5] == 0.0 &&
               box[6] == 0.0 && box[7] == 0.0) ? 1 : 0;
    const int K0 = km[0], K1 = km[1], K2 = km[2], Kzh = K2 / 2 + 1;
    D.fchunk[0] = pick_lines(K2, ces, line_target);
    D.fchunk[1] = pick_lines(K1, ces, line_target);
    D.fchunk[2] = pick_lines(K0, ces, line_target);
    D.atom0 = P.off[kSegAtom][s];
    D.n = n_atoms[s];
    D.cell0 = P.off[kSegCell][s];
    D.ncell = D.cg.ncell;
    D.mesh0 = P.off[kSegMesh][s];
    P.off[kSegAtom][s + 1] = D.atom0 + D.n;
    P.off[kSegCell][s + 1] = D.cell0 + D.ncell;
    P.off[kSegMesh][s + 1] = D.mesh0 + (int64_t)K0 * K1 * K2;
    P.off[kSegFftZ][s + 1] = P.off[kSegFftZ][s] + ((int64_t)K0 * K1 + 2 * D.fchunk[0] - 1) / (2 * D.fchunk[0]);
    const size_t lds_plane = plane_lds_bytes(K0, K1, ces);
    D.plane = plane_fits(K0, K1, D.rad[0], D.rad[1], lds_plane) ? 1 : 0;
    // line layout: rows padded to whole 128-byte lines, so that the line passes'
    // chunks of consecutive kz do not straddle cache lines; every system's
    // spectrum starts on a 128-byte boundary
    const int line_elems = (int)(128 / ces);
    D.spec0 = spec;
    D.kzp = D.plane ? Kzh : (Kzh + line_elems - 1) / line_elems * line_elems;
    spec += (int64_t)K0 * K1 * D.kzp;
    spec = (spec + line_elems - 1) / line_elems * line_elems;
    P.off[kSegFftY][s + 1] =
        P.off[kSegFftY][s] + (D.plane ? 0 : (int64_t)K0 * ((Kzh + D.fchunk[1] - 1) / D.fchunk[1]));
    const int64_t xrows = D.plane ? Kzh : (int64_t)K1 * ((Kzh + D.fchunk[2] - 1) / D.fchunk[2]);
    P.off[kSegFftX][s + 1] = P.off[kSegFftX][s] + xrows;
    P.off[kSegPlane][s + 1] = P.off[kSegPlane][s] + (D.plane ? xrows : 0);
    P.off[kSegColsX][s + 1] = P.off[kSegColsX][s] + (D.plane ? 0 : xrows);
    P.off[kSegGather][s + 1] = P.off[kSegGather][s] + (D.n + kGatherAtomsPerBlock - 1) / kGatherAtomsPerBlock;
    const int64_t nbrk = (int64_t)D.nbr[0] * D.nbr[1] * D.nbr[2];
    P.off[kSegBrick][s + 1] = P.off[kSegBrick][s] + (D.brick == 1 ? nbrk : 0);
    P.off[kSegBrickZ][s + 1] = P.off[kSegBrickZ][s] + (D.brick == kBrickWaves ? nbrk : 0);
    P.lds_fft[0] = std::max(P.lds_fft[0], (size_t)(K2 + 2 * K2 * D.fchunk[0]) * ces);
    if (D.plane) {
      P.lds_plane = std::max(P.lds_plane, lds_plane);
      P.any_plane = 1;
    } else {
      P.lds_fft[1] = std::max(P.lds_fft[1], (size_t)(K1 + 2 * K1 * D.fchunk[1]) * ces);
      P.lds_fft[2] = std::max(P.lds_fft[2], (size_t)(K0 + 2 * K0 * D.fchunk[2]) * ces);
      P.any_cols = 1;
    }
    if (D.brick) P.any_brick = 1;
    else {
      P.any_nonbrick = 1;
      if (D.tg.ok) P.lds_tile = std::max(P.lds_tile, (size_t)D.tg.ntile * sizeof(double));
      else P.any_untiled = 1;
    }
  }
  P.ntot = P.off[kSegAtom][nsys];
  P.ctot = P.off[kSegCell][nsys];
  P.mtot = P.off[kSegMesh][nsys];
  P.stot = spec;
  // real-space splits: the batch's fewest (enough work-groups for small
  // batches), more for systems whose cells hold many candidate tests
  // (home atoms x half-stencil candidates per cell at the mean occupancy)
  P.nsplit = real_splits(P.ctot);
  for (int s = 0; s < nsys; ++s) {
    SysDesc& D = P.sys[s];
    const CellGrid& cg = D.cg;
    int nst = 1;
    for (int d = 0; d < 3; ++d) nst *= cg.nc[d] >= 3 ? 3 : 2 * cg.nimg[d] + 1;
    const double per_cell = (double)D.n / (double)std::max<int64_t>(D.ncell, 1);
    const double tests = per_cell * per_cell * (double)((nst - 1) / 2 + 1);
    const int want = (int)std::min<double>(std::ceil(tests / kRealBudget), (double)kRealSplitMax);
    D.nsplit = std::max(P.nsplit, want);
    P.off[kSegReal][s + 1] = P.off[kSegReal][s] + D.ncell * D.nsplit;
  }
  std::memset(&P.cg_all, 0, sizeof(P.cg_all));
  P.cg_all.ncell = (int)std::min<int64_t>(P.ctot, INT32_MAX);
  nb_layout(P.ntot, P.cg_all, &P.nb);
  size_t off = 0;
  auto take = [&](size_t b) {
    const size_t o = off;
    off += al(b);
    return o;
  };
  const int64_t rows_x = P.off[kSegFftX][nsys], rows_g = P.off[kSegGather][nsys];
  P.o_tab = take(sizeof(SysDesc) * nsys);
  P.o_off = take(sizeof(int64_t) * (nsys + 1) * kSegCount);
  P.o_nb = take(P.nb.bytes);
  P.o_sys_of = take(sizeof(int) * std::max<int64_t>(P.ntot, 1));
  P.o_ps = take(sizeof(P4<double>) * std::max<int64_t>(P.ntot, 1));
  P.o_mref = take(sizeof(int2) * std::max<int64_t>(P.ntot, 1));
  P.o_stm = take(sizeof(int4) * std::max<int64_t>(P.ntot, 1));
  P.o_stw = take(es * kStencilVals * std::max<int64_t>(P.ntot, 1));
  P.o_grad = take(sizeof(double) * 3 * std::max<int64_t>(P.ntot, 1));
  P.o_mesh = take(es * std::max<int64_t>(P.mtot, 1));
  P.o_spec = take(ces * std::max<int64_t>(P.stot, 1));
  P.o_rpart = take(sizeof(double) * kRealNV * std::max<int64_t>(P.off[kSegReal][nsys], 1));
  P.o_cpart = take(sizeof(double) * 10 * std::max<int64_t>(rows_x, 1));
  P.o_gpart = take(sizeof(double) * kGatherNV * std::max<int64_t>(rows_g, 1));
  P.o_red = take(sizeof(double) * kSysValues * nsys);
  P.o_tickets = take(sizeof(int) * kTicketStride * nsys);
  P.bytes = off;
  return P;
}

static void free_host(void* p) { delete[] static_cast<char*>(p); }

// Per device: a side stream (default priority) on which the real-space kernel
// runs concurrently with the reciprocal-space chain, and a ring of fork/join
// event pairs, one pair per call (a wait takes the event's latest record at
// the time it is enqueued, so reusing a pair 64 calls later is safe).  Made at
// the first call on the device that is not under stream capture; until then,
// and if making it fails, calls run every kernel on the caller's stream.
namespace {
constexpr int kForkRing = 64;
struct SideStream {
  hipStream_t s = nullptr;
  hipEvent_t fork[kForkRing] = {}, join[kForkRing] = {};
  std::atomic<unsigned> next{0};
  int state = 0;  // 0 not made, 1 ready, -1 failed
};
SideStream g_side[64];
std::mutex g_side_mu;

SideStream* side_stream(hipStream_t caller) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
  SideStream& x = g_side[dev];
  std::lock_guard<std::mutex> g(g_side_mu);
  if (x.state == 0) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(caller, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
      return nullptr;
    x.state = -1;
    if (hipStreamCreateWithFlags(&x.s, hipStreamNonBlocking) != hipSuccess) return nullptr;
    // both streams are on this device: the events need no system-scope fence
    // (measured 2.5 us less per 24k-atom step); the caller's stream still ends
    // each call with a kernel whose completion releases at system scope
    const unsigned ef = hipEventDisableTiming | hipEventDisableSystemFence;
    for (int k = 0; k < kForkRing; ++k)
      if (hipEventCreateWithFlags(&x.fork[k], ef) != hipSuccess ||
          hipEventCreateWithFlags(&x.join[k], ef) != hipSuccess)
        return nullptr;
    x.state = 1;
  }
  return x.state == 1 ? &x : nullptr;
}
}  // namespace

// Copy the system table and segment offsets into the workspace.  With
// `sync` the copy completes before returning (the host buffer is freed at
// once); otherwise a host function frees it when the stream reaches it.
int upload_table(const BatchPlan& P, char* ws, hipStream_t s, bool sync) {
  const int nsys = P.nsys;
  const size_t tab_bytes = sizeof(SysDesc) * nsys;
  const size_t off_bytes = sizeof(int64_t) * (nsys + 1) * kSegCount;
  char* host = new char[tab_bytes + off_bytes];
  std::memcpy(host, P.sys.data(), tab_bytes);
  for (int k = 0; k < kSegCount; ++k)
    std::memcpy(host + tab_bytes + sizeof(int64_t) * (nsys + 1) * k, P.off[k].data(),
                sizeof(int64_t) * (nsys + 1));
  if (hipMemcpyAsync(ws + P.o_tab, host, tab_bytes, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(ws + P.o_off, host + tab_bytes, off_bytes, hipMemcpyHostToDevice, s) !=
          hipSuccess) {
    delete[] host;
    return ADMP_ERR_HIP;
  }
  if (sync) {
    const hipError_t e = hipStreamSynchronize(s);
    delete[] host;
    return e == hipSuccess ? ADMP_OK : ADMP_ERR_HIP;
  }
  return hipLaunchHostFunc(s, free_host, host) == hipSuccess ? ADMP_OK : ADMP_ERR_HIP;
}

// Table (written by the prologue for a batch of one, uploaded otherwise),
// prologue, binning and the cell-order pass (sorted slots, packed positions
// and charges, mesh reference points).
template <typename T>
static int batch_prep(const BatchPlan& P, const double* pos, const double* q, char* ws,
                      hipStream_t s, bool upload) {
  const int nsys = P.nsys;
  SysDesc* tab = (SysDesc*)(ws + P.o_tab);
  int64_t* offs = (int64_t*)(ws + P.o_off);
  OneSystem one;
  std::memset(&one, 0, sizeof(one));
  if (nsys == 1) {
    // the prologue writes the table from its arguments
    one.d = P.sys[0];
    for (int k = 0; k < kSegCount; ++k) one.tot[k] = P.off[k][1];
  } else if (upload) {
    const int st = upload_table(P, ws, s, false);
    if (st != ADMP_OK) return st;
  }
  const int64_t ntot = P.ntot;
  char* nbws = ws + P.o_nb;
  const NbLayout& L = P.nb;
  int* count = (int*)(nbws + L.count);
  int* start = (int*)(nbws + L.start);
  const int64_t ncount = 2 * (P.ctot * kSub + 1);  // counters and scatter cursors
  const int64_t nmesh = P.any_nonbrick ? P.mtot : 0;
  const int64_t nfill = std::max(std::max(ncount, nmesh), 3 * ntot);
  k_b_prologue<T><<<(unsigned)std::min<int64_t>((nfill / 4 + kBlock) / kBlock, 8192), kBlock, 0, s>>>(
      nsys == 1, one, tab, offs, count, ncount, (T*)(ws + P.o_mesh), nmesh,
      (double*)(ws + P.o_grad), 3 * ntot, (int*)(ws + P.o_tickets), kTicketStride * nsys);
  if (ntot > 0)
    k_b_bin<<<(unsigned)std::max<int64_t>((ntot + kBlock - 1) / kBlock, 1), kBlock, 0, s>>>(
        ntot, nsys, offs + (int64_t)(nsys + 1) * kSegAtom, tab, pos, (double*)(nbws + L.w),
        (int*)(nbws + L.cell_of), (int*)(ws + P.o_sys_of), count, (int*)(nbws + L.unsorted));
  // cell starts (scanned in the scatter kernel when the counters fit LDS), scatter + pack
  const int64_t nkeys = P.ctot * kSub;
  const bool scan_inline = nkeys <= kScatterScanMax;
  if (!scan_inline) launch_scan<int>(count, nkeys, start, (int*)(nbws + L.scan), s);
  k_b_scatter_pack<<<(unsigned)std::max<int64_t>((ntot + kBlock - 1) / kBlock, 1), kBlock, 0,
                     s>>>(ntot, (int)std::min<int64_t>(nkeys, INT32_MAX), count, start,
                          scan_inline ? nullptr : start, (const int*)(nbws + L.unsorted),
                          (const int*)(nbws + L.cell_of), (const double*)(nbws + L.w), q,
                          (int*)(nbws + L.sorted), (P4<double>*)(ws + P.o_ps),
                          (const int*)(ws + P.o_sys_of), tab, (int2*)(ws + P.o_mref));
  return ADMP_OK;
}

// Spread of the charges q (per atom; the brick path reads them through qv when
// given, else from the packed ps): mesh bricks (every point stored, also with
// no atoms: the mesh is not zeroed for them), LDS tiles added to the mesh, or
// per-atom atomics.  Writes every atom's stencil record.
template <typename T>
static void batch_spread(const BatchPlan& P, const double* pos, const double* q,
                         const double* qv, char* ws, hipStream_t s) {
  const int nsys = P.nsys;
  const SysDesc* tab = (const SysDesc*)(ws + P.o_tab);
  const int64_t* offs = (const int64_t*)(ws + P.o_off);
  auto seg = [&](int k) { return offs + (int64_t)(nsys + 1) * k; };
  char* nbws = ws + P.o_nb;
  const NbLayout& L = P.nb;
  const int* start = (const int*)(nbws + L.start);
  const int* sorted = (const int*)(nbws + L.sorted);
  int4* stm = (int4*)(ws + P.o_stm);
  T* stw = (T*)(ws + P.o_stw);
  T* mesh = (T*)(ws + P.o_mesh);
  const int64_t ntot = P.ntot;
  const int64_t nbricks = P.off[kSegBrick][nsys], ndeep = P.off[kSegBrickZ][nsys];
  if (P.any_brick && nbricks > 0)
    k_b_spread_bricks<T, 1><<<(unsigned)nbricks, kBrickThreads, 0, s>>>(
        nsys, seg(kSegBrick), tab, (const P4<double>*)(ws + P.o_ps),
        (const int2*)(ws + P.o_mref), sorted, start, qv, stm, stw, mesh);
  if (P.any_brick && ndeep > 0)
    k_b_spread_bricks<T, kBrickWaves><<<(unsigned)ndeep, kBrickThreads, 0, s>>>(
        nsys, seg(kSegBrickZ), tab, (const P4<double>*)(ws + P.o_ps),
        (const int2*)(ws + P.o_mref), sorted, start, qv, stm, stw, mesh);
  if (ntot > 0) {
    if (P.lds_tile > 0)
      k_b_spread_tiles<T><<<(unsigned)P.ctot, kSpreadThreads, P.lds_tile, s>>>(
          nsys, seg(kSegCell), tab, pos, q, sorted, start, stm, stw, mesh);
    if (P.any_untiled)
      k_b_spread_atoms<T><<<(unsigned)((ntot * 8 + kBlock - 1) / kBlock), kBlock, 0, s>>>(
          ntot, (const int*)(ws + P.o_sys_of), tab, pos, q, stm, stw, mesh);
  }
}

// r2c FFT, influence function (energy and dE/dA partials), c2r FFT: the mesh
// becomes the potential mesh dE/dQ.
template <typename T>
static void batch_fft(const BatchPlan& P, char* ws, hipStream_t s) {
  using CT = cplx<T>;
  const int nsys = P.nsys;
  const SysDesc* tab = (const SysDesc*)(ws + P.o_tab);
  const int64_t* offs = (const int64_t*)(ws + P.o_off);
  auto seg = [&](int k) { return offs + (int64_t)(nsys + 1) * k; };
  T* mesh = (T*)(ws + P.o_mesh);
  CT* spec = (CT*)(ws + P.o_spec);
  const int64_t rows_x = P.off[kSegFftX][nsys];
  const unsigned bz = (unsigned)P.off[kSegFftZ][nsys], by = (unsigned)P.off[kSegFftY][nsys],
                 bx = (unsigned)rows_x;
  (void)bx;
  double* cpart = (double*)(ws + P.o_cpart);
  k_b_r2c<T><<<bz, kBlock, P.lds_fft[0], s>>>(nsys, seg(kSegFftZ), tab, mesh, spec);
  if (by > 0)
    k_b_cols_y<T, -1><<<by, kBlock, P.lds_fft[1], s>>>(nsys, seg(kSegFftY), tab, spec);
  const unsigned ncols = (unsigned)P.off[kSegColsX][nsys], nplane = (unsigned)P.off[kSegPlane][nsys];
  if (P.any_cols && ncols > 0)
    k_b_cols_x<T><<<ncols, kBlock, P.lds_fft[2], s>>>(nsys, seg(kSegColsX), seg(kSegFftX), tab,
                                                      spec, cpart, rows_x);
  if (P.any_plane && nplane > 0)
    k_b_plane<T><<<nplane, kPlaneThreads, P.lds_plane, s>>>(nsys, seg(kSegPlane), seg(kSegFftX),
                                                            tab, spec, cpart, rows_x);
  if (by > 0)
    k_b_cols_y<T, 1><<<by, kBlock, P.lds_fft[1], s>>>(nsys, seg(kSegFftY), tab, spec);
  k_b_c2r<T><<<bz, kBlock, P.lds_fft[0], s>>>(nsys, seg(kSegFftZ), tab, spec, mesh);
}

template <typename T>
static int run_batch_t(const BatchPlan& P, const double* pos, const double* q, double* energy,
                       double* forces, double* dEdbox, char* ws, hipStream_t s, hipEvent_t* marks,
                       bool upload, bool fork) {
  const int nsys = P.nsys;
  const SysDesc* tab = (const SysDesc*)(ws + P.o_tab);
  const int64_t* offs = (const int64_t*)(ws + P.o_off);
  auto seg = [&](int k) { return offs + (int64_t)(nsys + 1) * k; };
  auto mark = [&](int k) {
    if (marks) (void)hipEventRecord(marks[k], s);
  };
  char* nbws = ws + P.o_nb;
  const NbLayout& L = P.nb;
  const int* start = (const int*)(nbws + L.start);
  double* grad = (double*)(ws + P.o_grad);
  T* mesh = (T*)(ws + P.o_mesh);
  int4* stm = (int4*)(ws + P.o_stm);
  T* stw = (T*)(ws + P.o_stw);
  int* tickets = (int*)(ws + P.o_tickets);
  const int64_t rows_real = P.off[kSegReal][nsys];
  const int64_t rows_x = P.off[kSegFftX][nsys], rows_g = P.off[kSegGather][nsys];

  mark(0);
  // ---- prologue, binning, in-cell order
  const int st = batch_prep<T>(P, pos, q, ws, s, upload);
  if (st != ADMP_OK) return st;
  mark(1);
  // ---- real space (also for empty batches: its partial rows feed the reductions),
  // forked onto the side stream unless stages are being timed or the caller asks not to
  double* rpart = (double*)(ws + P.o_rpart);
  SideStream* side = (P.ctot > 0 && !marks && fork) ? side_stream(s) : nullptr;
  hipStream_t rs = s;
  hipEvent_t join = nullptr;
  if (side) {
    const unsigned k = side->next.fetch_add(1) % kForkRing;
    if (hipEventRecord(side->fork[k], s) == hipSuccess &&
        hipStreamWaitEvent(side->s, side->fork[k], 0) == hipSuccess) {
      rs = side->s;
      join = side->join[k];
    }
  }
  // Enqueued after spread when forked: the stream order is unchanged, but a
  // captured graph then replays real space beside the FFT chain (0.123 against
  // 0.144 ms a 24k-atom step) instead of ahead of it.
  auto launch_real = [&]() -> bool {
    if (P.ctot > 0) {
      k_b_real<T><<<(unsigned)rows_real, kRealBlock, 0, rs>>>(
          nsys, seg(kSegReal), tab, (const P4<double>*)(ws + P.o_ps),
          (const int*)(nbws + L.sorted), start, grad, rpart, rows_real);
    }
    return !join || hipEventRecord(join, rs) == hipSuccess;
  };
  if (!join && !launch_real()) return ADMP_ERR_HIP;
  mark(2);
  // ---- spread
  batch_spread<T>(P, pos, q, nullptr, ws, s);
  if (join && !launch_real()) return ADMP_ERR_HIP;
  mark(3);
  // ---- FFT, influence function, inverse FFT
  batch_fft<T>(P, ws, s);
  mark(4);
  if (join && hipStreamWaitEvent(s, join, 0) != hipSuccess) return ADMP_ERR_HIP;
  // ---- gather with the final forces
  double* gpart = (double*)(ws + P.o_gpart);
  if (rows_g > 0)
    k_b_gather<T><<<(unsigned)rows_g, kBlock, 0, s>>>(nsys, seg(kSegGather), tab, pos, q, mesh,
                                                      grad, stm, stw, forces, gpart, rows_g);
  mark(5);
  // ---- per-system reductions and outputs
  PartialSets sets;
  sets.real = rpart;
  sets.recip = (double*)(ws + P.o_cpart);
  sets.gath = gpart;
  sets.rows_real = rows_real;
  sets.rows_recip = rows_x;
  sets.rows_gath = rows_g;
  sets.nsplit = P.nsplit;
  if (nsys >= kFinalizePerSystem)
    k_b_reduce_finalize_sys<<<nsys, kBlock, 0, s>>>(nsys, offs, tab, sets, energy, dEdbox);
  else
    k_b_reduce_finalize<<<dim3(nsys, kSysValues), kBlock, 0, s>>>(
        nsys, offs, tab, sets, (double*)(ws + P.o_red), tickets, energy, dEdbox);
  mark(6);
  return hipGetLastError() == hipSuccess ? ADMP_OK : ADMP_ERR_HIP;
}

// ------------------------------------------------------------------ QEq pieces
// out[i] += sum_g phi(g) theta_i(g): the reciprocal-space potential of every
// atom from the potential mesh and the stencil records (8 lanes per atom, as
// the gather).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_b_potential(int nsys, const int64_t* __restrict__ off,
                                                        const SysDesc* __restrict__ tab,
                                                        const T* __restrict__ meshes,
                                                        const int4* __restrict__ stm,
                                                        const T* __restrict__ stw,
                                                        double* __restrict__ out) {
  const int sid = find_seg(off, nsys, blockIdx.x);
  const SysDesc* D = tab + sid;
  const int a = threadIdx.x & 7;
  const int64_t atom =
      D->atom0 + (blockIdx.x - off[sid]) * kGatherAtomsPerBlock + (threadIdx.x >> 3);
  const bool valid = atom < D->atom0 + D->n;
  T sx = 0, sy = 0, sz = 0, sp = 0;
  if (valid)
    gather_lane_rec<T>(D->g.k[0], D->g.k[1], D->g.k[2], stm[atom],
                       stw + (int64_t)kStencilVals * atom, a, meshes + D->mesh0, &sx, &sy, &sz,
                       &sp);
  sp = group8_sum(sp);  // every lane takes part
  if (valid && a == 0) out[atom] += (double)sp;
}

int batch_geometry(const BatchPlan& P, const double* pos, const double* q, char* ws,
                   hipStream_t s) {
  if (P.status != ADMP_OK) return P.status;
  const int st = P.precision == ADMP_F32 ? batch_prep<float>(P, pos, q, ws, s, true)
                                         : batch_prep<double>(P, pos, q, ws, s, true);
  if (st != ADMP_OK) return st;
  return hipGetLastError() == hipSuccess ? ADMP_OK : ADMP_ERR_HIP;
}

template <typename T>
static int batch_potential_t(const BatchPlan& P, const double* pos, const double* v, double* out,
                             char* ws, hipStream_t s) {
  const int nsys = P.nsys;
  T* mesh = (T*)(ws + P.o_mesh);
  if (P.any_nonbrick)
    k_fill<T><<<(unsigned)std::max<int64_t>(std::min<int64_t>((P.mtot + kBlock - 1) / kBlock, 4096),
                                            1),
                kBlock, 0, s>>>(mesh, P.mtot, T(0));
  batch_spread<T>(P, pos, v, v, ws, s);
  batch_fft<T>(P, ws, s);
  const int64_t rows_g = P.off[kSegGather][nsys];
  if (rows_g > 0)
    k_b_potential<T><<<(unsigned)rows_g, kBlock, 0, s>>>(
        nsys, (const int64_t*)(ws + P.o_off) + (int64_t)(nsys + 1) * kSegGather,
        (const SysDesc*)(ws + P.o_tab), mesh, (const int4*)(ws + P.o_stm),
        (const T*)(ws + P.o_stw), out);
  return hipGetLastError() == hipSuccess ? ADMP_OK : ADMP_ERR_HIP;
}

int batch_potential(const BatchPlan& P, const double* pos, const double* v, double* out, char* ws,
                    hipStream_t s) {
  if (P.status != ADMP_OK) return P.status;
  return P.precision == ADMP_F32 ? batch_potential_t<float>(P, pos, v, out, ws, s)
                                 : batch_potential_t<double>(P, pos, v, out, ws, s);
}

int run_batch(const BatchPlan& plan, const double* pos, const double* q, double* energy,
              double* forces, double* dEdbox, char* ws, hipStream_t s, hipEvent_t* marks,
              bool upload, bool fork) {
  if (plan.status != ADMP_OK) return plan.status;
  return plan.precision == ADMP_F32
             ? run_batch_t<float>(plan, pos, q, energy, forces, dEdbox, ws, s, marks, upload, fork)
             : run_batch_t<double>(plan, pos, q, energy, forces, dEdbox, ws, s, marks, upload,
                                   fork);
}

}  // namespace admp
