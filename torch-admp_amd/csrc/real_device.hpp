// Device building blocks of the cell-list real-space kernels, shared by the
// single-system and batched paths.
#pragma once

#include <climits>

#include "nblist_kernels.hpp"

namespace admp {

// Stencil entries along one dimension: cell index and image shift.
__device__ __forceinline__ int stencil_len(const CellGrid& cg, int d) {
  return cg.nc[d] >= 3 ? 3 : 2 * cg.nimg[d] + 1;
}
__device__ __forceinline__ void stencil_at(const CellGrid& cg, int d, int ci, int o, int* cell,
                                           int* shift) {
  if (cg.nc[d] >= 3) {
    int c = ci + o - 1;
    int s = 0;
    if (c < 0) { c += cg.nc[d]; s = -1; }
    else if (c >= cg.nc[d]) { c -= cg.nc[d]; s = 1; }
    *cell = c;
    *shift = s;
  } else {
    *cell = 0;
    *shift = o - cg.nimg[d];
  }
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T>
struct PairAcc {
  double e;
  T vxx, vyy, vzz, vxy, vxz, vyz;
};

// Staged atom: position relative to the home-cell origin as an unevaluated sum
// hi + lo of the bulk type (exact to float64 for float), plus the charge.
template <typename T>
struct Staged {
  T xh, yh, zh, q, xl, yl, zl, pad;
};

template <typename T>
__device__ __forceinline__ Staged<T> stage(double x, double y, double z, double q) {
  Staged<T> s;
  s.xh = (T)x;
  s.yh = (T)y;
  s.zh = (T)z;
  s.xl = (T)(x - (double)s.xh);
  s.yl = (T)(y - (double)s.yh);
  s.zl = (T)(z - (double)s.zh);
  s.q = (T)q;
  s.pad = T(0);
  return s;
}

// One unordered pair per lane of a wavefront (`valid` lanes only): adds dE/dr_i
// to the home accumulator, the energy and the pair virial, and dE/dr_j to the
// candidate accumulator.  Lanes holding the same candidate form contiguous
// runs (queue order), so their candidate terms are summed by a segmented
// wavefront reduction and added by one LDS atomic per run instead of
// colliding on one address.
template <typename T>
__device__ __forceinline__ void pair_eval(const Staged<T>& pi, const Staged<T>& pj, bool valid,
                                          int kx, T kappa, T eps, T c_exp, T k2, double* facc_i,
                                          double* cacc, PairAcc<T>& acc) {
  double gx = 0.0, gy = 0.0, gz = 0.0;
  if (valid) {
    // hi parts of nearby atoms subtract exactly; the lo parts restore float64 accuracy
    const T dx = (pj.xh - pi.xh) + (pj.xl - pi.xl);
    const T dy = (pj.yh - pi.yh) + (pj.yl - pi.yl);
    const T dz = (pj.zh - pi.zh) + (pj.zl - pi.zl);
    const T r2 = dx * dx + dy * dy + dz * dz;
    const T r = sqrt(r2);
    const T ir = T(1) / r;
    const T qq = eps * pi.q * pj.q;
    const T ec = erfc(kappa * r);
    const T fs = -qq * (c_exp * exp(-k2 * r2) + ec * ir) * ir * ir;  // (de/dr) / r
    acc.e += (double)(qq * ec * ir);
    gx = (double)(fs * dx);
    gy = (double)(fs * dy);
    gz = (double)(fs * dz);
    atomicAdd(facc_i + 0, -gx);
    atomicAdd(facc_i + 1, -gy);
    atomicAdd(facc_i + 2, -gz);
    acc.vxx += fs * dx * dx;
    acc.vyy += fs * dy * dy;
    acc.vzz += fs * dz * dz;
    acc.vxy += fs * dx * dy;
    acc.vxz += fs * dx * dz;
    acc.vyz += fs * dy * dz;
  }
  const int lane = threadIdx.x & 63;
  const int key = valid ? kx : -1 - lane;  // invalid lanes form runs of their own
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const double tx = __shfl_down(gx, off, 64);
    const double ty = __shfl_down(gy, off, 64);
    const double tz = __shfl_down(gz, off, 64);
    const int tk = __shfl_down(key, off, 64);
    if (lane + off < 64 && tk == key) {
      gx += tx;
      gy += ty;
      gz += tz;
    }
  }
  const int prev = __shfl_up(key, 1, 64);
  if (valid && (lane == 0 || prev != key)) {
    atomicAdd(cacc + 3 * kx, gx);
    atomicAdd(cacc + 3 * kx + 1, gy);
    atomicAdd(cacc + 3 * kx + 2, gz);
  }
}

// Stencil entries [st0, st0 + kMaxStencil) of home cell ci: first slot, atom
// count (prefix-summed into st_pre) and image shift of every entry.
// Returns the number of entries of this chunk.
__device__ __forceinline__ int load_stencil_chunk(const CellGrid& cg, int64_t cell0,
                                                  const int ci[3], int st0, int nst,
                                                  const int* __restrict__ start, int* st_pre,
                                                  int* st_beg, double* st_sh) {
  const int l1 = stencil_len(cg, 1), l2 = stencil_len(cg, 2);
  const int m = min(kMaxStencil, nst - st0);
  const int lane = threadIdx.x & 63;
  for (int e = threadIdx.x; e < m; e += blockDim.x) {
    const int st = st0 + e;
    const int o2 = st % l2, o1 = (st / l2) % l1, o0 = st / (l1 * l2);
    int c0, s0, c1, s1, c2, s2;
    stencil_at(cg, 0, ci[0], o0, &c0, &s0);
    stencil_at(cg, 1, ci[1], o1, &c1, &s1);
    stencil_at(cg, 2, ci[2], o2, &c2, &s2);
    const int64_t cc = cell0 + (c0 * cg.nc[1] + c1) * cg.nc[2] + c2;
    st_beg[e] = start[cc * kSub];
    st_pre[e + 1] = start[(cc + 1) * kSub] - start[cc * kSub];
#pragma unroll
    for (int d = 0; d < 3; ++d) st_sh[e * 3 + d] = s0 * cg.h[d] + s1 * cg.h[3 + d] + s2 * cg.h[6 + d];
  }
  __syncthreads();
  if (threadIdx.x < 64) {
    // inclusive prefix of the counts st_pre[1..m] by wavefront scan, 64 at a time
    int carry = 0;
    for (int b0 = 1; b0 <= m; b0 += 64) {
      const int idx = b0 + lane;
      int v = idx <= m ? st_pre[idx] : 0;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(v, off, 64);
        if (lane >= off) v += u;
      }
      if (idx <= m) st_pre[idx] = v + carry;
      carry += __shfl(v, 63, 64);
    }
    if (lane == 0) st_pre[0] = 0;
  }
  __syncthreads();
  return m;
}

// One work-group's share of the cell-staged real-space sum, each unordered
// pair once (Newton's third law): home cell lc of a grid whose cells start at
// global index cell0, against the half stencil -- the entries after the centre
// in lexicographic offset order, plus the centre itself with slot j > slot i --
// split over gridDim.y work-groups (candidate k goes to split k mod S).
// Candidates are staged in LDS as hi/lo pairs relative to the home-cell
// origin; lanes of a wavefront filter (home, candidate) pairs against the
// cutoff and compact the survivors into a per-wavefront queue evaluated 64
// pairs at a time.  dE/dr of home and candidate atoms gather in float64 LDS
// accumulators and are added to grad[i][3] with float64 global atomics
// (grad zeroed beforehand).  The block's energy, pair virial, sum r (x) dE/dr,
// sum q^2 and sum q (kRealNV values) go to partial[k * pstride].
template <typename T>
__device__ __forceinline__ void real_cell_body(const CellGrid& cg, int64_t cell0, int lc,
                                               T kappa, T eps,
                                               const P4<double>* __restrict__ ps,
                                               const int* __restrict__ sorted,
                                               const int* __restrict__ start,
                                               const double* __restrict__ pos,
                                               double* __restrict__ grad,
                                               double* __restrict__ partial, int64_t pstride) {
  __shared__ Staged<T> cand[kCandCap];
  __shared__ int cslot[kCandCap];
  __shared__ double cacc[kCandCap * 3];
  __shared__ Staged<T> home[kHomeCap];
  __shared__ double facc[kHomeCap * 3];  // float64 LDS atomics in both precisions
  __shared__ int queue[(kBlock / 64) * kQueueCap];
  __shared__ int st_pre[kMaxStencil + 1];
  __shared__ int st_beg[kMaxStencil];
  __shared__ double st_sh[kMaxStencil * 3];
  __shared__ double red[kRealNV * (kBlock / 64)];

  const int split = blockIdx.y, S = gridDim.y;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ci[3] = {lc / (cg.nc[1] * cg.nc[2]), (lc / cg.nc[2]) % cg.nc[1], lc % cg.nc[2]};
  const int l0 = stencil_len(cg, 0), l1 = stencil_len(cg, 1), l2 = stencil_len(cg, 2);
  const int nst = l0 * l1 * l2;
  const int st_c = ((l0 / 2) * l1 + l1 / 2) * l2 + l2 / 2;  // the zero offset
  double o3[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double o = 0.0;
#pragma unroll
    for (int e = 0; e < 3; ++e) o += ((double)ci[e] / cg.nc[e]) * cg.h[e * 3 + d];
    o3[d] = o;
  }
  const double ox = o3[0], oy = o3[1], oz = o3[2];
  const int64_t c = cell0 + lc;
  const int hb = start[c * kSub], he = start[(c + 1) * kSub];
  const T rc2 = (T)cg.rc2, rmin2 = (T)1e-4;
  const T c_exp = (T)(2.0 / kSqrtPi) * kappa, k2 = kappa * kappa;
  PairAcc<T> pacc;
  pacc.e = 0.0;
  pacc.vxx = pacc.vyy = pacc.vzz = pacc.vxy = pacc.vxz = pacc.vyz = T(0);
  double acc[kRealNV];
#pragma unroll
  for (int k = 0; k < kRealNV; ++k) acc[k] = 0.0;
  int* q = queue + wave * kQueueCap;
  const int nhalf = nst - st_c;
  const bool one_chunk = nhalf <= kMaxStencil;
  int m = 0;
  if (one_chunk) m = load_stencil_chunk(cg, cell0, ci, st_c, nst, start, st_pre, st_beg, st_sh);

  for (int h0 = hb; h0 < he; h0 += kHomeCap) {
    const int hc = min(kHomeCap, he - h0);
    for (int t = threadIdx.x; t < hc; t += blockDim.x) {
      const P4<double> p = ps[h0 + t];
      home[t] = stage<T>(p.x - ox, p.y - oy, p.z - oz, p.q);
      facc[3 * t] = facc[3 * t + 1] = facc[3 * t + 2] = 0.0;
    }
    for (int st0 = st_c; st0 < nst; st0 += kMaxStencil) {
      if (!one_chunk) {
        __syncthreads();
        m = load_stencil_chunk(cg, cell0, ci, st0, nst, start, st_pre, st_beg, st_sh);
      }
      // the centre entry (home cell, no shift) leads the first chunk
      const int ncentre = st0 == st_c ? st_pre[1] : 0;
      const int ctot = st_pre[m];
      const int cs = ctot > split ? (ctot - split + S - 1) / S : 0;
      for (int k0 = 0; k0 < cs; k0 += kCandCap) {
        const int cc = min(kCandCap, cs - k0);
        for (int kk = threadIdx.x; kk < cc; kk += blockDim.x) {
          const int kg = split + S * (k0 + kk);
          int lo = 0, hi = m;  // last entry with st_pre[entry] <= kg
          while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (st_pre[mid] <= kg) lo = mid;
            else hi = mid;
          }
          const int slot = st_beg[lo] + (kg - st_pre[lo]);
          const P4<double> p = ps[slot];
          cand[kk] = stage<T>((p.x + st_sh[lo * 3]) - ox, (p.y + st_sh[lo * 3 + 1]) - oy,
                              (p.z + st_sh[lo * 3 + 2]) - oz, p.q);
          // centre candidates pair only with home atoms of lower slot
          cslot[kk] = kg < ncentre ? slot : INT_MAX;
          cacc[3 * kk] = cacc[3 * kk + 1] = cacc[3 * kk + 2] = 0.0;
        }
        __syncthreads();
        for (int g0 = 0; g0 < hc; g0 += 64) {
          const int hg = min(64, hc - g0);
          const int nph = 64 / hg;
          const int hl = lane % hg, ph = lane / hg;
          const bool active = ph < nph;
          const Staged<T> pi = home[g0 + hl];
          const int my_slot = h0 + g0 + hl;
          const int stride = nph * (kBlock / 64);
          const int nit = (cc + stride - 1) / stride;
          int qn = 0;
          for (int j0 = 0; j0 < nit; j0 += kFilterUnroll) {
            bool pred[kFilterUnroll];
            int kks[kFilterUnroll];
#pragma unroll
            for (int u = 0; u < kFilterUnroll; ++u) {
              const int kk = ph + nph * wave + stride * (j0 + u);
              kks[u] = kk;
              pred[u] = false;
              if (active && kk < cc) {
                // cutoff test on the hi parts only (a 1e-6 A band at rc in float32)
                const Staged<T>& pj = cand[kk];
                const T dx = pj.xh - pi.xh, dy = pj.yh - pi.yh, dz = pj.zh - pi.zh;
                const T r2 = dx * dx + dy * dy + dz * dz;
                pred[u] = r2 < rc2 && r2 > rmin2 && cslot[kk] > my_slot;
              }
            }
#pragma unroll
            for (int u = 0; u < kFilterUnroll; ++u) {
              const unsigned long long msk = __ballot(pred[u]);
              if (pred[u]) {
                const int at = qn + __popcll(msk & ((1ull << lane) - 1ull));
                q[at] = ((g0 + hl) << 16) | kks[u];
              }
              qn += __popcll(msk);
            }
            while (qn >= 64) {
              wave_sync();
              const int ent = q[lane];
              const int hx = ent >> 16, kx = ent & 0xffff;
              pair_eval<T>(home[hx], cand[kx], true, kx, kappa, eps, c_exp, k2, facc + 3 * hx,
                           cacc, pacc);
              const int rem = qn - 64;
              // move the tail (at most kQueueCap - 64 entries) to the front
              int tails[kFilterUnroll];
#pragma unroll
              for (int u = 0; u < kFilterUnroll; ++u)
                tails[u] = (64 * u + lane < rem) ? q[64 + 64 * u + lane] : 0;
              wave_sync();
#pragma unroll
              for (int u = 0; u < kFilterUnroll; ++u)
                if (64 * u + lane < rem) q[64 * u + lane] = tails[u];
              wave_sync();
              qn = rem;
            }
          }
          wave_sync();
          if (qn > 0) {
            const bool valid = lane < qn;
            const int ent = valid ? q[lane] : 0;
            const int hx = ent >> 16, kx = ent & 0xffff;
            pair_eval<T>(home[hx], cand[kx], valid, kx, kappa, eps, c_exp, k2, facc + 3 * hx,
                         cacc, pacc);
          }
          wave_sync();
        }
        __syncthreads();
        // candidate gradients of this chunk
        for (int kk = threadIdx.x; kk < cc; kk += blockDim.x) {
          const double g[3] = {cacc[3 * kk], cacc[3 * kk + 1], cacc[3 * kk + 2]};
          if (g[0] == 0.0 && g[1] == 0.0 && g[2] == 0.0) continue;
          const int kg = split + S * (k0 + kk);
          int lo = 0, hi = m;
          while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (st_pre[mid] <= kg) lo = mid;
            else hi = mid;
          }
          const int j = sorted[st_beg[lo] + (kg - st_pre[lo])];
          unsafeAtomicAdd(grad + 3 * j, g[0]);
          unsafeAtomicAdd(grad + 3 * j + 1, g[1]);
          unsafeAtomicAdd(grad + 3 * j + 2, g[2]);
          const double r[3] = {pos[3 * j], pos[3 * j + 1], pos[3 * j + 2]};
#pragma unroll
          for (int a2 = 0; a2 < 3; ++a2)
#pragma unroll
            for (int b2 = 0; b2 < 3; ++b2) acc[7 + a2 * 3 + b2] += r[a2] * g[b2];
        }
      }
    }
    __syncthreads();
    // home gradients of this split
    for (int t = threadIdx.x; t < hc; t += blockDim.x) {
      const int i = sorted[h0 + t];
      const double g[3] = {facc[3 * t], facc[3 * t + 1], facc[3 * t + 2]};
      if (g[0] != 0.0 || g[1] != 0.0 || g[2] != 0.0) {
        unsafeAtomicAdd(grad + 3 * i, g[0]);
        unsafeAtomicAdd(grad + 3 * i + 1, g[1]);
        unsafeAtomicAdd(grad + 3 * i + 2, g[2]);
        const double r[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
#pragma unroll
        for (int a2 = 0; a2 < 3; ++a2)
#pragma unroll
          for (int b2 = 0; b2 < 3; ++b2) acc[7 + a2 * 3 + b2] += r[a2] * g[b2];
      }
      if (split == 0) {
        const double qi = ps[h0 + t].q;
        acc[16] += qi * qi;
        acc[17] += qi;
      }
    }
    __syncthreads();
  }
  acc[0] = pacc.e;
  acc[1] = (double)pacc.vxx;
  acc[2] = (double)pacc.vyy;
  acc[3] = (double)pacc.vzz;
  acc[4] = (double)pacc.vxy;
  acc[5] = (double)pacc.vxz;
  acc[6] = (double)pacc.vyz;
  block_sum<kRealNV>(acc, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < kRealNV; ++k) partial[k * pstride] = acc[k];
}

}  // namespace admp
