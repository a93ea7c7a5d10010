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

// Staged atom: position relative to the home-cell origin as an unevaluated sum
// hi + lo of the bulk type (exact to float64 for float), the charge, and a tag
// in the bits of the fourth hi word (for candidates: the sorted slot of a
// centre-cell atom, INT_MAX otherwise), so the cutoff filter reads one
// 16-byte word (32 in float64) per candidate.
template <typename T>
struct Staged {
  typedef T V4 __attribute__((ext_vector_type(4)));
  V4 hi;  // x, y, z hi parts, tag
  V4 lo;  // x, y, z lo parts, charge
};

__device__ __forceinline__ float tag_bits(int v, float) { return __int_as_float(v); }
__device__ __forceinline__ double tag_bits(int v, double) {
  return __longlong_as_double((long long)v);
}
__device__ __forceinline__ int tag_of(float v) { return __float_as_int(v); }
__device__ __forceinline__ int tag_of(double v) { return (int)__double_as_longlong(v); }

template <typename T>
__device__ __forceinline__ Staged<T> stage(double x, double y, double z, double q, int tag) {
  Staged<T> s;
  s.hi.x = (T)x;
  s.hi.y = (T)y;
  s.hi.z = (T)z;
  s.hi.w = tag_bits(tag, T(0));
  s.lo.x = (T)(x - (double)s.hi.x);
  s.lo.y = (T)(y - (double)s.hi.y);
  s.lo.z = (T)(z - (double)s.hi.z);
  s.lo.w = (T)q;
  return s;
}

// Screened Coulomb pair term at squared distance r2: e = qq erfc(kappa r) / r and
// fs = (de/dr) / r.  Constants: kappa, kappa^2 as hi + lo, 2 kappa / sqrt(pi).
template <typename T>
struct CoulConst {
  T kappa, k2h, k2l, c_exp;
};

__device__ __forceinline__ void coul_pair(double r2, double qq, const CoulConst<double>& C,
                                          double* e, double* fs) {
  const double r = sqrt(r2);
  const double ir = 1.0 / r;
  const double ec = erfc(C.kappa * r);
  *e = qq * ec * ir;
  *fs = -qq * (C.c_exp * exp(-C.k2h * r2) + ec * ir) * ir * ir;
}

// Float32: erfc(x) = exp(-x^2) t S(t) with t = 1 / (1 + p x), the form of
// Abramowitz & Stegun 7.1.26, S of degree 10 fitted to erfc(x) exp(x^2) / t for
// all x >= 0 by weighted least squares (relative error 4.5e-8 in exact
// arithmetic with these float32 coefficients).  exp(-kappa^2 r^2) is one
// hardware exp2 with the rounding of kappa^2 r^2 and of its log2(e) scaling
// carried as a first-order correction, and serves both the energy and the
// force.  Relative error of erfc evaluated this way in float32 (r 0.01-8 A,
// kappa 0.3-0.7): below 3.5e-7, mean 7.4e-8, bias 3.6e-9.
__device__ __forceinline__ void coul_pair(float r2, float qq, const CoulConst<float>& C, float* e,
                                          float* fs) {
  constexpr float kLog2eH = 1.44269502163f, kLog2eL = 1.925963033e-8f, kLn2 = 0.693147182f;
  const float ir = __builtin_amdgcn_rsqf(r2);
  const float r = r2 * ir;
  const float y = C.k2h * r2;
  const float yl = fmaf(C.k2h, r2, -y) + C.k2l * r2;
  const float a = y * kLog2eH;
  const float al = fmaf(y, kLog2eH, -a) + fmaf(y, kLog2eL, yl * kLog2eH);
  float ex = __builtin_amdgcn_exp2f(-a);
  ex = fmaf(-ex, al * kLn2, ex);
  const float t = __builtin_amdgcn_rcpf(fmaf(0.47f, C.kappa * r, 1.0f));
  float s = 3.685952723e-02f;
  s = fmaf(s, t, -2.302833945e-01f);
  s = fmaf(s, t, 5.820118189e-01f);
  s = fmaf(s, t, -7.192687988e-01f);
  s = fmaf(s, t, 4.181051552e-01f);
  s = fmaf(s, t, -1.793078780e-01f);
  s = fmaf(s, t, 1.588312238e-01f);
  s = fmaf(s, t, 1.654326916e-01f);
  s = fmaf(s, t, 2.373880148e-01f);
  s = fmaf(s, t, 2.650590837e-01f);
  s = fmaf(s, t, 2.651726007e-01f);
  const float ec_ir = ex * t * s * ir;  // erfc(kappa r) / r
  *e = qq * ec_ir;
  *fs = -qq * fmaf(C.c_exp, ex, ec_ir) * ir * ir;
}

// Lane i receives lane i - K of its row of 16 (rows do not mix); lanes with no
// such source get `old`.
template <int K>
__device__ __forceinline__ int row_shr(int old, int v) {
  return __builtin_amdgcn_update_dpp(old, v, 0x110 + K, 0xf, 0xf, false);
}
template <int K>
__device__ __forceinline__ float row_shr(float v) {
  return __int_as_float(row_shr<K>(0, __float_as_int(v)));
}
template <int K>
__device__ __forceinline__ double row_shr(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = row_shr<K>(0, (int)b), hi = row_shr<K>(0, (int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
// Lane i receives lane i + 1 of its row; the last lane of a row gets `old`.
__device__ __forceinline__ int row_shl1(int old, int v) {
  return __builtin_amdgcn_update_dpp(old, v, 0x101, 0xf, 0xf, false);
}

// Per-lane energy and pair virial, float64 in both precisions.
struct PairAcc {
  double e, vxx, vyy, vzz, vxy, vxz, vyz;
};

// One unordered pair per lane of a wavefront (`valid` lanes only): adds dE/dr_i
// to the home accumulator, the energy and the pair virial, and dE/dr_j to
// candidate kx's accumulator.  `run` identifies the filter ballot the pair came
// from (home group and candidate); lanes of one run are contiguous in queue
// order and no run id recurs elsewhere in the wavefront, so within each row of
// 16 lanes the candidate terms of a run are summed by a segmented inclusive
// scan over DPP row shifts, and the last lane of each run piece adds the sum
// with one LDS atomic.
template <typename T>
__device__ __forceinline__ void pair_eval(const Staged<T>& pi, const Staged<T>& pj, bool valid,
                                          int kx, int run, T eps, const CoulConst<T>& C,
                                          double* facc_i,
                                          double* cacc, PairAcc& acc) {
  T gx = T(0), gy = T(0), gz = T(0);
  if (valid) {
    // hi parts of nearby atoms subtract exactly; the lo parts restore float64 accuracy
    const T dx = (pj.hi.x - pi.hi.x) + (pj.lo.x - pi.lo.x);
    const T dy = (pj.hi.y - pi.hi.y) + (pj.lo.y - pi.lo.y);
    const T dz = (pj.hi.z - pi.hi.z) + (pj.lo.z - pi.lo.z);
    const T r2 = dx * dx + dy * dy + dz * dz;
    T e, fs;
    coul_pair(r2, eps * pi.lo.w * pj.lo.w, C, &e, &fs);
    acc.e += (double)e;
    gx = fs * dx;
    gy = fs * dy;
    gz = fs * dz;
    atomicAdd(facc_i + 0, -(double)gx);
    atomicAdd(facc_i + 1, -(double)gy);
    atomicAdd(facc_i + 2, -(double)gz);
    acc.vxx += (double)(gx * dx);
    acc.vyy += (double)(gy * dy);
    acc.vzz += (double)(gz * dz);
    acc.vxy += (double)(gx * dy);
    acc.vxz += (double)(gx * dz);
    acc.vyz += (double)(gy * dz);
  }
  const int lane = threadIdx.x & 63;
  const int key = valid ? run + 1 : -1 - lane;  // invalid lanes form runs of their own; 0 = none
#define ADMP_SEG_STEP(K)                            \
  {                                                 \
    const int tk = row_shr<K>(0, key);              \
    const T tx = row_shr<K>(gx), ty = row_shr<K>(gy), tz = row_shr<K>(gz); \
    const bool same = tk == key;                    \
    gx += same ? tx : T(0);                         \
    gy += same ? ty : T(0);                         \
    gz += same ? tz : T(0);                         \
  }
  ADMP_SEG_STEP(1)
  ADMP_SEG_STEP(2)
  ADMP_SEG_STEP(4)
  ADMP_SEG_STEP(8)
#undef ADMP_SEG_STEP
  const int next = row_shl1(0, key);
  if (valid && next != key) {
    atomicAdd(cacc + 3 * kx, (double)gx);
    atomicAdd(cacc + 3 * kx + 1, (double)gy);
    atomicAdd(cacc + 3 * kx + 2, (double)gz);
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
// split S ways (candidate k goes to split k mod S).
// Candidates are staged in LDS as hi/lo pairs relative to the home-cell
// origin; lanes of a wavefront filter (home, candidate) pairs against the
// cutoff and compact the survivors into a per-wavefront queue evaluated 64
// pairs at a time.  dE/dr of home and candidate atoms gather in float64 LDS
// accumulators and are added to grad[i][3] with float64 global atomics
// (grad zeroed beforehand).  The block's energy, pair virial, sum q^2 and
// sum q (kRealNV values) go to partial[k * pstride]; the term sum r (x) dE/dr
// of the virial is formed by the gather from the finished gradient rows.
template <typename T>
__device__ __forceinline__ void real_cell_body(const CellGrid& cg, int64_t cell0, int lc,
                                               int split, int S, double kappa, T eps,
                                               const P4<double>* __restrict__ ps,
                                               const int* __restrict__ sorted,
                                               const int* __restrict__ start,
                                               double* __restrict__ grad,
                                               double* __restrict__ partial, int64_t pstride) {
  constexpr int kCandCap = cand_cap<T>();
  __shared__ Staged<T> cand[kCandCap];
  __shared__ double cacc[kCandCap * 3];
  __shared__ Staged<T> home[kHomeCap];
  __shared__ double facc[kHomeCap * 3];  // float64 LDS atomics in both precisions
  __shared__ int queue[(kRealBlock / 64) * kQueueRing];
  __shared__ int st_pre[kMaxStencil + 1];
  __shared__ int st_beg[kMaxStencil];
  __shared__ double st_sh[kMaxStencil * 3];
  __shared__ double red[kRealNV * (kRealBlock / 64)];
  __shared__ int cslot[kCandCap];            // sorted slot of each staged candidate
  __shared__ int wcnt[kRealBlock / 64];      // staged candidates per wavefront
  static_assert(kCandCap == kRealBlock, "one candidate of a chunk per thread");

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
  // rectangular cells: candidates farther than rc from the home cell's box
  // pair with no home atom and are not staged (home atoms lie in the box up to
  // the binning's rounding; 1e-4 A of slack)
  const bool prune = cg.h[1] == 0.0 && cg.h[2] == 0.0 && cg.h[3] == 0.0 && cg.h[5] == 0.0 &&
                     cg.h[6] == 0.0 && cg.h[7] == 0.0;
  const double ex = cg.h[0] / cg.nc[0], ey = cg.h[4] / cg.nc[1], ez = cg.h[8] / cg.nc[2];
  const double prune_r = sqrt(cg.rc2) + 1e-4, prune_r2 = prune_r * prune_r;
  const int64_t c = cell0 + lc;
  const int hb = start[c * kSub], he = start[(c + 1) * kSub];
  const T rc2 = (T)cg.rc2, rmin2 = (T)1e-4;
  CoulConst<T> C;
  C.kappa = (T)kappa;
  C.k2h = (T)(kappa * kappa);
  C.k2l = (T)(kappa * kappa - (double)C.k2h);
  C.c_exp = (T)(2.0 / kSqrtPi * kappa);
  PairAcc pacc;
  pacc.e = pacc.vxx = pacc.vyy = pacc.vzz = pacc.vxy = pacc.vxz = pacc.vyz = 0.0;
  double qsq = 0.0, qsum = 0.0;
  int* q = queue + wave * kQueueRing;
  const int nhalf = nst - st_c;
  const bool one_chunk = nhalf <= kMaxStencil;
  int m = 0;
  if (one_chunk) m = load_stencil_chunk(cg, cell0, ci, st_c, nst, start, st_pre, st_beg, st_sh);

  for (int h0 = hb; h0 < he; h0 += kHomeCap) {
    const int hc = min(kHomeCap, he - h0);
    for (int t = threadIdx.x; t < hc; t += blockDim.x) {
      const P4<double> p = ps[h0 + t];
      home[t] = stage<T>(p.x - ox, p.y - oy, p.z - oz, p.q, 0);
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
        // stage the chunk's candidates (thread kk takes candidate kk), keeping
        // those that can reach the home cell, compacted in candidate order
        const int kk = threadIdx.x;
        bool keep = false;
        int slot = 0, kg = 0;
        double sx = 0.0, sy = 0.0, sz = 0.0, sq = 0.0;
        if (kk < min(kCandCap, cs - k0)) {
          kg = split + S * (k0 + kk);
          int lo = 0, hi = m;
          while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (st_pre[mid] <= kg) lo = mid;
            else hi = mid;
          }
          slot = st_beg[lo] + (kg - st_pre[lo]);
          const P4<double> p = ps[slot];
          sx = (p.x + st_sh[lo * 3]) - ox;
          sy = (p.y + st_sh[lo * 3 + 1]) - oy;
          sz = (p.z + st_sh[lo * 3 + 2]) - oz;
          sq = p.q;
          keep = true;
          if (prune) {
            const double dx = fmax(0.0, fmax(-sx, sx - ex));
            const double dy = fmax(0.0, fmax(-sy, sy - ey));
            const double dz = fmax(0.0, fmax(-sz, sz - ez));
            keep = dx * dx + dy * dy + dz * dz < prune_r2;
          }
        }
        const unsigned long long kmask = __ballot(keep);
        if (lane == 0) wcnt[wave] = __popcll(kmask);
        __syncthreads();
        int cc = 0, at = __popcll(kmask & ((1ull << lane) - 1ull));
#pragma unroll
        for (int w = 0; w < kRealBlock / 64; ++w) {
          at += w < wave ? wcnt[w] : 0;
          cc += wcnt[w];
        }
        if (keep) {
          // centre candidates pair only with home atoms of lower slot
          cand[at] = stage<T>(sx, sy, sz, sq, kg < ncentre ? slot : INT_MAX);
          cslot[at] = slot;
          cacc[3 * at] = cacc[3 * at + 1] = cacc[3 * at + 2] = 0.0;
        }
        __syncthreads();
        // per-wavefront ring of (home, candidate) pairs, entries [qh, qt); it runs
        // on across home groups and is drained once per candidate chunk
        int qh = 0, qt = 0;
        for (int g0 = 0; g0 < hc; g0 += 64) {
          const int hg = min(64, hc - g0);
          const int nph = 64 / hg;
          const int hl = lane % hg, ph = lane / hg;
          const bool active = ph < nph;
          const Staged<T> pi = home[g0 + hl];
          const int my_slot = h0 + g0 + hl;
          const int stride = nph * (kRealBlock / 64);
          const int nit = (cc + stride - 1) / stride;
          for (int j0 = 0; j0 < nit; j0 += kFilterUnroll) {
            typename Staged<T>::V4 cw[kFilterUnroll];
            int kks[kFilterUnroll];
#pragma unroll
            for (int u = 0; u < kFilterUnroll; ++u) {
              kks[u] = ph + nph * wave + stride * (j0 + u);
              cw[u] = cand[min(kks[u], max(cc, 1) - 1)].hi;
            }
            bool pred[kFilterUnroll];
#pragma unroll
            for (int u = 0; u < kFilterUnroll; ++u) {
              // cutoff test on the hi parts only (a 1e-6 A band at rc in float32)
              const T dx = cw[u].x - pi.hi.x, dy = cw[u].y - pi.hi.y, dz = cw[u].z - pi.hi.z;
              const T r2 = dx * dx + dy * dy + dz * dz;
              pred[u] = active & (kks[u] < cc) & (r2 < rc2) & (r2 > rmin2) &
                        (tag_of(cw[u].w) > my_slot);
            }
#pragma unroll
            for (int u = 0; u < kFilterUnroll; ++u) {
              const unsigned long long msk = __ballot(pred[u]);
              if (pred[u]) {
                const int at = qt + __popcll(msk & ((1ull << lane) - 1ull));
                q[at & (kQueueRing - 1)] = ((g0 + hl) << 16) | kks[u];
              }
              qt += __popcll(msk);
            }
            while (qt - qh >= 64) {
              wave_sync();
              const int ent = q[(qh + lane) & (kQueueRing - 1)];
              const int hx = ent >> 16, kx = ent & 0xffff;
              pair_eval<T>(home[hx], cand[kx], true, kx, ((hx >> 6) << 16) | kx, eps, C,
                           facc + 3 * hx, cacc, pacc);
              qh += 64;
            }
          }
        }
        wave_sync();
        if (qt > qh) {
          const bool valid = lane < qt - qh;
          const int ent = valid ? q[(qh + lane) & (kQueueRing - 1)] : 0;
          const int hx = ent >> 16, kx = ent & 0xffff;
          pair_eval<T>(home[hx], cand[kx], valid, kx, ((hx >> 6) << 16) | kx, eps, C,
                       facc + 3 * hx, cacc, pacc);
        }
        wave_sync();
        __syncthreads();
        // candidate gradients of this chunk
        for (int kc = threadIdx.x; kc < cc; kc += blockDim.x) {
          const double g[3] = {cacc[3 * kc], cacc[3 * kc + 1], cacc[3 * kc + 2]};
          if (g[0] == 0.0 && g[1] == 0.0 && g[2] == 0.0) continue;
          const int j = sorted[cslot[kc]];
          unsafeAtomicAdd(grad + 3 * j, g[0]);
          unsafeAtomicAdd(grad + 3 * j + 1, g[1]);
          unsafeAtomicAdd(grad + 3 * j + 2, g[2]);
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
      }
      if (split == 0) {
        const double qi = ps[h0 + t].q;
        qsq += qi * qi;
        qsum += qi;
      }
    }
    __syncthreads();
  }
  double acc[kRealNV] = {pacc.e,   pacc.vxx, pacc.vyy, pacc.vzz, pacc.vxy,
                         pacc.vxz, pacc.vyz, qsq,      qsum};
  block_sum<kRealNV>(acc, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < kRealNV; ++k) partial[k * pstride] = acc[k];
}

}  // namespace admp
