// Shared device helpers for the admp_amd HIP library (gfx950 / CDNA4).
//
// Conventions
//   * box h is row-major with the lattice vectors as rows; A = h^-1.
//   * fractional coordinates s = r A (row vector times matrix).
//   * mesh index of grid point (i0, i1, i2) is (i0 * K1 + i1) * K2 + i2.
//   * "real" is the bulk arithmetic type (float or double); every energy,
//     virial and charge reduction is accumulated in double.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/admp_amd.h"

namespace admp {

constexpr int kOrder = 6;              // order-6 cardinal B-spline only
constexpr int kBlock = 256;            // default work-group size (4 wavefronts)
constexpr int kRedBlocks = 512;        // fixed grid for deterministic reductions
constexpr double kPi = 3.14159265358979323846;
constexpr double kSqrtPi = 1.77245385090551602730;

struct Geom {
  double h[9];    // box rows
  double a[9];    // inverse box
  double vol;     // det(h)
  int k[3];       // mesh size
};

template <typename T> struct cplx_t;
template <> struct cplx_t<float> { using type = float2; };
template <> struct cplx_t<double> { using type = double2; };
template <typename T> using cplx = typename cplx_t<T>::type;

template <typename C> __device__ __forceinline__ C cmk(double re, double im) {
  C c; c.x = re; c.y = im; return c;
}
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { a.x += b.x; a.y += b.y; return a; }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { a.x -= b.x; a.y -= b.y; return a; }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
  C c; c.x = a.x * b.x - a.y * b.y; c.y = a.x * b.y + a.y * b.x; return c;
}

__host__ __device__ inline int pmod(long long m, int k) {
  long long r = m % k;
  return (int)(r < 0 ? r + k : r);
}

// 32-bit positive modulo (one integer division)
__host__ __device__ inline int pmod32(int m, int k) {
  const int r = m % k;
  return r < 0 ? r + k : r;
}

// m >= 0 reduced into [0, k) by subtraction (m is a few k at most)
__host__ __device__ inline int wrapk(int m, int k) {
  while (m >= k) m -= k;
  return m;
}

// Order-6 cardinal B-spline on the uniform knots by the Cox-de Boor
// recursion: w[j] = M6(f + j), dw[j] = M6'(f + j) for f in [0, 1), j = 0..5.
// M6'(u) = M5(u) - M5(u - 1).
template <typename T>
__device__ __forceinline__ void bspline6(T f, T w[6], T dw[6]) {
  T m[6];
  m[0] = f;
  m[1] = T(1) - f;
#pragma unroll
  for (int k = 3; k <= 5; ++k) {
    const T inv = T(1) / T(k - 1);
    T prev = T(0);
#pragma unroll
    for (int j = 0; j < k; ++j) {
      const T cur = (j < k - 1) ? m[j] : T(0);
      const T nv = ((f + T(j)) * cur + (T(k) - f - T(j)) * prev) * inv;
      prev = cur;
      m[j] = nv;
    }
  }
  // m[0..4] = M5(f + j)
  T prev = T(0);
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    const T cur = (j < 5) ? m[j] : T(0);
    dw[j] = cur - prev;
    w[j] = ((f + T(j)) * cur + (T(6) - f - T(j)) * prev) * T(0.2);
    prev = cur;
  }
}

// Weights and derivatives evaluated in float64 and rounded once to T.  The
// float32 recursion multiplies by rounded constants (1/3, 1/5), which biases
// every weight in the same direction and the reciprocal energy by ~1e-7.
template <typename T>
__device__ __forceinline__ void bspline6_rounded(double f, T w[6], T dw[6]) {
  double wd[6], dd[6];
  bspline6<double>(f, wd, dd);
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    w[j] = (T)wd[j];
    dw[j] = (T)dd[j];
  }
}

// Reference point of an atom on the mesh: R = K (r A); m0 = ceil(R);
// f = m0 - R in [0, 1).  Grid point m0 - 3 + j carries weight M6(f + j).
__device__ __forceinline__ void mesh_ref(const Geom& g, double x, double y, double z,
                                         int m0[3], double f[3]) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double s = x * g.a[0 * 3 + d] + y * g.a[1 * 3 + d] + z * g.a[2 * 3 + d];
    const double R = s * g.k[d];
    const double c = ceil(R);
    m0[d] = (int)c;
    f[d] = c - R;
  }
}

// ---- wavefront / block reductions (64-wide wavefronts, every lane active) ----
// DPP moves: lane i receives the lane that CTRL names; lanes of the rows (16
// lanes each) outside ROWS receive 0.
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_move(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xf, false);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_move(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = dpp_move<CTRL, ROWS>((int)b), hi = dpp_move<CTRL, ROWS>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_move(float v) {
  return __int_as_float(dpp_move<CTRL, ROWS>(__float_as_int(v)));
}

// Sum over each aligned group of 8 lanes, in every lane of the group: pairs
// (quad_perm 1,0,3,2), quads (2,3,0,1), then lane i adds lane 7 - i of its half
// row (row_half_mirror).
template <typename T>
__device__ __forceinline__ T group8_sum(T v) {
  v += dpp_move<0xb1, 0xf>(v);
  v += dpp_move<0x4e, 0xf>(v);
  v += dpp_move<0x141, 0xf>(v);
  return v;
}

// Sum over each aligned group of 16 lanes (one DPP row), in every lane of the
// row: the 8-lane sums, then lane i adds lane 15 - i (row_mirror).
template <typename T>
__device__ __forceinline__ T group16_sum(T v) {
  v = group8_sum(v);
  v += dpp_move<0x140, 0xf>(v);
  return v;
}

__device__ __forceinline__ double read_lane(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Work-groups are dealt to the 8 XCDs round-robin by index (each XCD has its
// own L2).  Renumbering them so that XCD x takes one contiguous range of
// logical indices keeps neighbouring work (rows of nearby cells) in one L2.
// A bijection of [0, nb).
constexpr int kNumXcd = 8;
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t nb) {
  const int64_t x = b % kNumXcd, per = nb / kNumXcd, rem = nb % kNumXcd;
  return x * per + (x < rem ? x : rem) + b / kNumXcd;
}

// Sum over the 64 lanes, returned in every lane, in a fixed order: pairs
// (quad_perm 1,0,3,2), quads (2,3,0,1), half rows (row_half_mirror), rows
// (row_mirror), then row_bcast:15 adds row 0 into row 1 and row 2 into row 3,
// and row_bcast:31 adds rows 0-1 into row 3, whose lane 63 holds the total.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_move<0xb1, 0xf>(v);
  v += dpp_move<0x4e, 0xf>(v);
  v += dpp_move<0x141, 0xf>(v);
  v += dpp_move<0x140, 0xf>(v);
  v += dpp_move<0x142, 0xa>(v);
  v += dpp_move<0x143, 0xc>(v);
  return read_lane(v, 63);
}

// Sum NV doubles over the block; result valid in thread 0.  `smem` holds
// NV * (blockDim/64) doubles.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* smem) {
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = wave_sum(v[i]);
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) smem[wid * NV + i] = v[i];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < nw; ++w) {
#pragma unroll
      for (int i = 0; i < NV; ++i) v[i] += smem[w * NV + i];
    }
  }
  __syncthreads();
}

// Deterministic reduction of per-block partials in a fixed order, one
// wavefront per value: value 0 -> first[0], values 1.. -> rest[v - 1]; either
// pointer may be null.  Launch with kReduceThreads threads.
constexpr int kReduceThreads = 1024;
template <int UNUSED = 0>
__global__ void __launch_bounds__(kReduceThreads) k_reduce_partials_t(
    const double* __restrict__ p, int nblocks, int nv, double scale, double* __restrict__ first,
    double* __restrict__ rest) {
  const int lane = threadIdx.x & 63;
  const int nw = blockDim.x >> 6;
  for (int v = threadIdx.x >> 6; v < nv; v += nw) {
    double s = 0.0;
    for (int bi = lane; bi < nblocks; bi += 64) s += p[(int64_t)bi * nv + v];
    s = wave_sum(s);
    if (lane == 0) {
      if (v == 0 && first) first[0] = scale * s;
      if (v > 0 && rest) rest[v - 1] = scale * s;
    }
  }
}
#define k_reduce_partials k_reduce_partials_t<0>

// Fill kernel (graph-safe replacement for memset).
template <typename T>
__global__ void k_fill(T* __restrict__ p, int64_t n, T v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Energies and dE/dh of one system from its reduced partial sums:
//   rr[0..17]  real space: E, strain derivative D (6), sum r (x) g (9), sum q^2, sum q
//   rk[0..9]   reciprocal: E, dE/dA through k (9)
//   rg[0..8]   gather: dE/dA through s = r A (9)
// energy[5] = real, reciprocal, self, non-neutral, total; dEdbox (may be null)
// = A^T (D - R^T g) - A^T dA A^T - (E_recip + E_nn) A^T.
__device__ inline void finalize_system(const double* rr, const double* rk, const double* rg,
                                       const Geom& g, double kappa, double eps, double* energy,
                                       double* dEdbox) {
  const double e_real = rr[0];
  const double e_recip = rk[0];
  const double e_self = -eps * kappa / kSqrtPi * rr[16];
  const double e_nn = -eps * kPi / (2.0 * g.vol * kappa * kappa) * rr[17] * rr[17];
  energy[0] = e_real;
  energy[1] = e_recip;
  energy[2] = e_self;
  energy[3] = e_nn;
  energy[4] = e_real + e_recip + e_self + e_nn;
  if (!dEdbox) return;
  const double* A = g.a;
  const double D[9] = {rr[1], rr[4], rr[5], rr[4], rr[2], rr[6], rr[5], rr[6], rr[3]};
  double M[9], dA[9];
  for (int i = 0; i < 9; ++i) {
    M[i] = D[i] - rr[7 + i];
    dA[i] = rk[1 + i] + rg[i];
  }
  for (int c = 0; c < 3; ++c)
    for (int b = 0; b < 3; ++b) {
      double gr = 0.0, gk = 0.0;
      for (int a = 0; a < 3; ++a) gr += A[a * 3 + c] * M[a * 3 + b];
      // -(A^T dA A^T)[c][b] = -sum_{a,e} A[a][c] dA[a][e] A[b][e]
      for (int a = 0; a < 3; ++a)
        for (int e = 0; e < 3; ++e) gk -= A[a * 3 + c] * dA[a * 3 + e] * A[b * 3 + e];
      dEdbox[c * 3 + b] = gr + gk - (e_recip + e_nn) * A[b * 3 + c];
    }
}

}  // namespace admp
