// Device building blocks of the LDS Stockham FFT and the PME influence
// function, shared by the single-system and batched kernels.
#pragma once

#include "admp_common.hpp"
#include "pme_kernels.hpp"

namespace admp {

// Work-group size of every kernel that runs the out-of-place Stockham
// passes and the line bodies below (r2c_rows, c2r_rows, strided_lines,
// x_convolve_lines): all of them are launched with kBlock threads.  A
// constant stride lets their loops run without reading the work-group size,
// which the compiler loads from the dispatch's hidden arguments with a vector
// memory load before each loop.
constexpr int kFftThreads = kBlock;


// ---------------------------------------------------------------------------
// Stockham FFT on LDS.  C interleaved lines of length L, element e of line c
// at buf[e * C + c].  One pass of radix r (p = product of the earlier radices,
// n = L / r): for butterfly i, k = i mod p,
//   u_j = x[i + j n] * w^(sign k j / (p r)),  U = DFT_r(u),  y[(i - k) r + k + j p] = U_j.
// tw[t] = exp(-2 pi i t / L).

template <typename C> __device__ __forceinline__ C mul_i(C a, int sign) {
  // a * (sign * i)
  C r;
  if (sign > 0) { r.x = -a.y; r.y = a.x; }
  else { r.x = a.y; r.y = -a.x; }
  return r;
}

template <typename T, int SIGN, int R>
__device__ __forceinline__ void dft_r(cplx<T>* u) {
  using C = cplx<T>;
  if constexpr (R == 2) {
    const C a = u[0], b = u[1];
    u[0] = cadd(a, b);
    u[1] = csub(a, b);
  } else if constexpr (R == 3) {
    const T h = T(0.86602540378443864676);
    const C s = cadd(u[1], u[2]);
    const C d = csub(u[1], u[2]);
    C m;
    m.x = u[0].x - T(0.5) * s.x;
    m.y = u[0].y - T(0.5) * s.y;
    C t = mul_i(d, SIGN);
    t.x *= h;
    t.y *= h;
    u[0] = cadd(u[0], s);
    u[1] = cadd(m, t);
    u[2] = csub(m, t);
  } else if constexpr (R == 4) {
    const C t0 = cadd(u[0], u[2]), t1 = csub(u[0], u[2]);
    const C t2 = cadd(u[1], u[3]), t3 = mul_i(csub(u[1], u[3]), SIGN);
    u[0] = cadd(t0, t2);
    u[2] = csub(t0, t2);
    u[1] = cadd(t1, t3);
    u[3] = csub(t1, t3);
  } else if constexpr (R == 5) {
    const T c1 = T(0.30901699437494742410), c2 = T(-0.80901699437494742410);
    const T s1 = T(0.95105651629515357212), s2 = T(0.58778525229247312917);
    const C a1 = cadd(u[1], u[4]), b1 = csub(u[1], u[4]);
    const C a2 = cadd(u[2], u[3]), b2 = csub(u[2], u[3]);
    C p1, p2, v1, v2;
    p1.x = u[0].x + c1 * a1.x + c2 * a2.x;
    p1.y = u[0].y + c1 * a1.y + c2 * a2.y;
    p2.x = u[0].x + c2 * a1.x + c1 * a2.x;
    p2.y = u[0].y + c2 * a1.y + c1 * a2.y;
    v1.x = s1 * b1.x + s2 * b2.x;
    v1.y = s1 * b1.y + s2 * b2.y;
    v2.x = s2 * b1.x - s1 * b2.x;
    v2.y = s2 * b1.y - s1 * b2.y;
    const C q1 = mul_i(v1, SIGN);
    const C q2 = mul_i(v2, SIGN);
    u[0] = cadd(u[0], cadd(a1, a2));
    u[1] = cadd(p1, q1);
    u[4] = csub(p1, q1);
    u[2] = cadd(p2, q2);
    u[3] = csub(p2, q2);
  } else if constexpr (R == 8) {
    C e[4] = {u[0], u[2], u[4], u[6]};
    C o[4] = {u[1], u[3], u[5], u[7]};
    dft_r<T, SIGN, 4>(e);
    dft_r<T, SIGN, 4>(o);
    const T h = T(0.70710678118654752440);
    C w1, w3;  // exp(SIGN 2 pi i k / 8) for k = 1, 3
    w1.x = h;
    w1.y = SIGN * h;
    w3.x = -h;
    w3.y = SIGN * h;
    const C o1 = cmul(o[1], w1);
    const C o2 = mul_i(o[2], SIGN);
    const C o3 = cmul(o[3], w3);
    u[0] = cadd(e[0], o[0]);
    u[4] = csub(e[0], o[0]);
    u[1] = cadd(e[1], o1);
    u[5] = csub(e[1], o1);
    u[2] = cadd(e[2], o2);
    u[6] = csub(e[2], o2);
    u[3] = cadd(e[3], o3);
    u[7] = csub(e[3], o3);
  }
}

template <typename T>
__device__ __forceinline__ cplx<T> tw_get(const cplx<T>* tw, int t, int sign) {
  cplx<T> w = tw[t];
  if (sign > 0) w.y = -w.y;
  return w;
}

// q = x / d, r = x mod d with the multiplier m = div_magic(d) (x, d < 2^16;
// m = 0 for d = 1): one high multiply instead of an integer division.
__device__ __forceinline__ void divmod_m(int x, int d, uint32_t m, int& q, int& r) {
  q = m ? (int)__umulhi((uint32_t)x, m) : x;
  r = x - q * d;
}

// q = x / d, r = x mod d (x >= 0, d > 0 and the same in every lane): shifts
// when d is a power of two (the usual mesh sizes), else a division.
__device__ __forceinline__ void divmod_u(int x, int d, int& q, int& r) {
  if ((d & (d - 1)) == 0) {
    const int sh = __ffs(d) - 1;
    q = x >> sh;
    r = x & (d - 1);
  } else {
    q = x / d;
    r = x - q * d;
  }
}

// One Stockham pass with a radix known at compile time (registers only).
template <typename T, int SIGN, int R>
__device__ __forceinline__ void pass_r(const cplx<T>* __restrict__ a, cplx<T>* __restrict__ b,
                                       int L, int C, int p, uint32_t mp, const cplx<T>* tw) {
  using CT = cplx<T>;
  const int nb = L / R;
  const int tstep = L / (p * R);
  for (int idx = threadIdx.x; idx < nb * C; idx += kFftThreads) {
    int c, i, k, ip;
    divmod_u(idx, C, i, c);
    divmod_m(i, p, mp, ip, k);
    CT u[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
      CT v = a[(i + j * nb) * C + c];
      // k j tstep <= (p - 1)(R - 1) L / (p R) < L
      if (j > 0 && k > 0) v = cmul(v, tw_get<T>(tw, k * j * tstep, SIGN));
      u[j] = v;
    }
    dft_r<T, SIGN, R>(u);
    const int base = (i - k) * R + k;
#pragma unroll
    for (int j = 0; j < R; ++j) b[(base + j * p) * C + c] = u[j];
  }
}

// Generic odd radix r (the primes >= 7 that factorize leaves): inputs j and
// r - j enter output q as x_j w^(jq) + x_(r-j) w^(-jq) = c a_j + i s b_j with
// a_j = x_j + x_(r-j), b_j = x_j - x_(r-j) and w^(jq) = c + i s, and output
// r - q as c a_j - i s b_j.  A thread takes kGenOut values q0.. of one
// butterfly and produces the outputs q and r - q of each (plus output 0 for
// the first group): per pair (j, r - j) two LDS loads, the inter-pass
// twiddles, and per q two real-times-complex products into A_q and B_q,
// w^(j(q+1)) from w^(jq) by one product with w^j.  Against a direct sum over
// all r inputs per output this is half the loads and a quarter of the
// multiplies.  Twiddle indices advance by additions (k j tstep and j rstep
// stay below L; j q0 rstep wraps once at most per step).
constexpr int kGenOut = 4;

template <typename T, int SIGN>
__device__ void pass_generic(const cplx<T>* __restrict__ a, cplx<T>* __restrict__ b, int L,
                             int C, int p, int r, uint32_t mp, uint32_t mnb,
                             const cplx<T>* tw) {
  using CT = cplx<T>;
  const int nb = L / r;
  const int tstep = L / (p * r);
  const int rstep = L / r;
  const int h = (r - 1) / 2;
  const int ng = (h + kGenOut - 1) / kGenOut;
  for (int idx = threadIdx.x; idx < nb * ng * C; idx += kFftThreads) {
    int c, rest, i, g, k, ip;
    divmod_u(idx, C, rest, c);
    divmod_m(rest, nb, mnb, g, i);
    divmod_m(i, p, mp, ip, k);
    const int q0 = 1 + g * kGenOut;
    const int nq = min(kGenOut, h + 1 - q0);
    const int dk = k * tstep;      // twiddle index of input j is k j tstep
    const int tr = k * r * tstep;  // k L / p < L; input r - j takes tr - k j tstep
    const int dq = q0 * rstep;     // < L
    const CT x0 = a[i * C + c];
    CT A[kGenOut], B[kGenOut];
#pragma unroll
    for (int m = 0; m < kGenOut; ++m) {
      A[m] = cmk<CT>(0, 0);
      B[m] = cmk<CT>(0, 0);
    }
    CT s0 = cmk<CT>(0, 0);
    int tj = 0, tq = 0, tf = 0;
    for (int j = 1; j <= h; ++j) {
      tj += dk;
      tq += dq;
      if (tq >= L) tq -= L;
      tf += rstep;
      CT xj = a[(i + j * nb) * C + c];
      CT xm = a[(i + (r - j) * nb) * C + c];
      if (k > 0) {
        xj = cmul(xj, tw_get<T>(tw, tj, SIGN));
        xm = cmul(xm, tw_get<T>(tw, tr - tj, SIGN));
      }
      const CT aj = cadd(xj, xm), bj = csub(xj, xm);
      s0 = cadd(s0, aj);
      CT w = tw_get<T>(tw, tq, SIGN);
      const CT f = tw_get<T>(tw, tf, SIGN);
#pragma unroll
      for (int m = 0; m < kGenOut; ++m) {
        A[m].x += w.x * aj.x;
        A[m].y += w.x * aj.y;
        B[m].x += w.y * bj.x;
        B[m].y += w.y * bj.y;
        if (m + 1 < kGenOut) w = cmul(w, f);
      }
    }
    const int base = (i - k) * r + k;
    if (g == 0) b[base * C + c] = cadd(x0, s0);
#pragma unroll
    for (int m = 0; m < kGenOut; ++m)
      if (m < nq) {
        const int q = q0 + m;
        CT u, v;  // x0 + A +- i B
        u.x = x0.x + A[m].x - B[m].y;
        u.y = x0.y + A[m].y + B[m].x;
        v.x = x0.x + A[m].x + B[m].y;
        v.y = x0.y + A[m].y - B[m].x;
        b[(base + q * p) * C + c] = u;
        b[(base + (r - q) * p) * C + c] = v;
      }
  }
}

template <typename T, int SIGN>
__device__ __forceinline__ cplx<T>* stockham(cplx<T>* a, cplx<T>* b, int L, int C, const Radix& rd,
                             const cplx<T>* tw) {
  int p = 1;
  for (int s = 0; s < rd.n; ++s) {
    const int r = rd.r[s];
    const uint32_t mp = rd.mp[s];
    switch (r) {
      case 2: pass_r<T, SIGN, 2>(a, b, L, C, p, mp, tw); break;
      case 3: pass_r<T, SIGN, 3>(a, b, L, C, p, mp, tw); break;
      case 4: pass_r<T, SIGN, 4>(a, b, L, C, p, mp, tw); break;
      case 5: pass_r<T, SIGN, 5>(a, b, L, C, p, mp, tw); break;
      case 8: pass_r<T, SIGN, 8>(a, b, L, C, p, mp, tw); break;
      default: pass_generic<T, SIGN>(a, b, L, C, p, r, mp, rd.mnb[s], tw); break;
    }
    __syncthreads();
    cplx<T>* t = a;
    a = b;
    b = t;
    p *= r;
  }
  return a;
}

// nt: the work-group's thread count (kFftThreads in the line passes).
template <typename T>
__device__ void build_twiddles(cplx<T>* tw, int L, int nt) {
  for (int t = threadIdx.x; t < L; t += nt) {
    double s, c;
    sincospi(2.0 * t / L, &s, &c);
    tw[t].x = (T)c;
    tw[t].y = (T)(-s);
  }
}

// ---------------------------------------------------------------------------
// Pass 3: along x, forward FFT, influence function, energy and dE/dA
// reduction, multiply, inverse FFT.  Block (chunk, y) holds K0 x C elements.
struct InflParams {
  Geom g;
  double kappa;
  double phi_scale;  // multiplies 2 * eps * G
  double eps;
  int conv;          // ADMP_KCONV_*
  int ortho;         // box diagonal: |k| of the Hermitian partner equals |k|
};

__device__ __forceinline__ double recip(double x) { return 1.0 / x; }
__device__ __forceinline__ float recip(float x) { return __builtin_amdgcn_rcpf(x); }  // 1 ulp

__device__ __forceinline__ int fft_m(int idx, int K) { return idx < (K + 1) / 2 ? idx : idx - K; }

__device__ __forceinline__ double theta1(int m, int K) {
  // sum_{j=-2..2} M6(j + 3) cos(2 pi j m / K) with M6(1..5) = 1, 26, 66, 26, 1 (/120)
  const double c1 = cospi(2.0 * m / K), c2 = cospi(4.0 * m / K);
  return (66.0 + 52.0 * c1 + 2.0 * c2) / 120.0;
}

// k vector of integer index m: rows (k = 2 pi m A) or columns (k = 2 pi A m)
__device__ __forceinline__ void kvec(const InflParams& P, int mx, int my, int mz, double kv[3]) {
  const double m[3] = {(double)mx, (double)my, (double)mz};
  const double* A = P.g.a;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    double s = 0.0;
    if (P.conv == ADMP_KCONV_REFERENCE) {
#pragma unroll
      for (int i = 0; i < 3; ++i) s += m[i] * A[i * 3 + j];
    } else {
#pragma unroll
      for (int i = 0; i < 3; ++i) s += A[j * 3 + i] * m[i];
    }
    kv[j] = 2.0 * kPi * s;
  }
}

// G = 2 pi / V exp(-k^2 / 4 kappa^2) / (k^2 theta^2) and g' = (dG/dk_j) / k_j
// = -2 G (1/(4 kappa^2) + 1/k^2); zero at k = 0.
__device__ __forceinline__ double gfun(const InflParams& P, const double kv[3], double th,
                                       double* gp) {
  const double k2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
  if (k2 == 0.0) {
    *gp = 0.0;
    return 0.0;
  }
  const double inv4k2 = 0.25 / (P.kappa * P.kappa);
  const double G = 2.0 * kPi / P.g.vol * exp(-k2 * inv4k2) / (k2 * th * th);
  *gp = -2.0 * G * (inv4k2 + 1.0 / k2);
  return G;
}

__device__ __forceinline__ void add_dEdA(double acc[10], const InflParams& P, int mx, int my,
                                         int mz, const double kv[3], double gp, double wS2) {
  // dE/dk_j = wS2 * gp * k_j ;  rows: dE/dA[i][j] += dE/dk_j 2 pi m_i
  //                             columns: dE/dA[j][i] += dE/dk_j 2 pi m_i
  const double m[3] = {(double)mx, (double)my, (double)mz};
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double v = wS2 * gp * kv[j] * 2.0 * kPi * m[i];
      if (P.conv == ADMP_KCONV_REFERENCE) acc[1 + i * 3 + j] += v;
      else acc[1 + j * 3 + i] += v;
    }
}


// ---------------------------------------------------------------------------
// Block bodies of the five FFT passes.  Each runs one work-group's lines with
// the Stockham buffers in dynamic LDS: twiddles (L) then two L x C buffers.

// Pass 1: real-to-complex along z for rows [row0, row0 + 2C).  Rows 2c and
// 2c+1 are packed as the real and imaginary parts of one complex line; after
// the FFT A_k = (Z_k + conj Z_{L-k}) / 2 and B_k = (Z_k - conj Z_{L-k}) / (2i).
// load(row, z) supplies the real input of row `row`, element z.
// kz_major: store element (row, k) at out[k * nrows + row] (planes of one k
// contiguous) instead of out[row * pitch + k] (pitch 0: L / 2 + 1).
template <typename T, typename Load>
__device__ __forceinline__ void r2c_rows(Load load, cplx<T>* __restrict__ out, int64_t nrows,
                                         int64_t row0, int L, int C, const Radix& rd,
                                         bool kz_major = false, int pitch = 0) {
  using CT = cplx<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  CT* tw = reinterpret_cast<CT*>(smem_raw);
  CT* a = tw + L;
  CT* b = a + (size_t)L * C;
  build_twiddles<T>(tw, L, kFftThreads);
  for (int idx = threadIdx.x; idx < 2 * C * L; idx += kFftThreads) {
    int rl, z;
    divmod_m(idx, L, rd.mlen, rl, z);
    const int64_t row = row0 + rl;
    const T v = row < nrows ? load(row, z) : T(0);
    T* dst = reinterpret_cast<T*>(&a[z * C + (rl >> 1)]);
    dst[rl & 1] = v;
  }
  __syncthreads();
  CT* res = stockham<T, -1>(a, b, L, C, rd, tw);
  const int Lh = L / 2 + 1;
  for (int idx = threadIdx.x; idx < 2 * C * Lh; idx += kFftThreads) {
    // consecutive threads take consecutive k (row-major store) or rows (kz-major)
    int rl, k;
    if (kz_major) divmod_u(idx, 2 * C, k, rl);
    else divmod_m(idx, Lh, rd.mhalf, rl, k);
    const int64_t row = row0 + rl;
    if (row >= nrows) continue;
    const CT zk = res[k * C + (rl >> 1)];
    const CT zm = res[(k ? L - k : 0) * C + (rl >> 1)];
    CT o;
    if ((rl & 1) == 0) {
      o.x = T(0.5) * (zk.x + zm.x);
      o.y = T(0.5) * (zk.y - zm.y);
    } else {
      o.x = T(0.5) * (zk.y + zm.y);
      o.y = T(-0.5) * (zk.x - zm.x);
    }
    out[kz_major ? (int64_t)k * nrows + row : row * (pitch ? pitch : Lh) + k] = o;
  }
}

// Pass 5: complex-to-real along z (unnormalised inverse) for rows
// [row0, row0 + 2C).  Z_k = A_k + i B_k with A, B extended by Hermitian
// symmetry; the inverse transform gives z = a + i b.
template <typename T>
__device__ __forceinline__ void c2r_rows(const cplx<T>* __restrict__ in, T* __restrict__ out,
                                         int64_t nrows, int64_t row0, int L, int C,
                                         const Radix& rd, bool kz_major = false, int pitch = 0) {
  using CT = cplx<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  CT* tw = reinterpret_cast<CT*>(smem_raw);
  CT* a = tw + L;
  CT* b = a + (size_t)L * C;
  build_twiddles<T>(tw, L, kFftThreads);
  const int Lh = L / 2 + 1, Lp = pitch ? pitch : Lh;
  for (int idx = threadIdx.x; idx < C * L; idx += kFftThreads) {
    int c, k;
    if (kz_major) divmod_u(idx, C, k, c);
    else divmod_m(idx, L, rd.mlen, c, k);
    const int64_t ra = row0 + 2 * c, rb = ra + 1;
    const int kk = k < Lh ? k : L - k;
    const T sg = k < Lh ? T(1) : T(-1);
    CT A = cmk<CT>(0, 0), B = cmk<CT>(0, 0);
    if (ra < nrows) A = in[kz_major ? (int64_t)kk * nrows + ra : ra * Lp + kk];
    if (rb < nrows) B = in[kz_major ? (int64_t)kk * nrows + rb : rb * Lp + kk];
    A.y *= sg;
    B.y *= sg;
    CT z;
    z.x = A.x - B.y;
    z.y = A.y + B.x;
    a[k * C + c] = z;
  }
  __syncthreads();
  CT* res = stockham<T, 1>(a, b, L, C, rd, tw);
  for (int idx = threadIdx.x; idx < 2 * C * L; idx += kFftThreads) {
    int rl, z;
    divmod_m(idx, L, rd.mlen, rl, z);
    const int64_t row = row0 + rl;
    if (row >= nrows) continue;
    const T* src = reinterpret_cast<const T*>(&res[z * C + (rl >> 1)]);
    out[row * L + z] = src[rl & 1];
  }
}

// Passes 2 and 4: cmax (<= C) complex lines of length L along a strided axis,
// starting at base; line elements are `stride` apart, lines are contiguous.
template <typename T, int SIGN>
__device__ __forceinline__ void strided_lines(cplx<T>* __restrict__ base, int L, int64_t stride,
                                              int cmax, int C, const Radix& rd) {
  using CT = cplx<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  CT* tw = reinterpret_cast<CT*>(smem_raw);
  CT* a = tw + L;
  CT* b = a + (size_t)L * C;
  build_twiddles<T>(tw, L, kFftThreads);
  for (int idx = threadIdx.x; idx < L * C; idx += kFftThreads) {
    int e, c;
    divmod_u(idx, C, e, c);
    a[idx] = c < cmax ? base[e * stride + c] : cmk<CT>(0, 0);
  }
  __syncthreads();
  CT* res = stockham<T, SIGN>(a, b, L, C, rd, tw);
  for (int idx = threadIdx.x; idx < L * C; idx += kFftThreads) {
    int e, c;
    divmod_u(idx, C, e, c);
    if (c < cmax) base[e * stride + c] = res[idx];
  }
}

// Pass 3: forward FFT along x of the lines (iy, z0 .. z0 + C), the
// influence-function product and the inverse FFT, in LDS.  Accumulates
// E = sum_m w |S(m)|^2 Gsym(m) and dE/dA (10 values, without the factor eps)
// into partial_dst when it is not null.  Gsym = (G(m) + G(m')) / 2 with m'
// the index of -m, which makes the result exact for either k convention; the
// half-spectrum weight w is 2 for interior kz.
// pitch: elements between consecutive (x, y) rows of the spectrum (0: K2/2 + 1).
template <typename T>
__device__ __forceinline__ void x_convolve_lines(cplx<T>* __restrict__ data, int iy, int z0,
                                                 int C, const Radix& rd, const InflParams& P,
                                                 double* __restrict__ partial_dst, int pitch = 0) {
  using CT = cplx<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ double red[10 * (kBlock / 64)];
  const int K0 = P.g.k[0], K1 = P.g.k[1], K2 = P.g.k[2];
  const int Kzh = K2 / 2 + 1, Kzp = pitch ? pitch : Kzh;
  const int L = K0;
  const int64_t stride = (int64_t)K1 * Kzp;
  CT* tw = reinterpret_cast<CT*>(smem_raw);
  CT* a = tw + L;
  CT* b = a + (size_t)L * C;
  build_twiddles<T>(tw, L, kFftThreads);
  CT* base = data + (int64_t)iy * Kzp + z0;
  const int cmax = min(C, Kzh - z0);
  for (int idx = threadIdx.x; idx < L * C; idx += kFftThreads) {
    int e, c;
    divmod_u(idx, C, e, c);
    a[idx] = c < cmax ? base[e * stride + c] : cmk<CT>(0, 0);
  }
  __syncthreads();
  CT* res = stockham<T, -1>(a, b, L, C, rd, tw);
  CT* oth = (res == a) ? b : a;
  // B-spline moduli theta(m) = prod_d theta_d(m_d) from per-dimension factors
  double* thx = reinterpret_cast<double*>(oth);  // the idle Stockham buffer holds the x factors
  for (int ix = threadIdx.x; ix < L; ix += kFftThreads) thx[ix] = theta1(fft_m(ix, K0), K0);
  __syncthreads();
  double acc[10];
#pragma unroll
  for (int i = 0; i < 10; ++i) acc[i] = 0.0;
  const int my = fft_m(iy, K1);
  const int myp = fft_m((K1 - iy) % K1, K1);
  const double thy = theta1(my, K1);
  for (int idx = threadIdx.x; idx < L * C; idx += kFftThreads) {
    int ix, c;
    divmod_u(idx, C, ix, c);
    if (c >= cmax) continue;
    const int iz = z0 + c;
    const int mx = fft_m(ix, K0), mz = fft_m(iz, K2);
    const int mxp = fft_m((K0 - ix) % K0, K0), mzp = fft_m((K2 - iz) % K2, K2);
    const double th = thx[ix] * thy * theta1(mz, K2);
    double kv[3], kvp[3], gp, gpp;
    kvec(P, mx, my, mz, kv);
    kvec(P, mxp, myp, mzp, kvp);
    const double G = gfun(P, kv, th, &gp);
    double Gp;
    if (P.ortho) {
      Gp = G;  // |k(m')| = |k(m)| for a diagonal box
      gpp = gp;
    } else {
      Gp = gfun(P, kvp, th, &gpp);
    }
    const double Gs = 0.5 * (G + Gp);
    CT s = res[idx];
    if (partial_dst) {  // energy and dE/dA sums only when a caller reads them
      const double w = (iz > 0 && 2 * iz != K2) ? 2.0 : 1.0;
      const double S2 = (double)s.x * (double)s.x + (double)s.y * (double)s.y;
      acc[0] += w * Gs * S2;
      add_dEdA(acc, P, mx, my, mz, kv, gp, 0.5 * w * S2);
      add_dEdA(acc, P, mxp, myp, mzp, kvp, gpp, 0.5 * w * S2);
    }
    const T f = (T)(2.0 * P.eps * P.phi_scale * Gs);
    s.x *= f;
    s.y *= f;
    res[idx] = s;
  }
  __syncthreads();
  CT* out = stockham<T, 1>(res, oth, L, C, rd, tw);
  for (int idx = threadIdx.x; idx < L * C; idx += kFftThreads) {
    int e, c;
    divmod_u(idx, C, e, c);
    if (c < cmax) base[e * stride + c] = out[idx];
  }
  if (partial_dst) {
    block_sum<10>(acc, red);
    if (threadIdx.x == 0)
      for (int i = 0; i < 10; ++i) partial_dst[i] = acc[i];
  }
}


// ---------------------------------------------------------------------------
// One kz plane of the half spectrum held whole in LDS (element (x, y) at
// buf[y * pitch + x], pitch = K0 + 1 so that lines along y and along x both
// fall on distinct banks).  Forward FFT along y and x, influence function,
// inverse along x and y, in place: each Stockham pass first reads the inputs
// of a thread's butterflies into registers, the block synchronises, then the
// outputs are written.
constexpr int kPlaneThreads = 512;  // measured faster than 1024 and 256 at 64^2
constexpr int kPlaneB = 1;  // butterflies per thread and pass
constexpr int kPlaneLoad = 8;  // plane elements per thread (K0 K1 <= kPlaneLoad kPlaneThreads)

template <typename T, int SIGN, int R>
__device__ __forceinline__ void pass_inplace(cplx<T>* buf, int L, int C, int es, int cs, int p,
                                             const cplx<T>* tw) {
  using CT = cplx<T>;
  const int nb = L / R;
  const int tstep = L / (p * R);
  const int nbf = nb * C;
  CT u[kPlaneB][R];
#pragma unroll
  for (int b = 0; b < kPlaneB; ++b) {
    const int idx = threadIdx.x + b * blockDim.x;
    if (idx < nbf) {
      int c, i, k, ip;
      divmod_u(idx, C, i, c);
      divmod_u(i, p, ip, k);
#pragma unroll
      for (int j = 0; j < R; ++j) {
        CT v = buf[(i + j * nb) * es + c * cs];
        if (j > 0 && k > 0) v = cmul(v, tw_get<T>(tw, k * j * tstep, SIGN));  // < L
        u[b][j] = v;
      }
      dft_r<T, SIGN, R>(u[b]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < kPlaneB; ++b) {
    const int idx = threadIdx.x + b * blockDim.x;
    if (idx < nbf) {
      int c, i, k, ip;
      divmod_u(idx, C, i, c);
      divmod_u(i, p, ip, k);
      const int base = (i - k) * R + k;
#pragma unroll
      for (int j = 0; j < R; ++j) buf[(base + j * p) * es + c * cs] = u[b][j];
    }
  }
  __syncthreads();
}

// In-place pass with up to B butterflies per thread: all reads into
// registers, a barrier, then all writes.
template <typename T, int SIGN, int R, int B>
__device__ __forceinline__ void pass_ip(cplx<T>* buf, int L, int C, int es, int cs, int p,
                                        const cplx<T>* tw) {
  using CT = cplx<T>;
  const int nb = L / R;
  const int tstep = L / (p * R);
  const int nbf = nb * C;
  CT u[B][R];
#pragma unroll
  for (int b = 0; b < B; ++b) {
    const int idx = threadIdx.x + b * blockDim.x;
    if (idx < nbf) {
      int c, i, k, ip;
      divmod_u(idx, C, i, c);
      divmod_u(i, p, ip, k);
#pragma unroll
      for (int j = 0; j < R; ++j) {
        CT v = buf[(i + j * nb) * es + c * cs];
        if (j > 0 && k > 0) v = cmul(v, tw_get<T>(tw, k * j * tstep, SIGN));  // < L
        u[b][j] = v;
      }
      dft_r<T, SIGN, R>(u[b]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < B; ++b) {
    const int idx = threadIdx.x + b * blockDim.x;
    if (idx < nbf) {
      int c, i, k, ip;
      divmod_u(idx, C, i, c);
      divmod_u(i, p, ip, k);
      const int base = (i - k) * R + k;
#pragma unroll
      for (int j = 0; j < R; ++j) buf[(base + j * p) * es + c * cs] = u[b][j];
    }
  }
  __syncthreads();
}

// Butterflies per thread and pass that pass_ip_any handles for radix R (at
// most 16 complex values in registers); the plan checks (L / R) C against it.
__host__ __device__ constexpr int ip_max_b(int R) {
  return R == 8 ? 2 : (R == 5 || R == 4) ? 3 : R == 3 ? 4 : 5;
}

template <typename T, int SIGN, int R>
__device__ __forceinline__ void pass_ip_any(cplx<T>* buf, int L, int C, int es, int cs, int p,
                                            const cplx<T>* tw) {
  constexpr int kMax = ip_max_b(R);
  const int bpt = ((L / R) * C + (int)blockDim.x - 1) / (int)blockDim.x;
  if (bpt <= 1) pass_ip<T, SIGN, R, 1>(buf, L, C, es, cs, p, tw);
  else if (bpt == 2 || kMax == 2) pass_ip<T, SIGN, R, 2>(buf, L, C, es, cs, p, tw);
  else if (bpt == 3 || kMax == 3) pass_ip<T, SIGN, R, (kMax >= 3 ? 3 : 2)>(buf, L, C, es, cs, p, tw);
  else if (bpt == 4 || kMax == 4) pass_ip<T, SIGN, R, (kMax >= 4 ? 4 : 2)>(buf, L, C, es, cs, p, tw);
  else pass_ip<T, SIGN, R, kMax>(buf, L, C, es, cs, p, tw);
}

// C lines of length L in place (element stride es, line stride cs); radices
// 2, 3, 4, 5, 8 and at most ip_max_b(r) butterflies per thread in every pass.
template <typename T, int SIGN>
__device__ __forceinline__ void stockham_ip(cplx<T>* buf, int L, int C, int es, int cs,
                                            const Radix& rd, const cplx<T>* tw) {
  int p = 1;
  for (int s = 0; s < rd.n; ++s) {
    const int r = rd.r[s];
    switch (r) {
      case 2: pass_ip_any<T, SIGN, 2>(buf, L, C, es, cs, p, tw); break;
      case 3: pass_ip_any<T, SIGN, 3>(buf, L, C, es, cs, p, tw); break;
      case 4: pass_ip_any<T, SIGN, 4>(buf, L, C, es, cs, p, tw); break;
      case 5: pass_ip_any<T, SIGN, 5>(buf, L, C, es, cs, p, tw); break;
      default: pass_ip_any<T, SIGN, 8>(buf, L, C, es, cs, p, tw); break;
    }
    p *= r;
  }
}

// C lines of length L, element stride es, line stride cs; radices 2, 3, 4, 5,
// 8 with (L / r) * C <= kPlaneB * blockDim.x (checked when planning).
template <typename T, int SIGN>
__device__ __forceinline__ void stockham_inplace(cplx<T>* buf, int L, int C, int es, int cs, const Radix& rd,
                                 const cplx<T>* tw) {
  int p = 1;
  for (int s = 0; s < rd.n; ++s) {
    const int r = rd.r[s];
    switch (r) {
      case 2: pass_inplace<T, SIGN, 2>(buf, L, C, es, cs, p, tw); break;
      case 3: pass_inplace<T, SIGN, 3>(buf, L, C, es, cs, p, tw); break;
      case 4: pass_inplace<T, SIGN, 4>(buf, L, C, es, cs, p, tw); break;
      case 5: pass_inplace<T, SIGN, 5>(buf, L, C, es, cs, p, tw); break;
      default: pass_inplace<T, SIGN, 8>(buf, L, C, es, cs, p, tw); break;
    }
    p *= r;
  }
}

// Bytes of dynamic LDS the plane body needs.
__host__ __device__ inline size_t plane_lds_bytes(int K0, int K1, size_t celem) {
  return (size_t)K1 * (K0 + 1) * celem + (size_t)(K0 + K1) * celem +
         (size_t)3 * (K0 + K1) * sizeof(double);
}

// The plane kz of the kz-major half spectrum `spec` (element (x, y, kz) at
// spec[(kz * K0 + x) * K1 + y]): forward FFT along y and x, the same
// influence-function product, energy and dE/dA sums as x_convolve_lines,
// inverse FFT along x and y.  For a diagonal box exp(-k^2 / 4 kappa^2) / theta^2
// factors over the three axes and is tabulated per axis.
template <typename T>
__device__ __forceinline__ void plane_convolve(cplx<T>* __restrict__ spec, int kz,
                                               const Radix& rx, const Radix& ry,
                                               const InflParams& P,
                                               double* __restrict__ partial_dst) {
  using CT = cplx<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ double red[10 * (kPlaneThreads / 64)];
  const int K0 = P.g.k[0], K1 = P.g.k[1], K2 = P.g.k[2];
  const int pitch = K0 + 1;
  CT* buf = reinterpret_cast<CT*>(smem_raw);
  CT* twy = buf + (size_t)K1 * pitch;
  CT* twx = twy + K1;
  double* thx = reinterpret_cast<double*>(twx + K0);
  double* thy = thx + K0;
  T* exx = reinterpret_cast<T*>(thy + K1);  // diagonal box: exp(-k_x^2 / 4 kappa^2) / theta_x^2
  T* exy = exx + K0;
  T* kvx = exy + K1;  // diagonal box: k_x
  T* kvy = kvx + K0;
  CT* plane = spec + (int64_t)kz * K0 * K1;
  const int ne = K0 * K1;
  // global loads first (into registers), the tables while they are in flight
  CT ld[kPlaneLoad];
#pragma unroll
  for (int k = 0; k < kPlaneLoad; ++k) {
    const int e = threadIdx.x + k * blockDim.x;
    if (e < ne) ld[k] = plane[e];
  }
  const double inv4k2 = 0.25 / (P.kappa * P.kappa);
  // tasks: twiddles along y, along x, x tables, y tables (2 (K0 + K1) <= blockDim.x)
  {
    const int t = threadIdx.x;
    if (t < K1 + K0) {
      const int L = t < K1 ? K1 : K0, i = t < K1 ? t : t - K1;
      double sn, cs;
      sincospi(2.0 * i / L, &sn, &cs);
      CT w;
      w.x = (T)cs;
      w.y = (T)(-sn);
      (t < K1 ? twy : twx)[i] = w;
    } else if (t < 2 * K1 + 2 * K0) {
      const bool dx = t < K1 + 2 * K0;  // x tables at [K1 + K0, K1 + 2 K0), then y
      const int i = dx ? t - (K1 + K0) : t - (K1 + 2 * K0);
      const int K = dx ? K0 : K1;
      const double th = theta1(fft_m(i, K), K);
      const double k = 2.0 * kPi * fft_m(i, K) * P.g.a[dx ? 0 : 4];
      (dx ? thx : thy)[i] = th;
      (dx ? kvx : kvy)[i] = (T)k;
      (dx ? exx : exy)[i] = (T)(exp(-k * k * inv4k2) / (th * th));
    }
  }
  // the plane's own factors along z, also while the loads are in flight
  const int mz = fft_m(kz, K2), mzp = fft_m((K2 - kz) % K2, K2);
  const double thz = theta1(mz, K2);
  const double kz_v = 2.0 * kPi * mz * P.g.a[8];
  const double exz = exp(-kz_v * kz_v * inv4k2) / (thz * thz);
#pragma unroll
  for (int k = 0; k < kPlaneLoad; ++k) {
    const int e = threadIdx.x + k * blockDim.x;
    if (e < ne) {
      int x, y;
      divmod_u(e, K1, x, y);
      buf[y * pitch + x] = ld[k];
    }
  }
  __syncthreads();
  stockham_inplace<T, -1>(buf, K1, K0, pitch, 1, ry, twy);  // along y: lines x
  stockham_inplace<T, -1>(buf, K0, K1, 1, pitch, rx, twx);  // along x: lines y
  double acc[10];
#pragma unroll
  for (int i = 0; i < 10; ++i) acc[i] = 0.0;
  const double pref = 2.0 * kPi / P.g.vol;
  const double w = (kz > 0 && 2 * kz != K2) ? 2.0 : 1.0;
  // diagonal box: the partner m' of m is m with every component negated except
  // a Nyquist one (index K/2), so m' (x) k(m') = s_i s_j m (x) k(m) with
  // s_d = -1 (+1 at Nyquist), and |k(m')| = |k(m)|: the two dE/dA terms of
  // the general case merge into one with factor 1 + s_i s_j
  const bool nyq_z = 2 * kz == K2;
  if (P.ortho) {
    // per point in T; a thread's partials over its <= kPlaneLoad points in T,
    // then float64 across threads and planes
    T pa[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) pa[i] = T(0);
    const T kzt = (T)kz_v, kz2 = (T)(kz_v * kz_v), i4k2 = (T)inv4k2, pz = (T)(pref * exz);
    const T wt = (T)w, c2pi = (T)(2.0 * kPi), fscale = (T)(2.0 * P.eps * P.phi_scale);
    const T mzt = (T)mz;
#pragma unroll
    for (int k = 0; k < kPlaneLoad; ++k) {
      const int e = threadIdx.x + k * blockDim.x;
      if (e >= ne) break;
      int ix, iy;
      divmod_u(e, K0, iy, ix);
      CT s = buf[iy * pitch + ix];
      const T S2 = s.x * s.x + s.y * s.y;
      const T kx = kvx[ix], ky = kvy[iy];
      const T k2 = kx * kx + ky * ky + kz2;
      T G = T(0), gp = T(0);
      if (k2 != T(0)) {
        const T ik2 = recip(k2);
        G = pz * exx[ix] * exy[iy] * ik2;
        gp = T(-2) * G * (i4k2 + ik2);
      }
      pa[0] += wt * G * S2;
      const T c = wt * S2 * gp * c2pi;  // 2 x (0.5 w S2 gp 2 pi)
      const T m[3] = {(T)fft_m(ix, K0), (T)fft_m(iy, K1), mzt};
      const T kk[3] = {kx, ky, kzt};
      const bool ny[3] = {2 * ix == K0, 2 * iy == K1, nyq_z};
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          if (i != j && ny[i] != ny[j]) continue;  // 1 + s_i s_j = 0
          const T v = c * kk[j] * m[i];
          if (P.conv == ADMP_KCONV_REFERENCE) pa[1 + i * 3 + j] += v;
          else pa[1 + j * 3 + i] += v;
        }
      const T f = fscale * G;
      s.x *= f;
      s.y *= f;
      buf[iy * pitch + ix] = s;
    }
#pragma unroll
    for (int i = 0; i < 10; ++i) acc[i] = (double)pa[i];
  } else {
    for (int e = threadIdx.x; e < ne; e += blockDim.x) {
      int ix, iy;
      divmod_u(e, K0, iy, ix);
      CT s = buf[iy * pitch + ix];
      const double S2 = (double)s.x * (double)s.x + (double)s.y * (double)s.y;
      const int ixp = ix ? K0 - ix : 0, iyp = iy ? K1 - iy : 0;  // index of -m
      const int mx = fft_m(ix, K0), my = fft_m(iy, K1);
      const int mxp = fft_m(ixp, K0), myp = fft_m(iyp, K1);
      double kv[3], kvp[3], gp, gpp;
      const double th = thx[ix] * thy[iy] * thz;
      kvec(P, mx, my, mz, kv);
      kvec(P, mxp, myp, mzp, kvp);
      const double G = gfun(P, kv, th, &gp);
      const double Gp = gfun(P, kvp, th, &gpp);
      const double Gs = 0.5 * (G + Gp);
      acc[0] += w * Gs * S2;
      add_dEdA(acc, P, mx, my, mz, kv, gp, 0.5 * w * S2);
      add_dEdA(acc, P, mxp, myp, mzp, kvp, gpp, 0.5 * w * S2);
      const T f = (T)(2.0 * P.eps * P.phi_scale * Gs);
      s.x *= f;
      s.y *= f;
      buf[iy * pitch + ix] = s;
    }
  }
  __syncthreads();
  stockham_inplace<T, 1>(buf, K0, K1, 1, pitch, rx, twx);
  stockham_inplace<T, 1>(buf, K1, K0, pitch, 1, ry, twy);
  for (int e = threadIdx.x; e < K0 * K1; e += blockDim.x) {
    int x, y;
    divmod_u(e, K1, x, y);
    plane[e] = buf[y * pitch + x];
  }
  block_sum<10>(acc, red);
  if (threadIdx.x == 0)
    for (int i = 0; i < 10; ++i) partial_dst[i] = acc[i];
}

}  // namespace admp
