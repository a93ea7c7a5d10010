// Charge equilibration on gfx950: the Hessian-vector product of the QEq
// energy and a projected, Jacobi-preconditioned conjugate-gradient solve.
//
// E(q) = chi.q + sum hardness q^2 + E_coul(q) + E_damp(q) is quadratic in q
// (src/torch_admp/qeq.py:33-209, 408-458), so
//   H v = hdiag .* v + J_pair v + J_recip v + J_global v
// with
//   hdiag_i    = 2 hardness_i + [kspace] (-2 eps kappa / sqrt(pi)) + [damping] eps / (sqrt(pi) eta_i)
//   J_pair     = per pair row c_p (real-space erfc and Gaussian-damping terms, built once)
//   J_recip v  = gather(convolve(spread(v)))        (the potential of the mesh charges)
//   J_global v = non-neutral -eps pi / (V kappa^2) sum(v) and the slab-correction terms.
// The solve minimises E on {A q = b} (src/torch_admp/qeq.py:517-765) with
// P = I - A^T (A A^T)^-1 A; all scalars (alpha, beta, norms) stay on the
// device and the host reads the residual every `check_every` iterations.
#include <cmath>
#include <vector>

#include "nblist_kernels.hpp"
#include "pme_kernels.hpp"
#include "qeq.hpp"
#include "tiles.hpp"

namespace admp {

// ------------------------------------------------------------------ pairs
__global__ void __launch_bounds__(kBlock) k_pair_coeff(int64_t np, const int64_t* __restrict__ pairs,
                                                       const double* __restrict__ ds,
                                                       const double* __restrict__ bs,
                                                       const double* __restrict__ eta, int coul,
                                                       int damp, double kappa, double eps,
                                                       double* __restrict__ c) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const double d = ds[p], b = bs[p];
  double v = 0.0;
  if (b != 0.0) {
    const double inv4 = fabs(d) > 1e-4 ? 1.0 / d : 0.0;
    if (coul == 1) v += erfc(kappa * d) * inv4;
    if (coul == 2) v += fabs(d) > 1e-8 ? 1.0 / d : 0.0;
    if (damp) {
      const int64_t i = pairs[2 * p], j = pairs[2 * p + 1];
      double eij = sqrt((eta[i] * eta[i] + eta[j] * eta[j]) * 0.5);
      if (eij == 0.0) eij = 1e-10;
      v -= erfc(d / (2.0 * eij)) * inv4;
    }
    v *= eps * b;
  }
  c[p] = v;
}

__global__ void __launch_bounds__(kBlock) k_pair_matvec(int64_t np, const int64_t* __restrict__ pairs,
                                                        const double* __restrict__ c,
                                                        const double* __restrict__ v,
                                                        double* __restrict__ out) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const double cp = c[p];
  if (cp == 0.0) return;
  const int64_t i = pairs[2 * p], j = pairs[2 * p + 1];
  unsafeAtomicAdd(&out[i], cp * v[j]);
  unsafeAtomicAdd(&out[j], cp * v[i]);
}

// Row lists of the nonzero pair coefficients, built once per geometry: every
// pair (i, j, c) enters row i as (j, c) and row j as (i, c).  The order inside
// a row follows the atomics of the fill.
__global__ void __launch_bounds__(kBlock) k_csr_count(int64_t np, const int64_t* __restrict__ pairs,
                                                      const double* __restrict__ c,
                                                      int* __restrict__ cnt) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np || c[p] == 0.0) return;
  atomicAdd(&cnt[pairs[2 * p]], 1);
  atomicAdd(&cnt[pairs[2 * p + 1]], 1);
}

__global__ void __launch_bounds__(kBlock) k_csr_fill(int64_t np, const int64_t* __restrict__ pairs,
                                                     const double* __restrict__ c,
                                                     const int* __restrict__ rowptr,
                                                     int* __restrict__ cur, int* __restrict__ col,
                                                     double* __restrict__ val) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const double cp = c[p];
  if (cp == 0.0) return;
  const int64_t i = pairs[2 * p], j = pairs[2 * p + 1];
  const int ki = rowptr[i] + atomicAdd(&cur[i], 1);
  col[ki] = (int)j;
  val[ki] = cp;
  const int kj = rowptr[j] + atomicAdd(&cur[j], 1);
  col[kj] = (int)i;
  val[kj] = cp;
}

// Terms of the Hessian product that depend on sums over v (sums[0..2] =
// sum v, sum v z, sum v z^2): the neutralising-background term of the
// reciprocal sum and the slab correction.
struct GlobalTerms {
  const double* z;
  const double* sums;
  double c_nn, pre_slab, lz2;
  int slab, on;
};

__device__ __forceinline__ double global_term(const GlobalTerms& gt, int64_t i) {
  const double Q = gt.sums[0];
  double t = gt.c_nn * Q;
  if (gt.slab) {
    const double M = gt.sums[1], S = gt.sums[2], zi = gt.z[i];
    t += gt.pre_slab * (2.0 * M * zi - S - Q * zi * zi - Q * gt.lz2 / 6.0);
  }
  return t;
}

// out[i] = hdiag[i] v[i] + sum over row i of c v[col] (+ the global terms)
// (overwrite): 16 lanes per row, DPP row sum, no atomics.
__global__ void __launch_bounds__(kBlock) k_csr_matvec(int64_t n, const int* __restrict__ rowptr,
                                                       const int* __restrict__ col,
                                                       const double* __restrict__ val,
                                                       const double* __restrict__ d,
                                                       const double* __restrict__ v,
                                                       GlobalTerms gt,
                                                       double* __restrict__ out) {
  const int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4;
  const int sub = threadIdx.x & 15;
  double acc = 0.0;
  if (row < n) {
    const int e = rowptr[row + 1];
    for (int k = rowptr[row] + sub; k < e; k += 16) acc += val[k] * v[col[k]];
  }
  acc = group16_sum(acc);  // every lane of the wavefront takes part
  if (row < n && sub == 0) {
    double o = d[row] * v[row] + acc;
    if (gt.on) o += global_term(gt, row);
    out[row] = o;
  }
}

// out = hdiag .* v (overwrite)
__global__ void k_diag_mul(int64_t n, const double* __restrict__ d, const double* __restrict__ v,
                           double* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = d[i] * v[i];
}

// ------------------------------------------------------------------ reductions
// Per-block partial sums of up to kQeqMaxDots dot products
//   slot 0:            x . y            (or |x|^2 when y == x)
//   slots 1..nc:       A_c . x
//   slots nc+1..nc+3:  sum x, sum x z, sum x z^2 (global terms; when z != null)
__global__ void __launch_bounds__(kBlock) k_dots(int64_t n, const double* __restrict__ x,
                                                 const double* __restrict__ y,
                                                 const double* __restrict__ A, int nc,
                                                 const double* __restrict__ z,
                                                 double* __restrict__ partials, int nv,
                                                 double* __restrict__ dst,
                                                 int* __restrict__ ticket) {
  __shared__ double red[kQeqMaxDots * (kBlock / 64)];
  double acc[kQeqMaxDots];
#pragma unroll
  for (int k = 0; k < kQeqMaxDots; ++k) acc[k] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double xi = x[i];
    if (y) acc[0] += xi * y[i];
    for (int c = 0; c < nc; ++c) acc[1 + c] += A[(int64_t)c * n + i] * xi;
    if (z) {
      const double zi = z[i];
      acc[nc + 1] += xi;
      acc[nc + 2] += xi * zi;
      acc[nc + 3] += xi * zi * zi;
    }
  }
  block_sum<kQeqMaxDots>(acc, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < kQeqMaxDots; ++k) partials[(int64_t)blockIdx.x * kQeqMaxDots + k] = acc[k];
  // the last work-group to finish (ticket) sums values 0..nv-1 over the blocks
  // in block order into dst[0..nv-1] and re-arms the ticket
  __threadfence();
  __syncthreads();
  __shared__ int last;
  if (threadIdx.x == 0) last = atomicAdd(ticket, 1) == (int)gridDim.x - 1;
  __syncthreads();
  if (!last) return;
  __threadfence();
  const int lane = threadIdx.x & 63;
  for (int v = threadIdx.x >> 6; v < nv; v += blockDim.x >> 6) {
    double t = 0.0;
    for (int bi = lane; bi < (int)gridDim.x; bi += 64)
      t += __hip_atomic_load(partials + (int64_t)bi * kQeqMaxDots + v, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
    t = wave_sum(t);
    if (lane == 0) dst[v] = t;
  }
  if (threadIdx.x == 0) *ticket = 0;
}

// Gram matrix A A^T (nc x nc), per-block partials
__global__ void __launch_bounds__(kBlock) k_gram(int64_t n, const double* __restrict__ A, int nc,
                                                 double* __restrict__ partials) {
  __shared__ double red[kQeqMaxCons * kQeqMaxCons * (kBlock / 64)];
  double acc[kQeqMaxCons * kQeqMaxCons];
#pragma unroll
  for (int k = 0; k < kQeqMaxCons * kQeqMaxCons; ++k) acc[k] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    for (int a = 0; a < nc; ++a)
      for (int b = 0; b < nc; ++b) acc[a * kQeqMaxCons + b] += A[(int64_t)a * n + i] * A[(int64_t)b * n + i];
  block_sum<kQeqMaxCons * kQeqMaxCons>(acc, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < kQeqMaxCons * kQeqMaxCons; ++k)
      partials[(int64_t)blockIdx.x * kQeqMaxCons * kQeqMaxCons + k] = acc[k];
}

// G = (A A^T)^-1 by Gauss-Jordan with partial pivoting (single thread, nc <= 8)
__global__ void k_invert_small(const double* __restrict__ gram, int nc, double* __restrict__ G,
                               int* __restrict__ status) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double a[kQeqMaxCons][2 * kQeqMaxCons];
  for (int r = 0; r < nc; ++r)
    for (int c = 0; c < 2 * nc; ++c)
      a[r][c] = c < nc ? gram[r * kQeqMaxCons + c] : (c - nc == r ? 1.0 : 0.0);
  // Pivots below 1e-12 of the largest Gram entry mark dependent constraint rows
  // (a duplicated row eliminates to an exact zero): the projection is then
  // undefined and the host reports the problem as invalid.
  double gmax = 0.0;
  for (int r = 0; r < nc; ++r)
    for (int c = 0; c < nc; ++c) gmax = fmax(gmax, fabs(a[r][c]));
  const double pivot_floor = fmax(1e-12 * gmax, 1e-300);
  int ok = 1;
  for (int col = 0; col < nc; ++col) {
    int piv = col;
    for (int r = col + 1; r < nc; ++r)
      if (fabs(a[r][col]) > fabs(a[piv][col])) piv = r;
    if (!(fabs(a[piv][col]) > pivot_floor)) {
      ok = 0;
      break;
    }
    for (int c = 0; c < 2 * nc; ++c) {
      const double t = a[col][c];
      a[col][c] = a[piv][c];
      a[piv][c] = t;
    }
    const double inv = 1.0 / a[col][col];
    for (int c = 0; c < 2 * nc; ++c) a[col][c] *= inv;
    for (int r = 0; r < nc; ++r)
      if (r != col) {
        const double f = a[r][col];
        for (int c = 0; c < 2 * nc; ++c) a[r][c] -= f * a[col][c];
      }
  }
  for (int r = 0; r < nc; ++r)
    for (int c = 0; c < nc; ++c) G[r * kQeqMaxCons + c] = a[r][nc + c];
  status[0] = ok;
}

// y = G (Ax - shift) for the projection, evaluated per thread from the reduced values
__device__ __forceinline__ void proj_coeff(const double* __restrict__ G, const double* Ax,
                                           const double* shift, int nc, double* y) {
  for (int a = 0; a < nc; ++a) {
    double s = 0.0;
    for (int b = 0; b < nc; ++b) s += G[a * kQeqMaxCons + b] * (Ax[b] - (shift ? shift[b] : 0.0));
    y[a] = s;
  }
}

// x <- x - A^T G (A x - b): projection onto {A x = b} (b = shift, may be null -> 0)
__global__ void k_project(int64_t n, double* __restrict__ x, const double* __restrict__ A, int nc,
                          const double* __restrict__ G, const double* __restrict__ Ax,
                          const double* __restrict__ shift) {
  double y[kQeqMaxCons];
  proj_coeff(G, Ax, shift, nc, y);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int c = 0; c < nc; ++c) s += A[(int64_t)c * n + i] * y[c];
    x[i] -= s;
  }
}

// global (rank-one / slab) terms of J v from the reduced sums of v
//   sums = (sum v, sum v z, sum v z^2)
__global__ void k_global_terms(int64_t n, const double* __restrict__ z,
                               const double* __restrict__ sums, double c_nn, double pre_slab,
                               double lz2, int slab, double* __restrict__ out) {
  const GlobalTerms gt{z, sums, c_nn, pre_slab, lz2, slab, 1};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] += global_term(gt, i);
}

// g <- g + chi ; r <- -g (projection applied by the caller)
__global__ void k_init_residual(int64_t n, const double* __restrict__ g,
                                const double* __restrict__ chi, double* __restrict__ r) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    r[i] = -(g[i] + chi[i]);
}

// s = M^-1 r (M = precond, identity when null)
__global__ void k_precond(int64_t n, const double* __restrict__ r, const double* __restrict__ m,
                          double* __restrict__ s) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    s[i] = m ? r[i] / m[i] : r[i];
}

// CG step: alpha = rz / (p.w); q += alpha p; r -= alpha P w
// (sc: [0] rz, [1] p.w, [2..] A w)
__global__ void k_cg_step(int64_t n, double* __restrict__ q, double* __restrict__ r,
                          const double* __restrict__ p, const double* __restrict__ w,
                          const double* __restrict__ A, int nc, const double* __restrict__ G,
                          const double* __restrict__ rz, const double* __restrict__ red_pw) {
  const double pw = red_pw[0];
  const double alpha = pw != 0.0 ? rz[0] / pw : 0.0;
  double y[kQeqMaxCons];
  proj_coeff(G, red_pw + 1, nullptr, nc, y);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double pw_i = w[i];
    for (int c = 0; c < nc; ++c) pw_i -= A[(int64_t)c * n + i] * y[c];
    q[i] += alpha * p[i];
    r[i] -= alpha * pw_i;
  }
}

// p <- z + beta p with beta = rz_new / rz_old
__global__ void k_cg_dir(int64_t n, double* __restrict__ p, const double* __restrict__ z,
                         const double* __restrict__ rz_new, const double* __restrict__ rz_old) {
  const double beta = rz_old[0] != 0.0 ? rz_new[0] / rz_old[0] : 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = z[i] + beta * p[i];
}

__global__ void k_copy(int64_t n, const double* __restrict__ a, double* __restrict__ b) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

__global__ void k_extract_z(int64_t n, const double* __restrict__ pos, int axis,
                            double* __restrict__ z) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    z[i] = pos[3 * i + axis];
}

// ------------------------------------------------------------------ small systems
// One work-group carries the vector part of a PCG iteration for n <= kPcgSmallN
// (the Hessian product w = H p has been formed): w.p and A w, the step
// (alpha, projected w, q and r), |r|^2 for the host's residual check, the
// projected preconditioned residual z, r.z and the new direction p, with
// work-group barriers between the reductions instead of kernel boundaries.
// Fixed-order reductions (deterministic); same arithmetic as k_cg_step,
// k_precond, k_project and k_cg_dir.
constexpr int kPcgThreads = 1024;
constexpr int64_t kPcgSmallN = 16384;

// Sums over the work-group in every thread: wavefront sums into smem, then
// thread i < NV adds value i over the wavefronts in order into res[i].  res
// lies outside every call's partial slots, so the next call (whose first
// barrier every thread reaches only after reading res) needs no barrier here.
template <int NV>
__device__ __forceinline__ void block_allsum(double (&v)[NV], double* smem, double* res) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = wave_sum(v[i]);
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < NV; ++i) smem[wid * NV + i] = v[i];
  __syncthreads();
  if (threadIdx.x < NV) {
    double t = 0.0;
    for (int u = 0; u < nw; ++u) t += smem[u * NV + threadIdx.x];
    res[threadIdx.x] = t;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = res[i];
}

__global__ void __launch_bounds__(kPcgThreads) k_pcg_vector(
    int64_t n, double* __restrict__ q, double* __restrict__ r, double* __restrict__ p,
    const double* __restrict__ w, double* __restrict__ z, const double* __restrict__ A, int nc,
    const double* __restrict__ G, const double* __restrict__ m, double* __restrict__ sc,
    int rz_old, int rz_new, int rr_slot, const double* __restrict__ zc, int sums_slot) {
  __shared__ double red[(kQeqMaxCons + 1) * (kPcgThreads / 64)];  // wavefront partials
  __shared__ double res[kQeqMaxCons + 1];                          // the sums
  const int64_t t0 = threadIdx.x, st = blockDim.x;
  // w.p and A w
  double d[kQeqMaxCons + 1];
#pragma unroll
  for (int c = 0; c <= kQeqMaxCons; ++c) d[c] = 0.0;
  for (int64_t i = t0; i < n; i += st) {
    const double wi = w[i];
    d[0] += wi * p[i];
    for (int c = 0; c < nc; ++c) d[1 + c] += A[(int64_t)c * n + i] * wi;
  }
  block_allsum<kQeqMaxCons + 1>(d, red, res);
  const double pw = d[0];
  const double alpha = pw != 0.0 ? sc[rz_old] / pw : 0.0;
  double y[kQeqMaxCons];
  proj_coeff(G, d + 1, nullptr, nc, y);
  // step; |r|^2
  double e[1] = {0.0};
  for (int64_t i = t0; i < n; i += st) {
    double pw_i = w[i];
    for (int c = 0; c < nc; ++c) pw_i -= A[(int64_t)c * n + i] * y[c];
    q[i] += alpha * p[i];
    const double ri = r[i] - alpha * pw_i;
    r[i] = ri;
    e[0] += ri * ri;
  }
  block_allsum<1>(e, red, res);
  if (threadIdx.x == 0) sc[rr_slot] = e[0];
  // z = M^-1 r, projected
#pragma unroll
  for (int c = 0; c <= kQeqMaxCons; ++c) d[c] = 0.0;
  for (int64_t i = t0; i < n; i += st) {
    const double zi = m ? r[i] / m[i] : r[i];
    z[i] = zi;
    for (int c = 0; c < nc; ++c) d[c] += A[(int64_t)c * n + i] * zi;
  }
  if (nc > 0) {
    block_allsum<kQeqMaxCons + 1>(d, red, res);
    proj_coeff(G, d, nullptr, nc, y);
  }
  // r.z of the projected z
  e[0] = 0.0;
  for (int64_t i = t0; i < n; i += st) {
    double zi = z[i];
    if (nc > 0) {
      double s = 0.0;
      for (int c = 0; c < nc; ++c) s += A[(int64_t)c * n + i] * y[c];
      zi -= s;
      z[i] = zi;
    }
    e[0] += r[i] * zi;
  }
  block_allsum<1>(e, red, res);
  const double rzn = e[0], rzo = sc[rz_old];
  const double beta = rzo != 0.0 ? rzn / rzo : 0.0;
  // new direction; with zc, also (sum p, sum p z, sum p z^2) for the next
  // Hessian product's global terms (sc[sums_slot..])
  double g3[3] = {0.0, 0.0, 0.0};
  for (int64_t i = t0; i < n; i += st) {
    const double pi = z[i] + beta * p[i];
    p[i] = pi;
    if (zc) {
      const double zi = zc[i];
      g3[0] += pi;
      g3[1] += pi * zi;
      g3[2] += pi * zi * zi;
    }
  }
  if (zc) {
    block_allsum<3>(g3, red, res);
    if (threadIdx.x == 0)
      for (int k = 0; k < 3; ++k) sc[sums_slot + k] = g3[k];
  }
  __syncthreads();  // every thread has read sc[rz_old] before it can be overwritten
  if (threadIdx.x == 0) sc[rz_new] = rzn;
}

// ------------------------------------------------------------------ point spread
// The Hessian product spreads a new charge vector on fixed positions every
// iteration.  Once per geometry each atom's stencil is stored: base b_d =
// (m0_d - 3) mod K_d (mesh_ref) and weights w_d[a] = M6(f_d + a), so that
// point (b + a) mod K receives q wx[a] wy[b] wz[c]; atoms are binned by base
// in 4-point bins.  The spread then runs over mesh points: a work-group of
// 4 x 4 x 16 points lists the bins within reach (bases p - 5 .. p), loads
// their extents in parallel (one bin per thread), stages the candidates
// (atom, base) in LDS, and each point sums the weight at offset (p - b) mod K
// of every candidate with all three offsets < 6 (unique for K >= 6).  Each
// point is written once: no atomics and no zero fill.
constexpr int kPtBin = 4;                         // mesh points per bin and axis (a shift)
constexpr int kPtB0 = 16, kPtB1 = 8, kPtB2 = 16;  // points per work-group and axis
constexpr int kPtThreads = 256;                   // 8 points along x per thread
constexpr int kPtPer = kPtB0 * kPtB1 * kPtB2 / kPtThreads;
constexpr int kPtL0 = 7, kPtL1 = 5, kPtL2 = 7;    // bins per axis: ceil((B + 5) / 4) + 1
constexpr int kPtChunk = 128;                     // candidates staged in LDS at a time
static_assert(kPtL0 * kPtL1 * kPtL2 <= kPtThreads, "one bin per thread");
static_assert(kPtPer == 8 && kPtB2 == 16 && kPtB1 == 8, "thread to point map below");

template <typename T>
__global__ void __launch_bounds__(kBlock) k_point_rec(const double* __restrict__ pos, int64_t n,
                                                      Geom g, int nb1, int nb2,
                                                      int4* __restrict__ base,
                                                      T* __restrict__ w,
                                                      int* __restrict__ cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int m0[3];
  double f[3];
  mesh_ref(g, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], m0, f);
  int b[3];
  T* wi = w + 18 * i;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    b[d] = pmod32(m0[d] - 3, g.k[d]);
    T wd[6], dd[6];
    bspline6_rounded<T>(f[d], wd, dd);
#pragma unroll
    for (int a = 0; a < 6; ++a) wi[6 * d + a] = wd[a];
  }
  const int bin = ((b[0] / kPtBin) * nb1 + b[1] / kPtBin) * nb2 + b[2] / kPtBin;
  base[i] = make_int4(b[0], b[1], b[2], bin);
  atomicAdd(&cnt[bin], 1);
}

__global__ void __launch_bounds__(kBlock) k_point_fill(int64_t n, const int4* __restrict__ base,
                                                       const int* __restrict__ start,
                                                       int* __restrict__ cur,
                                                       int* __restrict__ atoms) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int bin = base[i].w;
  atoms[start[bin] + atomicAdd(&cur[bin], 1)] = (int)i;
}

// bins (each once) holding bases o - 5 .. o + B - 1 (mod K); returns the count
__device__ __forceinline__ int point_bins(int o, int B, int K, int* lst) {
  int m = 0;
  for (int j = 0; j < B + 5; ++j) {
    int v = o - 5 + j;
    v += v < 0 ? K : 0;
    v = wrapk(v, K);
    const int bin = v / kPtBin;
    bool seen = false;
    for (int u = 0; u < m; ++u) seen |= lst[u] == bin;
    if (!seen) lst[m++] = bin;
  }
  return m;
}

template <typename T>
__global__ void __launch_bounds__(kPtThreads) k_point_spread(
    const int4* __restrict__ base, const T* __restrict__ w, const int* __restrict__ start,
    const int* __restrict__ atoms, const double* __restrict__ q, int K0, int K1, int K2, int nb1,
    int nb2, int g1, int g2, T* __restrict__ mesh) {
  __shared__ int l0[kPtL0 + 4], l1[kPtL1 + 4], l2[kPtL2 + 4], nl[3];
  __shared__ int pre[kPtThreads], bst[kPtThreads];  // inclusive count prefix, bin start
  __shared__ int wtot[kPtThreads / 64];
  __shared__ int4 cb[kPtChunk];       // staged candidates: base,
  __shared__ double cq[kPtChunk];     // charge
  __shared__ T cw[kPtChunk][18];      // and weights
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int blk = blockIdx.x;
  const int c2 = blk % g2, c1 = (blk / g2) % g1, c0 = blk / (g1 * g2);
  const int o0 = c0 * kPtB0, o1 = c1 * kPtB1, o2 = c2 * kPtB2;
  if (tid == 0) nl[0] = point_bins(o0, kPtB0, K0, l0);
  if (tid == 64) nl[1] = point_bins(o1, kPtB1, K1, l1);
  if (tid == 128) nl[2] = point_bins(o2, kPtB2, K2, l2);
  __syncthreads();
  const int n0 = nl[0], n1 = nl[1], n2 = nl[2], nbins = n0 * n1 * n2;
  // bin extents, one per thread, and the inclusive prefix of their counts
  int cnt = 0, st = 0;
  if (tid < nbins) {
    const int u2 = tid % n2, u1 = (tid / n2) % n1, u0 = tid / (n1 * n2);
    const int bin = (l0[u0] * nb1 + l1[u1]) * nb2 + l2[u2];
    st = start[bin];
    cnt = start[bin + 1] - st;
  }
  int inc = cnt;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int up = __shfl_up(inc, off, 64);
    if (lane >= off) inc += up;
  }
  if (lane == 63) wtot[wid] = inc;
  __syncthreads();
  for (int u = 0; u < wid; ++u) inc += wtot[u];
  pre[tid] = inc;
  bst[tid] = st;
  __syncthreads();
  const int total = pre[kPtThreads - 1];
  // thread -> z = tid % 16, y = (tid / 16) % 8, x = 8 (tid / 128) + k
  const int p2 = o2 + (tid & 15), p1 = o1 + ((tid >> 4) & 7), px = o0 + kPtPer * (tid >> 7);
  double acc[kPtPer];
#pragma unroll
  for (int k = 0; k < kPtPer; ++k) acc[k] = 0.0;
  for (int c0k = 0; c0k < total; c0k += kPtChunk) {
    const int c = c0k + tid;
    if (tid < kPtChunk && c < total) {
      // the bin slot holding candidate c: first slot with pre > c
      int lo = 0, hi = nbins - 1;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (pre[mid] > c) hi = mid; else lo = mid + 1;
      }
      const int first = lo ? pre[lo - 1] : 0;  // candidates before this slot
      const int a = atoms[bst[lo] + (c - first)];
      cb[tid] = base[a];
      cq[tid] = q[a];
      const T* wa = w + 18 * (int64_t)a;
#pragma unroll
      for (int e = 0; e < 18; ++e) cw[tid][e] = wa[e];
    }
    __syncthreads();
    const int m = min(kPtChunk, total - c0k);
    for (int j = 0; j < m; ++j) {
      const int4 b = cb[j];
      int d1 = p1 - b.y, d2 = p2 - b.z;
      d1 += d1 < 0 ? K1 : 0;
      d2 += d2 < 0 ? K2 : 0;
      if (d1 < 6 && d2 < 6) {
        const double qyz = cq[j] * ((double)cw[j][6 + d1] * (double)cw[j][12 + d2]);
#pragma unroll
        for (int k = 0; k < kPtPer; ++k) {
          int d0 = px + k - b.x;
          d0 += d0 < 0 ? K0 : 0;
          if (d0 < 6) acc[k] += qyz * (double)cw[j][d0];
        }
      }
    }
    __syncthreads();
  }
  if (p1 < K1 && p2 < K2) {
#pragma unroll
    for (int k = 0; k < kPtPer; ++k)
      if (px + k < K0) mesh[((int64_t)(px + k) * K1 + p1) * K2 + p2] = (T)acc[k];
  }
}

// ------------------------------------------------------------------ host side
static unsigned grid_for(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + kBlock - 1) / kBlock, kRedBlocks));
}

QeqLayout qeq_layout(const QeqProblem& P) {
  QeqLayout L;
  size_t off = 0;
  auto take = [&](size_t b) {
    const size_t o = off;
    off += (b + 255) & ~size_t(255);
    return o;
  };
  const int64_t n = std::max<int64_t>(P.n, 1);
  const size_t vec = sizeof(double) * n;
  L.c = take(sizeof(double) * std::max<int64_t>(P.n_pairs, 1));
  const int64_t ent = 2 * std::max<int64_t>(P.n_pairs, 1);
  L.csr = ent < (int64_t)INT32_MAX && n < (int64_t)INT32_MAX;
  if (L.csr) {
    L.rowcnt = take(sizeof(int) * n);
    L.rowcur = take(sizeof(int) * n);
    L.rowptr = take(sizeof(int) * (n + 1));
    L.col = take(sizeof(int) * ent);
    L.val = take(sizeof(double) * ent);
    L.scan = take(sizeof(int) * (n / kBlock + 2));
  } else {
    L.rowcnt = L.rowcur = L.rowptr = L.col = L.val = L.scan = 0;
  }
  L.r = take(vec);
  L.z = take(vec);
  L.p = take(vec);
  L.w = take(vec);
  L.s = take(vec);
  L.zc = take(vec);
  L.dots = take(sizeof(double) * kRedBlocks * std::max(kQeqMaxDots, kQeqMaxCons * kQeqMaxCons));
  L.sc = take(sizeof(double) * 256);  // device scalars, see kSc* below
  L.status = take(sizeof(int) * 8);
  L.ticket = take(sizeof(int) * 4);
  const size_t es = P.precision == ADMP_F32 ? 4 : 8;
  if (P.kspace) {
    L.cg = make_grid(P.box, P.cutoff, false);
    nb_layout(P.n, L.cg, &L.nb);
    L.nbws = take(L.nb.bytes);
    L.tg = make_tile_geom(L.cg, P.kmesh, es);
    int64_t nbins = 1;
    for (int d = 0; d < 3; ++d) {
      L.nbin[d] = (P.kmesh[d] + kPtBin - 1) / kPtBin;
      nbins *= L.nbin[d];
    }
    L.pt = P.kmesh[0] >= 6 && P.kmesh[1] >= 6 && P.kmesh[2] >= 6 && nbins < (int64_t)INT32_MAX &&
           n < (int64_t)INT32_MAX;
    if (L.pt) {
      L.prec_base = take(sizeof(int4) * n);
      L.prec_w = take(sizeof(double) * 18 * n);
      L.pbcnt = take(sizeof(int) * nbins);
      L.pbcur = take(sizeof(int) * nbins);
      L.pbstart = take(sizeof(int) * (nbins + 1));
      L.pbatoms = take(sizeof(int) * n);
      L.pbscan = take(sizeof(int) * (nbins / kBlock + 2));
    }
    L.mesh = take((size_t)P.kmesh[0] * P.kmesh[1] * P.kmesh[2] * es);
    L.spec = take(spec_elems(P.kmesh, 2 * es) * 2 * es);
    L.ppart = take(sizeof(double) * (std::max(convolve_partials_count(P.kmesh), (size_t)kRedBlocks * 9) + 16));
  } else {
    L.tg.ok = 0;
    L.pt = 0;
    L.nbws = L.mesh = L.spec = L.ppart = 0;
  }
  L.bytes = off;
  return L;
}

// sc layout (doubles): 0 rz (even iterations), 10 rz (odd), 20.. reduced dots,
// 32.. G (kQeqMaxCons^2), 112.. A q / A g, 128.. global sums, 160 |r|^2,
// 176.. Gram matrix A A^T
constexpr int kScRz0 = 0, kScRz1 = 10, kScDots = 20, kScG = 32, kScAx = 112, kScSums = 128,
              kScRr = 160, kScGram = 176;

// out[v] = sum_b p[b * stride + v] for v < nv, fixed order (one wavefront per value)
__global__ void __launch_bounds__(kReduceThreads) k_reduce_cols(const double* __restrict__ p,
                                                                int nblocks, int stride, int nv,
                                                                double* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int nw = blockDim.x >> 6;
  for (int v = threadIdx.x >> 6; v < nv; v += nw) {
    double s = 0.0;
    for (int bi = lane; bi < nblocks; bi += 64) s += p[(int64_t)bi * stride + v];
    s = wave_sum(s);
    if (lane == 0) out[v] = s;
  }
}

// sums_ready: sc[kScSums..] already holds the sums over v of the global terms
template <typename T>
static hipError_t matvec(const QeqProblem& P, const QeqLayout& L, char* ws, const Geom& g,
                         const double* v, double* out, hipStream_t s, bool sums_ready = false) {
  const int64_t n = P.n;
  double* sc = (double*)(ws + L.sc);
  const bool global = P.kspace || P.slab;
  GlobalTerms gt{(const double*)(ws + L.zc), sc + kScSums,
                 (P.kspace && P.kappa > 0.0) ? -P.eps * kPi / (g.vol * P.kappa * P.kappa) : 0.0,
                 2.0 * kPi / g.vol * P.eps, P.lz2, P.slab, 0};
  if (global && !sums_ready) {
    // slots 1..3 hold (sum v, sum v z, sum v z^2) -> sc[kScSums..]
    k_dots<<<grid_for(n), kBlock, 0, s>>>(n, v, nullptr, nullptr, 0, gt.z,
                                          (double*)(ws + L.dots), 4, sc + kScSums - 1,
                                          (int*)(ws + L.ticket));
  }
  if (L.csr) {
    gt.on = global ? 1 : 0;  // the row kernel adds the global terms as it writes
    k_csr_matvec<<<(unsigned)((16 * n + kBlock - 1) / kBlock), kBlock, 0, s>>>(
        n, (const int*)(ws + L.rowptr), (const int*)(ws + L.col), (const double*)(ws + L.val),
        P.hdiag, v, gt, out);
  } else {
    k_diag_mul<<<grid_for(n), kBlock, 0, s>>>(n, P.hdiag, v, out);
    if (P.n_pairs > 0)
      k_pair_matvec<<<(unsigned)((P.n_pairs + kBlock - 1) / kBlock), kBlock, 0, s>>>(
          P.n_pairs, P.pairs, (const double*)(ws + L.c), v, out);
    if (global)
      k_global_terms<<<grid_for(n), kBlock, 0, s>>>(n, gt.z, gt.sums, gt.c_nn, gt.pre_slab,
                                                    gt.lz2, gt.slab, out);
  }
  if (P.kspace) {
    T* mesh = (T*)(ws + L.mesh);
    hipError_t e = hipSuccess;
    if (L.pt) {
      const int g0 = (g.k[0] + kPtB0 - 1) / kPtB0, g1 = (g.k[1] + kPtB1 - 1) / kPtB1,
                g2 = (g.k[2] + kPtB2 - 1) / kPtB2;
      k_point_spread<T><<<(unsigned)((int64_t)g0 * g1 * g2), kPtThreads, 0, s>>>(
          (const int4*)(ws + L.prec_base), (const T*)(ws + L.prec_w),
          (const int*)(ws + L.pbstart), (const int*)(ws + L.pbatoms), v, g.k[0], g.k[1], g.k[2],
          L.nbin[1], L.nbin[2], g1, g2, mesh);
    } else if (L.tg.ok) {
      e = launch_spread_tiled<T>(P.pos, v, n, g, L.tg, L.cg.ncell,
                                 (const int*)(ws + L.nbws + L.nb.sorted),
                                 (const int*)(ws + L.nbws + L.nb.start), mesh, s);
    } else {
      e = launch_spread<T>(P.pos, v, n, g, mesh, s);
    }
    if (e != hipSuccess) return e;
    e = launch_convolve<T>(mesh, g, P.convention, P.kappa, P.eps, 1.0, mesh, nullptr, nullptr,
                           (cplx<T>*)(ws + L.spec), (double*)(ws + L.ppart), s);
    if (e != hipSuccess) return e;
    e = launch_gather<T>(P.pos, v, n, g, mesh, nullptr, out, nullptr, 1,
                         (double*)(ws + L.ppart), s);
    if (e != hipSuccess) return e;
  }
  return hipGetLastError();
}

// dots of (x, y) and A x into sc[dst..]: dst = x.y, dst+1.. = A x
static void dots_into(const QeqProblem& P, const QeqLayout& L, char* ws, const double* x,
                      const double* y, int dst, hipStream_t s) {
  const unsigned gb = grid_for(P.n);
  double* sc = (double*)(ws + L.sc);
  k_dots<<<gb, kBlock, 0, s>>>(P.n, x, y, P.cons, P.n_cons, nullptr, (double*)(ws + L.dots),
                               1 + P.n_cons, sc + dst, (int*)(ws + L.ticket));
}

template <typename T>
static int setup(const QeqProblem& P, const QeqLayout& L, char* ws, Geom* g, hipStream_t s) {
  *g = Geom{};
  if (P.box) {
    const CellGrid cg = make_grid(P.box, 1.0, false);
    for (int i = 0; i < 9; ++i) {
      g->h[i] = cg.h[i];
      g->a[i] = cg.a[i];
    }
    const double* h = P.box;
    g->vol = h[0] * (h[4] * h[8] - h[5] * h[7]) - h[1] * (h[3] * h[8] - h[5] * h[6]) +
             h[2] * (h[3] * h[7] - h[4] * h[6]);
  }
  for (int d = 0; d < 3; ++d) g->k[d] = P.kmesh ? P.kmesh[d] : 1;
  if (P.n_pairs > 0)
    k_pair_coeff<<<(unsigned)((P.n_pairs + kBlock - 1) / kBlock), kBlock, 0, s>>>(
        P.n_pairs, P.pairs, P.ds, P.bs, P.eta, P.coul, P.damping, P.kappa, P.eps,
        (double*)(ws + L.c));
  k_fill<int><<<1, 64, 0, s>>>((int*)(ws + L.ticket), 4, 0);  // re-armed by each k_dots
  if (L.csr && P.n > 0) {
    int* cnt = (int*)(ws + L.rowcnt);
    int* cur = (int*)(ws + L.rowcur);
    int* rowptr = (int*)(ws + L.rowptr);
    k_fill<int><<<grid_for(P.n), kBlock, 0, s>>>(cnt, P.n, 0);
    k_fill<int><<<grid_for(P.n), kBlock, 0, s>>>(cur, P.n, 0);
    const unsigned gp = (unsigned)((P.n_pairs + kBlock - 1) / kBlock);
    if (P.n_pairs > 0) k_csr_count<<<gp, kBlock, 0, s>>>(P.n_pairs, P.pairs, (const double*)(ws + L.c), cnt);
    launch_scan<int>(cnt, P.n, rowptr, (int*)(ws + L.scan), s);
    if (P.n_pairs > 0)
      k_csr_fill<<<gp, kBlock, 0, s>>>(P.n_pairs, P.pairs, (const double*)(ws + L.c), rowptr, cur,
                                       (int*)(ws + L.col), (double*)(ws + L.val));
  }
  if (L.pt) {
    const int64_t nbins = (int64_t)L.nbin[0] * L.nbin[1] * L.nbin[2];
    int* cnt = (int*)(ws + L.pbcnt);
    int* cur = (int*)(ws + L.pbcur);
    int* start = (int*)(ws + L.pbstart);
    const unsigned gbin = (unsigned)std::min<int64_t>((nbins + kBlock - 1) / kBlock, 4096);
    k_fill<int><<<gbin, kBlock, 0, s>>>(cnt, nbins, 0);
    k_fill<int><<<gbin, kBlock, 0, s>>>(cur, nbins, 0);
    if (P.n > 0)
      k_point_rec<T><<<(unsigned)((P.n + kBlock - 1) / kBlock), kBlock, 0, s>>>(
          P.pos, P.n, *g, L.nbin[1], L.nbin[2], (int4*)(ws + L.prec_base), (T*)(ws + L.prec_w),
          cnt);
    launch_scan<int>(cnt, nbins, start, (int*)(ws + L.pbscan), s);
    if (P.n > 0)
      k_point_fill<<<(unsigned)((P.n + kBlock - 1) / kBlock), kBlock, 0, s>>>(
          P.n, (const int4*)(ws + L.prec_base), start, cur, (int*)(ws + L.pbatoms));
  } else if (P.kspace && launch_bin(P.pos, P.n, L.cg, ws + L.nbws, L.nb, s) != hipSuccess) {
    return ADMP_ERR_HIP;
  }
  if (P.slab || P.kspace)
    k_extract_z<<<grid_for(P.n), kBlock, 0, s>>>(P.n, P.pos, P.slab ? P.slab_axis : 0,
                                                  (double*)(ws + L.zc));
  return hipGetLastError() == hipSuccess ? ADMP_OK : ADMP_ERR_HIP;
}

template <typename T>
int qeq_matvec_impl(const QeqProblem& P, const double* v, double* out, void* workspace,
                    hipStream_t s) {
  QeqLayout L = qeq_layout(P);
  L.csr = 0;  // one product: atomics cost less than building row lists and stencils
  L.pt = 0;
  char* ws = (char*)workspace;
  Geom g;
  int st = setup<T>(P, L, ws, &g, s);
  if (st != ADMP_OK) return st;
  if (matvec<T>(P, L, ws, g, v, out, s) != hipSuccess) return ADMP_ERR_HIP;
  return hipGetLastError() == hipSuccess ? ADMP_OK : ADMP_ERR_HIP;
}

// A pinned host word and an event per device for reading the residual without
// draining the stream: one holder per thread, created on first use and freed
// when the thread exits (a serving worker thread ends with its holder).
struct ResidualReader {
  static constexpr int kDev = 16;
  double* host = nullptr;
  hipEvent_t ev[kDev] = {};
  ~ResidualReader() {
    for (int d = 0; d < kDev; ++d)
      if (ev[d]) (void)hipEventDestroy(ev[d]);
    if (host) (void)hipHostFree(host);
  }
};

// Returns false when the word or the event cannot be made.
static bool residual_reader(double** host, hipEvent_t* ev) {
  static thread_local ResidualReader rr;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= ResidualReader::kDev) return false;
  if (!rr.host && hipHostMalloc((void**)&rr.host, 64, hipHostMallocDefault) != hipSuccess) {
    rr.host = nullptr;
    return false;
  }
  if (!rr.ev[dev] && hipEventCreateWithFlags(&rr.ev[dev], hipEventDisableTiming) != hipSuccess) {
    rr.ev[dev] = nullptr;
    return false;
  }
  *host = rr.host;
  *ev = rr.ev[dev];
  return true;
}

template <typename T>
int qeq_solve_impl(const QeqProblem& P, const double* q0, double tol, int max_iter,
                   int check_every, double* q, int* iters, double* resid, void* workspace,
                   hipStream_t s) {
  const QeqLayout L = qeq_layout(P);
  char* ws = (char*)workspace;
  Geom g;
  int st = setup<T>(P, L, ws, &g, s);
  if (st != ADMP_OK) return st;
  const int64_t n = P.n;
  const int nc = P.n_cons;
  double* sc = (double*)(ws + L.sc);
  double* r = (double*)(ws + L.r);
  double* z = (double*)(ws + L.z);
  double* p = (double*)(ws + L.p);
  double* w = (double*)(ws + L.w);
  double* sv = (double*)(ws + L.s);
  double* G = sc + kScG;
  const unsigned gb = grid_for(n);
  // (A A^T)^-1
  if (nc > 0) {
    k_gram<<<gb, kBlock, 0, s>>>(n, P.cons, nc, (double*)(ws + L.dots));
    k_reduce_cols<<<1, kReduceThreads, 0, s>>>((double*)(ws + L.dots), gb,
                                               kQeqMaxCons * kQeqMaxCons,
                                               kQeqMaxCons * kQeqMaxCons, sc + kScGram);
    k_invert_small<<<1, 64, 0, s>>>(sc + kScGram, nc, G, (int*)(ws + L.status));
  }
  // feasible start: q = q0 - A^T G (A q0 - b)
  k_copy<<<gb, kBlock, 0, s>>>(n, q0, q);
  if (nc > 0) {
    dots_into(P, L, ws, q, nullptr, kScAx - 1, s);
    k_project<<<gb, kBlock, 0, s>>>(n, q, P.cons, nc, G, sc + kScAx, P.cons_vals_dev);
  }
  // r = -P (H q + chi)
  if (matvec<T>(P, L, ws, g, q, w, s) != hipSuccess) return ADMP_ERR_HIP;
  k_init_residual<<<gb, kBlock, 0, s>>>(n, w, P.chi, r);
  if (nc > 0) {
    dots_into(P, L, ws, r, nullptr, kScAx - 1, s);
    k_project<<<gb, kBlock, 0, s>>>(n, r, P.cons, nc, G, sc + kScAx, nullptr);
  }
  // z = P M^-1 r, p = z, rz = r.z
  auto precond_project = [&](int rz_slot) {
    k_precond<<<gb, kBlock, 0, s>>>(n, r, P.precond, z);
    if (nc > 0) {
      dots_into(P, L, ws, z, nullptr, kScAx - 1, s);
      k_project<<<gb, kBlock, 0, s>>>(n, z, P.cons, nc, G, sc + kScAx, nullptr);
    }
    dots_into(P, L, ws, r, z, rz_slot, s);
  };
  precond_project(kScRz0);
  k_copy<<<gb, kBlock, 0, s>>>(n, z, p);
  int it = 0;
  double res = 0.0;
  bool converged = false;
  // residual of the starting point, and the status of the Gram inversion
  dots_into(P, L, ws, r, r, kScRr, s);
  int gram_ok = 1;
  if (hipMemcpyAsync(&res, sc + kScRr, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
      (nc > 0 && hipMemcpyAsync(&gram_ok, ws + L.status, sizeof(int), hipMemcpyDeviceToHost, s) !=
                     hipSuccess) ||
      hipStreamSynchronize(s) != hipSuccess)
    return ADMP_ERR_HIP;
  if (!gram_ok) return ADMP_ERR_INVALID;  // linearly dependent constraint rows
  res = std::sqrt(std::max(res, 0.0)) / (double)std::max<int64_t>(n, 1);
  converged = res < tol;
  double* rr_host = nullptr;
  hipEvent_t rr_ev = nullptr;
  const bool ahead = n <= kPcgSmallN && residual_reader(&rr_host, &rr_ev);
  bool product_queued = false;  // this iteration's H p was enqueued ahead of a residual read
  while (!converged && it < max_iter) {
    ++it;
    const int rz_old = (it & 1) ? kScRz0 : kScRz1;
    const int rz_new = (it & 1) ? kScRz1 : kScRz0;
    const bool small = n <= kPcgSmallN;
    const bool global = P.kspace || P.slab;
    // after the first iteration the vector step has summed p for the global terms
    if (!product_queued && matvec<T>(P, L, ws, g, p, w, s, small && global && it > 1) != hipSuccess)
      return ADMP_ERR_HIP;
    product_queued = false;
    if (small && ahead) {
      k_pcg_vector<<<1, kPcgThreads, 0, s>>>(n, q, r, p, w, z, P.cons, nc, G, P.precond, sc,
                                              rz_old, rz_new, kScRr,
                                              global ? (const double*)(ws + L.zc) : nullptr,
                                              kScSums);
      if (it % check_every == 0 || it == max_iter) {
        // copy |r|^2 to pinned memory, queue the next product (it only writes w),
        // then wait for the copy alone: the GPU stays busy while the host decides
        if (hipMemcpyAsync(rr_host, sc + kScRr, sizeof(double), hipMemcpyDeviceToHost, s) !=
                hipSuccess ||
            hipEventRecord(rr_ev, s) != hipSuccess)
          return ADMP_ERR_HIP;
        if (it < max_iter) {
          if (matvec<T>(P, L, ws, g, p, w, s, global) != hipSuccess) return ADMP_ERR_HIP;
          product_queued = true;
        }
        if (hipEventSynchronize(rr_ev) != hipSuccess) return ADMP_ERR_HIP;
        res = std::sqrt(std::max(*rr_host, 0.0)) / (double)std::max<int64_t>(n, 1);
        if (res < tol) {
          converged = true;
          break;
        }
      }
      continue;
    }
    if (small) {
      // the whole vector part in one work-group; |r|^2 lands in sc[kScRr]
      k_pcg_vector<<<1, kPcgThreads, 0, s>>>(n, q, r, p, w, z, P.cons, nc, G, P.precond, sc,
                                              rz_old, rz_new, kScRr,
                                              global ? (const double*)(ws + L.zc) : nullptr,
                                              kScSums);
      if (it % check_every == 0 || it == max_iter) {
        if (hipMemcpyAsync(&res, sc + kScRr, sizeof(double), hipMemcpyDeviceToHost, s) !=
                hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
          return ADMP_ERR_HIP;
        res = std::sqrt(std::max(res, 0.0)) / (double)std::max<int64_t>(n, 1);
        if (res < tol) {
          converged = true;
          break;
        }
      }
      continue;
    }
    dots_into(P, L, ws, w, p, kScDots, s);  // w.p and A w (A acts on the first argument)
    k_cg_step<<<gb, kBlock, 0, s>>>(n, q, r, p, w, P.cons, nc, G, sc + rz_old, sc + kScDots);
    if (it % check_every == 0 || it == max_iter) {
      dots_into(P, L, ws, r, r, kScRr, s);
      if (hipMemcpyAsync(&res, sc + kScRr, sizeof(double), hipMemcpyDeviceToHost, s) !=
              hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess)
        return ADMP_ERR_HIP;
      res = std::sqrt(std::max(res, 0.0)) / (double)std::max<int64_t>(n, 1);
      if (res < tol) {
        converged = true;
        break;
      }
    }
    precond_project(rz_new);
    k_cg_dir<<<gb, kBlock, 0, s>>>(n, p, z, sc + rz_new, sc + rz_old);
  }
  (void)sv;
  *iters = converged ? it : -1;
  *resid = res;
  return hipGetLastError() == hipSuccess ? ADMP_OK : ADMP_ERR_HIP;
}

// ------------------------------------------------------------------ batched solves
// Several systems' solves advance together: one Hessian product for the whole
// batch per iteration (pair rows over the concatenated atoms, the batch
// engine's spread / FFT / influence function / potential gather), then one
// work-group per system for the vector step.  With the single constraint
// sum q = Q the projection is x - mean(x).  Each system stops on its own: the
// vector step compares |r| / n with tol and writes the iteration into a
// device flag; a converged system's charges stay as they are while the others
// go on, and the host reads the flags every check_every iterations.
//
// Per-system scalars (sc + 8 s): rz (even / odd iterations), |r|^2, sum of the
// vector the next product multiplies, the neutralising coefficient
// -eps pi / (V kappa^2), the total charge, the tolerance.
constexpr int kQbRz0 = 0, kQbRz1 = 1, kQbRr = 2, kQbSumV = 3, kQbCnn = 4, kQbQ = 5, kQbTol = 6,
              kQbStride = 8;

__global__ void __launch_bounds__(kBlock) k_qb_pair_coeff(
    int64_t np, const int64_t* __restrict__ pairs, const double* __restrict__ ds,
    const double* __restrict__ bs, const double* __restrict__ eta, const int* __restrict__ sys_of,
    const SysDesc* __restrict__ tab, int damp, double eps, double* __restrict__ c) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const double d = ds[p], b = bs[p];
  double v = 0.0;
  if (b != 0.0) {
    const int64_t i = pairs[2 * p], j = pairs[2 * p + 1];
    const double kappa = tab[sys_of[i]].kappa;
    const double inv4 = fabs(d) > 1e-4 ? 1.0 / d : 0.0;
    v += erfc(kappa * d) * inv4;
    if (damp) {
      double eij = sqrt((eta[i] * eta[i] + eta[j] * eta[j]) * 0.5);
      if (eij == 0.0) eij = 1e-10;
      v -= erfc(d / (2.0 * eij)) * inv4;
    }
    v *= eps * b;
  }
  c[p] = v;
}

// out[row] = hdiag v + sum over row of c v[col] + c_nn(s) sum(v over s) (overwrite)
__global__ void __launch_bounds__(kBlock) k_qb_rows(int64_t n, const int* __restrict__ rowptr,
                                                    const int* __restrict__ col,
                                                    const double* __restrict__ val,
                                                    const double* __restrict__ d,
                                                    const double* __restrict__ v,
                                                    const int* __restrict__ sys_of,
                                                    const double* __restrict__ sc,
                                                    double* __restrict__ out) {
  const int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4;
  const int sub = threadIdx.x & 15;
  double acc = 0.0;
  if (row < n) {
    const int e = rowptr[row + 1];
    for (int k = rowptr[row] + sub; k < e; k += 16) acc += val[k] * v[col[k]];
  }
  acc = group16_sum(acc);  // every lane of the wavefront takes part
  if (row < n && sub == 0) {
    const double* ss = sc + (int64_t)kQbStride * sys_of[row];
    out[row] = d[row] * v[row] + acc + ss[kQbCnn] * ss[kQbSumV];
  }
}

// Feasible start of system blockIdx.x: q = q0 - (sum q0 - Q) / n
__global__ void __launch_bounds__(kPcgThreads) k_qb_start(const int64_t* __restrict__ aoff,
                                                          const double* __restrict__ q0,
                                                          double* __restrict__ q,
                                                          double* __restrict__ sc) {
  __shared__ double red[kPcgThreads / 64];
  __shared__ double res[1];
  const int sid = blockIdx.x;
  const int64_t a0 = aoff[sid], a1 = aoff[sid + 1], n = a1 - a0;
  double* ss = sc + (int64_t)kQbStride * sid;
  double t[1] = {0.0};
  for (int64_t i = a0 + threadIdx.x; i < a1; i += blockDim.x) t[0] += q0[i];
  block_allsum<1>(t, red, res);
  const double shift = n > 0 ? (t[0] - ss[kQbQ]) / (double)n : 0.0;
  for (int64_t i = a0 + threadIdx.x; i < a1; i += blockDim.x) q[i] = q0[i] - shift;
  if (threadIdx.x == 0) ss[kQbSumV] = ss[kQbQ];
}

// r = -P (H q + chi), z = P M^-1 r, p = z, rz, |r|^2, sum p; converged at
// iteration 0 when |r| / n < tol
__global__ void __launch_bounds__(kPcgThreads) k_qb_init(
    const int64_t* __restrict__ aoff, const double* __restrict__ w, const double* __restrict__ chi,
    const double* __restrict__ m, double* __restrict__ r, double* __restrict__ z,
    double* __restrict__ p, double* __restrict__ sc, int* __restrict__ done) {
  __shared__ double red[3 * (kPcgThreads / 64)];
  __shared__ double res[3];
  const int sid = blockIdx.x;
  const int64_t a0 = aoff[sid], a1 = aoff[sid + 1], n = a1 - a0;
  const double inv_n = n > 0 ? 1.0 / (double)n : 0.0;
  double* ss = sc + (int64_t)kQbStride * sid;
  double t[1] = {0.0};
  for (int64_t i = a0 + threadIdx.x; i < a1; i += blockDim.x) {
    const double ri = -(w[i] + chi[i]);
    r[i] = ri;
    t[0] += ri;
  }
  block_allsum<1>(t, red, res);
  const double mr = t[0] * inv_n;
  t[0] = 0.0;
  for (int64_t i = a0 + threadIdx.x; i < a1; i += blockDim.x) {
    const double ri = r[i] - mr;
    r[i] = ri;
    const double zi = m ? ri / m[i] : ri;
    z[i] = zi;
    t[0] += zi;
  }
  block_allsum<1>(t, red, res);
  const double mz = t[0] * inv_n;
  double u[3] = {0.0, 0.0, 0.0};
  for (int64_t i = a0 + threadIdx.x; i < a1; i += blockDim.x) {
    const double zi = z[i] - mz, ri = r[i];
    z[i] = zi;
    p[i] = zi;
    u[0] += ri * zi;
    u[1] += ri * ri;
    u[2] += zi;
  }
  block_allsum<3>(u, red, res);
  if (threadIdx.x == 0) {
    ss[kQbRz0] = u[0];
    ss[kQbRr] = u[1];
    ss[kQbSumV] = u[2];
    done[sid] = sqrt(fmax(u[1], 0.0)) * inv_n < ss[kQbTol] ? 0 : -1;
  }
}

// One projected PCG iteration of system blockIdx.x (w = H p formed), the
// arithmetic of k_pcg_vector with A = ones: alpha, q, r, |r|^2; at |r| / n <
// tol the flag takes the iteration and the system stops; else z = P M^-1 r,
// r.z, beta and the new direction p with its sum.
__global__ void __launch_bounds__(kPcgThreads) k_qb_step(
    const int64_t* __restrict__ aoff, double* __restrict__ q, double* __restrict__ r,
    double* __restrict__ p, const double* __restrict__ w, double* __restrict__ z,
    const double* __restrict__ m, double* __restrict__ sc, int* __restrict__ done, int it,
    int rz_old, int rz_new) {
  __shared__ double red[3 * (kPcgThreads / 64)];
  __shared__ double res[3];
  const int sid = blockIdx.x;
  if (done[sid] >= 0) return;  // the whole work-group: the flag is only written below
  const int64_t a0 = aoff[sid], a1 = aoff[sid + 1], n = a1 - a0;
  const double inv_n = n > 0 ? 1.0 / (double)n : 0.0;
  double* ss = sc + (int64_t)kQbStride * sid;
  const double rzo = ss[rz_old];
  double u[2] = {0.0, 0.0};
  for (int64_t i = a0 + threadIdx.x; i < a1; i += blockDim.x) {
    const double wi = w[i];
    u[0] += wi * p[i];
    u[1] += wi;
  }
  block_allsum<2>(u, red, res);
  const double alpha = u[0] != 0.0 ? rzo / u[0] : 0.0;
  const double mw = u[1] * inv_n;
  double e[1] = {0.0};
  for (int64_t i = a0 + threadIdx.x; i < a1; i += blockDim.x) {
    q[i] += alpha * p[i];
    const double ri = r[i] - alpha * (w[i] - mw);
    r[i] = ri;
    e[0] += ri * ri;
  }
  block_allsum<1>(e, red, res);
  const double rr = e[0];
  if (sqrt(fmax(rr, 0.0)) * inv_n < ss[kQbTol]) {
    __syncthreads();
    if (threadIdx.x == 0) {
      ss[kQbRr] = rr;
      done[sid] = it;
    }
    return;
  }
  double t[1] = {0.0};
  for (int64_t i = a0 + threadIdx.x; i < a1; i += blockDim.x) {
    const double zi = m ? r[i] / m[i] : r[i];
    z[i] = zi;
    t[0] += zi;
  }
  block_allsum<1>(t, red, res);
  const double mz = t[0] * inv_n;
  e[0] = 0.0;
  for (int64_t i = a0 + threadIdx.x; i < a1; i += blockDim.x) {
    const double zi = z[i] - mz;
    z[i] = zi;
    e[0] += r[i] * zi;
  }
  block_allsum<1>(e, red, res);
  const double rzn = e[0];
  const double beta = rzo != 0.0 ? rzn / rzo : 0.0;
  t[0] = 0.0;
  for (int64_t i = a0 + threadIdx.x; i < a1; i += blockDim.x) {
    const double pi = z[i] + beta * p[i];
    p[i] = pi;
    t[0] += pi;
  }
  block_allsum<1>(t, red, res);
  __syncthreads();  // every thread has read sc[rz_old] before it can be overwritten
  if (threadIdx.x == 0) {
    ss[rz_new] = rzn;
    ss[kQbRr] = rr;
    ss[kQbSumV] = t[0];
  }
}

namespace {
struct QbLayout {
  size_t c, rowcnt, rowcur, rowptr, col, val, scan, r, z, p, w, sc, done, bytes;
  bool ok;
};
QbLayout qb_layout(const BatchPlan& P, int64_t n_pairs) {
  QbLayout L;
  size_t off = P.bytes;
  auto take = [&](size_t b) {
    const size_t o = off;
    off += (b + 255) & ~size_t(255);
    return o;
  };
  const int64_t n = std::max<int64_t>(P.ntot, 1), np = std::max<int64_t>(n_pairs, 1);
  L.ok = 2 * np < (int64_t)INT32_MAX && n < (int64_t)INT32_MAX;
  L.c = take(sizeof(double) * np);
  L.rowcnt = take(sizeof(int) * n);
  L.rowcur = take(sizeof(int) * n);
  L.rowptr = take(sizeof(int) * (n + 1));
  L.col = take(sizeof(int) * 2 * np);
  L.val = take(sizeof(double) * 2 * np);
  L.scan = take(sizeof(int) * (n / kBlock + 2));
  L.r = take(sizeof(double) * n);
  L.z = take(sizeof(double) * n);
  L.p = take(sizeof(double) * n);
  L.w = take(sizeof(double) * n);
  L.sc = take(sizeof(double) * kQbStride * P.nsys);
  L.done = take(sizeof(int) * P.nsys);
  L.bytes = off;
  return L;
}
}  // namespace

size_t qeq_batch_extra_bytes(const BatchPlan& P, int64_t n_pairs) {
  const QbLayout L = qb_layout(P, n_pairs);
  return L.ok ? L.bytes - P.bytes : 0;
}

int qeq_solve_batch(const BatchPlan& P, const QeqBatch& B, const double* q0, const double* tol,
                    int max_iter, int check_every, double* q, int32_t* iters, double* resid,
                    char* ws, hipStream_t s) {
  const QbLayout L = qb_layout(P, B.n_pairs);
  if (!L.ok) return ADMP_ERR_UNSUPPORTED;
  const int nsys = P.nsys;
  const int64_t n = P.ntot;
  for (int k = 0; k < nsys; ++k) {
    iters[k] = -1;
    resid[k] = 0.0;
  }
  // geometry: table, binning, cell order and mesh reference points
  int st = batch_geometry(P, B.pos, B.chi, ws, s);
  if (st != ADMP_OK) return st;
  const SysDesc* tab = (const SysDesc*)(ws + P.o_tab);
  const int64_t* aoff = (const int64_t*)(ws + P.o_off) + (int64_t)(nsys + 1) * kSegAtom;
  const int* sys_of = (const int*)(ws + P.o_sys_of);
  double* c = (double*)(ws + L.c);
  int* rowptr = (int*)(ws + L.rowptr);
  int* col = (int*)(ws + L.col);
  double* val = (double*)(ws + L.val);
  double* r = (double*)(ws + L.r);
  double* z = (double*)(ws + L.z);
  double* p = (double*)(ws + L.p);
  double* w = (double*)(ws + L.w);
  double* sc = (double*)(ws + L.sc);
  int* done = (int*)(ws + L.done);
  // pair rows (both orientations of every pair), built once
  const unsigned gp = (unsigned)std::max<int64_t>((B.n_pairs + kBlock - 1) / kBlock, 1);
  const unsigned gn = grid_for(n);
  if (B.n_pairs > 0)
    k_qb_pair_coeff<<<gp, kBlock, 0, s>>>(B.n_pairs, B.pairs, B.ds, B.bs, B.eta, sys_of, tab,
                                          B.damping, P.sys.empty() ? 0.0 : P.sys[0].eps, c);
  k_fill<int><<<gn, kBlock, 0, s>>>((int*)(ws + L.rowcnt), n, 0);
  k_fill<int><<<gn, kBlock, 0, s>>>((int*)(ws + L.rowcur), n, 0);
  if (B.n_pairs > 0)
    k_csr_count<<<gp, kBlock, 0, s>>>(B.n_pairs, B.pairs, c, (int*)(ws + L.rowcnt));
  launch_scan<int>((int*)(ws + L.rowcnt), n, rowptr, (int*)(ws + L.scan), s);
  if (B.n_pairs > 0)
    k_csr_fill<<<gp, kBlock, 0, s>>>(B.n_pairs, B.pairs, c, rowptr, (int*)(ws + L.rowcur), col,
                                     val);
  // per-system constants
  std::vector<double> sch((size_t)kQbStride * nsys, 0.0);
  for (int k = 0; k < nsys; ++k) {
    const SysDesc& D = P.sys[k];
    sch[(size_t)kQbStride * k + kQbCnn] = -D.eps * kPi / (D.g.vol * D.kappa * D.kappa);
    sch[(size_t)kQbStride * k + kQbQ] = B.total_charge[k];
    sch[(size_t)kQbStride * k + kQbTol] = tol[k];
  }
  if (hipMemcpyAsync(sc, sch.data(), sizeof(double) * sch.size(), hipMemcpyHostToDevice, s) !=
      hipSuccess)
    return ADMP_ERR_HIP;
  auto product = [&](const double* v, double* out) -> int {
    k_qb_rows<<<(unsigned)((16 * n + kBlock - 1) / kBlock), kBlock, 0, s>>>(
        n, rowptr, col, val, B.hdiag, v, sys_of, sc, out);
    return batch_potential(P, B.pos, v, out, ws, s);
  };
  // start: q feasible, r = -P (H q + chi), z, p
  k_qb_start<<<nsys, kPcgThreads, 0, s>>>(aoff, q0, q, sc);
  if ((st = product(q, w)) != ADMP_OK) return st;
  k_qb_init<<<nsys, kPcgThreads, 0, s>>>(aoff, w, B.chi, B.precond, r, z, p, sc, done);
  std::vector<int> flags(nsys);
  auto all_done = [&]() -> int {
    if (hipMemcpyAsync(flags.data(), done, sizeof(int) * nsys, hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return -1;
    for (int k = 0; k < nsys; ++k)
      if (flags[k] < 0) return 0;
    return 1;
  };
  int ad = all_done();
  if (ad < 0) return ADMP_ERR_HIP;
  for (int it = 1; it <= max_iter && !ad; ++it) {
    if ((st = product(p, w)) != ADMP_OK) return st;
    k_qb_step<<<nsys, kPcgThreads, 0, s>>>(aoff, q, r, p, w, z, B.precond, sc, done, it,
                                            (it & 1) ? kQbRz0 : kQbRz1,
                                            (it & 1) ? kQbRz1 : kQbRz0);
    if (it % check_every == 0 || it == max_iter) {
      ad = all_done();
      if (ad < 0) return ADMP_ERR_HIP;
    }
  }
  // flags and residuals
  if (hipMemcpyAsync(flags.data(), done, sizeof(int) * nsys, hipMemcpyDeviceToHost, s) !=
          hipSuccess ||
      hipMemcpyAsync(sch.data(), sc, sizeof(double) * sch.size(), hipMemcpyDeviceToHost, s) !=
          hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return ADMP_ERR_HIP;
  for (int k = 0; k < nsys; ++k) {
    const double nk = (double)std::max<int64_t>(P.sys[k].n, 1);
    iters[k] = flags[k];
    resid[k] = std::sqrt(std::max(sch[(size_t)kQbStride * k + kQbRr], 0.0)) / nk;
  }
  return hipGetLastError() == hipSuccess ? ADMP_OK : ADMP_ERR_HIP;
}

template int qeq_matvec_impl<float>(const QeqProblem&, const double*, double*, void*, hipStream_t);
template int qeq_matvec_impl<double>(const QeqProblem&, const double*, double*, void*, hipStream_t);
template int qeq_solve_impl<float>(const QeqProblem&, const double*, double, int, int, double*,
                                   int*, double*, void*, hipStream_t);
template int qeq_solve_impl<double>(const QeqProblem&, const double*, double, int, int, double*,
                                    int*, double*, void*, hipStream_t);

}  // namespace admp
