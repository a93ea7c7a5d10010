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
                                                 double* __restrict__ partials) {
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
  int ok = 1;
  for (int col = 0; col < nc; ++col) {
    int piv = col;
    for (int r = col + 1; r < nc; ++r)
      if (fabs(a[r][col]) > fabs(a[piv][col])) piv = r;
    if (fabs(a[piv][col]) < 1e-300) {
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
  const double Q = sums[0], M = sums[1], S = sums[2];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double v = c_nn * Q;
    if (slab) {
      const double zi = z[i];
      v += pre_slab * (2.0 * M * zi - S - Q * zi * zi - Q * lz2 / 6.0);
    }
    out[i] += v;
  }
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
  L.r = take(vec);
  L.z = take(vec);
  L.p = take(vec);
  L.w = take(vec);
  L.s = take(vec);
  L.zc = take(vec);
  L.dots = take(sizeof(double) * kRedBlocks * std::max(kQeqMaxDots, kQeqMaxCons * kQeqMaxCons));
  L.sc = take(sizeof(double) * 256);  // device scalars, see kSc* below
  L.status = take(sizeof(int) * 8);
  const size_t es = P.precision == ADMP_F32 ? 4 : 8;
  if (P.kspace) {
    L.cg = make_grid(P.box, P.cutoff, false);
    nb_layout(P.n, L.cg, &L.nb);
    L.nbws = take(L.nb.bytes);
    L.tg = make_tile_geom(L.cg, P.kmesh, es);
    L.mesh = take((size_t)P.kmesh[0] * P.kmesh[1] * P.kmesh[2] * es);
    L.spec = take((size_t)P.kmesh[0] * P.kmesh[1] * (P.kmesh[2] / 2 + 1) * 2 * es);
    L.ppart = take(sizeof(double) * (std::max(convolve_partials_count(P.kmesh), (size_t)kRedBlocks * 9) + 16));
  } else {
    L.tg.ok = 0;
    L.nbws = L.mesh = L.spec = L.ppart = 0;
  }
  L.bytes = off;
  return L;
}

// sc layout (doubles): 0 rz (even iterations), 1 rz (odd), 16.. reduced dots,
// 32.. G (kQeqMaxCons^2), 112.. A q / A g, 128.. global sums, 160 |r|^2
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

template <typename T>
static hipError_t matvec(const QeqProblem& P, const QeqLayout& L, char* ws, const Geom& g,
                         const double* v, double* out, hipStream_t s) {
  const int64_t n = P.n;
  k_diag_mul<<<grid_for(n), kBlock, 0, s>>>(n, P.hdiag, v, out);
  if (P.n_pairs > 0)
    k_pair_matvec<<<(unsigned)((P.n_pairs + kBlock - 1) / kBlock), kBlock, 0, s>>>(
        P.n_pairs, P.pairs, (const double*)(ws + L.c), v, out);
  if (P.kspace) {
    T* mesh = (T*)(ws + L.mesh);
    hipError_t e =
        L.tg.ok ? launch_spread_tiled<T>(P.pos, v, n, g, L.tg, L.cg.ncell,
                                         (const int*)(ws + L.nbws + L.nb.sorted),
                                         (const int*)(ws + L.nbws + L.nb.start), mesh, s)
                : launch_spread<T>(P.pos, v, n, g, mesh, s);
    if (e != hipSuccess) return e;
    e = launch_convolve<T>(mesh, g, P.convention, P.kappa, P.eps, 1.0, mesh, nullptr, nullptr,
                           (cplx<T>*)(ws + L.spec), (double*)(ws + L.ppart), s);
    if (e != hipSuccess) return e;
    e = launch_gather<T>(P.pos, v, n, g, mesh, nullptr, out, nullptr, 1,
                         (double*)(ws + L.ppart), s);
    if (e != hipSuccess) return e;
  }
  if (P.kspace || P.slab) {
    double* sc = (double*)(ws + L.sc);
    const unsigned gb = grid_for(n);
    k_dots<<<gb, kBlock, 0, s>>>(n, v, nullptr, nullptr, 0, (const double*)(ws + L.zc),
                                 (double*)(ws + L.dots));
    // slots 1..3 hold (sum v, sum v z, sum v z^2) -> sc[kScSums..]
    k_reduce_cols<<<1, kReduceThreads, 0, s>>>((double*)(ws + L.dots), gb, kQeqMaxDots, 4,
                                               sc + kScSums - 1);
    const double c_nn = (P.kspace && P.kappa > 0.0) ? -P.eps * kPi / (g.vol * P.kappa * P.kappa) : 0.0;
    const double pre_slab = 2.0 * kPi / g.vol * P.eps;
    k_global_terms<<<grid_for(n), kBlock, 0, s>>>(n, (const double*)(ws + L.zc), sc + kScSums, c_nn,
                                                  pre_slab, P.lz2, P.slab, out);
  }
  return hipGetLastError();
}

// dots of (x, y) and A x into sc[dst..]: dst = x.y, dst+1.. = A x
static void dots_into(const QeqProblem& P, const QeqLayout& L, char* ws, const double* x,
                      const double* y, int dst, hipStream_t s) {
  const unsigned gb = grid_for(P.n);
  double* sc = (double*)(ws + L.sc);
  k_dots<<<gb, kBlock, 0, s>>>(P.n, x, y, P.cons, P.n_cons, nullptr, (double*)(ws + L.dots));
  k_reduce_cols<<<1, kReduceThreads, 0, s>>>((double*)(ws + L.dots), gb, kQeqMaxDots,
                                             1 + P.n_cons, sc + dst);
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
  if (P.kspace && launch_bin(P.pos, P.n, L.cg, ws + L.nbws, L.nb, s) != hipSuccess)
    return ADMP_ERR_HIP;
  if (P.slab || P.kspace)
    k_extract_z<<<grid_for(P.n), kBlock, 0, s>>>(P.n, P.pos, P.slab ? P.slab_axis : 0,
                                                  (double*)(ws + L.zc));
  return hipGetLastError() == hipSuccess ? ADMP_OK : ADMP_ERR_HIP;
}

template <typename T>
int qeq_matvec_impl(const QeqProblem& P, const double* v, double* out, void* workspace,
                    hipStream_t s) {
  const QeqLayout L = qeq_layout(P);
  char* ws = (char*)workspace;
  Geom g;
  int st = setup<T>(P, L, ws, &g, s);
  if (st != ADMP_OK) return st;
  if (matvec<T>(P, L, ws, g, v, out, s) != hipSuccess) return ADMP_ERR_HIP;
  return hipGetLastError() == hipSuccess ? ADMP_OK : ADMP_ERR_HIP;
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
  // residual of the starting point
  dots_into(P, L, ws, r, r, kScRr, s);
  if (hipMemcpyAsync(&res, sc + kScRr, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return ADMP_ERR_HIP;
  res = std::sqrt(std::max(res, 0.0)) / (double)std::max<int64_t>(n, 1);
  converged = res < tol;
  while (!converged && it < max_iter) {
    ++it;
    const int rz_old = (it & 1) ? kScRz0 : kScRz1;
    const int rz_new = (it & 1) ? kScRz1 : kScRz0;
    if (matvec<T>(P, L, ws, g, p, w, s) != hipSuccess) return ADMP_ERR_HIP;
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

template int qeq_matvec_impl<float>(const QeqProblem&, const double*, double*, void*, hipStream_t);
template int qeq_matvec_impl<double>(const QeqProblem&, const double*, double*, void*, hipStream_t);
template int qeq_solve_impl<float>(const QeqProblem&, const double*, double, int, int, double*,
                                   int*, double*, void*, hipStream_t);
template int qeq_solve_impl<double>(const QeqProblem&, const double*, double, int, int, double*,
                                    int*, double*, void*, hipStream_t);

}  // namespace admp
