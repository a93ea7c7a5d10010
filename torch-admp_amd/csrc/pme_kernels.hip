// Smooth-PME reciprocal-space kernels for gfx950: order-6 spread, the 3D
// real-to-complex FFT (LDS Stockham passes), the influence-function
// convolution with energy / dE/dA reduction, and back-interpolation.
//
// Replaces the torch-op chains of src/torch_admp/recip.py:13-252 (spread),
// torch.fft.fftn at src/torch_admp/pme.py:434, the k-space sum of
// pme.py:410-443 and 661-683, and the autograd transpose of the spread.
#include <atomic>
#include "fft_device.hpp"
#include "mesh_device.hpp"
#include "pme_kernels.hpp"

namespace admp {

// ---------------------------------------------------------------------------
// Spread: Q[(m0 - 3 + a, m0 - 3 + b, m0 - 3 + c) mod K] += q wx[a] wy[b] wz[c].
// Eight lanes per atom; lane a < 6 owns the x offset a and issues 36 adds.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_spread(const double* __restrict__ pos,
                                                   const double* __restrict__ q, int64_t n,
                                                   Geom g, T* __restrict__ mesh) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t atom = t >> 3;
  const int a = (int)(t & 7);
  if (atom >= n || a >= kOrder) return;
  spread_lane<T>(g, pos[3 * atom], pos[3 * atom + 1], pos[3 * atom + 2], q[atom], a, mesh);
}

// ---------------------------------------------------------------------------
// Gather: for a mesh field phi = dE/dQ,
//   dE/dR_d = -q sum_g phi(g) M6'(u_d) prod_{e != d} M6(u_e)   (u = f + j, du/dR = -1)
//   dE/dr_j = sum_d K_d A[j][d] dE/dR_d,  dE/dq = sum_g phi(g) theta(g)
//   dE/dA[j][d] = sum_atoms dE/dR_d K_d r_j  (per-block partials, 9 values)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_gather(const double* __restrict__ pos,
                                                   const double* __restrict__ q, int64_t n,
                                                   Geom g, const T* __restrict__ phi,
                                                   double* __restrict__ grad_pos,
                                                   double* __restrict__ dEdq,
                                                   double* __restrict__ partials, int accumulate) {
  __shared__ double red[9 * (kBlock / 64)];
  const int a = threadIdx.x & 7;
  const int64_t groups_per_grid = (int64_t)gridDim.x * (blockDim.x >> 3);
  double acc[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) acc[i] = 0.0;
  for (int64_t atom = (int64_t)blockIdx.x * (blockDim.x >> 3) + (threadIdx.x >> 3); atom < n;
       atom += groups_per_grid) {
    const double x = pos[3 * atom], y = pos[3 * atom + 1], z = pos[3 * atom + 2];
    T sx, sy, sz, sp;
    gather_lane<T>(g, x, y, z, a, phi, &sx, &sy, &sz, &sp);
    // reduce across the 8 lanes of this atom
#pragma unroll
    for (int off = 1; off < 8; off <<= 1) {
      sx += __shfl_xor(sx, off, 64);
      sy += __shfl_xor(sy, off, 64);
      sz += __shfl_xor(sz, off, 64);
      sp += __shfl_xor(sp, off, 64);
    }
    if (a == 0) {
      const double qi = q[atom];
      const double dR[3] = {-qi * (double)sx, -qi * (double)sy, -qi * (double)sz};
      if (grad_pos) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          double gj = 0.0;
#pragma unroll
          for (int d = 0; d < 3; ++d) gj += g.k[d] * g.a[j * 3 + d] * dR[d];
          if (accumulate) grad_pos[3 * atom + j] += gj;
          else grad_pos[3 * atom + j] = gj;
        }
      }
      if (dEdq) {
        if (accumulate) dEdq[atom] += (double)sp;
        else dEdq[atom] = (double)sp;
      }
      const double r[3] = {x, y, z};
#pragma unroll
      for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int d = 0; d < 3; ++d) acc[j * 3 + d] += dR[d] * g.k[d] * r[j];
    }
  }
  if (partials) {
    block_sum<9>(acc, red);
    if (threadIdx.x == 0)
      for (int i = 0; i < 9; ++i) partials[blockIdx.x * 9 + i] = acc[i];
  }
}

// ---------------------------------------------------------------------------
// FFT passes of one system (bodies in fft_device.hpp)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_fft_r2c_z(const T* __restrict__ in,
                                                      cplx<T>* __restrict__ out, int64_t nrows,
                                                      int L, int C, Radix rd, int pitch) {
  r2c_rows<T>([=](int64_t row, int z) { return in[row * L + z]; }, out, nrows,
              (int64_t)blockIdx.x * 2 * C, L, C, rd, false, pitch);
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_fft_c2r_z(const cplx<T>* __restrict__ in,
                                                      T* __restrict__ out, int64_t nrows, int L,
                                                      int C, Radix rd, int pitch) {
  c2r_rows<T>(in, out, nrows, (int64_t)blockIdx.x * 2 * C, L, C, rd, false, pitch);
}

template <typename T, int SIGN>
__global__ void __launch_bounds__(kBlock) k_fft_cols(cplx<T>* __restrict__ data, int L,
                                                     int64_t stride, int n_inner,
                                                     int64_t outer_stride, int C, Radix rd) {
  strided_lines<T, SIGN>(data + (int64_t)blockIdx.y * outer_stride + (int64_t)blockIdx.x * C, L,
                         stride, min(C, n_inner - (int)blockIdx.x * C), C, rd);
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_fft_x_convolve(cplx<T>* __restrict__ data,
                                                           int C, Radix rd, InflParams P,
                                                           double* __restrict__ partials,
                                                           int pitch) {
  x_convolve_lines<T>(data, blockIdx.y, blockIdx.x * C, C, rd, P,
                      partials ? partials + ((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * 10
                               : nullptr,
                      pitch);
}

// ---------------------------------------------------------------------------
// Passes 1-2 (and 4-5) for one x slab whose half spectrum K1 x (K2/2 + 1)
// fits in LDS: one work-group per slab packs real rows 2p and 2p + 1 into
// complex line p, transforms along z, untangles both rows' half spectra in
// place (line p holds row 2p at [0, Kzh) and row 2p + 1 at [Kzh, 2 Kzh), so
// the buffer becomes the [y][kz] plane with pitch Kzh), transforms along y
// and stores the slab.  The inverse runs the same steps backwards, with
// Z_k = A_k + i B_k and A, B extended by Hermitian symmetry (c2r_rows).
// Same arithmetic as r2c_rows / strided_lines / c2r_rows.
constexpr int kYzThreads = 1024;
constexpr int kYzPack = 8;   // packed elements per thread in the inverse (checked when planning)
constexpr int kYzSplit = 4;  // untangled element pairs per thread in the forward

template <typename T>
__global__ void __launch_bounds__(kYzThreads) k_fft_yz_fwd(const T* __restrict__ mesh,
                                                            cplx<T>* __restrict__ spec, int K1,
                                                            int K2, Radix ry, Radix rz) {
  using CT = cplx<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int Kzh = K2 / 2 + 1, np = K1 / 2, pitch2 = 2 * Kzh;
  CT* buf = reinterpret_cast<CT*>(smem_raw);
  CT* twz = buf + (size_t)K1 * Kzh;
  CT* twy = twz + K2;
  build_twiddles<T>(twz, K2, blockDim.x);
  build_twiddles<T>(twy, K1, blockDim.x);
  const T* src = mesh + (int64_t)blockIdx.x * K1 * K2;
  for (int idx = threadIdx.x; idx < np * K2; idx += blockDim.x) {
    int pr, z;
    divmod_u(idx, K2, pr, z);
    CT v;
    v.x = src[(int64_t)(2 * pr) * K2 + z];
    v.y = src[(int64_t)(2 * pr + 1) * K2 + z];
    buf[pr * pitch2 + z] = v;
  }
  __syncthreads();
  stockham_ip<T, -1>(buf, K2, np, 1, pitch2, rz, twz);
  // untangle: row 2p <- (Z_k + conj Z_-k) / 2, row 2p + 1 <- (Z_k - conj Z_-k) / 2i
  CT ea[kYzSplit], eb[kYzSplit];
#pragma unroll
  for (int u = 0; u < kYzSplit; ++u) {
    const int idx = threadIdx.x + u * blockDim.x;
    if (idx < np * Kzh) {
      int pr, k;
      divmod_u(idx, Kzh, pr, k);
      const CT zk = buf[pr * pitch2 + k];
      const CT zm = buf[pr * pitch2 + (k ? K2 - k : 0)];
      ea[u].x = T(0.5) * (zk.x + zm.x);
      ea[u].y = T(0.5) * (zk.y - zm.y);
      eb[u].x = T(0.5) * (zk.y + zm.y);
      eb[u].y = T(-0.5) * (zk.x - zm.x);
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kYzSplit; ++u) {
    const int idx = threadIdx.x + u * blockDim.x;
    if (idx < np * Kzh) {
      int pr, k;
      divmod_u(idx, Kzh, pr, k);
      buf[pr * pitch2 + k] = ea[u];
      buf[pr * pitch2 + Kzh + k] = eb[u];
    }
  }
  __syncthreads();
  stockham_ip<T, -1>(buf, K1, Kzh, Kzh, 1, ry, twy);  // along y: lines kz
  CT* dst = spec + (int64_t)blockIdx.x * K1 * Kzh;
  for (int idx = threadIdx.x; idx < K1 * Kzh; idx += blockDim.x) dst[idx] = buf[idx];
}

template <typename T>
__global__ void __launch_bounds__(kYzThreads) k_fft_yz_inv(const cplx<T>* __restrict__ spec,
                                                            T* __restrict__ mesh, int K1, int K2,
                                                            Radix ry, Radix rz) {
  using CT = cplx<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int Kzh = K2 / 2 + 1, np = K1 / 2, pitch2 = 2 * Kzh;
  CT* buf = reinterpret_cast<CT*>(smem_raw);
  CT* twz = buf + (size_t)K1 * Kzh;
  CT* twy = twz + K2;
  build_twiddles<T>(twz, K2, blockDim.x);
  build_twiddles<T>(twy, K1, blockDim.x);
  const CT* src = spec + (int64_t)blockIdx.x * K1 * Kzh;
  for (int idx = threadIdx.x; idx < K1 * Kzh; idx += blockDim.x) buf[idx] = src[idx];
  __syncthreads();
  stockham_ip<T, 1>(buf, K1, Kzh, Kzh, 1, ry, twy);  // inverse along y
  // pack: Z_k = A_k + i B_k for k < K2 (A, B of rows 2p, 2p + 1, Hermitian-extended)
  CT ez[kYzPack];
#pragma unroll
  for (int u = 0; u < kYzPack; ++u) {
    const int idx = threadIdx.x + u * blockDim.x;
    if (idx < np * K2) {
      int pr, k;
      divmod_u(idx, K2, pr, k);
      const int kk = k < Kzh ? k : K2 - k;
      const T sg = k < Kzh ? T(1) : T(-1);
      CT A = buf[pr * pitch2 + kk], B = buf[pr * pitch2 + Kzh + kk];
      A.y *= sg;
      B.y *= sg;
      ez[u].x = A.x - B.y;
      ez[u].y = A.y + B.x;
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kYzPack; ++u) {
    const int idx = threadIdx.x + u * blockDim.x;
    if (idx < np * K2) {
      int pr, k;
      divmod_u(idx, K2, pr, k);
      buf[pr * pitch2 + k] = ez[u];
    }
  }
  __syncthreads();
  stockham_ip<T, 1>(buf, K2, np, 1, pitch2, rz, twz);  // inverse along z
  T* dst = mesh + (int64_t)blockIdx.x * K1 * K2;
  for (int idx = threadIdx.x; idx < np * K2; idx += blockDim.x) {
    int pr, z;
    divmod_u(idx, K2, pr, z);
    const CT v = buf[pr * pitch2 + z];
    dst[(int64_t)(2 * pr) * K2 + z] = v.x;
    dst[(int64_t)(2 * pr + 1) * K2 + z] = v.y;
  }
}

// ---------------------------------------------------------------------------
// host-side launchers

Radix factorize(int L) {
  Radix rd;
  rd.n = 0;
  int n = L;
  const int pref[5] = {8, 4, 2, 3, 5};
  for (int pi = 0; pi < 5; ++pi) {
    const int r = pref[pi];
    while (n % r == 0 && rd.n < kMaxPasses) {
      rd.r[rd.n++] = r;
      n /= r;
    }
  }
  for (int p = 7; n > 1 && rd.n < kMaxPasses; p += 2) {
    while (n % p == 0 && rd.n < kMaxPasses) {
      rd.r[rd.n++] = p;
      n /= p;
    }
  }
  if (n != 1) rd.n = -1;
  int p = 1;
  for (int s = 0; s < rd.n; ++s) {
    rd.mp[s] = div_magic((uint32_t)p);
    rd.mnb[s] = div_magic((uint32_t)(L / rd.r[s]));
    p *= rd.r[s];
  }
  rd.mlen = div_magic((uint32_t)L);
  rd.mhalf = div_magic((uint32_t)(L / 2 + 1));
  return rd;
}

bool radix_ok(const Radix& rd) { return rd.n >= 0; }

static int pick_chunk(int L, int64_t lines, size_t elem) {
  // lines per block: keep 2 * L * C complex plus the twiddles within 64 KiB
  int C = 16;
  while (C > 1 && (size_t)(2 * L * C + L) * elem > 65536) C >>= 1;
  while (C > 1 && lines / C < 512) C >>= 1;
  return C;
}

// Zero n 16-byte words of a 16-byte aligned buffer.
__global__ void __launch_bounds__(kBlock) k_zero16(int4* __restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = make_int4(0, 0, 0, 0);
}

template <typename T>
hipError_t launch_spread(const double* pos, const double* q, int64_t n, const Geom& g, T* mesh,
                         hipStream_t s) {
  const int64_t nm = (int64_t)g.k[0] * g.k[1] * g.k[2];
  // 16-byte stores over the whole words of an aligned mesh, element stores for the rest
  const int64_t n16 = (reinterpret_cast<uintptr_t>(mesh) & 15) ? 0 : nm * (int64_t)sizeof(T) / 16;
  if (n16 > 0)
    k_zero16<<<(int)std::min<int64_t>((n16 + kBlock - 1) / kBlock, 8192), kBlock, 0, s>>>(
        reinterpret_cast<int4*>(mesh), n16);
  const int64_t done = n16 * 16 / (int64_t)sizeof(T);
  if (done < nm)
    k_fill<T><<<(int)std::min<int64_t>((nm - done + kBlock - 1) / kBlock, 4096), kBlock, 0, s>>>(
        mesh + done, nm - done, T(0));
  if (n > 0) {
    const int64_t threads = n * 8;
    k_spread<T><<<(unsigned)((threads + kBlock - 1) / kBlock), kBlock, 0, s>>>(pos, q, n, g, mesh);
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_gather(const double* pos, const double* q, int64_t n, const Geom& g,
                         const T* phi, double* grad_pos, double* dEdq, double* dEdA,
                         int accumulate, double* partials, hipStream_t s) {
  int blocks = (int)std::min<int64_t>((n * 8 + kBlock - 1) / kBlock, kRedBlocks);
  if (blocks < 1) blocks = 1;
  k_gather<T><<<blocks, kBlock, 0, s>>>(pos, q, n, g, phi, grad_pos, dEdq,
                                        dEdA ? partials : nullptr, accumulate);
  if (dEdA) k_reduce_partials<<<1, kReduceThreads, 0, s>>>(partials, blocks, 9, 1.0, dEdA, dEdA + 1);
  return hipGetLastError();
}

// Dynamic LDS bytes of the slab kernels for this mesh, or 0 when they do not
// apply: even K1 and K2, radices 2/3/4/5/8, at most ip_max_b(r) butterflies
// and kYzPack / kYzSplit elements per thread, and the device's per-work-group
// LDS.
static size_t yz_lds(int K1, int K2, size_t elem) {
  if ((K1 & 1) || (K2 & 1)) return 0;
  const Radix ry = factorize(K1), rz = factorize(K2);
  if (!radix_ok(ry) || !radix_ok(rz)) return 0;
  const int Kzh = K2 / 2 + 1, np = K1 / 2;
  auto fits = [](const Radix& rd, int L, int C) {
    for (int i = 0; i < rd.n; ++i) {
      const int r = rd.r[i];
      if (r != 2 && r != 3 && r != 4 && r != 5 && r != 8) return false;
      if ((int64_t)(L / r) * C > (int64_t)ip_max_b(r) * kYzThreads) return false;
    }
    return true;
  };
  if (!fits(rz, K2, np) || !fits(ry, K1, Kzh)) return 0;
  if ((int64_t)np * K2 > (int64_t)kYzPack * kYzThreads) return 0;
  if ((int64_t)np * Kzh > (int64_t)kYzSplit * kYzThreads) return 0;
  const size_t lds = ((size_t)K1 * Kzh + K1 + K2) * elem;
  static std::atomic<int> max_lds[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
  int lim = max_lds[dev].load(std::memory_order_relaxed);
  if (lim == 0) {
    if (hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
            hipSuccess || lim <= 0)
      return 0;
    max_lds[dev].store(lim, std::memory_order_relaxed);
  }
  return lds <= (size_t)lim ? lds : 0;
}

template <typename T>
hipError_t launch_convolve(const T* mesh, const Geom& g, int conv, double kappa, double eps,
                           double phi_scale, T* phi, double* energy, double* dEdA,
                           cplx<T>* spec, double* partials, hipStream_t s) {
  using CT = cplx<T>;
  const int K0 = g.k[0], K1 = g.k[1], K2 = g.k[2];
  const int Kzh = K2 / 2 + 1;
  const Radix r0 = factorize(K0), r1 = factorize(K1), r2 = factorize(K2);
  const int64_t nrows = (int64_t)K0 * K1;
  const size_t yz = yz_lds(K1, K2, sizeof(CT));
  // the slab kernels store whole [y][kz] slabs (pitch Kzh); the line passes use
  // rows padded to 128-byte lines
  const int Kzp = yz ? Kzh : spec_row_pitch(K2, sizeof(CT));
  if (yz) {
    // passes 1-2 in one work-group per x slab
    k_fft_yz_fwd<T><<<K0, kYzThreads, yz, s>>>(mesh, spec, K1, K2, r1, r2);
  } else {
  // pass 1: r2c along z
  {
    const int C = pick_chunk(K2, (nrows + 1) / 2, sizeof(CT));
    const unsigned blocks = (unsigned)((nrows + 2 * C - 1) / (2 * C));
    const size_t sh = (size_t)(K2 + 2 * K2 * C) * sizeof(CT);
    k_fft_r2c_z<T><<<blocks, kBlock, sh, s>>>(mesh, spec, nrows, K2, C, r2, Kzp);
  }
  // pass 2: forward along y
  {
    const int C = pick_chunk(K1, (int64_t)K0 * Kzh, sizeof(CT));
    dim3 grid((Kzh + C - 1) / C, K0);
    const size_t sh = (size_t)(K1 + 2 * K1 * C) * sizeof(CT);
    k_fft_cols<T, -1><<<grid, kBlock, sh, s>>>(spec, K1, Kzp, Kzh, (int64_t)K1 * Kzp, C, r1);
  }
  }
  // pass 3: along x with the influence function
  InflParams P;
  P.g = g;
  P.kappa = kappa;
  P.phi_scale = phi_scale;
  P.eps = eps;
  P.conv = conv;
  P.ortho = (g.h[1] == 0.0 && g.h[2] == 0.0 && g.h[3] == 0.0 && g.h[5] == 0.0 && g.h[6] == 0.0 &&
             g.h[7] == 0.0) ? 1 : 0;
  const int C3 = pick_chunk(K0, (int64_t)K1 * Kzh, sizeof(CT));
  dim3 grid3((Kzh + C3 - 1) / C3, K1);
  {
    const size_t sh = (size_t)(K0 + 2 * K0 * C3) * sizeof(CT);
    k_fft_x_convolve<T><<<grid3, kBlock, sh, s>>>(spec, C3, r0, P,
                                                  (energy || dEdA) ? partials : nullptr, Kzp);
  }
  const int nb3 = grid3.x * grid3.y;
  if (energy || dEdA)
    k_reduce_partials<<<1, kReduceThreads, 0, s>>>(partials, nb3, 10, eps, energy, dEdA);
  if (phi && yz) {
    // passes 4-5 in one work-group per x slab
    k_fft_yz_inv<T><<<K0, kYzThreads, yz, s>>>(spec, phi, K1, K2, r1, r2);
  } else if (phi) {
    // pass 4: inverse along y
    {
      const int C = pick_chunk(K1, (int64_t)K0 * Kzh, sizeof(CT));
      dim3 grid((Kzh + C - 1) / C, K0);
      const size_t sh = (size_t)(K1 + 2 * K1 * C) * sizeof(CT);
      k_fft_cols<T, 1><<<grid, kBlock, sh, s>>>(spec, K1, Kzp, Kzh, (int64_t)K1 * Kzp, C, r1);
    }
    // pass 5: c2r along z
    {
      const int C = pick_chunk(K2, (nrows + 1) / 2, sizeof(CT));
      const unsigned blocks = (unsigned)((nrows + 2 * C - 1) / (2 * C));
      const size_t sh = (size_t)(K2 + 2 * K2 * C) * sizeof(CT);
      k_fft_c2r_z<T><<<blocks, kBlock, sh, s>>>(spec, phi, nrows, K2, C, r2, Kzp);
    }
  }
  return hipGetLastError();
}

// dE/dbox of the reciprocal energy from its two dE/dA parts (k vectors and
// the gather's s = r A term, 9 values each, row-major) and the energy (the
// volume term): with A = box^-1, dE/dbox = -A^T (dA1 + dA2) A^T - E A^T.  One
// thread; every input is on the device, so no stream synchronisation.
__global__ void k_box_grad(Geom g, const double* __restrict__ dA1, const double* __restrict__ dA2,
                           const double* __restrict__ energy, double* __restrict__ out) {
  if (threadIdx.x != 0) return;
  double d[9];
  for (int k = 0; k < 9; ++k) d[k] = dA1[k] + (dA2 ? dA2[k] : 0.0);
  const double e = energy[0];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      // (A^T d A^T)[r][c] = sum_kl A[k][r] d[k][l] A[c][l]
      double acc = 0.0;
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) acc += g.a[3 * k + r] * d[3 * k + l] * g.a[3 * c + l];
      out[3 * r + c] = -acc - e * g.a[3 * c + r];
    }
}

hipError_t launch_box_grad(const Geom& g, const double* dA1, const double* dA2,
                           const double* energy, double* out, hipStream_t s) {
  k_box_grad<<<1, 64, 0, s>>>(g, dA1, dA2, energy, out);
  return hipGetLastError();
}

// Self and non-neutral Ewald terms (src/torch_admp/pme.py:448-493) of nf
// frames of n charges with device boxes (row-major, 9 per frame):
//   e_self = -eps kappa / sqrt(pi) sum q^2,  e_nn = -eps pi Q^2 / (2 V kappa^2),
// Q = sum q, V = det(box).  With g_self / g_nn given (the upstream gradients,
// either may be NULL for zero) the same kernel writes the backward instead:
//   dE/dq_i = g_self (-2 eps kappa / sqrt(pi)) q_i + g_nn (-eps pi Q / (V kappa^2)),
//   dE/dbox = g_nn (-e_nn / V) cof(box)   (d det / d box = the cofactor matrix).
// One work-group of kSelfThreads per frame; the two sums in a fixed order.
constexpr int kSelfThreads = 1024;
__global__ void __launch_bounds__(kSelfThreads)
    k_self_terms(const double* __restrict__ q, int64_t n, const double* __restrict__ box,
                 double kappa, double eps, double* __restrict__ e_self,
                 double* __restrict__ e_nn, const double* __restrict__ g_self,
                 const double* __restrict__ g_nn, double* __restrict__ dq,
                 double* __restrict__ dbox) {
  __shared__ double red[2 * (kSelfThreads / 64)];
  __shared__ double bcast[2];
  const int64_t f = blockIdx.x;
  const double* qf = q + f * n;
  double acc[2] = {0.0, 0.0};
#pragma unroll 4
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    const double v = qf[i];
    acc[0] += v * v;
    acc[1] += v;
  }
  block_sum<2>(acc, red);
  const double* h = box + 9 * f;
  double cof[9];
  cof[0] = h[4] * h[8] - h[5] * h[7];
  cof[1] = h[5] * h[6] - h[3] * h[8];
  cof[2] = h[3] * h[7] - h[4] * h[6];
  cof[3] = h[2] * h[7] - h[1] * h[8];
  cof[4] = h[0] * h[8] - h[2] * h[6];
  cof[5] = h[1] * h[6] - h[0] * h[7];
  cof[6] = h[1] * h[5] - h[2] * h[4];
  cof[7] = h[2] * h[3] - h[0] * h[5];
  cof[8] = h[0] * h[4] - h[1] * h[3];
  const double vol = h[0] * cof[0] + h[1] * cof[1] + h[2] * cof[2];
  const double cs = -eps * kappa / kSqrtPi;
  const double cn = -eps * kPi / (2.0 * vol * kappa * kappa);
  if (threadIdx.x == 0) {
    bcast[0] = acc[0];
    bcast[1] = acc[1];
  }
  __syncthreads();
  const double Q = bcast[1];
  if (!dq) {
    if (threadIdx.x == 0) {
      e_self[f] = cs * bcast[0];
      e_nn[f] = cn * Q * Q;
    }
    return;
  }
  const double gs = g_self ? g_self[f] : 0.0, gn = g_nn ? g_nn[f] : 0.0;
  const double a = gs * 2.0 * cs, b = gn * 2.0 * cn * Q;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) dq[f * n + i] = a * qf[i] + b;
  if (dbox && threadIdx.x < 9) dbox[9 * f + threadIdx.x] = gn * (-(cn * Q * Q) / vol) * cof[threadIdx.x];
}

hipError_t launch_self_terms(const double* q, int64_t nf, int64_t n, const double* box,
                             double kappa, double eps, double* e_self, double* e_nn,
                             const double* g_self, const double* g_nn, double* dq, double* dbox,
                             hipStream_t s) {
  if (nf > 0)
    k_self_terms<<<(unsigned)nf, kSelfThreads, 0, s>>>(q, n, box, kappa, eps, e_self, e_nn,
                                                       g_self, g_nn, dq, dbox);
  return hipGetLastError();
}

size_t convolve_partials_count(const int32_t* k) {
  const int Kzh = k[2] / 2 + 1;
  // the chunk picked for pass 3 is at least 1, so this bounds the block count
  return (size_t)Kzh * k[1] * 10 + 16;
}

template hipError_t launch_spread<float>(const double*, const double*, int64_t, const Geom&,
                                         float*, hipStream_t);
template hipError_t launch_spread<double>(const double*, const double*, int64_t, const Geom&,
                                          double*, hipStream_t);
template hipError_t launch_gather<float>(const double*, const double*, int64_t, const Geom&,
                                         const float*, double*, double*, double*, int, double*,
                                         hipStream_t);
template hipError_t launch_gather<double>(const double*, const double*, int64_t, const Geom&,
                                          const double*, double*, double*, double*, int,
                                          double*, hipStream_t);
template hipError_t launch_convolve<float>(const float*, const Geom&, int, double, double,
                                           double, float*, double*, double*, float2*, double*,
                                           hipStream_t);
template hipError_t launch_convolve<double>(const double*, const Geom&, int, double, double,
                                            double, double*, double*, double*, double2*,
                                            double*, hipStream_t);

}  // namespace admp
