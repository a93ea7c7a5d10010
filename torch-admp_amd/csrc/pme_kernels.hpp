// Internal declarations shared by the kernel files and the C-ABI layer.
#pragma once

#include <algorithm>

#include "admp_common.hpp"

namespace admp {

constexpr int kMaxPasses = 12;  // a length <= kMaxMesh has at most 7 factors
constexpr int kMaxMesh = 1024;  // largest mesh dimension (LDS line limit)

// Row pitch of a half spectrum whose rows are read in chunks of consecutive
// kz (line layout): K2/2 + 1 complex elements of celem bytes rounded up to whole
// 128-byte lines, so that no chunk straddles two cache lines.
inline int spec_row_pitch(int K2, size_t celem) {
  const int e = (int)(128 / celem);
  return (K2 / 2 + 1 + e - 1) / e * e;
}
// Complex elements of a K0 x K1 x K2 mesh's half spectrum at that pitch.
inline size_t spec_elems(const int32_t* k, size_t celem) {
  return (size_t)k[0] * k[1] * spec_row_pitch(k[2], celem);
}

// Factorisation of an FFT length into Stockham passes: radices 8, 4, 2, 3, 5
// first, then any remaining primes (handled by a direct-sum pass).  With it
// the multipliers for dividing by the pass strides without a division
// instruction: q = mulhi(x, m) = floor(x / d) with m = floor(2^32 / d) + 1,
// exact for x, d < 2^16 (x m / 2^32 exceeds x / d by at most 2^-16 < 1/d);
// m = 0 stands for d = 1.
struct Radix {
  int n;
  int r[kMaxPasses];
  uint32_t mp[kMaxPasses];   // for p = r[0] ... r[s - 1], the product before pass s
  uint32_t mnb[kMaxPasses];  // for L / r[s], the butterflies of pass s
  uint32_t mlen;             // for L itself
  uint32_t mhalf;            // for L / 2 + 1
};
inline uint32_t div_magic(uint32_t d) {
  return d <= 1 ? 0u : (uint32_t)((((uint64_t)1) << 32) / d + 1);
}

Radix factorize(int L);
bool radix_ok(const Radix& rd);

template <typename T>
hipError_t launch_spread(const double* pos, const double* q, int64_t n, const Geom& g, T* mesh,
                         hipStream_t s);

template <typename T>
hipError_t launch_gather(const double* pos, const double* q, int64_t n, const Geom& g,
                         const T* phi, double* grad_pos, double* dEdq, double* dEdA,
                         int accumulate, double* partials, hipStream_t s);

template <typename T>
hipError_t launch_convolve(const T* mesh, const Geom& g, int conv, double kappa, double eps,
                           double phi_scale, T* phi, double* energy, double* dEdA,
                           cplx<T>* spec, double* partials, hipStream_t s);

// number of doubles the convolution needs for its per-block partials
// out (9) = -A^T (dA1 + dA2) A^T - energy A^T with A = box^-1 (dA2 may be NULL)
hipError_t launch_box_grad(const Geom& g, const double* dA1, const double* dA2,
                           const double* energy, double* out, hipStream_t s);
// self and non-neutral Ewald terms per frame (dq == NULL), or their
// backward into dq (nf, n) and dbox (nf, 9; may be NULL)
hipError_t launch_self_terms(const double* q, int64_t nf, int64_t n, const double* box,
                             double kappa, double eps, double* e_self, double* e_nn,
                             const double* g_self, const double* g_nn, double* dq, double* dbox,
                             hipStream_t s);
size_t convolve_partials_count(const int32_t* k);

}  // namespace admp
