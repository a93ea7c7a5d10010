// Internal declarations for the neighbour-list and real-space kernels.
#pragma once

#include <algorithm>

#include "admp_common.hpp"

namespace admp {

constexpr int kRealNV = 18;               // fused real-space partial values per block
constexpr int kPairBlocksPerFrame = 256;  // fixed grid of the pair-energy reduction

struct CellGrid {
  double h[9];
  double a[9];
  double rc2;
  int nc[3];
  int nimg[3];
  int ncell;
  int obc;
};

template <typename T>
struct alignas(4 * sizeof(T)) P4 {
  T x, y, z, q;
};

// byte offsets of the neighbour-list workspace regions
struct NbLayout {
  size_t w, cell_of, count, start, cursor, sorted, unsorted, pcount, offsets, scan, bytes;
};

CellGrid make_grid(const double* box, double cutoff, bool obc);
void nb_layout(int64_t n, const CellGrid& cg, NbLayout* L);

hipError_t launch_bin(const double* pos, int64_t n, const CellGrid& cg, char* ws,
                      const NbLayout& L, hipStream_t s);
// mode 0: count pairs and scan the per-atom counts (total at offsets[n]);
// mode 1: write the pairs
hipError_t launch_pairs(int mode, int64_t n, const CellGrid& cg, char* ws, const NbLayout& L,
                        int64_t* pairs, int64_t capacity, hipStream_t s);
hipError_t launch_pair_ds(const double* pos, const int64_t* pairs, int64_t np,
                          const CellGrid& cg, double* ds, double* dr, hipStream_t s);
hipError_t launch_pair_ds_bwd(const double* gds, const int64_t* pairs, int64_t np,
                              const double* pos, const CellGrid& cg, const double* dr,
                              const double* ds, double* gpos, double* gbox, hipStream_t s);
hipError_t launch_pair_energy(int mode, int64_t nf, int64_t n, int64_t np, const double* q,
                              const double* eta, const int64_t* pairs, const double* ds,
                              const double* bs, double kappa, double eps, double* energy,
                              double* gds, double* gq, double* partials, hipStream_t s);
template <typename T>
hipError_t launch_pack(const double* q, int64_t n, char* ws, const NbLayout& L, P4<T>* ps,
                       hipStream_t s);
// per-block partials of the fused real-space kernel: ceil(n / kBlock) rows of kRealNV
template <typename T>
hipError_t launch_real_fused(const double* pos, int64_t n, const CellGrid& cg, double kappa,
                             double eps, char* ws, const NbLayout& L, const P4<T>* ps,
                             double* grad, double* partials, hipStream_t s);

}  // namespace admp
