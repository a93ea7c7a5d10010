// Internal declarations for the neighbour-list and real-space kernels.
#pragma once

#include <algorithm>

#include "admp_common.hpp"

namespace admp {

constexpr int kRealNV = 9;                // fused real-space partial values per block
constexpr int kPairBlocksPerFrame = 4096;  // most work-groups (partial sums) per pair-energy frame
constexpr int kSub = 8;                   // sub-counters per cell (atomic contention)
constexpr int64_t kMaxCells = 1 << 21;    // cap on the cell-grid size
// staged candidates per pass of the cell kernel (32 and 64 bytes each; 512
// float32 candidates measured no faster)
template <typename T>
constexpr int cand_cap() { return 256; }
constexpr int kHomeCap = 128;     // home atoms per pass of the cell kernel
constexpr int kMaxStencil = 128;  // stencil entries the cell kernel supports
constexpr int kRealBlock = 256;  // threads of a real-space work-group (whole wavefronts)
constexpr int kRealSplit = 8;     // most candidate splits (work-groups) per home cell
// real-space work-groups aimed for: at 24k atoms 2 splits of 343 cells fill
// two thirds of the resident slots and measured fastest; 3 splits (1029
// groups) spill past the 1024 resident slots, and a persistent grid pulling
// (cell, split) items from a counter measured 10-50% slower at every item count
constexpr int64_t kRealTargetBlocks = 512;
// A batch system whose cells hold many candidate tests (few cells: a box
// under three cutoffs wide is one cell with its periodic images as the
// stencil) takes more splits, up to kRealSplitMax, so that no work-group
// tests more than about kRealBudget (home, candidate) pairs: one such
// work-group otherwise sets the kernel's duration.
constexpr int kRealSplitMax = 64;
constexpr double kRealBudget = 262144.0;
constexpr int kFilterUnroll = 4;  // candidates tested per lane between queue checks
// per-wavefront pair ring: holds < 64 + 64 * kFilterUnroll entries, and new
// entries never reach the slots of the 64 being evaluated
constexpr int kQueueRing = 512;
static_assert(2 * 64 + 64 * kFilterUnroll <= kQueueRing, "pair ring too small");

// cell c holds the sorted slots [start[c * kSub], start[(c + 1) * kSub])
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
  size_t w, cell_of, count, start, cursor, sorted, unsorted, pcount, offsets, scan, entry, bytes;
};

CellGrid make_grid(const double* box, double cutoff, bool obc);
// with_entries: also room for the per-row stencil-entry starts of the
// reverse index (n * (entries + 1) int32)
void nb_layout(int64_t n, const CellGrid& cg, NbLayout* L, bool with_entries = false);

hipError_t launch_bin(const double* pos, int64_t n, const CellGrid& cg, char* ws,
                      const NbLayout& L, hipStream_t s);
// out[0..n] = exclusive scan of in[0..n-1] with out[n] = total (I = int or int64_t)
template <typename I>
void launch_scan(const I* in, int64_t n, I* out, I* scratch, hipStream_t s);
// counting sort of binned atoms (cell_of and counts filled): start, sorted
hipError_t launch_cell_sort(int64_t n, int64_t ncell, char* ws, const NbLayout& L, hipStream_t s);
// its first half: start and the unordered cell lists in `unsorted`
hipError_t launch_cell_scatter(int64_t n, int64_t ncell, char* ws, const NbLayout& L,
                               hipStream_t s);
// mode 0: count pairs per atom and scan the counts in atom order (row
// starts at offsets[i], total at offsets[n]); mode 1: write the pairs, and
// with `keys` (int32 per pair) their in-row order keys and with `scales`
// (int64 per pair) the buffer scales i < j
hipError_t launch_pairs(int mode, int64_t n, const CellGrid& cg, char* ws, const NbLayout& L,
                        int64_t* pairs, int64_t capacity, int* keys, int64_t* scales,
                        hipStream_t s);
// stencil entries of the cell grid (3 per dimension of 3+ cells, else the
// images -nimg..nimg)
int nb_stencil_entries(const CellGrid& cg);
// whether the pair keys can encode the list (n < 2^24 atoms, <= 127 stencil
// entries, P < 2^31)
bool nb_reverse_supported(int64_t n, const CellGrid& cg, int64_t np);
// after mode 1 with keys: row_ptr (n + 1, may be NULL) = the row starts, and
// rev[o] (may be NULL) = index of the reverse pair of o, -1 if absent (every
// entry -1 when keys is NULL, the layout has no entry region or the list is
// not supported)
hipError_t launch_nb_reverse(int64_t n, const CellGrid& cg, char* ws, const NbLayout& L,
                             const int* keys, int64_t np, int64_t* row_ptr, int* rev,
                             hipStream_t s);
hipError_t launch_pair_ds(const double* pos, const int64_t* pairs, int64_t np,
                          const CellGrid& cg, double* ds, double* dr, hipStream_t s);
// out[0] = max_i |pos_i - ref_i| (minimum image under PBC), on the stream
hipError_t launch_max_disp(const double* pos, const double* ref, int64_t n, const CellGrid& cg,
                           double* out, hipStream_t s);
// workspace of launch_pair_ds_bwd (per-block partial sums of dL/dbox)
size_t pair_ds_bwd_scratch_bytes(int64_t np);
// row_ptr (n + 1) and rev given: the row form (k_pair_ds_bwd_rows), else
// any pair order (k_pair_ds_bwd)
hipError_t launch_pair_ds_bwd(const double* gds, const int64_t* pairs, int64_t np,
                              const double* pos, const CellGrid& cg, const double* dr,
                              const double* ds, double* gpos, double* gbox, double* partial,
                              const int64_t* row_ptr, const int* rev, int64_t n,
                              hipStream_t s);
// bs float64, or bsl int64 when bs is null
hipError_t launch_pair_energy(int mode, int64_t nf, int64_t n, int64_t np, const double* q,
                              const double* eta, const int64_t* pairs, const double* ds,
                              const double* bs, const int64_t* bsl, double kappa, double eps,
                              double* energy, double* gds, double* gq, double* geta,
                              double* partials, hipStream_t s);
// buffer scales (i < j, and lo < ds < hi when ds is given) as int64
hipError_t launch_pair_scales(const int64_t* pairs, int64_t np, const double* ds, double lo,
                              double hi, int64_t* out, hipStream_t s);
hipError_t launch_pack(const double* q, int64_t n, char* ws, const NbLayout& L, P4<double>* ps,
                       hipStream_t s);
}  // namespace admp
