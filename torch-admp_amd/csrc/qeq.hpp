// QEq Hessian-vector product and projected CG solve (see qeq.hip).
#pragma once

#include "admp_common.hpp"
#include "nblist_kernels.hpp"
#include "tiles.hpp"
#include "batch.hpp"

namespace admp {

constexpr int kQeqMaxCons = 8;                   // constraint rows
constexpr int kQeqMaxDots = 1 + kQeqMaxCons + 3;  // values of one k_dots pass

struct QeqProblem {
  const double* pos;
  int64_t n;
  const double* box;  // host, null for open boundaries
  double cutoff;
  const int32_t* kmesh;
  int precision;
  int convention;
  double kappa;
  double eps;
  int kspace;
  int coul;     // pair Coulomb: 0 none, 1 erfc (Ewald real space), 2 bare (open boundary)
  int damping;  // Gaussian damping pairs
  int slab;
  int slab_axis;
  double lz2;   // squared length of the slab-axis box vector
  const int64_t* pairs;
  const double* ds;
  const double* bs;
  int64_t n_pairs;
  const double* eta;
  const double* hdiag;
  const double* chi;
  const double* precond;
  const double* cons;
  const double* cons_vals_dev;
  int n_cons;
};

struct QeqLayout {
  size_t c, r, z, p, w, s, zc, dots, sc, status, ticket, nbws, mesh, spec, ppart, bytes;
  // row lists of the pair coefficients (both orientations of every pair):
  // rowcnt/rowcur n ints, rowptr n + 1 ints, col/val 2 n_pairs entries, and the
  // scan's block sums; csr = 0 when 2 n_pairs does not fit an int32
  size_t rowcnt, rowcur, rowptr, col, val, scan;
  int csr;
  // point spread (each mesh point sums the atoms whose stencil covers it):
  // per atom the wrapped stencil base and bin (int4) and 18 B-spline weights,
  // atoms binned by base in 4-point bins (nbin per axis); pt = 0 when off
  size_t prec_base, prec_w, pbcnt, pbcur, pbstart, pbatoms, pbscan;
  int pt, nbin[3];
  CellGrid cg;
  NbLayout nb;
  TileGeom tg;
};

QeqLayout qeq_layout(const QeqProblem& P);
template <typename T>
int qeq_matvec_impl(const QeqProblem& P, const double* v, double* out, void* workspace,
                    hipStream_t s);
template <typename T>
int qeq_solve_impl(const QeqProblem& P, const double* q0, double tol, int max_iter,
                   int check_every, double* q, int* iters, double* resid, void* workspace,
                   hipStream_t s);

// Batched solves: independent periodic systems side by side (atoms back to
// back, pairs with batch-wide atom indices), each with Ewald real and
// reciprocal Coulomb terms, optional Gaussian damping and the one constraint
// sum q = total_charge[s].  The batch plan gives boxes, kappas and meshes.
struct QeqBatch {
  const double* pos;
  const int64_t* pairs;
  const double* ds;
  const double* bs;
  int64_t n_pairs;
  const double* eta;
  const double* hdiag;
  const double* chi;
  const double* precond;       // NULL: no preconditioner
  const double* total_charge;  // host, one per system
  int damping;
};
// Workspace bytes after the batch plan's own (0 when unsupported: row lists
// beyond 32-bit indices).
size_t qeq_batch_extra_bytes(const BatchPlan& plan, int64_t n_pairs);
int qeq_solve_batch(const BatchPlan& plan, const QeqBatch& B, const double* q0, const double* tol,
                    int max_iter, int check_every, double* q, int32_t* iters, double* resid,
                    char* ws, hipStream_t s);

}  // namespace admp
