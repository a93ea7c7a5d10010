/*
 * admp_amd -- C ABI of the MI355X (gfx950) electrostatics hot path.
 *
 * Every entry point takes plain device pointers, sizes, a host copy of the
 * 3x3 box (row vectors, row-major doubles) and a hipStream_t passed as void*.
 * The caller owns every buffer: inputs, outputs and a scratch workspace whose
 * size the matching *_workspace_bytes query returns.  No entry point allocates
 * device memory, and only admp_nblist_count synchronises the stream (to return
 * the pair count the caller must allocate for).
 *
 * Return value: ADMP_OK (0); ADMP_OVERFLOW (>0) when an output capacity is too
 * small (the required size is written to the out-param); a negative code for an
 * invalid argument or a HIP error.  No C++ exception crosses this boundary.
 *
 * Precision codes: ADMP_F64 = float64 throughout; ADMP_F32 = float32 bulk
 * arithmetic (mesh, FFT, pair terms) with float64 accumulation of energies,
 * virials and per-atom charge sums.
 *
 * Reference interfaces are cited as path:line relative to the torch-admp 1.1.5
 * source tree.  The reference is pure PyTorch; each entry point replaces the
 * torch-op sequence named in its comment.
 */
#ifndef ADMP_AMD_H
#define ADMP_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ADMP_OK 0
#define ADMP_OVERFLOW 1
#define ADMP_ERR_INVALID (-1)
#define ADMP_ERR_HIP (-2)
#define ADMP_ERR_UNSUPPORTED (-3)
#define ADMP_ERR_SCRATCH (-4)  /* more than ADMP_NOALLOC_SCRATCH_BYTES needed */

#define ADMP_F64 0
#define ADMP_F32 1

/* k-vector convention: ADMP_KCONV_REFERENCE reproduces k = 2 pi m A (rows of
 * A = box^-1, src/torch_admp/recip.py:304); ADMP_KCONV_COLUMNS uses the
 * reciprocal lattice vectors proper, k = 2 pi A m.  Identical for orthorhombic
 * and symmetric boxes. */
#define ADMP_KCONV_REFERENCE 0
#define ADMP_KCONV_COLUMNS 1

/* Library version (major * 10000 + minor * 100 + patch). */
int admp_version(void);

/* Human-readable text for a status code (static storage). */
const char* admp_status_string(int status);

/* ------------------------------------------------------------------------ */
/* Neighbour list: replaces TorchNeighborList.forward / forward_pbc /
 * wrap_positions (src/torch_admp/nblist.py:304-470).  Pairs (i, j) are kept
 * when 0.01 < |image_j - wrapped_i| < cutoff with both orderings listed, as
 * nblist.py:439-445.  box == NULL selects open boundaries (nblist.py:472-490,
 * same set without the 0.01 floor). Pairs are int64 (P, 2), grouped by i.  */

/* Workspace for a neighbour-list build of n atoms. */
int admp_nblist_workspace_bytes(int64_t n, const double* box, double cutoff, size_t* bytes);

/* Bins atoms into cells and counts pairs; synchronises `stream` and writes the
 * total number of ordered pairs to *n_pairs (host).  The workspace must be
 * passed unchanged to the following admp_nblist_fill. */
int admp_nblist_count(const double* pos, int64_t n, const double* box, double cutoff,
                      void* workspace, size_t ws_bytes, int64_t* n_pairs, void* stream);

/* Writes the pairs counted by admp_nblist_count; capacity is in pairs.  The
 * pairs come in n rows, one per atom in cell order (rows may be empty), each
 * row's pairs sharing their first atom.  Optional outputs (each may be NULL):
 * row_ptr (int64, n + 1) receives the row starts (row r = pairs
 * [row_ptr[r], row_ptr[r + 1])); reverse (int32, capacity) receives for each
 * pair (i, j) the index of (j, i), -1 when the list cannot be indexed
 * (n >= 2^24 atoms or P >= 2^31), and then needs scratch (int32, capacity);
 * buffer_scales (int64, capacity) receives 1 where i < j else 0, the
 * reference's pairs_buffer_scales (src/torch_admp/nblist.py:492-515) written
 * by the same pass.  row_ptr and reverse let admp_pair_ds_backward run
 * without per-pair atomics; they have no reference counterpart
 * (TorchNeighborList returns the pairs alone, nblist.py:304-333). */
int admp_nblist_fill(const double* pos, int64_t n, const double* box, double cutoff,
                     void* workspace, size_t ws_bytes, int64_t* pairs, int64_t capacity,
                     int64_t* row_ptr, int32_t* reverse, int32_t* scratch,
                     int64_t* buffer_scales, void* stream);

/* Buffer scales of an ordered pair list (P, 2) into scales (int64, P): 1
 * where i < j (src/torch_admp/nblist.py:492-515), and, when ds (P, float64)
 * is given, lo < ds < hi, else 0.  A list kept across steps (a Verlet skin)
 * masks its rows outside the cutoff this way.  No synchronisation. */
int admp_pair_scales(const int64_t* pairs, int64_t n_pairs, const double* ds, double lo,
                     double hi, int64_t* scales, void* stream);

/* Largest displacement max_i |pos_i - ref_i| (minimum image when box is
 * given) into out[0] (device, float64), for a Verlet-skin neighbour list: a
 * list built at cutoff + skin from ref holds every pair within cutoff while
 * the largest displacement stays below skin / 2.  The reference rebuilds its
 * list on every call (src/torch_admp/nblist.py:304-333); this has no
 * counterpart there. */
int admp_max_displacement(const double* pos, const double* ref, int64_t n, const double* box,
                          double* out, void* stream);

/* Pair distances |pbc_shift(r_i - r_j)| (src/torch_admp/nblist.py:517-551,
 * src/torch_admp/spatial.py:15-36); box == NULL gives |r_j - r_i|.  dr (may
 * be NULL) receives the minimum-image displacement (P, 3). */
int admp_pair_ds(const double* pos, const int64_t* pairs, int64_t n_pairs, const double* box,
                 double* ds, double* dr, void* stream);

/* Backward of admp_pair_ds: grad_pos (n, 3) += dL/dds * d ds / d r;
 * grad_box (9) += dL/dds * d ds / d box (may be NULL).  grad_pos and grad_box
 * must be zeroed by the caller.  dr may be NULL (the displacement is then
 * recomputed from pos).  With row_ptr and reverse from admp_nblist_fill (both
 * or neither; n_atoms = the number of rows) each row sums its pairs' first-atom terms and
 * their reverses' second-atom terms, one float64 atomic per row component;
 * otherwise runs of equal first atoms are summed in registers before the
 * atomics and the second atom adds per pair.  The box term goes through
 * per-work-group partials in the workspace (admp_pair_ds_backward_workspace_bytes,
 * needed when grad_box and box are given) and a fixed-order reduction. */
int admp_pair_ds_backward_workspace_bytes(int64_t n_pairs, size_t* bytes);
int admp_pair_ds_backward(const double* grad_ds, const int64_t* pairs, int64_t n_pairs,
                          const double* pos, const double* box, const double* dr,
                          const double* ds, double* grad_pos, double* grad_box,
                          const int64_t* row_ptr, const int32_t* reverse, int64_t n_atoms,
                          void* workspace, size_t ws_bytes, void* stream);

/* ------------------------------------------------------------------------ */
/* Pair energy over a caller-supplied pair list, nf frames of P pairs each.
 * mode 0: Ewald real space eps*erfc(kappa d) qi qj safe_inv(d, 1e-4) bs
 *         (src/torch_admp/pme.py:312-349)
 * mode 1: open boundary eps*qi qj safe_inv(d, 1e-8) bs (src/torch_admp/pme.py:529-561)
 * mode 2: Gaussian damping -eps*erfc(d / (2 eta_ij)) qi qj safe_inv(d, 1e-4) bs,
 *         eta_ij = sqrt((eta_i^2 + eta_j^2)/2) (src/torch_admp/qeq.py:106-121)
 * q is (nf, n); eta (nf, n) for mode 2 else NULL; pairs (nf, P, 2); ds, bs (nf, P).
 * Writes energy (nf) and, when non-NULL, dE/dds (nf, P), dE/dq (nf, n) and,
 * for mode 2 only, dE/deta (nf, n) (dE/dq and dE/deta are overwritten).
 * bs is float64, or int64 (the neighbour list's buffer scales as they come,
 * no conversion pass) when mode carries ADMP_PAIR_BS_INT64. */
#define ADMP_PAIR_BS_INT64 0x100
int admp_pair_energy(int32_t mode, int64_t nf, int64_t n, int64_t n_pairs, const double* q,
                     const double* eta, const int64_t* pairs, const double* ds,
                     const void* bs, double kappa, double dielectric, double* energy,
                     double* grad_ds, double* grad_q, double* grad_eta, void* workspace,
                     size_t ws_bytes, void* stream);
int admp_pair_energy_workspace_bytes(int64_t nf, size_t* bytes);

/* ------------------------------------------------------------------------ */
/* Smooth PME reciprocal space, order-6 B-splines, one frame per call.
 * kmesh = (K1, K2, K3); the mesh is real (K1*K2*K3) of the precision type.  */

int admp_pme_workspace_bytes(const int32_t* kmesh, int32_t precision, int64_t n,
                             size_t* bytes);

/* Charge spreading Q = sum_i q_i theta_i (src/torch_admp/recip.py:216-252).
 * Overwrites mesh.  With the workspace of admp_pme_workspace_bytes for these n
 * atoms, the atoms are binned into cells of about 8 mesh points per axis and
 * each cell's stencils are summed in an LDS tile added once to the mesh;
 * without it (workspace NULL) every atom adds its 216 weights to the mesh. */
int admp_pme_spread(const double* pos, const double* q, int64_t n, const double* box,
                    const int32_t* kmesh, int32_t precision, void* mesh, void* workspace,
                    size_t ws_bytes, void* stream);

/* Reciprocal energy and potential mesh (src/torch_admp/pme.py:410-443,
 * 661-683; torch.fft.fftn at pme.py:434 replaced by an in-library r2c/c2r FFT).
 * energy[0] = eps sum_k G(k) |S(k)|^2; dEdA (9, may be NULL) = dE/dA through
 * the k vectors at fixed A-independent structure factor; phi (may be NULL) =
 * dE/dQ on the mesh, scaled by phi_scale.  energy and dEdA are overwritten. */
int admp_pme_convolve(const void* mesh, const double* box, const int32_t* kmesh,
                      int32_t precision, int32_t convention, double kappa, double dielectric,
                      double phi_scale, void* phi, double* energy, double* dEdA,
                      void* workspace, size_t ws_bytes, void* stream);

/* Back-interpolation, the transpose of admp_pme_spread (the autograd
 * transpose of recip.py:94-99,206).  For a mesh field phi = dE/dQ:
 *   grad_pos (n,3) = dE/dr,  dEdq (n) = dE/dq,  dEdA (9) = dE/dA via s = r A.
 * Any output may be NULL.  accumulate != 0 adds into grad_pos / dEdq instead
 * of overwriting; dEdA is always overwritten. */
int admp_pme_gather(const double* pos, const double* q, int64_t n, const double* box,
                    const int32_t* kmesh, int32_t precision, const void* phi,
                    double* grad_pos, double* dEdq, double* dEdA, int32_t accumulate,
                    void* workspace, size_t ws_bytes, void* stream);

/* Self and non-neutral Ewald terms of nf frames of n charges
 * (src/torch_admp/pme.py:448-493): e_self[f] = -dielectric kappa / sqrt(pi)
 * sum q^2 and e_nonneutral[f] = -dielectric pi Q^2 / (2 V kappa^2), Q = sum q,
 * V = det(box_f); box is on the device (nf x 9, row vectors), so no host copy
 * is needed.  The backward writes grad_q (nf, n) and grad_box (nf x 9, may be
 * NULL) for upstream gradients g_self, g_nonneutral (nf each, device, either
 * may be NULL for zero). */
int admp_ewald_self_terms(const double* q, int64_t nf, int64_t n, const double* box,
                          double kappa, double dielectric, double* e_self, double* e_nonneutral,
                          void* stream);
int admp_ewald_self_terms_backward(const double* q, int64_t nf, int64_t n, const double* box,
                                   double kappa, double dielectric, const double* g_self,
                                   const double* g_nonneutral, double* grad_q, double* grad_box,
                                   void* stream);

/* dE/dbox of the reciprocal energy (the autograd box derivative through
 * recip.py:255-308 and the volume in pme.py:410-443): with A = box^-1,
 *   grad_box (9) = -A^T (dEdA_k + dEdA_gather) A^T - energy[0] A^T,
 * dEdA_k and energy from admp_pme_convolve, dEdA_gather (may be NULL) from
 * admp_pme_gather.  Device inputs and output; no synchronisation. */
int admp_pme_box_grad(const double* box, const double* dEdA_k, const double* dEdA_gather,
                      const double* energy, double* grad_box, void* stream);

/* ------------------------------------------------------------------------ */
/* Fused Ewald/PME energy + forces for one periodic frame straight from
 * positions: cell-list real space (no pair list materialised), spread, FFT,
 * influence function, gather, self and non-neutral terms.  Equivalent to
 * CoulombForceModule.forward on the TorchNeighborList pair list followed by
 * calc_grads (src/torch_admp/pme.py:261-310, src/torch_admp/utils.py:84-112)
 * whenever every box height is at least 2 cutoff.  Below that a pair has
 * several images inside the cutoff; this entry point sums each of them (the
 * lattice sum, as the extended-coordinate lists of dp_nblist do,
 * src/torch_admp/nblist.py:36-95,166-240), where the minimum-image pair
 * distances of TorchNeighborList (nblist.py:517-551) repeat the nearest image.
 *   energy (5, device): real, reciprocal, self, non-neutral, total
 *   forces (n, 3, device): -dE/dr (overwritten)
 *   dEdbox (9, device, may be NULL): dE/dh at fixed Cartesian positions
 * Workspace from admp_pme_fused_workspace_bytes. */
int admp_pme_fused_workspace_bytes(int64_t n, const double* box, double cutoff,
                                   const int32_t* kmesh, int32_t precision, size_t* bytes);
int admp_pme_fused(const double* pos, const double* q, int64_t n, const double* box,
                   double cutoff, double kappa, const int32_t* kmesh, int32_t precision,
                   int32_t convention, double dielectric, double* energy, double* forces,
                   double* dEdbox, void* workspace, size_t ws_bytes, void* stream);

/* admp_pme_fused without a workspace argument.  No counterpart in the
 * reference: it serves callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results as admp_pme_fused; the scratch is a static
 * device array of ADMP_NOALLOC_SCRATCH_BYTES per GPU, held from library load to
 * process end.  A problem whose admp_pme_fused_workspace_bytes exceeds it gets
 * ADMP_ERR_SCRATCH (the 24,000-atom water box at 64^3 needs 8.4 MiB in float32
 * and 13.8 MiB in float64).  It calls no allocating routine and launches every
 * kernel on `stream` (no side stream).  Calls that overlap on one GPU share
 * that GPU's scratch and race: serialise them. */
#define ADMP_NOALLOC_SCRATCH_BYTES (64u << 20)
int admp_pme_fused_noalloc(const double* pos, const double* q, int64_t n, const double* box,
                           double cutoff, double kappa, const int32_t* kmesh, int32_t precision,
                           int32_t convention, double dielectric, double* energy, double* forces,
                           double* dEdbox, void* stream);

/* ------------------------------------------------------------------------ */
/* Batched form of admp_pme_fused: n_sys independent periodic systems of any
 * mix of atom counts, boxes, cutoffs and meshes in one launch sequence (each
 * stage is one launch over the whole batch).  Per system the outputs equal
 * those of admp_pme_fused to float64 rounding, and they do not depend on the
 * other systems of the batch.  The reference evaluates one frame per
 * CoulombForceModule.forward call (src/torch_admp/pme.py:261-310); this entry
 * point has no single counterpart there and serves callers that hold many
 * systems at once.
 *   host arrays: n_atoms[n_sys], boxes[n_sys][9], cutoffs[n_sys],
 *                kappas[n_sys], kmesh[n_sys][3]
 *   device: pos (sum n_atoms, 3) and q (sum n_atoms), systems back to back;
 *           energy (n_sys, 5); forces (sum n_atoms, 3); dEdbox (n_sys, 9) or NULL
 * Returns ADMP_ERR_UNSUPPORTED when the batch exceeds 2^30 atoms or 2^28 cells. */
int admp_pme_batch_workspace_bytes(int32_t n_sys, const int64_t* n_atoms, const double* boxes,
                                   const double* cutoffs, const int32_t* kmesh,
                                   int32_t precision, size_t* bytes);
int admp_pme_batch(int32_t n_sys, const int64_t* n_atoms, const double* boxes,
                   const double* cutoffs, const double* kappas, const int32_t* kmesh,
                   int32_t precision, int32_t convention, double dielectric, const double* pos,
                   const double* q, double* energy, double* forces, double* dEdbox,
                   void* workspace, size_t ws_bytes, void* stream);

/* bytes[i] = admp_pme_batch_workspace_bytes of system i alone, for n_sys
 * systems in one call: the per-request workspace share an admission loop
 * charges against its device-memory budget, without one host call (and one
 * batch plan) per request.  Same arguments and errors as
 * admp_pme_batch_workspace_bytes.  No counterpart in the reference (it
 * evaluates one frame per call and admits nothing). */
int admp_pme_batch_system_bytes(int32_t n_sys, const int64_t* n_atoms, const double* boxes,
                                const double* cutoffs, const int32_t* kmesh, int32_t precision,
                                size_t* bytes);

/* admp_pme_batch with flags.  ADMP_BATCH_NO_OVERLAP keeps every kernel on
 * `stream`: admp_pme_batch forks the real-space kernel onto one side stream
 * per device, which callers running batches on several streams at once would
 * share (their batches then wait on each other's real space); such callers
 * (a serving loop with batches in flight on two streams) pass this flag and
 * get their overlap from the concurrent batches instead.  No counterpart in
 * the reference. */
#define ADMP_BATCH_NO_OVERLAP 1
int admp_pme_batch_ex(int32_t n_sys, const int64_t* n_atoms, const double* boxes,
                      const double* cutoffs, const double* kappas, const int32_t* kmesh,
                      int32_t precision, int32_t convention, double dielectric, const double* pos,
                      const double* q, double* energy, double* forces, double* dEdbox,
                      int32_t flags, void* workspace, size_t ws_bytes, void* stream);

/* Graph capture of a batch (no counterpart in the reference; for callers that
 * replay one batch composition many times under a HIP graph).  admp_pme_batch
 * copies the system table from host memory on every call, which a captured
 * graph cannot repeat.  Instead: admp_pme_batch_upload writes the table into
 * the workspace once (synchronously), and admp_pme_batch_run (same arguments
 * as admp_pme_batch) launches only kernels, reading that table; it may be
 * captured and replayed with new positions and charges in the same buffers.
 * A batch of one needs no upload: its first kernel writes the table. */
int admp_pme_batch_upload(int32_t n_sys, const int64_t* n_atoms, const double* boxes,
                          const double* cutoffs, const double* kappas, const int32_t* kmesh,
                          int32_t precision, int32_t convention, double dielectric,
                          void* workspace, size_t ws_bytes, void* stream);
int admp_pme_batch_run(int32_t n_sys, const int64_t* n_atoms, const double* boxes,
                       const double* cutoffs, const double* kappas, const int32_t* kmesh,
                       int32_t precision, int32_t convention, double dielectric,
                       const double* pos, const double* q, double* energy, double* forces,
                       double* dEdbox, void* workspace, size_t ws_bytes, void* stream);

/* ------------------------------------------------------------------------ */
/* Charge equilibration.  The QEq energy (src/torch_admp/qeq.py:33-209,408-458)
 *   E(q) = chi.q + sum hardness q^2 + E_coulomb(q) + E_damping(q)
 * is quadratic in q.  The problem below describes its Hessian H once per
 * geometry; admp_qeq_matvec applies it (the matrix-free replacement of the
 * autograd double backward of calc_hessian, qeq.py:460-515) and admp_qeq_solve
 * minimises E on {A q = b} by projected, Jacobi-preconditioned conjugate
 * gradients (replacing the solvers of qeq.py:609-1114 and the projection of
 * src/torch_admp/utils.py:115-210) to |P grad E| / N < tol. */
#define ADMP_QEQ_KSPACE 1   /* reciprocal, self and non-neutral Coulomb terms */
#define ADMP_QEQ_RSPACE 2   /* real-space erfc pairs (PBC) or bare pairs (box NULL) */
#define ADMP_QEQ_DAMPING 4  /* Gaussian-damping pairs (the per-atom term is in hdiag) */
#define ADMP_QEQ_SLAB 8     /* slab correction along slab_axis */
#define ADMP_QEQ_MAX_CONS 8

typedef struct admp_qeq_problem {
  const double* pos;        /* (n, 3) device */
  int64_t n;
  const double* box;        /* host (9) or NULL for open boundaries */
  double cutoff;            /* real-space cutoff, sets the cell grid of the spread */
  int32_t kmesh[3];
  int32_t precision;        /* ADMP_F64 / ADMP_F32 for the reciprocal mesh */
  int32_t convention;       /* ADMP_KCONV_* */
  double kappa;
  double dielectric;
  int32_t flags;            /* ADMP_QEQ_* */
  int32_t slab_axis;
  const int64_t* pairs;     /* (n_pairs, 2) device */
  const double* ds;         /* (n_pairs) device */
  const double* bs;         /* (n_pairs) device, buffer scales */
  int64_t n_pairs;
  const double* eta;        /* (n) device, Gaussian widths (damping) */
  const double* hdiag;      /* (n) device, diagonal terms: 2 hardness + self terms */
  const double* chi;        /* (n) device, linear term (solve only) */
  const double* precond;    /* (n) device Jacobi diagonal, or NULL (solve only) */
  const double* cons;       /* (n_cons, n) device constraint matrix A (solve only) */
  const double* cons_vals;  /* (n_cons) device right-hand side b (solve only) */
  int32_t n_cons;
} admp_qeq_problem;

int admp_qeq_workspace_bytes(const admp_qeq_problem* problem, size_t* bytes);
/* out = H v */
int admp_qeq_matvec(const admp_qeq_problem* problem, const double* v, double* out,
                    void* workspace, size_t ws_bytes, void* stream);
/* q = argmin E on {A q = b} from the start q0 (projected onto the constraints);
 * synchronises every check_every iterations to test the residual.  *iters is
 * the iteration count at convergence or -1; *residual = |P grad E| / n.
 * Returns ADMP_ERR_INVALID when the rows of A are linearly dependent (a pivot
 * of A A^T below 1e-12 of its largest entry), where the reference's
 * torch.inverse of A A^T raises (src/torch_admp/utils.py:151-171). */
int admp_qeq_solve(const admp_qeq_problem* problem, const double* q0, double tol,
                   int32_t max_iter, int32_t check_every, double* q, int32_t* iters,
                   double* residual, void* workspace, size_t ws_bytes, void* stream);

/* Batched charge equilibration: n_sys independent periodic systems (atoms back
 * to back in pos, pairs with batch-wide atom indices), each with Ewald
 * real-space and reciprocal Coulomb terms (kappas, kmesh per system), Gaussian
 * damping when flags has ADMP_QEQ_DAMPING, diagonal terms hdiag (2 hardness +
 * self terms) and the one constraint sum q = total_charge[s] (host, n_sys).
 * Every iteration forms the Hessian products of all systems in one launch
 * sequence (pair rows, then the batched spread / FFT / influence function /
 * gather of admp_pme_batch), then one projected-CG step per system; a system
 * stops on its own at |P grad E| / n_s < tols[s] (host array; a device flag,
 * the others go on)
 * and the host reads the flags every check_every iterations.  iters[s] is the
 * iteration at convergence or -1, residuals[s] = |P grad E| / n_s (host
 * arrays).  The same fixed point as admp_qeq_solve on each system alone.  No
 * counterpart in the reference, which solves one frame per call
 * (src/torch_admp/qeq.py:609-765); for serving loops that solve many systems
 * at once. */
int admp_qeq_solve_batch_workspace_bytes(int32_t n_sys, const int64_t* n_atoms,
                                         const double* boxes, const double* cutoffs,
                                         const int32_t* kmesh, int32_t precision,
                                         int64_t n_pairs, size_t* bytes);
int admp_qeq_solve_batch(int32_t n_sys, const int64_t* n_atoms, const double* boxes,
                         const double* cutoffs, const double* kappas, const int32_t* kmesh,
                         int32_t precision, int32_t convention, double dielectric, int32_t flags,
                         const double* pos, const int64_t* pairs, const double* ds,
                         const double* bs, int64_t n_pairs, const double* eta,
                         const double* hdiag, const double* chi, const double* precond,
                         const double* total_charge, const double* q0, const double* tols,
                         int32_t max_iter, int32_t check_every, double* q, int32_t* iters,
                         double* residuals, void* workspace, size_t ws_bytes, void* stream);

/* ------------------------------------------------------------------------ */
/* Stage timing of admp_pme_fused (a profiling aid with no counterpart in the
 * reference; single-threaded use).  admp_profile_begin arms hipEvent pairs
 * around the stages of the next max_calls fused calls; admp_profile_collect
 * synchronises, writes the summed GPU milliseconds per stage and the number of
 * calls recorded, and disarms.  Stages: 0 cell binning and packing, 1 real-space
 * pair kernel, 2 charge spreading, 3 FFT passes with the influence function,
 * 4 gather, 5 reductions and finalisation.  admp_pme_batch calls are timed
 * the same way. */
#define ADMP_PROFILE_STAGES 6
int admp_profile_begin(int32_t max_calls);
int admp_profile_collect(double* stage_ms, int32_t* n_calls);

#ifdef __cplusplus
}
#endif

#endif /* ADMP_AMD_H */
