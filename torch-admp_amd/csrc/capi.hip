// C ABI of the admp_amd library: argument validation, workspace carving and
// kernel sequencing.  No entry point allocates device memory; only
// admp_nblist_count synchronises its stream.
#include <cmath>
#include <cstring>
#include <mutex>

#include "batch.hpp"
#include "nblist_kernels.hpp"
#include "pme_kernels.hpp"
#include "qeq.hpp"
#include "tiles.hpp"

using namespace admp;

namespace {

inline int hip_status(hipError_t e) { return e == hipSuccess ? ADMP_OK : ADMP_ERR_HIP; }

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

bool box_ok(const double* box) {
  if (!box) return false;
  for (int i = 0; i < 9; ++i)
    if (!std::isfinite(box[i])) return false;
  const double* h = box;
  const double det = h[0] * (h[4] * h[8] - h[5] * h[7]) - h[1] * (h[3] * h[8] - h[5] * h[6]) +
                     h[2] * (h[3] * h[7] - h[4] * h[6]);
  return det != 0.0 && std::isfinite(det);
}

bool kmesh_ok(const int32_t* k) {
  if (!k) return false;
  for (int d = 0; d < 3; ++d)
    if (k[d] < 1 || k[d] > kMaxMesh) return false;
  return true;
}

Geom make_geom(const double* box, const int32_t* kmesh) {
  Geom g;
  const CellGrid cg = make_grid(box, 1.0, false);
  for (int i = 0; i < 9; ++i) {
    g.h[i] = cg.h[i];
    g.a[i] = cg.a[i];
  }
  const double* h = box;
  g.vol = h[0] * (h[4] * h[8] - h[5] * h[7]) - h[1] * (h[3] * h[8] - h[5] * h[6]) +
          h[2] * (h[3] * h[7] - h[4] * h[6]);
  for (int d = 0; d < 3; ++d) g.k[d] = kmesh ? kmesh[d] : 1;
  return g;
}

size_t elem_size(int32_t precision) { return precision == ADMP_F32 ? 4 : 8; }

// Cell grid of the tiled spread of admp_pme_spread: about kSpreadCellPts mesh
// points per cell along each axis, from the mesh alone (the box only sets the
// fractional coordinates).
constexpr int kSpreadCellPts = 8;

CellGrid spread_grid(const double* box, const int32_t* k) {
  CellGrid cg = make_grid(box, 1.0, box == nullptr);
  int64_t ncell = 1;
  for (int d = 0; d < 3; ++d) {
    cg.nc[d] = std::max(1, std::min(1024, k[d] / kSpreadCellPts));
    cg.nimg[d] = 1;
    ncell *= cg.nc[d];
  }
  cg.ncell = (int)ncell;
  return cg;
}

// PME workspace: spectrum (K1*K2*(K3/2+1) complex) + partials, then the cell
// binning of n atoms used by the tiled spread.
struct PmeLayout {
  size_t spec, partials, nb, bytes;
  NbLayout nbl;
};

PmeLayout pme_layout(const int32_t* k, int32_t precision, int64_t n = 0) {
  PmeLayout L;
  const size_t nspec = spec_elems(k, 2 * elem_size(precision));
  L.spec = 0;
  size_t off = align256(nspec * 2 * elem_size(precision));
  L.partials = off;
  const size_t np = std::max(convolve_partials_count(k), (size_t)kRedBlocks * 9 + 16);
  off += align256(np * sizeof(double));
  L.nb = off;
  CellGrid cg;
  cg.ncell = 1;
  for (int d = 0; d < 3; ++d) cg.ncell *= std::max(1, std::min(1024, k[d] / kSpreadCellPts));
  nb_layout(n, cg, &L.nbl);
  off += align256(L.nbl.bytes);
  L.bytes = off;
  return L;
}

// Stage timing (profiling aid): hipEvent pairs around the fused stages.
struct Profiler {
  bool on = false;
  int max_calls = 0;
  int calls = 0;
  hipEvent_t* ev = nullptr;
};
Profiler g_prof;

struct StageMarks {
  hipStream_t s;
  hipEvent_t* ev = nullptr;
  void mark(int k) {
    if (ev) (void)hipEventRecord(ev[k], s);  // a failed mark shows as a failed elapsed time
  }
};

StageMarks stage_marks(hipStream_t s) {
  StageMarks m;
  m.s = s;
  if (g_prof.on && g_prof.calls < g_prof.max_calls) {
    m.ev = g_prof.ev + (size_t)g_prof.calls * (ADMP_PROFILE_STAGES + 1);
    ++g_prof.calls;
  }
  return m;
}

bool qeq_problem(const admp_qeq_problem* pr, bool need_solve, QeqProblem* P) {
  if (!pr || pr->n < 0 || pr->n_pairs < 0) return false;
  if (pr->n > 0 && (!pr->pos || !pr->hdiag)) return false;
  if (pr->n_pairs > 0 && (!pr->pairs || !pr->ds || !pr->bs)) return false;
  if ((pr->flags & ADMP_QEQ_DAMPING) && pr->n_pairs > 0 && !pr->eta) return false;
  if (pr->box && !box_ok(pr->box)) return false;
  const bool pbc = pr->box != nullptr;
  const bool kspace = pbc && (pr->flags & ADMP_QEQ_KSPACE);
  if (kspace && (!kmesh_ok(pr->kmesh) || !(pr->kappa > 0.0) || !(pr->cutoff > 0.0))) return false;
  if (pr->precision != ADMP_F32 && pr->precision != ADMP_F64) return false;
  if (pr->convention != ADMP_KCONV_REFERENCE && pr->convention != ADMP_KCONV_COLUMNS) return false;
  if (pr->n_cons < 0 || pr->n_cons > ADMP_QEQ_MAX_CONS) return false;
  if (need_solve && (!pr->chi || (pr->n_cons > 0 && (!pr->cons || !pr->cons_vals)))) return false;
  const bool slab = pbc && (pr->flags & ADMP_QEQ_SLAB);
  if (slab && (pr->slab_axis < 0 || pr->slab_axis > 2)) return false;
  P->pos = pr->pos;
  P->n = pr->n;
  P->box = pr->box;
  P->cutoff = pr->cutoff > 0.0 ? pr->cutoff : 1.0;
  P->kmesh = pr->kmesh;
  P->precision = pr->precision;
  P->convention = pr->convention;
  P->kappa = pr->kappa;
  P->eps = pr->dielectric;
  P->kspace = kspace ? 1 : 0;
  P->coul = (pr->flags & ADMP_QEQ_RSPACE) ? (pbc ? 1 : 2) : 0;
  P->damping = (pr->flags & ADMP_QEQ_DAMPING) ? 1 : 0;
  P->slab = slab ? 1 : 0;
  P->slab_axis = slab ? pr->slab_axis : 0;
  P->lz2 = 0.0;
  if (slab)
    for (int d = 0; d < 3; ++d) P->lz2 += pr->box[3 * pr->slab_axis + d] * pr->box[3 * pr->slab_axis + d];
  P->pairs = pr->pairs;
  P->ds = pr->ds;
  P->bs = pr->bs;
  P->n_pairs = pr->n_pairs;
  P->eta = pr->eta;
  P->hdiag = pr->hdiag;
  P->chi = pr->chi;
  P->precond = pr->precond;
  P->cons = pr->cons;
  P->cons_vals_dev = pr->cons_vals;
  P->n_cons = pr->n_cons;
  return true;
}

}  // namespace

// Scratch of admp_pme_fused_noalloc: one array per GPU, held from the library's
// load until the process ends.  Host code never passes `scratch` itself (its
// host-side name is not a device address): each device's address comes from
// hipGetSymbolAddress once, with that device current, and only that pointer
// reaches kernels.
static __device__ __align__(256) unsigned char scratch[ADMP_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kScratchDevices = 64;
void* g_scratch_ptr[kScratchDevices] = {};
std::mutex g_scratch_mu;

// Device address of the current device's scratch, looked up on the first call
// on that device; nullptr when the runtime refuses.
void* scratch_of_current_device() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kScratchDevices) return nullptr;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}
}  // namespace

extern "C" {

int admp_version(void) { return 100; }

int admp_profile_begin(int32_t max_calls) {
  if (max_calls < 1 || max_calls > 100000) return ADMP_ERR_INVALID;
  if (g_prof.ev) {
    for (int i = 0; i < g_prof.max_calls * (ADMP_PROFILE_STAGES + 1); ++i)
      (void)hipEventDestroy(g_prof.ev[i]);
    delete[] g_prof.ev;
  }
  const int nev = max_calls * (ADMP_PROFILE_STAGES + 1);
  g_prof.ev = new hipEvent_t[nev];
  for (int i = 0; i < nev; ++i)
    if (hipEventCreate(&g_prof.ev[i]) != hipSuccess) return ADMP_ERR_HIP;
  g_prof.max_calls = max_calls;
  g_prof.calls = 0;
  g_prof.on = true;
  return ADMP_OK;
}

int admp_profile_collect(double* stage_ms, int32_t* n_calls) {
  if (!stage_ms || !n_calls) return ADMP_ERR_INVALID;
  for (int k = 0; k < ADMP_PROFILE_STAGES; ++k) stage_ms[k] = 0.0;
  *n_calls = 0;
  if (!g_prof.ev) return ADMP_ERR_INVALID;
  g_prof.on = false;
  int status = ADMP_OK;
  for (int c = 0; c < g_prof.calls; ++c) {
    hipEvent_t* ev = g_prof.ev + (size_t)c * (ADMP_PROFILE_STAGES + 1);
    if (hipEventSynchronize(ev[ADMP_PROFILE_STAGES]) != hipSuccess) status = ADMP_ERR_HIP;
    for (int k = 0; k < ADMP_PROFILE_STAGES; ++k) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, ev[k], ev[k + 1]) != hipSuccess) status = ADMP_ERR_HIP;
      stage_ms[k] += ms;
    }
  }
  *n_calls = g_prof.calls;
  for (int i = 0; i < g_prof.max_calls * (ADMP_PROFILE_STAGES + 1); ++i)
    (void)hipEventDestroy(g_prof.ev[i]);
  delete[] g_prof.ev;
  g_prof.ev = nullptr;
  g_prof.calls = g_prof.max_calls = 0;
  return status;
}

const char* admp_status_string(int status) {
  switch (status) {
    case ADMP_OK: return "ok";
    case ADMP_OVERFLOW: return "output capacity too small";
    case ADMP_ERR_INVALID: return "invalid argument";
    case ADMP_ERR_HIP: return "HIP runtime error";
    case ADMP_ERR_UNSUPPORTED: return "unsupported configuration";
    case ADMP_ERR_SCRATCH:
      return "the problem needs more than the static scratch of admp_pme_fused_noalloc "
             "(ADMP_NOALLOC_SCRATCH_BYTES); use admp_pme_fused with a workspace";
    default: return "unknown status";
  }
}

// ------------------------------------------------------------------ nblist
int admp_nblist_workspace_bytes(int64_t n, const double* box, double cutoff, size_t* bytes) {
  if (n < 0 || !(cutoff > 0.0) || !bytes) return ADMP_ERR_INVALID;
  if (box && !box_ok(box)) return ADMP_ERR_INVALID;
  NbLayout L;
  nb_layout(n, make_grid(box, cutoff, box == nullptr), &L, true);
  *bytes = L.bytes;
  return ADMP_OK;
}

int admp_nblist_count(const double* pos, int64_t n, const double* box, double cutoff,
                      void* workspace, size_t ws_bytes, int64_t* n_pairs, void* stream) {
  if (n < 0 || !(cutoff > 0.0) || !n_pairs || (n > 0 && (!pos || !workspace)))
    return ADMP_ERR_INVALID;
  if (n > (int64_t)1 << 30) return ADMP_ERR_UNSUPPORTED;
  if (box && !box_ok(box)) return ADMP_ERR_INVALID;
  const CellGrid cg = make_grid(box, cutoff, box == nullptr);
  NbLayout L;
  nb_layout(n, cg, &L, true);
  if (ws_bytes < L.bytes) return ADMP_ERR_INVALID;
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)workspace;
  hipError_t e = launch_bin(pos, n, cg, ws, L, s);
  if (e != hipSuccess) return ADMP_ERR_HIP;
  e = launch_pairs(0, n, cg, ws, L, nullptr, 0, nullptr, nullptr, s);
  if (e != hipSuccess) return ADMP_ERR_HIP;
  int64_t total = 0;
  if (n > 0) {
    e = hipMemcpyAsync(&total, ws + L.offsets + sizeof(int64_t) * n, sizeof(int64_t),
                       hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return ADMP_ERR_HIP;
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return ADMP_ERR_HIP;
  }
  *n_pairs = total;
  return ADMP_OK;
}

int admp_nblist_fill(const double* pos, int64_t n, const double* box, double cutoff,
                     void* workspace, size_t ws_bytes, int64_t* pairs, int64_t capacity,
                     int64_t* row_ptr, int32_t* reverse, int32_t* scratch,
                     int64_t* buffer_scales, void* stream) {
  if (n < 0 || !(cutoff > 0.0) || capacity < 0 || (n > 0 && (!pos || !workspace)))
    return ADMP_ERR_INVALID;
  if (capacity > 0 && !pairs) return ADMP_ERR_INVALID;
  if (reverse && capacity > 0 && !scratch) return ADMP_ERR_INVALID;
  if (box && !box_ok(box)) return ADMP_ERR_INVALID;
  const CellGrid cg = make_grid(box, cutoff, box == nullptr);
  NbLayout L;
  nb_layout(n, cg, &L, true);
  if (ws_bytes < L.bytes) return ADMP_ERR_INVALID;
  (void)pos;
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)workspace;
  int* keys = reverse ? (int*)scratch : nullptr;
  hipError_t e = launch_pairs(1, n, cg, ws, L, pairs, capacity, keys, buffer_scales, s);
  if (e == hipSuccess && (row_ptr || reverse) && n > 0)
    e = launch_nb_reverse(n, cg, ws, L, keys, capacity, row_ptr, (int*)reverse, s);
  return hip_status(e);
}

int admp_max_displacement(const double* pos, const double* ref, int64_t n, const double* box,
                          double* out, void* stream) {
  if (n < 0 || !out || (n > 0 && (!pos || !ref))) return ADMP_ERR_INVALID;
  if (box && !box_ok(box)) return ADMP_ERR_INVALID;
  const CellGrid cg = make_grid(box, 1.0, box == nullptr);
  return hip_status(launch_max_disp(pos, ref, n, cg, out, (hipStream_t)stream));
}

int admp_pair_ds(const double* pos, const int64_t* pairs, int64_t n_pairs, const double* box,
                 double* ds, double* dr, void* stream) {
  if (n_pairs < 0 || (n_pairs > 0 && (!pos || !pairs || !ds))) return ADMP_ERR_INVALID;
  if (box && !box_ok(box)) return ADMP_ERR_INVALID;
  const CellGrid cg = make_grid(box, 1.0, box == nullptr);
  return hip_status(launch_pair_ds(pos, pairs, n_pairs, cg, ds, dr, (hipStream_t)stream));
}

int admp_pair_ds_backward_workspace_bytes(int64_t n_pairs, size_t* bytes) {
  if (n_pairs < 0 || !bytes) return ADMP_ERR_INVALID;
  *bytes = align256(pair_ds_bwd_scratch_bytes(n_pairs));
  return ADMP_OK;
}

int admp_pair_ds_backward(const double* grad_ds, const int64_t* pairs, int64_t n_pairs,
                          const double* pos, const double* box, const double* dr,
                          const double* ds, double* grad_pos, double* grad_box,
                          const int64_t* row_ptr, const int32_t* reverse, int64_t n_atoms,
                          void* workspace, size_t ws_bytes, void* stream) {
  if (n_pairs < 0 || (n_pairs > 0 && (!grad_ds || !pairs || !pos || !ds || !grad_pos)))
    return ADMP_ERR_INVALID;
  if ((row_ptr != nullptr) != (reverse != nullptr) || (row_ptr && n_atoms < 0))
    return ADMP_ERR_INVALID;
  if (box && !box_ok(box)) return ADMP_ERR_INVALID;
  if (grad_box && box && n_pairs > 0 &&
      (!workspace || ws_bytes < pair_ds_bwd_scratch_bytes(n_pairs)))
    return ADMP_ERR_INVALID;
  const CellGrid cg = make_grid(box, 1.0, box == nullptr);
  return hip_status(launch_pair_ds_bwd(grad_ds, pairs, n_pairs, pos, cg, dr, ds, grad_pos,
                                       grad_box, (double*)workspace, row_ptr,
                                       (const int*)reverse, n_atoms, (hipStream_t)stream));
}

// ------------------------------------------------------------- pair energy
int admp_pair_energy_workspace_bytes(int64_t nf, size_t* bytes) {
  if (nf < 0 || !bytes) return ADMP_ERR_INVALID;
  *bytes = align256((size_t)std::max<int64_t>(nf, 1) * kPairBlocksPerFrame * sizeof(double));
  return ADMP_OK;
}

int admp_pair_energy(int32_t mode, int64_t nf, int64_t n, int64_t n_pairs, const double* q,
                     const double* eta, const int64_t* pairs, const double* ds,
                     const void* bs, double kappa, double dielectric, double* energy,
                     double* grad_ds, double* grad_q, double* grad_eta, void* workspace,
                     size_t ws_bytes, void* stream) {
  const bool bs_i64 = (mode & ADMP_PAIR_BS_INT64) != 0;
  mode &= ~ADMP_PAIR_BS_INT64;
  if (mode < 0 || mode > 2 || nf < 1 || nf > 65535 || n < 0 || n_pairs < 0 || !energy ||
      !workspace)
    return ADMP_ERR_INVALID;
  if (n_pairs > 0 && (!q || !pairs || !ds || !bs)) return ADMP_ERR_INVALID;
  if (mode == 2 && n_pairs > 0 && !eta) return ADMP_ERR_INVALID;
  if (grad_eta && (mode != 2 || !eta)) return ADMP_ERR_INVALID;
  size_t need;
  admp_pair_energy_workspace_bytes(nf, &need);
  if (ws_bytes < need) return ADMP_ERR_INVALID;
  const double* bsd = bs_i64 ? nullptr : (const double*)bs;
  const int64_t* bsl = bs_i64 ? (const int64_t*)bs : nullptr;
  return hip_status(launch_pair_energy(mode, nf, n, n_pairs, q, eta, pairs, ds, bsd, bsl, kappa,
                                       dielectric, energy, grad_ds, grad_q, grad_eta,
                                       (double*)workspace, (hipStream_t)stream));
}

int admp_pair_scales(const int64_t* pairs, int64_t n_pairs, const double* ds, double lo,
                     double hi, int64_t* scales, void* stream) {
  if (n_pairs < 0 || (n_pairs > 0 && (!pairs || !scales))) return ADMP_ERR_INVALID;
  return hip_status(launch_pair_scales(pairs, n_pairs, ds, lo, hi, scales, (hipStream_t)stream));
}

// --------------------------------------------------------------------- PME
int admp_pme_workspace_bytes(const int32_t* kmesh, int32_t precision, int64_t n, size_t* bytes) {
  if (!kmesh_ok(kmesh) || (precision != ADMP_F32 && precision != ADMP_F64) || n < 0 || !bytes)
    return ADMP_ERR_INVALID;
  *bytes = pme_layout(kmesh, precision, n).bytes;
  return ADMP_OK;
}

int admp_pme_spread(const double* pos, const double* q, int64_t n, const double* box,
                    const int32_t* kmesh, int32_t precision, void* mesh, void* workspace,
                    size_t ws_bytes, void* stream) {
  if (!kmesh_ok(kmesh) || !box_ok(box) || n < 0 || !mesh || (n > 0 && (!pos || !q)))
    return ADMP_ERR_INVALID;
  if (precision != ADMP_F32 && precision != ADMP_F64) return ADMP_ERR_INVALID;
  const Geom g = make_geom(box, kmesh);
  hipStream_t s = (hipStream_t)stream;
  const size_t es = elem_size(precision);
  const PmeLayout L = pme_layout(kmesh, precision, n);
  const CellGrid cg = spread_grid(box, kmesh);
  const TileGeom tg = make_tile_geom(cg, kmesh, es);
  if (workspace && ws_bytes >= L.bytes && tg.ok && n > 0) {
    // bin the atoms into the spread cells, then one LDS tile per cell
    char* nbws = (char*)workspace + L.nb;
    if (launch_bin(pos, n, cg, nbws, L.nbl, s) != hipSuccess) return ADMP_ERR_HIP;
    const int* sorted = (const int*)(nbws + L.nbl.sorted);
    const int* start = (const int*)(nbws + L.nbl.start);
    if (precision == ADMP_F32)
      return hip_status(launch_spread_tiled<float>(pos, q, n, g, tg, cg.ncell, sorted, start,
                                                   (float*)mesh, s));
    return hip_status(launch_spread_tiled<double>(pos, q, n, g, tg, cg.ncell, sorted, start,
                                                  (double*)mesh, s));
  }
  if (precision == ADMP_F32) return hip_status(launch_spread<float>(pos, q, n, g, (float*)mesh, s));
  return hip_status(launch_spread<double>(pos, q, n, g, (double*)mesh, s));
}

int admp_pme_convolve(const void* mesh, const double* box, const int32_t* kmesh,
                      int32_t precision, int32_t convention, double kappa, double dielectric,
                      double phi_scale, void* phi, double* energy, double* dEdA,
                      void* workspace, size_t ws_bytes, void* stream) {
  if (!kmesh_ok(kmesh) || !box_ok(box) || !mesh || !(kappa > 0.0) || !workspace)
    return ADMP_ERR_INVALID;
  if (convention != ADMP_KCONV_REFERENCE && convention != ADMP_KCONV_COLUMNS)
    return ADMP_ERR_INVALID;
  if (precision != ADMP_F32 && precision != ADMP_F64) return ADMP_ERR_INVALID;
  const PmeLayout L = pme_layout(kmesh, precision);
  if (ws_bytes < L.bytes) return ADMP_ERR_INVALID;
  const Geom g = make_geom(box, kmesh);
  char* ws = (char*)workspace;
  hipStream_t s = (hipStream_t)stream;
  if (precision == ADMP_F32)
    return hip_status(launch_convolve<float>((const float*)mesh, g, convention, kappa, dielectric,
                                             phi_scale, (float*)phi, energy, dEdA,
                                             (float2*)(ws + L.spec), (double*)(ws + L.partials),
                                             s));
  return hip_status(launch_convolve<double>((const double*)mesh, g, convention, kappa, dielectric,
                                            phi_scale, (double*)phi, energy, dEdA,
                                            (double2*)(ws + L.spec), (double*)(ws + L.partials),
                                            s));
}

int admp_pme_gather(const double* pos, const double* q, int64_t n, const double* box,
                    const int32_t* kmesh, int32_t precision, const void* phi,
                    double* grad_pos, double* dEdq, double* dEdA, int32_t accumulate,
                    void* workspace, size_t ws_bytes, void* stream) {
  if (!kmesh_ok(kmesh) || !box_ok(box) || n < 0 || !phi || (n > 0 && (!pos || !q)))
    return ADMP_ERR_INVALID;
  if (precision != ADMP_F32 && precision != ADMP_F64) return ADMP_ERR_INVALID;
  const PmeLayout L = pme_layout(kmesh, precision);
  if (dEdA && (!workspace || ws_bytes < L.bytes)) return ADMP_ERR_INVALID;
  const Geom g = make_geom(box, kmesh);
  double* partials = workspace ? (double*)((char*)workspace + L.partials) : nullptr;
  hipStream_t s = (hipStream_t)stream;
  if (precision == ADMP_F32)
    return hip_status(launch_gather<float>(pos, q, n, g, (const float*)phi, grad_pos, dEdq, dEdA,
                                           accumulate, partials, s));
  return hip_status(launch_gather<double>(pos, q, n, g, (const double*)phi, grad_pos, dEdq, dEdA,
                                          accumulate, partials, s));
}

int admp_pme_box_grad(const double* box, const double* dEdA_k, const double* dEdA_gather,
                      const double* energy, double* grad_box, void* stream) {
  if (!box_ok(box) || !dEdA_k || !energy || !grad_box) return ADMP_ERR_INVALID;
  return hip_status(launch_box_grad(make_geom(box, nullptr), dEdA_k, dEdA_gather, energy,
                                    grad_box, (hipStream_t)stream));
}

int admp_ewald_self_terms(const double* q, int64_t nf, int64_t n, const double* box,
                          double kappa, double dielectric, double* e_self, double* e_nonneutral,
                          void* stream) {
  if (nf < 0 || n < 0 || (nf > 0 && (!box || !e_self || !e_nonneutral || (n > 0 && !q))))
    return ADMP_ERR_INVALID;
  return hip_status(launch_self_terms(q, nf, n, box, kappa, dielectric, e_self, e_nonneutral,
                                      nullptr, nullptr, nullptr, nullptr, (hipStream_t)stream));
}

int admp_ewald_self_terms_backward(const double* q, int64_t nf, int64_t n, const double* box,
                                   double kappa, double dielectric, const double* g_self,
                                   const double* g_nonneutral, double* grad_q, double* grad_box,
                                   void* stream) {
  if (nf < 0 || n < 0 || (nf > 0 && (!box || !grad_q || (n > 0 && !q)))) return ADMP_ERR_INVALID;
  return hip_status(launch_self_terms(q, nf, n, box, kappa, dielectric, nullptr, nullptr, g_self,
                                      g_nonneutral, grad_q, grad_box, (hipStream_t)stream));
}

// --------------------------------------------------------------------- QEq
int admp_qeq_workspace_bytes(const admp_qeq_problem* problem, size_t* bytes) {
  QeqProblem P;
  if (!bytes || !qeq_problem(problem, false, &P)) return ADMP_ERR_INVALID;
  *bytes = qeq_layout(P).bytes;
  return ADMP_OK;
}

int admp_qeq_matvec(const admp_qeq_problem* problem, const double* v, double* out,
                    void* workspace, size_t ws_bytes, void* stream) {
  QeqProblem P;
  if (!qeq_problem(problem, false, &P) || !workspace || (P.n > 0 && (!v || !out)))
    return ADMP_ERR_INVALID;
  if (ws_bytes < qeq_layout(P).bytes) return ADMP_ERR_INVALID;
  if (P.n == 0) return ADMP_OK;
  hipStream_t s = (hipStream_t)stream;
  return P.precision == ADMP_F32 ? qeq_matvec_impl<float>(P, v, out, workspace, s)
                                 : qeq_matvec_impl<double>(P, v, out, workspace, s);
}

int admp_qeq_solve(const admp_qeq_problem* problem, const double* q0, double tol,
                   int32_t max_iter, int32_t check_every, double* q, int32_t* iters,
                   double* residual, void* workspace, size_t ws_bytes, void* stream) {
  QeqProblem P;
  if (!qeq_problem(problem, true, &P) || !workspace || !iters || !residual || max_iter < 0 ||
      check_every < 1 || !(tol >= 0.0) || (P.n > 0 && (!q0 || !q)))
    return ADMP_ERR_INVALID;
  if (ws_bytes < qeq_layout(P).bytes) return ADMP_ERR_INVALID;
  *iters = 0;
  *residual = 0.0;
  if (P.n == 0) return ADMP_OK;
  hipStream_t s = (hipStream_t)stream;
  return P.precision == ADMP_F32
             ? qeq_solve_impl<float>(P, q0, tol, max_iter, check_every, q, iters, residual,
                                     workspace, s)
             : qeq_solve_impl<double>(P, q0, tol, max_iter, check_every, q, iters, residual,
                                      workspace, s);
}

// ------------------------------------------------------------------- fused
// A batch of one: the first kernel writes the system table from its arguments.
int admp_pme_fused_workspace_bytes(int64_t n, const double* box, double cutoff,
                                   const int32_t* kmesh, int32_t precision, size_t* bytes) {
  if (n < 0 || !box_ok(box) || !(cutoff > 0.0) || !kmesh_ok(kmesh) || !bytes)
    return ADMP_ERR_INVALID;
  return admp_pme_batch_workspace_bytes(1, &n, box, &cutoff, kmesh, precision, bytes);
}

int admp_pme_fused(const double* pos, const double* q, int64_t n, const double* box,
                   double cutoff, double kappa, const int32_t* kmesh, int32_t precision,
                   int32_t convention, double dielectric, double* energy, double* forces,
                   double* dEdbox, void* workspace, size_t ws_bytes, void* stream) {
  if (n < 0 || !box_ok(box) || !(cutoff > 0.0) || !(kappa > 0.0) || !kmesh_ok(kmesh) ||
      !energy || !workspace || (n > 0 && (!pos || !q || !forces)))
    return ADMP_ERR_INVALID;
  return admp_pme_batch(1, &n, box, &cutoff, &kappa, kmesh, precision, convention, dielectric,
                        pos, q, energy, forces, dEdbox, workspace, ws_bytes, stream);
}

// The fused evaluation in the library's static scratch: no workspace argument,
// no device allocation, every kernel on the caller's stream.  The launch
// sequence clears what it accumulates into (counters, cursors, mesh, gradient
// rows, tickets) before use, so the previous call's contents do not matter;
// calls that overlap on one GPU share its scratch and race.
int admp_pme_fused_noalloc(const double* pos, const double* q, int64_t n, const double* box,
                           double cutoff, double kappa, const int32_t* kmesh, int32_t precision,
                           int32_t convention, double dielectric, double* energy, double* forces,
                           double* dEdbox, void* stream) {
  if (n < 0 || !box_ok(box) || !(cutoff > 0.0) || !(kappa > 0.0) || !std::isfinite(kappa) ||
      !kmesh_ok(kmesh) || !energy || (n > 0 && (!pos || !q || !forces)) ||
      (precision != ADMP_F32 && precision != ADMP_F64) ||
      (convention != ADMP_KCONV_REFERENCE && convention != ADMP_KCONV_COLUMNS))
    return ADMP_ERR_INVALID;
  if (n > (int64_t)1 << 30) return ADMP_ERR_UNSUPPORTED;
  const BatchPlan P = plan_batch(1, &n, box, &cutoff, &kappa, kmesh, precision, convention,
                                 dielectric);
  if (P.status != ADMP_OK) return P.status;
  if (P.ctot * kSub + 1 >= ((int64_t)1 << 31)) return ADMP_ERR_UNSUPPORTED;
  if (P.bytes > (size_t)ADMP_NOALLOC_SCRATCH_BYTES) return ADMP_ERR_SCRATCH;
  char* ws = (char*)scratch_of_current_device();
  if (!ws) return ADMP_ERR_HIP;
  return run_batch(P, pos, q, energy, forces, dEdbox, ws, (hipStream_t)stream, nullptr, true,
                   false);
}

// ------------------------------------------------------------------- batch
static int batch_args_ok(int32_t n_sys, const int64_t* n_atoms, const double* boxes,
                         const double* cutoffs, const int32_t* kmesh, int32_t precision) {
  if (n_sys < 1 || n_sys > (1 << 20) || !n_atoms || !boxes || !cutoffs || !kmesh)
    return ADMP_ERR_INVALID;
  if (precision != ADMP_F32 && precision != ADMP_F64) return ADMP_ERR_INVALID;
  int64_t ntot = 0;
  for (int s = 0; s < n_sys; ++s) {
    if (n_atoms[s] < 0 || !box_ok(boxes + 9 * s) || !(cutoffs[s] > 0.0) ||
        !kmesh_ok(kmesh + 3 * s))
      return ADMP_ERR_INVALID;
    ntot += n_atoms[s];
  }
  if (ntot > (int64_t)1 << 30) return ADMP_ERR_UNSUPPORTED;
  return ADMP_OK;
}

static int batch_plan_ok(const BatchPlan& P) {
  if (P.status != ADMP_OK) return P.status;
  // cell ids and sorted slots are 32-bit
  if (P.ctot * kSub + 1 >= ((int64_t)1 << 31)) return ADMP_ERR_UNSUPPORTED;
  return ADMP_OK;
}

int admp_pme_batch_workspace_bytes(int32_t n_sys, const int64_t* n_atoms, const double* boxes,
                                   const double* cutoffs, const int32_t* kmesh,
                                   int32_t precision, size_t* bytes) {
  if (!bytes) return ADMP_ERR_INVALID;
  const int st = batch_args_ok(n_sys, n_atoms, boxes, cutoffs, kmesh, precision);
  if (st != ADMP_OK) return st;
  const BatchPlan P = plan_batch(n_sys, n_atoms, boxes, cutoffs, nullptr, kmesh, precision,
                                 ADMP_KCONV_REFERENCE, 1.0);
  const int ps = batch_plan_ok(P);
  if (ps != ADMP_OK) return ps;
  *bytes = P.bytes;
  return ADMP_OK;
}

int admp_pme_batch_system_bytes(int32_t n_sys, const int64_t* n_atoms, const double* boxes,
                                const double* cutoffs, const int32_t* kmesh, int32_t precision,
                                size_t* bytes) {
  if (!bytes) return ADMP_ERR_INVALID;
  const int st = batch_args_ok(n_sys, n_atoms, boxes, cutoffs, kmesh, precision);
  if (st != ADMP_OK) return st;
  for (int32_t s = 0; s < n_sys; ++s) {
    const BatchPlan P = plan_batch(1, n_atoms + s, boxes + 9 * (int64_t)s, cutoffs + s, nullptr,
                                   kmesh + 3 * (int64_t)s, precision, ADMP_KCONV_REFERENCE, 1.0);
    const int ps = batch_plan_ok(P);
    if (ps != ADMP_OK) return ps;
    bytes[s] = P.bytes;
  }
  return ADMP_OK;
}

static int batch_call(int32_t n_sys, const int64_t* n_atoms, const double* boxes,
                      const double* cutoffs, const double* kappas, const int32_t* kmesh,
                      int32_t precision, int32_t convention, double dielectric,
                      const double* pos, const double* q, double* energy, double* forces,
                      double* dEdbox, void* workspace, size_t ws_bytes, void* stream, int mode,
                      int32_t flags = 0) {
  // mode 0: upload and run; 1: upload only (synchronous); 2: run on the uploaded table
  const int st = batch_args_ok(n_sys, n_atoms, boxes, cutoffs, kmesh, precision);
  if (st != ADMP_OK) return st;
  if (!kappas || !workspace || (mode != 1 && !energy)) return ADMP_ERR_INVALID;
  if (convention != ADMP_KCONV_REFERENCE && convention != ADMP_KCONV_COLUMNS)
    return ADMP_ERR_INVALID;
  int64_t ntot = 0;
  for (int s = 0; s < n_sys; ++s) {
    if (!(kappas[s] > 0.0) || !std::isfinite(kappas[s])) return ADMP_ERR_INVALID;
    ntot += n_atoms[s];
  }
  if (mode != 1 && ntot > 0 && (!pos || !q || !forces)) return ADMP_ERR_INVALID;
  const BatchPlan P = plan_batch(n_sys, n_atoms, boxes, cutoffs, kappas, kmesh, precision,
                                 convention, dielectric);
  const int ps = batch_plan_ok(P);
  if (ps != ADMP_OK) return ps;
  if (ws_bytes < P.bytes) return ADMP_ERR_INVALID;
  hipStream_t s = (hipStream_t)stream;
  if (mode == 1) return n_sys > 1 ? upload_table(P, (char*)workspace, s, true) : ADMP_OK;
  if (flags & ~ADMP_BATCH_NO_OVERLAP) return ADMP_ERR_INVALID;
  StageMarks pm = stage_marks(s);
  return run_batch(P, pos, q, energy, forces, dEdbox, (char*)workspace, s, pm.ev, mode == 0,
                   (flags & ADMP_BATCH_NO_OVERLAP) == 0);
}

int admp_pme_batch(int32_t n_sys, const int64_t* n_atoms, const double* boxes,
                   const double* cutoffs, const double* kappas, const int32_t* kmesh,
                   int32_t precision, int32_t convention, double dielectric, const double* pos,
                   const double* q, double* energy, double* forces, double* dEdbox,
                   void* workspace, size_t ws_bytes, void* stream) {
  return batch_call(n_sys, n_atoms, boxes, cutoffs, kappas, kmesh, precision, convention,
                    dielectric, pos, q, energy, forces, dEdbox, workspace, ws_bytes, stream, 0);
}

int admp_pme_batch_ex(int32_t n_sys, const int64_t* n_atoms, const double* boxes,
                      const double* cutoffs, const double* kappas, const int32_t* kmesh,
                      int32_t precision, int32_t convention, double dielectric, const double* pos,
                      const double* q, double* energy, double* forces, double* dEdbox,
                      int32_t flags, void* workspace, size_t ws_bytes, void* stream) {
  return batch_call(n_sys, n_atoms, boxes, cutoffs, kappas, kmesh, precision, convention,
                    dielectric, pos, q, energy, forces, dEdbox, workspace, ws_bytes, stream, 0,
                    flags);
}

// ------------------------------------------------------------------- batched QEq
static int qeq_batch_args(int32_t n_sys, const int64_t* n_atoms, const double* boxes,
                          const double* cutoffs, const int32_t* kmesh, int32_t precision,
                          int64_t n_pairs, BatchPlan* P, size_t* bytes,
                          const double* kappas = nullptr, int32_t convention = ADMP_KCONV_REFERENCE,
                          double dielectric = 1.0) {
  const int st = batch_args_ok(n_sys, n_atoms, boxes, cutoffs, kmesh, precision);
  if (st != ADMP_OK) return st;
  if (n_pairs < 0) return ADMP_ERR_INVALID;
  *P = plan_batch(n_sys, n_atoms, boxes, cutoffs, kappas, kmesh, precision, convention,
                  dielectric);
  const int ps = batch_plan_ok(*P);
  if (ps != ADMP_OK) return ps;
  const size_t extra = qeq_batch_extra_bytes(*P, n_pairs);
  if (extra == 0) return ADMP_ERR_UNSUPPORTED;
  *bytes = P->bytes + extra;
  return ADMP_OK;
}

int admp_qeq_solve_batch_workspace_bytes(int32_t n_sys, const int64_t* n_atoms,
                                         const double* boxes, const double* cutoffs,
                                         const int32_t* kmesh, int32_t precision,
                                         int64_t n_pairs, size_t* bytes) {
  if (!bytes) return ADMP_ERR_INVALID;
  BatchPlan P;
  return qeq_batch_args(n_sys, n_atoms, boxes, cutoffs, kmesh, precision, n_pairs, &P, bytes);
}

int admp_qeq_solve_batch(int32_t n_sys, const int64_t* n_atoms, const double* boxes,
                         const double* cutoffs, const double* kappas, const int32_t* kmesh,
                         int32_t precision, int32_t convention, double dielectric, int32_t flags,
                         const double* pos, const int64_t* pairs, const double* ds,
                         const double* bs, int64_t n_pairs, const double* eta,
                         const double* hdiag, const double* chi, const double* precond,
                         const double* total_charge, const double* q0, const double* tols,
                         int32_t max_iter, int32_t check_every, double* q, int32_t* iters,
                         double* residuals, void* workspace, size_t ws_bytes, void* stream) {
  if (!kappas || !total_charge || !tols || !iters || !residuals || !workspace || max_iter < 0 ||
      check_every < 1 || !(dielectric > 0.0) ||
      (flags & ~ADMP_QEQ_DAMPING) != 0 ||
      (convention != ADMP_KCONV_REFERENCE && convention != ADMP_KCONV_COLUMNS))
    return ADMP_ERR_INVALID;
  for (int32_t k = 0; k < n_sys; ++k)
    if (!(kappas[k] > 0.0) || !std::isfinite(kappas[k]) || !std::isfinite(total_charge[k]) ||
        !(tols[k] >= 0.0))
      return ADMP_ERR_INVALID;
  BatchPlan P;
  size_t need = 0;
  const int st = qeq_batch_args(n_sys, n_atoms, boxes, cutoffs, kmesh, precision, n_pairs, &P,
                                &need, kappas, convention, dielectric);
  if (st != ADMP_OK) return st;
  if (ws_bytes < need) return ADMP_ERR_INVALID;
  if (P.ntot > 0 && (!pos || !hdiag || !chi || !q0 || !q || ((flags & ADMP_QEQ_DAMPING) && !eta)))
    return ADMP_ERR_INVALID;
  if (n_pairs > 0 && (!pairs || !ds || !bs)) return ADMP_ERR_INVALID;
  QeqBatch B;
  B.pos = pos;
  B.pairs = pairs;
  B.ds = ds;
  B.bs = bs;
  B.n_pairs = n_pairs;
  B.eta = eta;
  B.hdiag = hdiag;
  B.chi = chi;
  B.precond = precond;
  B.total_charge = total_charge;
  B.damping = (flags & ADMP_QEQ_DAMPING) ? 1 : 0;
  return qeq_solve_batch(P, B, q0, tols, max_iter, check_every, q, iters, residuals,
                         (char*)workspace, (hipStream_t)stream);
}


int admp_pme_batch_upload(int32_t n_sys, const int64_t* n_atoms, const double* boxes,
                          const double* cutoffs, const double* kappas, const int32_t* kmesh,
                          int32_t precision, int32_t convention, double dielectric,
                          void* workspace, size_t ws_bytes, void* stream) {
  return batch_call(n_sys, n_atoms, boxes, cutoffs, kappas, kmesh, precision, convention,
                    dielectric, nullptr, nullptr, nullptr, nullptr, nullptr, workspace, ws_bytes,
                    stream, 1);
}

int admp_pme_batch_run(int32_t n_sys, const int64_t* n_atoms, const double* boxes,
                       const double* cutoffs, const double* kappas, const int32_t* kmesh,
                       int32_t precision, int32_t convention, double dielectric,
                       const double* pos, const double* q, double* energy, double* forces,
                       double* dEdbox, void* workspace, size_t ws_bytes, void* stream) {
  return batch_call(n_sys, n_atoms, boxes, cutoffs, kappas, kmesh, precision, convention,
                    dielectric, pos, q, energy, forces, dEdbox, workspace, ws_bytes, stream, 2);
}

}  // extern "C"
