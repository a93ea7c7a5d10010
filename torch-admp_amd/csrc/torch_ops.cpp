// The admp operators (namespace admp) registered with the PyTorch dispatcher
// in C++: argument checks, output allocation, per-stream workspaces and the
// autograd formulas of every operator, each body one or two calls of the C
// ABI (include/admp_amd.h) on the caller's current stream.
//
// Python loads this library with torch.ops.load_library; a C++ program that
// runs a TorchScript model using these operators (torch::jit::load) dlopens
// it before loading the model.  The HIP library is found on the first
// operator call: $ADMP_AMD_LIB when set, else libadmp_amd.so in this
// library's directory.
//
// Derivatives.  Every operator that returns an energy has an autograd
// formula made of other admp operators, so gradients of gradients work where
// the reference's callers take them:
//  * pair_energy -> pair_energy_grads (dE/dq, dE/dds, dE/deta in one kernel).
//    Its own backward, for an upstream v on dE/dq: d/dq = the pair operator
//    applied to v (dE/dq at charges v, the operator being symmetric); d/dds
//    and d/deta by the polarisation identity [f(q + v) - f(q - v)] / 2, those
//    gradients being quadratic in the charges.
//  * pme_reciprocal -> pme_reciprocal_grads (gather and box terms).  Its own
//    backward for v on dE/dq: d/dq = the reciprocal potential of charges v at
//    the atoms; d/dr and d/dbox by the same polarisation identity.
//  * ewald_self_terms -> ewald_self_terms_backward, or the same formulas as
//    differentiable tensor expressions when a graph of the gradient is built.
// That covers the second derivative in the charges that the reference's
// QEqForceModule.calc_hessian takes (src/torch_admp/qeq.py:507-514), mixed
// with positions, box, ds and eta.  Differentiating forces, dE/dds or dE/deta
// again raises NotImplementedError rather than returning a partial gradient.
#include <dlfcn.h>

#include <ATen/ATen.h>
#include <ATen/core/dispatch/Dispatcher.h>
#include <c10/core/DeviceGuard.h>
#include <c10/hip/HIPStream.h>
#include <torch/csrc/autograd/custom_function.h>
#include <torch/library.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <optional>
#include <string>
#include <tuple>
#include <vector>

#include "admp_amd.h"

namespace {

using at::Tensor;
using std::optional;
using torch::autograd::AutogradContext;
using torch::autograd::variable_list;

// ------------------------------------------------------------------ the C ABI
struct Api {
  decltype(&admp_status_string) status_string = nullptr;
  decltype(&admp_nblist_workspace_bytes) nblist_workspace_bytes = nullptr;
  decltype(&admp_nblist_count) nblist_count = nullptr;
  decltype(&admp_nblist_fill) nblist_fill = nullptr;
  decltype(&admp_max_displacement) max_displacement = nullptr;
  decltype(&admp_pair_ds) pair_ds = nullptr;
  decltype(&admp_pair_ds_backward_workspace_bytes) pair_ds_backward_workspace_bytes = nullptr;
  decltype(&admp_pair_ds_backward) pair_ds_backward = nullptr;
  decltype(&admp_pair_scales) pair_scales = nullptr;
  decltype(&admp_pair_energy_workspace_bytes) pair_energy_workspace_bytes = nullptr;
  decltype(&admp_pair_energy) pair_energy = nullptr;
  decltype(&admp_ewald_self_terms) ewald_self_terms = nullptr;
  decltype(&admp_ewald_self_terms_backward) ewald_self_terms_backward = nullptr;
  decltype(&admp_pme_workspace_bytes) pme_workspace_bytes = nullptr;
  decltype(&admp_pme_spread) pme_spread = nullptr;
  decltype(&admp_pme_convolve) pme_convolve = nullptr;
  decltype(&admp_pme_gather) pme_gather = nullptr;
  decltype(&admp_pme_box_grad) pme_box_grad = nullptr;
};

Api g_api;
std::atomic<bool> g_bound{false};
std::mutex g_bind_mu;

template <class F>
void resolve(void* h, const char* name, F& fn) {
  fn = reinterpret_cast<F>(dlsym(h, name));
  TORCH_CHECK(fn != nullptr, "torch_admp_amd: the HIP library lacks ", name);
}

void bind_path(const std::string& path) {
  std::lock_guard<std::mutex> g(g_bind_mu);
  void* h = dlopen(path.c_str(), RTLD_NOW | RTLD_NOLOAD);
  if (!h) h = dlopen(path.c_str(), RTLD_NOW);
  TORCH_CHECK(h != nullptr, "torch_admp_amd: cannot open the HIP library ", path, ": ",
              dlerror(), " (build it with make -C torch-admp_amd/csrc)");
  Api a;
#define ADMP_RESOLVE(name) resolve(h, "admp_" #name, a.name)
  ADMP_RESOLVE(status_string);
  ADMP_RESOLVE(nblist_workspace_bytes);
  ADMP_RESOLVE(nblist_count);
  ADMP_RESOLVE(nblist_fill);
  ADMP_RESOLVE(max_displacement);
  ADMP_RESOLVE(pair_ds);
  ADMP_RESOLVE(pair_ds_backward_workspace_bytes);
  ADMP_RESOLVE(pair_ds_backward);
  ADMP_RESOLVE(pair_scales);
  ADMP_RESOLVE(pair_energy_workspace_bytes);
  ADMP_RESOLVE(pair_energy);
  ADMP_RESOLVE(ewald_self_terms);
  ADMP_RESOLVE(ewald_self_terms_backward);
  ADMP_RESOLVE(pme_workspace_bytes);
  ADMP_RESOLVE(pme_spread);
  ADMP_RESOLVE(pme_convolve);
  ADMP_RESOLVE(pme_gather);
  ADMP_RESOLVE(pme_box_grad);
#undef ADMP_RESOLVE
  g_api = a;
  g_bound.store(true, std::memory_order_release);
}

// $ADMP_AMD_LIB, else libadmp_amd.so beside this library.
std::string default_library_path() {
  const char* env = std::getenv("ADMP_AMD_LIB");
  if (env != nullptr && *env != '\0') return env;
  Dl_info info;
  if (dladdr(reinterpret_cast<void*>(&default_library_path), &info) != 0 &&
      info.dli_fname != nullptr) {
    std::string self(info.dli_fname);
    const size_t slash = self.rfind('/');
    if (slash != std::string::npos) return self.substr(0, slash) + "/libadmp_amd.so";
  }
  return "libadmp_amd.so";
}

const Api& api() {
  if (!g_bound.load(std::memory_order_acquire)) bind_path(default_library_path());
  return g_api;
}

void check(int status, const char* what) {
  if (status != ADMP_OK) {
    const char* msg = g_api.status_string != nullptr ? g_api.status_string(status) : "";
    TORCH_CHECK(false, what, " failed with status ", status, ": ", msg);
  }
}

// ------------------------------------------------------------------ helpers
void require_gpu(const Tensor& t) {
  TORCH_CHECK(!t.defined() || t.is_cuda(),
              "torch_admp_amd runs only on a ROCm GPU (no CPU fallback); got a tensor on ",
              t.device());
}

void require_gpu(const optional<Tensor>& t) {
  if (t.has_value()) require_gpu(*t);
}

bool has(const optional<Tensor>& t) { return t.has_value() && t->defined(); }

Tensor undef_or(const optional<Tensor>& t) { return has(t) ? *t : Tensor(); }

optional<Tensor> opt(const Tensor& t) {
  return t.defined() ? optional<Tensor>(t) : optional<Tensor>();
}

Tensor f64(const Tensor& t) { return t.to(at::kDouble).contiguous(); }

double* dptr(const Tensor& t) { return t.defined() ? t.data_ptr<double>() : nullptr; }

void* stream_of(const Tensor& t) {
  return reinterpret_cast<void*>(c10::hip::getCurrentHIPStream(t.device().index()).stream());
}

// The nine box values on the host: from box_host when the caller holds that
// copy, else copied from the (device) box, which waits for its stream.
struct HostBox {
  std::array<double, 9> v{};
  bool present = false;
  const double* ptr() const { return present ? v.data() : nullptr; }
};

HostBox host_box(const optional<Tensor>& box, const optional<Tensor>& box_host) {
  HostBox h;
  if (!has(box)) return h;
  const Tensor& src = has(box_host) ? *box_host : *box;
  const Tensor c = src.to(at::kCPU, at::kDouble).contiguous();
  TORCH_CHECK(c.numel() == 9, "box: 9 values (3, 3) expected, got ", c.numel());
  std::memcpy(h.v.data(), c.data_ptr<double>(), 9 * sizeof(double));
  h.present = true;
  return h;
}

// One workspace per (device, stream, kind): work on one stream runs in
// order, so a later call of the same kind on that stream reuses it; calls on
// other streams get their own.  Heap-held and never destroyed (the device
// runtime may be gone at process exit).
enum WsKind : int { kWsNblist = 0, kWsPairDs = 1, kWsPairEnergy = 2, kWsPme = 3 };
using WsKey = std::tuple<int, intptr_t, int>;
std::map<WsKey, Tensor>* g_ws = new std::map<WsKey, Tensor>();
std::mutex g_ws_mu;

Tensor workspace(const Tensor& like, void* stream, int kind, size_t nbytes) {
  std::lock_guard<std::mutex> g(g_ws_mu);
  const WsKey key{static_cast<int>(like.device().index()), reinterpret_cast<intptr_t>(stream),
                  kind};
  auto it = g_ws->find(key);
  if (it == g_ws->end()) {
    if (g_ws->size() >= 64) g_ws->clear();
    it = g_ws->emplace(key, Tensor()).first;
  }
  Tensor& w = it->second;
  if (!w.defined() || static_cast<size_t>(w.numel()) < nbytes)
    w = at::empty({static_cast<int64_t>(std::max<size_t>(nbytes, 1))},
                  like.options().dtype(at::kByte));
  return w;
}

const char* kSecondOrder =
    "second derivatives through %s (differentiating forces, dE/dds or dE/deta again) are "
    "not implemented; derivatives of dE/dq are";

void second_order(const char* what) {
  char buf[256];
  std::snprintf(buf, sizeof(buf), kSecondOrder, what);
  TORCH_CHECK_NOT_IMPLEMENTED(false, buf);
}

// ------------------------------------------------------------------ neighbour pairs
// Ordered pairs (P, 2) int64 with 0.01 < d < cutoff (box) or d < cutoff (box
// None), both orderings (src/torch_admp/nblist.py:335-490); with rows, also
// row_ptr (n + 1), reverse (P, int32) and the buffer scales (P, int64) of
// admp_nblist_fill.  One stream synchronisation returns P.
std::tuple<Tensor, Tensor, Tensor, Tensor> search(const Tensor& positions,
                                                  const optional<Tensor>& box, double cutoff,
                                                  const optional<Tensor>& box_host, bool rows) {
  require_gpu(positions);
  at::AutoDispatchBelowADInplaceOrView below;
  const Api& a = api();
  c10::DeviceGuard dg(positions.device());
  const Tensor pos = f64(positions);
  TORCH_CHECK(pos.dim() == 2 && pos.size(1) == 3, "positions: (n, 3) expected, got ",
              pos.sizes());
  const int64_t n = pos.size(0);
  const HostBox hb = host_box(box, box_host);
  void* s = stream_of(pos);
  size_t nbytes = 0;
  check(a.nblist_workspace_bytes(n, hb.ptr(), cutoff, &nbytes), "admp_nblist_workspace_bytes");
  const Tensor ws = workspace(pos, s, kWsNblist, nbytes);
  int64_t np = 0;
  check(a.nblist_count(pos.data_ptr<double>(), n, hb.ptr(), cutoff, ws.data_ptr(), nbytes, &np,
                       s),
        "admp_nblist_count");
  const auto lopt = pos.options().dtype(at::kLong);
  Tensor pairs = at::empty({np, 2}, lopt);
  Tensor row_ptr, reverse, scratch, scales;
  if (rows) {
    row_ptr = at::empty({n + 1}, lopt);
    reverse = at::empty({np}, pos.options().dtype(at::kInt));
    scratch = at::empty({np}, pos.options().dtype(at::kInt));
    scales = at::empty({np}, lopt);
    if (n == 0) row_ptr.zero_();
  } else {
    row_ptr = reverse = scales = at::empty({0}, lopt);
  }
  check(a.nblist_fill(pos.data_ptr<double>(), n, hb.ptr(), cutoff, ws.data_ptr(), nbytes,
                      pairs.data_ptr<int64_t>(), np,
                      rows ? row_ptr.data_ptr<int64_t>() : nullptr,
                      rows ? reverse.data_ptr<int32_t>() : nullptr,
                      rows ? scratch.data_ptr<int32_t>() : nullptr,
                      rows ? scales.data_ptr<int64_t>() : nullptr, s),
        "admp_nblist_fill");
  return {pairs, row_ptr, reverse, scales};
}

Tensor neighbor_pairs(const Tensor& positions, const optional<Tensor>& box, double cutoff,
                      const optional<Tensor>& box_host) {
  return std::get<0>(search(positions, box, cutoff, box_host, false));
}

std::tuple<Tensor, Tensor, Tensor, Tensor> neighbor_list(const Tensor& positions,
                                                         const optional<Tensor>& box,
                                                         double cutoff,
                                                         const optional<Tensor>& box_host) {
  return search(positions, box, cutoff, box_host, true);
}

// Buffer scales (P,) int64 of an ordered pair list (src/torch_admp/nblist.py:
// 492-515): 1 where i < j, and, with ds, lo < ds < hi; else 0.
Tensor pair_scales(const Tensor& pairs, const optional<Tensor>& ds, double lo, double hi) {
  require_gpu(pairs);
  require_gpu(ds);
  at::AutoDispatchBelowADInplaceOrView below;
  const Api& a = api();
  c10::DeviceGuard dg(pairs.device());
  const Tensor prs = pairs.to(at::kLong).contiguous();
  TORCH_CHECK(prs.dim() == 2 && prs.size(1) == 2, "pairs: (P, 2) expected, got ", prs.sizes());
  const int64_t np = prs.size(0);
  const Tensor d = has(ds) ? f64(*ds) : Tensor();
  TORCH_CHECK(!d.defined() || d.numel() == np, "ds: P values expected");
  Tensor out = at::empty({np}, prs.options());
  check(a.pair_scales(prs.data_ptr<int64_t>(), np, dptr(d), lo, hi, out.data_ptr<int64_t>(),
                      stream_of(prs)),
        "admp_pair_scales");
  return out;
}

// max_i |positions_i - reference_i| (minimum image when box is given), a
// 0-dim float64 device tensor.
Tensor max_displacement(const Tensor& positions, const Tensor& reference,
                        const optional<Tensor>& box, const optional<Tensor>& box_host) {
  require_gpu(positions);
  require_gpu(reference);
  at::AutoDispatchBelowADInplaceOrView below;
  const Api& a = api();
  c10::DeviceGuard dg(positions.device());
  const Tensor pos = f64(positions);
  const Tensor ref = f64(reference);
  TORCH_CHECK(pos.sizes() == ref.sizes() && pos.dim() == 2 && pos.size(1) == 3,
              "positions and reference: equal (n, 3) shapes expected");
  Tensor out = at::empty({}, pos.options());
  const HostBox hb = host_box(box, box_host);
  check(a.max_displacement(pos.data_ptr<double>(), ref.data_ptr<double>(), pos.size(0), hb.ptr(),
                           out.data_ptr<double>(), stream_of(pos)),
        "admp_max_displacement");
  return out;
}

// ------------------------------------------------------------------ pair distances
// ds = |pbc_shift(r_j - r_i)| (src/torch_admp/nblist.py:517-551,
// src/torch_admp/spatial.py:15-36), |r_j - r_i| when box is None; (P,) in the
// positions' dtype.  row_ptr and reverse (from neighbor_list, for these
// pairs) only change how the backward sums.
Tensor pair_distance(const Tensor& positions, const optional<Tensor>& box, const Tensor& pairs,
                     const optional<Tensor>& box_host, const optional<Tensor>& row_ptr,
                     const optional<Tensor>& reverse) {
  require_gpu(positions);
  require_gpu(pairs);
  at::AutoDispatchBelowADInplaceOrView below;
  const Api& a = api();
  c10::DeviceGuard dg(positions.device());
  const Tensor pos = f64(positions);
  const Tensor prs = pairs.to(at::kLong).contiguous();
  TORCH_CHECK(prs.dim() == 2 && prs.size(1) == 2, "pairs: (P, 2) expected, got ", prs.sizes());
  const int64_t np = prs.size(0);
  Tensor ds = at::empty({np}, pos.options());
  const HostBox hb = host_box(box, box_host);
  check(a.pair_ds(pos.data_ptr<double>(), prs.data_ptr<int64_t>(), np, hb.ptr(),
                  ds.data_ptr<double>(), nullptr, stream_of(pos)),
        "admp_pair_ds");
  (void)row_ptr;
  (void)reverse;
  return ds.to(positions.scalar_type());
}

// (dL/dpositions, dL/dbox) of L(ds); dL/dbox is zero when box is None and
// empty when want_box is false (the box term is then skipped).  The
// displacement is recomputed from the positions.  With row_ptr and reverse
// each atom's row is summed in one pass (no per-pair atomics).
std::tuple<Tensor, Tensor> pair_distance_backward(
    const Tensor& grad_ds, const Tensor& positions, const optional<Tensor>& box,
    const Tensor& pairs, const Tensor& ds, const optional<Tensor>& box_host,
    const optional<Tensor>& row_ptr, const optional<Tensor>& reverse, bool want_box) {
  require_gpu(grad_ds);
  require_gpu(positions);
  require_gpu(pairs);
  at::AutoDispatchBelowADInplaceOrView below;
  const Api& a = api();
  c10::DeviceGuard dg(positions.device());
  const Tensor pos = f64(positions);
  const int64_t n = pos.size(0);
  const bool rows = has(row_ptr) && has(reverse) && row_ptr->numel() == n + 1;
  const Tensor prs = pairs.to(at::kLong).contiguous();
  const int64_t np = prs.size(0);
  const Tensor g = f64(grad_ds);
  const Tensor ds64 = f64(ds);
  Tensor gpos = at::zeros_like(pos);
  Tensor gbox = want_box ? at::zeros({3, 3}, pos.options()) : at::empty({0}, pos.options());
  const HostBox hb = host_box(box, box_host);
  void* s = stream_of(pos);
  size_t nbytes = 0;
  check(a.pair_ds_backward_workspace_bytes(np, &nbytes), "admp_pair_ds_backward_workspace_bytes");
  const Tensor ws = workspace(pos, s, kWsPairDs, nbytes);
  Tensor rp, rv;
  if (rows) {
    rp = row_ptr->to(at::kLong).contiguous();
    rv = reverse->to(at::kInt).contiguous();
  }
  check(a.pair_ds_backward(g.data_ptr<double>(), prs.data_ptr<int64_t>(), np,
                           pos.data_ptr<double>(), hb.ptr(), nullptr, ds64.data_ptr<double>(),
                           gpos.data_ptr<double>(),
                           hb.present && want_box ? gbox.data_ptr<double>() : nullptr,
                           rows ? rp.data_ptr<int64_t>() : nullptr,
                           rows ? rv.data_ptr<int32_t>() : nullptr, n, ws.data_ptr(), nbytes, s),
        "admp_pair_ds_backward");
  if (has(box) && want_box) gbox = gbox.to(box->device(), box->scalar_type());
  return {gpos.to(positions.scalar_type()), gbox};
}

// ------------------------------------------------------------------ pair energies
constexpr int64_t kPairDamped = 2;

struct PairOut {
  Tensor energy, gq, gds, geta;
};

PairOut pair_call(const Tensor& q, const Tensor& ds, const Tensor& pairs, const Tensor& bs,
                  const optional<Tensor>& eta, int64_t mode, double kappa, double eps,
                  bool want_q, bool want_ds, bool want_eta) {
  require_gpu(q);
  require_gpu(ds);
  require_gpu(pairs);
  require_gpu(bs);
  require_gpu(eta);
  at::AutoDispatchBelowADInplaceOrView below;
  const Api& a = api();
  c10::DeviceGuard dg(q.device());
  TORCH_CHECK(q.dim() == 2, "q: (nf, n) expected, got ", q.sizes());
  TORCH_CHECK(pairs.dim() == 3 && pairs.size(2) == 2 && pairs.size(0) == q.size(0),
              "pairs: (nf, P, 2) expected, got ", pairs.sizes());
  const int64_t nf = q.size(0), n = q.size(1), np = pairs.size(1);
  TORCH_CHECK(ds.numel() == nf * np && bs.numel() == nf * np,
              "ds and bs: nf * P values expected");
  const Tensor q64 = f64(q), ds64 = f64(ds);
  // int64 scales (the neighbour list's) are read as they are
  const bool bs_i64 = bs.scalar_type() == at::kLong;
  const Tensor bsc = bs_i64 ? bs.contiguous() : f64(bs);
  const Tensor prs = pairs.to(at::kLong).contiguous();
  const Tensor eta64 = has(eta) ? f64(*eta) : Tensor();
  const auto o = q64.options();
  PairOut r;
  r.energy = at::empty({nf}, o);
  if (want_q) r.gq = at::empty({nf, n}, o);
  if (want_ds) r.gds = at::empty({nf, np}, o);
  if (want_eta && mode == kPairDamped && eta64.defined()) r.geta = at::empty({nf, n}, o);
  void* s = stream_of(q64);
  size_t nbytes = 0;
  check(a.pair_energy_workspace_bytes(nf, &nbytes), "admp_pair_energy_workspace_bytes");
  const Tensor ws = workspace(q64, s, kWsPairEnergy, nbytes);
  TORCH_CHECK(mode >= 0 && mode <= kPairDamped, "pair mode 0, 1 or 2 expected, got ", mode);
  const int32_t m = static_cast<int32_t>(mode) | (bs_i64 ? ADMP_PAIR_BS_INT64 : 0);
  check(a.pair_energy(m, nf, n, np, q64.data_ptr<double>(), dptr(eta64), prs.data_ptr<int64_t>(),
                      ds64.data_ptr<double>(), bsc.data_ptr(), kappa, eps,
                      r.energy.data_ptr<double>(), dptr(r.gds), dptr(r.gq), dptr(r.geta),
                      ws.data_ptr(), nbytes, s),
        "admp_pair_energy");
  return r;
}

// Per-frame pair energy over (nf, P) pair rows, float64 accumulation: mode 0
// eps erfc(kappa d) qi qj / d bs (src/torch_admp/pme.py:312-349), 1 eps qi qj
// / d bs (pme.py:529-561), 2 the Gaussian damping -eps erfc(d / (2 eta_ij))
// qi qj / d bs (src/torch_admp/qeq.py:106-121).
Tensor pair_energy(const Tensor& q, const Tensor& ds, const Tensor& pairs, const Tensor& bs,
                   const optional<Tensor>& eta, int64_t mode, double kappa, double eps) {
  return pair_call(q, ds, pairs, bs, eta, mode, kappa, eps, false, false, false)
      .energy.to(q.scalar_type());
}

// (dE/dq (nf, n), dE/dds (shape of ds), dE/deta (nf, n)) of pair_energy; the
// ones not wanted (and dE/deta outside mode 2) are zero.
// The gradients of a pair_call in the shapes and dtypes of q, ds and eta;
// those not computed are undefined.
std::tuple<Tensor, Tensor, Tensor> pair_grads_as_inputs(const PairOut& r, const Tensor& q,
                                                        const Tensor& ds,
                                                        const optional<Tensor>& eta) {
  at::AutoDispatchBelowADInplaceOrView below;
  Tensor gq = r.gq.defined() ? r.gq.to(q.scalar_type()) : Tensor();
  Tensor gds = r.gds.defined() ? r.gds.reshape(ds.sizes()).to(ds.scalar_type()) : Tensor();
  Tensor geta = r.geta.defined() ? r.geta.reshape(eta->sizes()).to(eta->scalar_type()) : Tensor();
  return {gq, gds, geta};
}

std::tuple<Tensor, Tensor, Tensor> pair_energy_grads(const Tensor& q, const Tensor& ds,
                                                     const Tensor& pairs, const Tensor& bs,
                                                     const optional<Tensor>& eta, int64_t mode,
                                                     double kappa, double eps, bool want_q,
                                                     bool want_ds, bool want_eta) {
  const PairOut r = pair_call(q, ds, pairs, bs, eta, mode, kappa, eps, want_q, want_ds, want_eta);
  auto [gq, gds, geta] = pair_grads_as_inputs(r, q, ds, eta);
  at::AutoDispatchBelowADInplaceOrView below;
  const int64_t nf = q.size(0), n = q.size(1);
  if (!gq.defined()) gq = at::zeros({nf, n}, q.options());
  if (!gds.defined()) gds = at::zeros(ds.sizes(), ds.options());
  if (!geta.defined()) geta = at::zeros({nf, n}, q.options());
  return {gq, gds, geta};
}

// ------------------------------------------------------------------ self and non-neutral terms
// Per frame (q (nf, n), box (nf, 3, 3) on the device): the self energy
// -eps kappa / sqrt(pi) sum q^2 and the non-neutral energy
// -eps pi Q^2 / (2 V kappa^2) (src/torch_admp/pme.py:448-493), one kernel,
// no host copy of the box.
std::tuple<Tensor, Tensor> ewald_self_terms(const Tensor& q, const Tensor& box, double kappa,
                                            double eps) {
  require_gpu(q);
  require_gpu(box);
  at::AutoDispatchBelowADInplaceOrView below;
  const Api& a = api();
  c10::DeviceGuard dg(q.device());
  const Tensor q64 = f64(q), b64 = f64(box);
  TORCH_CHECK(q64.dim() == 2 && b64.numel() == 9 * q64.size(0),
              "q (nf, n) and box (nf, 3, 3) expected");
  const int64_t nf = q64.size(0), n = q64.size(1);
  Tensor e_self = at::empty({nf}, q64.options());
  Tensor e_nn = at::empty({nf}, q64.options());
  check(a.ewald_self_terms(q64.data_ptr<double>(), nf, n, b64.data_ptr<double>(), kappa, eps,
                           e_self.data_ptr<double>(), e_nn.data_ptr<double>(), stream_of(q64)),
        "admp_ewald_self_terms");
  return {e_self.to(q.scalar_type()), e_nn.to(q.scalar_type())};
}

// (dL/dq (nf, n), dL/dbox (nf, 3, 3)) of ewald_self_terms for upstream
// gradients g_self, g_nn (nf each; None for zero).
std::tuple<Tensor, Tensor> ewald_self_terms_backward(const Tensor& q, const Tensor& box,
                                                     double kappa, double eps,
                                                     const optional<Tensor>& g_self,
                                                     const optional<Tensor>& g_nn) {
  require_gpu(q);
  require_gpu(box);
  at::AutoDispatchBelowADInplaceOrView below;
  const Api& a = api();
  c10::DeviceGuard dg(q.device());
  const Tensor q64 = f64(q), b64 = f64(box);
  TORCH_CHECK(q64.dim() == 2 && b64.numel() == 9 * q64.size(0),
              "q (nf, n) and box (nf, 3, 3) expected");
  const int64_t nf = q64.size(0), n = q64.size(1);
  Tensor dq = at::empty({nf, n}, q64.options());
  Tensor dbox = at::empty({nf, 3, 3}, q64.options());
  const Tensor gs = has(g_self) ? f64(*g_self).reshape({nf}) : Tensor();
  const Tensor gn = has(g_nn) ? f64(*g_nn).reshape({nf}) : Tensor();
  check(a.ewald_self_terms_backward(q64.data_ptr<double>(), nf, n, b64.data_ptr<double>(), kappa,
                                    eps, dptr(gs), dptr(gn), dq.data_ptr<double>(),
                                    dbox.data_ptr<double>(), stream_of(q64)),
        "admp_ewald_self_terms_backward");
  return {dq.to(q.scalar_type()), dbox.to(box.scalar_type())};
}

// ------------------------------------------------------------------ PME reciprocal space
std::array<int32_t, 3> mesh_of(at::IntArrayRef kmesh) {
  TORCH_CHECK(kmesh.size() == 3, "kmesh: 3 values expected, got ", kmesh.size());
  return {static_cast<int32_t>(kmesh[0]), static_cast<int32_t>(kmesh[1]),
          static_cast<int32_t>(kmesh[2])};
}

// One frame of the SPME reciprocal energy (src/torch_admp/pme.py:351-446):
// spread, r2c FFT, influence function, c2r FFT.  Returns (energy (0-dim, the
// positions' dtype), the potential mesh dE/dQ (empty unless need_potential),
// dE/dA through the k vectors (9,) float64).
std::tuple<Tensor, Tensor, Tensor> pme_reciprocal(const Tensor& positions, const Tensor& box,
                                                  const Tensor& q, at::IntArrayRef kmesh,
                                                  double kappa, double eps, int64_t precision,
                                                  int64_t convention, bool need_potential,
                                                  const optional<Tensor>& box_host) {
  require_gpu(positions);
  require_gpu(q);
  at::AutoDispatchBelowADInplaceOrView below;
  const Api& a = api();
  c10::DeviceGuard dg(positions.device());
  const Tensor pos = f64(positions);
  const Tensor q64 = f64(q).reshape({-1});
  const int64_t n = pos.size(0);
  TORCH_CHECK(pos.dim() == 2 && pos.size(1) == 3 && q64.numel() == n,
              "positions (n, 3) and n charges expected");
  const HostBox hb = host_box(box, box_host);
  const std::array<int32_t, 3> km = mesh_of(kmesh);
  const auto mesh_type = precision == ADMP_F32 ? at::kFloat : at::kDouble;
  Tensor mesh = at::empty({static_cast<int64_t>(km[0]) * km[1] * km[2]},
                          pos.options().dtype(mesh_type));
  void* s = stream_of(pos);
  size_t nbytes = 0;
  check(a.pme_workspace_bytes(km.data(), static_cast<int32_t>(precision), n, &nbytes),
        "admp_pme_workspace_bytes");
  const Tensor ws = workspace(pos, s, kWsPme, nbytes);
  check(a.pme_spread(pos.data_ptr<double>(), q64.data_ptr<double>(), n, hb.ptr(), km.data(),
                     static_cast<int32_t>(precision), mesh.data_ptr(), ws.data_ptr(), nbytes, s),
        "admp_pme_spread");
  Tensor energy = at::empty({}, pos.options());
  Tensor dEdA = at::empty({9}, pos.options());
  check(a.pme_convolve(mesh.data_ptr(), hb.ptr(), km.data(), static_cast<int32_t>(precision),
                       static_cast<int32_t>(convention), kappa, eps, 1.0,
                       need_potential ? mesh.data_ptr() : nullptr, energy.data_ptr<double>(),
                       dEdA.data_ptr<double>(), ws.data_ptr(), nbytes, s),
        "admp_pme_convolve");
  Tensor phi = need_potential ? mesh : mesh.new_empty({0});
  return {energy.to(positions.scalar_type()), phi, dEdA};
}

// (dE/dpositions (n, 3), dE/dq (shape of q), dE/dbox (3, 3)) of pme_reciprocal
// from its potential mesh: the gather (the transpose of the spread) plus the
// k-vector and volume terms of dE/dbox.
std::tuple<Tensor, Tensor, Tensor> pme_reciprocal_grads(
    const Tensor& positions, const Tensor& box, const Tensor& q, const Tensor& potential,
    const Tensor& dEdA_k, const Tensor& energy, at::IntArrayRef kmesh, double kappa, double eps,
    int64_t precision, int64_t convention, const optional<Tensor>& box_host) {
  require_gpu(positions);
  require_gpu(q);
  require_gpu(potential);
  at::AutoDispatchBelowADInplaceOrView below;
  const Api& a = api();
  c10::DeviceGuard dg(positions.device());
  (void)kappa;
  (void)eps;
  (void)convention;
  const Tensor pos = f64(positions);
  const Tensor q64 = f64(q).reshape({-1});
  const int64_t n = pos.size(0);
  TORCH_CHECK(pos.dim() == 2 && pos.size(1) == 3 && q64.numel() == n,
              "positions (n, 3) and n charges expected");
  const HostBox hb = host_box(box, box_host);
  const std::array<int32_t, 3> km = mesh_of(kmesh);
  const Tensor phi = potential.contiguous();
  TORCH_CHECK(phi.numel() == static_cast<int64_t>(km[0]) * km[1] * km[2] &&
                  phi.scalar_type() == (precision == ADMP_F32 ? at::kFloat : at::kDouble),
              "potential: the mesh of pme_reciprocal (need_potential=True) expected");
  Tensor gpos = at::empty({n, 3}, pos.options());
  Tensor gq = at::empty({n}, pos.options());
  Tensor dEdA = at::empty({9}, pos.options());
  void* s = stream_of(pos);
  size_t nbytes = 0;
  check(a.pme_workspace_bytes(km.data(), static_cast<int32_t>(precision), n, &nbytes),
        "admp_pme_workspace_bytes");
  const Tensor ws = workspace(pos, s, kWsPme, nbytes);
  check(a.pme_gather(pos.data_ptr<double>(), q64.data_ptr<double>(), n, hb.ptr(), km.data(),
                     static_cast<int32_t>(precision), phi.data_ptr(), gpos.data_ptr<double>(),
                     gq.data_ptr<double>(), dEdA.data_ptr<double>(), 0, ws.data_ptr(), nbytes, s),
        "admp_pme_gather");
  Tensor gbox = at::empty({3, 3}, pos.options());
  const Tensor dk = f64(dEdA_k);
  const Tensor e64 = f64(energy);
  check(a.pme_box_grad(hb.ptr(), dk.data_ptr<double>(), dEdA.data_ptr<double>(),
                       e64.data_ptr<double>(), gbox.data_ptr<double>(), s),
        "admp_pme_box_grad");
  return {gpos.to(positions.scalar_type()), gq.to(q.scalar_type()).reshape(q.sizes()),
          gbox.to(box.device(), box.scalar_type())};
}

// ------------------------------------------------------------------ dispatcher handles
using PdSig = Tensor(const Tensor&, const optional<Tensor>&, const Tensor&,
                     const optional<Tensor>&, const optional<Tensor>&, const optional<Tensor>&);
using PdbSig = std::tuple<Tensor, Tensor>(const Tensor&, const Tensor&, const optional<Tensor>&,
                                          const Tensor&, const Tensor&, const optional<Tensor>&,
                                          const optional<Tensor>&, const optional<Tensor>&, bool);
using PeSig = Tensor(const Tensor&, const Tensor&, const Tensor&, const Tensor&,
                     const optional<Tensor>&, int64_t, double, double);
using PegSig = std::tuple<Tensor, Tensor, Tensor>(const Tensor&, const Tensor&, const Tensor&,
                                                  const Tensor&, const optional<Tensor>&,
                                                  int64_t, double, double, bool, bool, bool);
using StSig = std::tuple<Tensor, Tensor>(const Tensor&, const Tensor&, double, double);
using StbSig = std::tuple<Tensor, Tensor>(const Tensor&, const Tensor&, double, double,
                                          const optional<Tensor>&, const optional<Tensor>&);
using PrSig = std::tuple<Tensor, Tensor, Tensor>(const Tensor&, const Tensor&, const Tensor&,
                                                 at::IntArrayRef, double, double, int64_t,
                                                 int64_t, bool, const optional<Tensor>&);
using PrgSig = std::tuple<Tensor, Tensor, Tensor>(const Tensor&, const Tensor&, const Tensor&,
                                                  const Tensor&, const Tensor&, const Tensor&,
                                                  at::IntArrayRef, double, double, int64_t,
                                                  int64_t, const optional<Tensor>&);

template <class Sig>
c10::TypedOperatorHandle<Sig> op_handle(const char* name) {
  return c10::Dispatcher::singleton().findSchemaOrThrow(name, "").typed<Sig>();
}

#define ADMP_OP(fn, sig, name)                                  \
  const c10::TypedOperatorHandle<sig>& fn() {                   \
    static const c10::TypedOperatorHandle<sig> h = op_handle<sig>(name); \
    return h;                                                   \
  }
ADMP_OP(op_pair_distance, PdSig, "admp::pair_distance")
ADMP_OP(op_pair_distance_backward, PdbSig, "admp::pair_distance_backward")
ADMP_OP(op_pair_energy, PeSig, "admp::pair_energy")
ADMP_OP(op_pair_energy_grads, PegSig, "admp::pair_energy_grads")
ADMP_OP(op_ewald_self_terms, StSig, "admp::ewald_self_terms")
ADMP_OP(op_ewald_self_terms_backward, StbSig, "admp::ewald_self_terms_backward")
ADMP_OP(op_pme_reciprocal, PrSig, "admp::pme_reciprocal")
ADMP_OP(op_pme_reciprocal_grads, PrgSig, "admp::pme_reciprocal_grads")
#undef ADMP_OP

// ------------------------------------------------------------------ autograd
// needs_input_grad counts the tensor arguments that had a value at the
// forward call (the graph edges), so the edge of an argument after an
// optional one depends on whether that one was given.

struct PairDistanceFn : public torch::autograd::Function<PairDistanceFn> {
  static Tensor forward(AutogradContext* ctx, const Tensor& positions, const optional<Tensor>& box,
                        const Tensor& pairs, const optional<Tensor>& box_host,
                        const optional<Tensor>& row_ptr, const optional<Tensor>& reverse) {
    at::AutoDispatchBelowADInplaceOrView below;
    Tensor ds = op_pair_distance().call(positions, box, pairs, box_host, row_ptr, reverse);
    ctx->save_for_backward({positions, undef_or(box), pairs, ds, undef_or(box_host),
                            undef_or(row_ptr), undef_or(reverse)});
    ctx->set_materialize_grads(false);
    return ds;
  }

  static variable_list backward(AutogradContext* ctx, variable_list grads) {
    variable_list out(6);
    if (!grads[0].defined()) return out;
    const variable_list s = ctx->get_saved_variables();
    const bool want_box = s[1].defined() && ctx->needs_input_grad(1);
    auto [gpos, gbox] = op_pair_distance_backward().call(grads[0], s[0], opt(s[1]), s[2],
                                                         s[3].detach(), opt(s[4]), opt(s[5]),
                                                         opt(s[6]), want_box);
    if (ctx->needs_input_grad(0)) out[0] = gpos;
    if (want_box) out[1] = gbox;
    return out;
  }
};

struct PairDistanceBackwardFn : public torch::autograd::Function<PairDistanceBackwardFn> {
  static variable_list forward(AutogradContext* ctx, const Tensor& grad_ds,
                               const Tensor& positions, const optional<Tensor>& box,
                               const Tensor& pairs, const Tensor& ds,
                               const optional<Tensor>& box_host, const optional<Tensor>& row_ptr,
                               const optional<Tensor>& reverse, bool want_box) {
    at::AutoDispatchBelowADInplaceOrView below;
    auto [gpos, gbox] = op_pair_distance_backward().call(grad_ds, positions, box, pairs, ds,
                                                         box_host, row_ptr, reverse, want_box);
    ctx->set_materialize_grads(false);
    return {gpos, gbox};
  }

  static variable_list backward(AutogradContext*, variable_list grads) {
    if (grads[0].defined() || grads[1].defined()) second_order("the pair distances");
    return variable_list(9);
  }
};

struct PairEnergyFn : public torch::autograd::Function<PairEnergyFn> {
  // keep: a backward may follow (grad mode on at the call).  The gradients
  // that backward would ask for then come out of the forward pass itself and
  // are kept; a backward building a graph recomputes them through
  // pair_energy_grads, which is differentiable.
  static Tensor forward(AutogradContext* ctx, const Tensor& q, const Tensor& ds,
                        const Tensor& pairs, const Tensor& bs, const optional<Tensor>& eta,
                        int64_t mode, double kappa, double eps, bool keep) {
    at::AutoDispatchBelowADInplaceOrView below;
    const bool want_q = keep && q.requires_grad();
    const bool want_ds = keep && ds.requires_grad();
    const bool want_eta = keep && has(eta) && eta->requires_grad() && mode == kPairDamped;
    Tensor e, gq, gds, geta;
    if (want_q || want_ds || want_eta) {
      const PairOut r = pair_call(q, ds, pairs, bs, eta, mode, kappa, eps, want_q, want_ds,
                                  want_eta);
      e = r.energy.to(q.scalar_type());
      std::tie(gq, gds, geta) = pair_grads_as_inputs(r, q, ds, eta);
    } else {
      e = op_pair_energy().call(q, ds, pairs, bs, eta, mode, kappa, eps);
    }
    ctx->save_for_backward({q, ds, pairs, bs, undef_or(eta), gq, gds, geta});
    ctx->saved_data["mode"] = mode;
    ctx->saved_data["kappa"] = kappa;
    ctx->saved_data["eps"] = eps;
    ctx->set_materialize_grads(false);
    return e;
  }

  static variable_list backward(AutogradContext* ctx, variable_list grads) {
    variable_list out(9);
    if (!grads[0].defined()) return out;
    const variable_list s = ctx->get_saved_variables();
    const int64_t mode = ctx->saved_data["mode"].toInt();
    const double kappa = ctx->saved_data["kappa"].toDouble();
    const double eps = ctx->saved_data["eps"].toDouble();
    const bool need_q = ctx->needs_input_grad(0);
    const bool need_ds = ctx->needs_input_grad(1);
    const bool want_eta = s[4].defined() && mode == kPairDamped && ctx->needs_input_grad(4);
    Tensor gq = s[5], gds = s[6], geta = s[7];
    const bool cached = (!need_q || gq.defined()) && (!need_ds || gds.defined()) &&
                        (!want_eta || geta.defined());
    if (at::GradMode::is_enabled() || !cached)
      std::tie(gq, gds, geta) = op_pair_energy_grads().call(s[0], s[1], s[2], s[3], opt(s[4]),
                                                            mode, kappa, eps, need_q, need_ds,
                                                            want_eta);
    const Tensor ge = grads[0].reshape({-1, 1});
    if (need_q) out[0] = ge * gq;
    if (need_ds) out[1] = ge * gds;
    if (want_eta) out[4] = ge * geta;
    return out;
  }
};

struct PairEnergyGradsFn : public torch::autograd::Function<PairEnergyGradsFn> {
  static variable_list forward(AutogradContext* ctx, const Tensor& q, const Tensor& ds,
                               const Tensor& pairs, const Tensor& bs,
                               const optional<Tensor>& eta, int64_t mode, double kappa,
                               double eps, bool want_q, bool want_ds, bool want_eta) {
    at::AutoDispatchBelowADInplaceOrView below;
    auto [gq, gds, geta] = op_pair_energy_grads().call(q, ds, pairs, bs, eta, mode, kappa, eps,
                                                       want_q, want_ds, want_eta);
    ctx->save_for_backward({q, ds, pairs, bs, undef_or(eta)});
    ctx->saved_data["mode"] = mode;
    ctx->saved_data["kappa"] = kappa;
    ctx->saved_data["eps"] = eps;
    ctx->set_materialize_grads(false);
    return {gq, gds, geta};
  }

  static variable_list backward(AutogradContext* ctx, variable_list grads) {
    if (grads[1].defined() || grads[2].defined())
      second_order("dE/dds or dE/deta of the pair energy");
    variable_list out(11);
    if (!grads[0].defined()) return out;
    const variable_list s = ctx->get_saved_variables();
    const int64_t mode = ctx->saved_data["mode"].toInt();
    const double kappa = ctx->saved_data["kappa"].toDouble();
    const double eps = ctx->saved_data["eps"].toDouble();
    const bool need_ds = ctx->needs_input_grad(1);
    const bool want_eta = s[4].defined() && mode == kPairDamped && ctx->needs_input_grad(4);
    if (ctx->needs_input_grad(0)) {
      // the pair operator is symmetric: v^T (C q) differentiates to C v
      out[0] = std::get<0>(op_pair_energy_grads().call(grads[0], s[1], s[2], s[3], opt(s[4]),
                                                       mode, kappa, eps, true, false, false));
    }
    if (need_ds || want_eta) {
      const Tensor v = grads[0].to(s[0].scalar_type());
      auto plus = op_pair_energy_grads().call(s[0] + v, s[1], s[2], s[3], opt(s[4]), mode, kappa,
                                              eps, false, need_ds, want_eta);
      auto minus = op_pair_energy_grads().call(s[0] - v, s[1], s[2], s[3], opt(s[4]), mode, kappa,
                                               eps, false, need_ds, want_eta);
      if (need_ds) out[1] = 0.5 * (std::get<1>(plus) - std::get<1>(minus));
      if (want_eta) out[4] = 0.5 * (std::get<2>(plus) - std::get<2>(minus));
    }
    return out;
  }
};

struct EwaldSelfTermsFn : public torch::autograd::Function<EwaldSelfTermsFn> {
  static variable_list forward(AutogradContext* ctx, const Tensor& q, const Tensor& box,
                               double kappa, double eps) {
    at::AutoDispatchBelowADInplaceOrView below;
    auto [e_self, e_nn] = op_ewald_self_terms().call(q, box, kappa, eps);
    ctx->save_for_backward({q, box});
    ctx->saved_data["kappa"] = kappa;
    ctx->saved_data["eps"] = eps;
    ctx->set_materialize_grads(false);
    return {e_self, e_nn};
  }

  static variable_list backward(AutogradContext* ctx, variable_list grads) {
    variable_list out(4);
    const Tensor& g_self = grads[0];
    const Tensor& g_nn = grads[1];
    if (!g_self.defined() && !g_nn.defined()) return out;
    const variable_list s = ctx->get_saved_variables();
    const Tensor& q = s[0];
    const Tensor& box = s[1];
    const double kappa = ctx->saved_data["kappa"].toDouble();
    const double eps = ctx->saved_data["eps"].toDouble();
    Tensor dq, dbox;
    if (at::GradMode::is_enabled()) {
      // a graph of the gradient is being built (second derivatives, as the
      // reference's Hessian takes them): the same formulas as tensor
      // expressions, differentiable in q, box and the upstream gradients
      const double cs = -eps * kappa / std::sqrt(M_PI);
      const Tensor ra = box.select(1, 0), rb = box.select(1, 1), rc = box.select(1, 2);
      // d det / d box row by row: b x c, c x a, a x b (the cofactor matrix)
      const Tensor dvol = at::stack(
          {at::linalg_cross(rb, rc, -1), at::linalg_cross(rc, ra, -1),
           at::linalg_cross(ra, rb, -1)},
          1);
      const Tensor vol = (ra * dvol.select(1, 0)).sum(-1);
      const Tensor qt = q.sum(-1);
      const Tensor cn = (-eps * M_PI) / (2.0 * vol * kappa * kappa);
      dq = at::zeros_like(q);
      dbox = at::zeros_like(box);
      if (g_self.defined()) dq = dq + (2.0 * cs) * g_self.unsqueeze(-1) * q;
      if (g_nn.defined()) {
        dq = dq + (2.0 * g_nn * cn * qt).unsqueeze(-1);
        dbox = dbox - (g_nn * cn * qt * qt / vol).reshape({-1, 1, 1}) * dvol;
      }
    } else {
      std::tie(dq, dbox) =
          op_ewald_self_terms_backward().call(q, box, kappa, eps, opt(g_self), opt(g_nn));
    }
    if (ctx->needs_input_grad(0)) out[0] = dq;
    if (ctx->needs_input_grad(1)) out[1] = dbox;
    return out;
  }
};

struct EwaldSelfTermsBackwardFn : public torch::autograd::Function<EwaldSelfTermsBackwardFn> {
  static variable_list forward(AutogradContext* ctx, const Tensor& q, const Tensor& box,
                               double kappa, double eps, const optional<Tensor>& g_self,
                               const optional<Tensor>& g_nn) {
    at::AutoDispatchBelowADInplaceOrView below;
    auto [dq, dbox] = op_ewald_self_terms_backward().call(q, box, kappa, eps, g_self, g_nn);
    ctx->set_materialize_grads(false);
    return {dq, dbox};
  }

  static variable_list backward(AutogradContext*, variable_list grads) {
    if (grads[0].defined() || grads[1].defined())
      second_order("the self and non-neutral terms without a graph");
    return variable_list(6);
  }
};

struct PmeReciprocalFn : public torch::autograd::Function<PmeReciprocalFn> {
  static variable_list forward(AutogradContext* ctx, const Tensor& positions, const Tensor& box,
                               const Tensor& q, at::IntArrayRef kmesh, double kappa, double eps,
                               int64_t precision, int64_t convention, bool need_potential,
                               const optional<Tensor>& box_host) {
    at::AutoDispatchBelowADInplaceOrView below;
    auto [e, phi, dA] = op_pme_reciprocal().call(positions, box, q, kmesh, kappa, eps, precision,
                                                 convention, need_potential, box_host);
    // the potential mesh and dE/dA feed the backward only
    ctx->mark_non_differentiable({phi, dA});
    ctx->save_for_backward({positions, box, q, phi, dA, e, undef_or(box_host)});
    ctx->saved_data["kmesh"] = std::vector<int64_t>(kmesh.begin(), kmesh.end());
    ctx->saved_data["kappa"] = kappa;
    ctx->saved_data["eps"] = eps;
    ctx->saved_data["precision"] = precision;
    ctx->saved_data["convention"] = convention;
    ctx->saved_data["need_potential"] = need_potential;
    ctx->set_materialize_grads(false);
    return {e, phi, dA};
  }

  static variable_list backward(AutogradContext* ctx, variable_list grads) {
    variable_list out(10);
    const Tensor& g_e = grads[0];
    if (!g_e.defined()) return out;
    TORCH_CHECK(ctx->saved_data["need_potential"].toBool(),
                "pme_reciprocal was called with need_potential=False; its energy cannot be "
                "differentiated");
    const variable_list s = ctx->get_saved_variables();
    const std::vector<int64_t> kmesh = ctx->saved_data["kmesh"].toIntVector();
    auto [gpos, gq, gbox] = op_pme_reciprocal_grads().call(
        s[0], s[1], s[2], s[3], s[4], s[5].detach(), kmesh, ctx->saved_data["kappa"].toDouble(),
        ctx->saved_data["eps"].toDouble(), ctx->saved_data["precision"].toInt(),
        ctx->saved_data["convention"].toInt(), opt(s[6]));
    if (ctx->needs_input_grad(0)) out[0] = g_e * gpos;
    if (ctx->needs_input_grad(1)) out[1] = g_e * gbox;
    if (ctx->needs_input_grad(2)) out[2] = g_e * gq;
    return out;
  }
};

struct PmeReciprocalGradsFn : public torch::autograd::Function<PmeReciprocalGradsFn> {
  static variable_list forward(AutogradContext* ctx, const Tensor& positions, const Tensor& box,
                               const Tensor& q, const Tensor& potential, const Tensor& dEdA_k,
                               const Tensor& energy, at::IntArrayRef kmesh, double kappa,
                               double eps, int64_t precision, int64_t convention,
                               const optional<Tensor>& box_host) {
    at::AutoDispatchBelowADInplaceOrView below;
    auto [gpos, gq, gbox] =
        op_pme_reciprocal_grads().call(positions, box, q, potential, dEdA_k, energy, kmesh,
                                       kappa, eps, precision, convention, box_host);
    ctx->save_for_backward({positions, box, q, undef_or(box_host)});
    ctx->saved_data["kmesh"] = std::vector<int64_t>(kmesh.begin(), kmesh.end());
    ctx->saved_data["kappa"] = kappa;
    ctx->saved_data["eps"] = eps;
    ctx->saved_data["precision"] = precision;
    ctx->saved_data["convention"] = convention;
    ctx->set_materialize_grads(false);
    return {gpos, gq, gbox};
  }

  static variable_list backward(AutogradContext* ctx, variable_list grads) {
    if (grads[0].defined() || grads[2].defined()) second_order("the reciprocal-space forces");
    variable_list out(12);
    if (!grads[1].defined()) return out;
    const variable_list s = ctx->get_saved_variables();
    const Tensor& positions = s[0];
    const Tensor& box = s[1];
    const Tensor& q = s[2];
    const optional<Tensor> bh = opt(s[3]);
    const std::vector<int64_t> kmesh = ctx->saved_data["kmesh"].toIntVector();
    const double kappa = ctx->saved_data["kappa"].toDouble();
    const double eps = ctx->saved_data["eps"].toDouble();
    const int64_t precision = ctx->saved_data["precision"].toInt();
    const int64_t convention = ctx->saved_data["convention"].toInt();
    auto eval_grads = [&](const Tensor& charges) {
      auto [e, phi, dA] = op_pme_reciprocal().call(positions, box, charges, kmesh, kappa, eps,
                                                   precision, convention, true, bh);
      return op_pme_reciprocal_grads().call(positions, box, charges, phi, dA, e, kmesh, kappa,
                                            eps, precision, convention, bh);
    };
    const Tensor v = grads[1].to(q.scalar_type()).reshape(q.sizes());
    if (ctx->needs_input_grad(2)) {
      // E is quadratic in q: v^T (dE/dq) differentiates to the potential of charges v
      out[2] = std::get<1>(eval_grads(v));
    }
    const bool need_pos = ctx->needs_input_grad(0), need_box = ctx->needs_input_grad(1);
    if (need_pos || need_box) {
      auto gp = eval_grads(q + v);
      auto gm = eval_grads(q - v);
      if (need_pos) out[0] = 0.5 * (std::get<0>(gp) - std::get<0>(gm));
      if (need_box) out[1] = 0.5 * (std::get<2>(gp) - std::get<2>(gm));
    }
    return out;
  }
};

// ------------------------------------------------------------------ Autograd-key kernels
Tensor pair_distance_ad(const Tensor& positions, const optional<Tensor>& box, const Tensor& pairs,
                        const optional<Tensor>& box_host, const optional<Tensor>& row_ptr,
                        const optional<Tensor>& reverse) {
  return PairDistanceFn::apply(positions, box, pairs, box_host, row_ptr, reverse);
}

std::tuple<Tensor, Tensor> pair_distance_backward_ad(
    const Tensor& grad_ds, const Tensor& positions, const optional<Tensor>& box,
    const Tensor& pairs, const Tensor& ds, const optional<Tensor>& box_host,
    const optional<Tensor>& row_ptr, const optional<Tensor>& reverse, bool want_box) {
  auto r = PairDistanceBackwardFn::apply(grad_ds, positions, box, pairs, ds, box_host, row_ptr,
                                         reverse, want_box);
  return {r[0], r[1]};
}

Tensor pair_energy_ad(const Tensor& q, const Tensor& ds, const Tensor& pairs, const Tensor& bs,
                      const optional<Tensor>& eta, int64_t mode, double kappa, double eps) {
  return PairEnergyFn::apply(q, ds, pairs, bs, eta, mode, kappa, eps,
                             at::GradMode::is_enabled());
}

std::tuple<Tensor, Tensor, Tensor> pair_energy_grads_ad(const Tensor& q, const Tensor& ds,
                                                        const Tensor& pairs, const Tensor& bs,
                                                        const optional<Tensor>& eta,
                                                        int64_t mode, double kappa, double eps,
                                                        bool want_q, bool want_ds,
                                                        bool want_eta) {
  auto r = PairEnergyGradsFn::apply(q, ds, pairs, bs, eta, mode, kappa, eps, want_q, want_ds,
                                    want_eta);
  return {r[0], r[1], r[2]};
}

std::tuple<Tensor, Tensor> ewald_self_terms_ad(const Tensor& q, const Tensor& box, double kappa,
                                               double eps) {
  auto r = EwaldSelfTermsFn::apply(q, box, kappa, eps);
  return {r[0], r[1]};
}

std::tuple<Tensor, Tensor> ewald_self_terms_backward_ad(const Tensor& q, const Tensor& box,
                                                        double kappa, double eps,
                                                        const optional<Tensor>& g_self,
                                                        const optional<Tensor>& g_nn) {
  auto r = EwaldSelfTermsBackwardFn::apply(q, box, kappa, eps, g_self, g_nn);
  return {r[0], r[1]};
}

std::tuple<Tensor, Tensor, Tensor> pme_reciprocal_ad(const Tensor& positions, const Tensor& box,
                                                     const Tensor& q, at::IntArrayRef kmesh,
                                                     double kappa, double eps, int64_t precision,
                                                     int64_t convention, bool need_potential,
                                                     const optional<Tensor>& box_host) {
  auto r = PmeReciprocalFn::apply(positions, box, q, kmesh, kappa, eps, precision, convention,
                                  need_potential, box_host);
  return {r[0], r[1], r[2]};
}

std::tuple<Tensor, Tensor, Tensor> pme_reciprocal_grads_ad(
    const Tensor& positions, const Tensor& box, const Tensor& q, const Tensor& potential,
    const Tensor& dEdA_k, const Tensor& energy, at::IntArrayRef kmesh, double kappa, double eps,
    int64_t precision, int64_t convention, const optional<Tensor>& box_host) {
  auto r = PmeReciprocalGradsFn::apply(positions, box, q, potential, dEdA_k, energy, kmesh, kappa,
                                       eps, precision, convention, box_host);
  return {r[0], r[1], r[2]};
}

}  // namespace

// Bind the operators to the HIP library at `path` (the one the caller has
// loaded); returns 0, or -1 with the reason in the thrown-away error.  Without
// this call the first operator binds to the default path.
extern "C" int admp_torch_ops_bind(const char* path) {
  try {
    bind_path(path);
    return 0;
  } catch (...) {
    return -1;
  }
}

TORCH_LIBRARY(admp, m) {
  m.def("neighbor_pairs(Tensor positions, Tensor? box, float cutoff, Tensor? box_host) -> Tensor");
  m.def("neighbor_list(Tensor positions, Tensor? box, float cutoff, Tensor? box_host) -> "
        "(Tensor, Tensor, Tensor, Tensor)");
  m.def("pair_scales(Tensor pairs, Tensor? ds, float lo, float hi) -> Tensor");
  m.def("max_displacement(Tensor positions, Tensor reference, Tensor? box, Tensor? box_host) -> "
        "Tensor");
  m.def("pair_distance(Tensor positions, Tensor? box, Tensor pairs, Tensor? box_host, "
        "Tensor? row_ptr=None, Tensor? reverse=None) -> Tensor");
  m.def("pair_distance_backward(Tensor grad_ds, Tensor positions, Tensor? box, Tensor pairs, "
        "Tensor ds, Tensor? box_host, Tensor? row_ptr=None, Tensor? reverse=None, "
        "bool want_box=True) -> (Tensor, Tensor)");
  m.def("pair_energy(Tensor q, Tensor ds, Tensor pairs, Tensor bs, Tensor? eta, int mode, "
        "float kappa, float eps) -> Tensor");
  m.def("pair_energy_grads(Tensor q, Tensor ds, Tensor pairs, Tensor bs, Tensor? eta, int mode, "
        "float kappa, float eps, bool want_q, bool want_ds, bool want_eta) -> "
        "(Tensor, Tensor, Tensor)");
  m.def("ewald_self_terms(Tensor q, Tensor box, float kappa, float eps) -> (Tensor, Tensor)");
  m.def("ewald_self_terms_backward(Tensor q, Tensor box, float kappa, float eps, "
        "Tensor? g_self, Tensor? g_nn) -> (Tensor, Tensor)");
  m.def("pme_reciprocal(Tensor positions, Tensor box, Tensor q, int[] kmesh, float kappa, "
        "float eps, int precision, int convention, bool need_potential, Tensor? box_host) -> "
        "(Tensor, Tensor, Tensor)");
  m.def("pme_reciprocal_grads(Tensor positions, Tensor box, Tensor q, Tensor potential, "
        "Tensor dEdA_k, Tensor energy, int[] kmesh, float kappa, float eps, int precision, "
        "int convention, Tensor? box_host) -> (Tensor, Tensor, Tensor)");
}

// The same bodies serve CPU tensors, which they reject: there is no CPU path.
#define ADMP_IMPLS(m)                                                 \
  m.impl("neighbor_pairs", &neighbor_pairs);                          \
  m.impl("neighbor_list", &neighbor_list);                            \
  m.impl("pair_scales", &pair_scales);                                \
  m.impl("max_displacement", &max_displacement);                      \
  m.impl("pair_distance", &pair_distance);                            \
  m.impl("pair_distance_backward", &pair_distance_backward);          \
  m.impl("pair_energy", &pair_energy);                                \
  m.impl("pair_energy_grads", &pair_energy_grads);                    \
  m.impl("ewald_self_terms", &ewald_self_terms);                      \
  m.impl("ewald_self_terms_backward", &ewald_self_terms_backward);    \
  m.impl("pme_reciprocal", &pme_reciprocal);                          \
  m.impl("pme_reciprocal_grads", &pme_reciprocal_grads)

TORCH_LIBRARY_IMPL(admp, CUDA, m) { ADMP_IMPLS(m); }
TORCH_LIBRARY_IMPL(admp, CPU, m) { ADMP_IMPLS(m); }
#undef ADMP_IMPLS

TORCH_LIBRARY_IMPL(admp, Autograd, m) {
  // integer outputs and the displacement check: nothing to differentiate
  m.impl("neighbor_pairs", torch::CppFunction::makeFallthrough());
  m.impl("neighbor_list", torch::CppFunction::makeFallthrough());
  m.impl("pair_scales", torch::CppFunction::makeFallthrough());
  m.impl("max_displacement", torch::CppFunction::makeFallthrough());
  m.impl("pair_distance", &pair_distance_ad);
  m.impl("pair_distance_backward", &pair_distance_backward_ad);
  m.impl("pair_energy", &pair_energy_ad);
  m.impl("pair_energy_grads", &pair_energy_grads_ad);
  m.impl("ewald_self_terms", &ewald_self_terms_ad);
  m.impl("ewald_self_terms_backward", &ewald_self_terms_backward_ad);
  m.impl("pme_reciprocal", &pme_reciprocal_ad);
  m.impl("pme_reciprocal_grads", &pme_reciprocal_grads_ad);
}
