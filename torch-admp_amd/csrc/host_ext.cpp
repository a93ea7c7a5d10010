// Python binding of the fused single-system call (admp_pme_fused) without
// ctypes: the argument checks, the workspace-size cache, one workspace per
// (device, stream), the output allocation and the C-ABI call in one native
// call.  The C ABI stays the boundary: admp_pme_fused and its helpers are
// looked up in the HIP library the package has already loaded.
#include <dlfcn.h>
#include <torch/extension.h>

#include <array>
#include <cstdint>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

namespace {

using fused_fn = int (*)(const double*, const double*, int64_t, const double*, double, double,
                         const int32_t*, int32_t, int32_t, double, double*, double*, double*,
                         void*, size_t, void*);
using bytes_fn = int (*)(int64_t, const double*, double, const int32_t*, int32_t, size_t*);
using status_fn = const char* (*)(int);

fused_fn g_fused = nullptr;
bytes_fn g_bytes = nullptr;
status_fn g_status = nullptr;

struct ShapeKey {
  int64_t n;
  std::array<double, 9> box;
  double rc;
  std::array<int32_t, 3> km;
  int32_t prec;
  bool operator<(const ShapeKey& o) const {
    return std::tie(n, box, rc, km, prec) < std::tie(o.n, o.box, o.rc, o.km, o.prec);
  }
};

// Heap-held and never destroyed: workspaces must not be freed after the
// device runtime has shut down at process exit.
std::map<ShapeKey, size_t>* g_sizes = new std::map<ShapeKey, size_t>();
std::map<std::pair<int64_t, int64_t>, at::Tensor>* g_ws =
    new std::map<std::pair<int64_t, int64_t>, at::Tensor>();
std::mutex g_mu;
constexpr size_t kMaxEntries = 256;

std::string status_text(int st) {
  return g_status ? std::string(g_status(st)) : std::to_string(st);
}

// Resolve the entry points in the HIP library at `path` (already loaded).
void bind(const std::string& path) {
  void* h = dlopen(path.c_str(), RTLD_NOW | RTLD_NOLOAD);
  if (!h) h = dlopen(path.c_str(), RTLD_NOW);
  if (!h) throw std::runtime_error("torch_admp_amd: cannot open " + path);
  g_fused = reinterpret_cast<fused_fn>(dlsym(h, "admp_pme_fused"));
  g_bytes = reinterpret_cast<bytes_fn>(dlsym(h, "admp_pme_fused_workspace_bytes"));
  g_status = reinterpret_cast<status_fn>(dlsym(h, "admp_status_string"));
  if (!g_fused || !g_bytes || !g_status)
    throw std::runtime_error("torch_admp_amd: " + path + " lacks the admp_pme_fused entry points");
}

// (energy[5], forces (n, 3), dE/dbox (3, 3)) of ops.pme_fused.  pos (n, 3) and
// q (n) float64 contiguous on the GPU; box a float64 contiguous host tensor of
// 9 values; stream the raw hipStream_t of the current stream.
std::vector<at::Tensor> pme_fused(const at::Tensor& pos, const at::Tensor& q, const at::Tensor& box,
                                  double rc, double kappa, int64_t k0, int64_t k1, int64_t k2,
                                  int64_t prec, int64_t conv, double dielectric, int64_t stream) {
  if (!g_fused) throw std::runtime_error("torch_admp_amd: host extension not bound");
  TORCH_CHECK(pos.is_cuda() && q.is_cuda(),
              "torch_admp_amd runs only on a ROCm GPU (no CPU fallback)");
  TORCH_CHECK(pos.scalar_type() == at::kDouble && pos.is_contiguous() && pos.dim() == 2 &&
                  pos.size(1) == 3,
              "positions: contiguous float64 (n, 3)");
  const int64_t n = pos.size(0);
  TORCH_CHECK(q.scalar_type() == at::kDouble && q.is_contiguous() && q.numel() == n,
              "charges: contiguous float64 (n)");
  TORCH_CHECK(!box.is_cuda() && box.scalar_type() == at::kDouble && box.is_contiguous() &&
                  box.numel() == 9,
              "box: 9 float64 values on the host");
  ShapeKey key;
  key.n = n;
  const double* bp = box.data_ptr<double>();
  for (int i = 0; i < 9; ++i) key.box[i] = bp[i];
  key.rc = rc;
  key.km = {(int32_t)k0, (int32_t)k1, (int32_t)k2};
  key.prec = (int32_t)prec;
  size_t nbytes = 0;
  at::Tensor ws;
  {
    std::lock_guard<std::mutex> g(g_mu);
    auto it = g_sizes->find(key);
    if (it == g_sizes->end()) {
      const int st = g_bytes(n, key.box.data(), rc, key.km.data(), key.prec, &nbytes);
      if (st != 0)
        throw std::runtime_error("admp_pme_fused_workspace_bytes failed: " + status_text(st));
      if (g_sizes->size() >= kMaxEntries) g_sizes->clear();
      g_sizes->emplace(key, nbytes);
    } else {
      nbytes = it->second;
    }
    // the caching allocator keeps a block made on this stream for this stream
    at::Tensor& w = (*g_ws)[{(int64_t)pos.get_device(), stream}];
    if (!w.defined() || (size_t)w.numel() < nbytes)
      w = at::empty({(int64_t)std::max<size_t>(nbytes, 1)}, pos.options().dtype(at::kByte));
    ws = w;
  }
  at::Tensor out = at::empty({14 + 3 * n}, pos.options());
  double* o = out.data_ptr<double>();
  const int st = g_fused(pos.data_ptr<double>(), q.data_ptr<double>(), n, key.box.data(), rc,
                         kappa, key.km.data(), key.prec, (int32_t)conv, dielectric, o, o + 14,
                         o + 5, ws.data_ptr(), nbytes, reinterpret_cast<void*>(stream));
  if (st != 0) throw std::runtime_error("admp_pme_fused failed: " + status_text(st));
  return {out.narrow(0, 0, 5), out.narrow(0, 14, 3 * n).view({n, 3}),
          out.narrow(0, 5, 9).view({3, 3})};
}

}  // namespace

PYBIND11_MODULE(_host_ext, m) {
  m.def("bind", &bind, "resolve the C-ABI entry points in the loaded HIP library");
  m.def("pme_fused", &pme_fused, "the fused single-system call");
}
