"""Python entry points over the HIP library.

* pair_distance, pair_energy, pme_reciprocal -- the differentiable pieces of
  the drop-in modules, as calls of the torch.ops.admp operators (_ops.py),
  whose backwards are the library's adjoint kernels;
* pme_fused -- energy, forces and dE/dbox of one frame in one native call;
* PMEBatchPlan / pme_batch -- many independent systems in one launch sequence;
* PMEGraph -- a batch captured in a HIP graph.
"""

from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import torch

from . import _lib
from ._ops import PAIR_BARE, PAIR_DAMPED, PAIR_EWALD  # noqa: F401
from .utils import DIELECTRIC


def _f64(t: torch.Tensor) -> torch.Tensor:
    return t.detach().to(torch.float64).contiguous()


def pair_distance(positions: torch.Tensor, box: Optional[torch.Tensor],
                  pairs: torch.Tensor) -> torch.Tensor:
    """ds = |pbc_shift(r_j - r_i)| (src/torch_admp/nblist.py:517-551); open
    boundaries when box is None.  Differentiable in positions and box."""
    return torch.ops.admp.pair_distance(positions, box, pairs, None)


def pair_energy(q: torch.Tensor, ds: torch.Tensor, pairs: torch.Tensor, bs: torch.Tensor,
                eta: Optional[torch.Tensor], mode: int, kappa: float, eps: float) -> torch.Tensor:
    """Per-frame pair energy over (nf, P) pair rows: mode 0
    (src/torch_admp/pme.py:312-349), 1 (pme.py:529-561), 2
    (src/torch_admp/qeq.py:106-121).  Differentiable in q, ds and eta, twice in q."""
    return torch.ops.admp.pair_energy(q, ds, pairs, bs, eta, int(mode), float(kappa),
                                      float(eps))


def pme_reciprocal(positions: torch.Tensor, box: torch.Tensor, q: torch.Tensor,
                   kmesh: Sequence[int], kappa: float, eps: float, precision: int,
                   convention: int) -> torch.Tensor:
    """One frame of the SPME reciprocal energy (src/torch_admp/pme.py:351-446).
    Differentiable in positions, box and charges, twice in the charges."""
    need = torch.is_grad_enabled() and (positions.requires_grad or box.requires_grad
                                        or q.requires_grad)
    return torch.ops.admp.pme_reciprocal(positions, box, q, [int(k) for k in kmesh],
                                         float(kappa), float(eps), int(precision),
                                         int(convention), need, None)[0]


# Host state reused across fused calls: per problem shape the workspace size and
# the ctypes box and mesh arrays, and one workspace per (device, stream).  Calls
# on one stream run in order, so reusing that stream's workspace is safe; the
# caching allocator keeps a workspace made on one stream for that stream.
_FUSED_SHAPES: dict = {}
_FUSED_WS: dict = {}
_FUSED_MAX_ENTRIES = 256


_EXT = None
_EXT_TRIED = False


def _host_ext():
    """The native binding of the fused call (host_ext.cpp), bound to the HIP
    library in use, or None when it has not been built."""
    global _EXT, _EXT_TRIED
    if not _EXT_TRIED:
        _EXT_TRIED = True
        try:
            from . import _host_ext as ext
        except ImportError:
            ext = None
        if ext is not None:
            _lib.load()
            ext.bind(str(_lib.library_path()))
        _EXT = ext
    return _EXT


def _as_f64(t: torch.Tensor) -> torch.Tensor:
    if t.dtype == torch.float64 and t.is_contiguous():
        return t
    return _f64(t)


def _fused_shape(lib, n: int, box: torch.Tensor, rcut: float, kmesh, prec: int):
    """(workspace bytes, ctypes box, ctypes mesh) of one problem shape, cached."""
    if box.device.type == "cpu" and box.dtype == torch.float64:
        boxv = tuple(box.detach().reshape(9).tolist())
    else:
        boxv = tuple(box.detach().reshape(9).to("cpu", torch.float64).tolist())
    km = tuple(int(k) for k in kmesh)
    key = (n, boxv, float(rcut), km, prec)
    hit = _FUSED_SHAPES.get(key)
    if hit is None:
        boxh = (ctypes.c_double * 9)(*boxv)
        kmh = (ctypes.c_int32 * 3)(*km)
        nbytes = ctypes.c_size_t()
        _lib.check(lib.admp_pme_fused_workspace_bytes(n, boxh, rcut, kmh, prec,
                                                      ctypes.byref(nbytes)),
                   "admp_pme_fused_workspace_bytes")
        if len(_FUSED_SHAPES) >= _FUSED_MAX_ENTRIES:
            _FUSED_SHAPES.clear()
        hit = _FUSED_SHAPES[key] = (nbytes.value, boxh, kmh)
    return hit


def pme_fused(positions: torch.Tensor, box: torch.Tensor, charges: torch.Tensor, rcut: float,
              kappa: float, kmesh: Sequence[int], precision: str = "float64",
              convention: str = "reference", dielectric: float = DIELECTRIC,
              workspace: Optional[torch.Tensor] = None):
    """Energy components, forces and dE/dbox of one periodic frame in one call.

    Returns (energy[5] = real, reciprocal, self, non-neutral, total; forces
    (N, 3); dE/dbox (3, 3)), all float64 on the device.  Matches
    CoulombForceModule on the TorchNeighborList pair list followed by
    calc_grads (src/torch_admp/pme.py:261-310).  Without `workspace` the call
    goes through the native binding (host_ext.cpp) when it is built, and reuses
    a workspace kept for the current stream in either case.
    """
    _lib.require_gpu(positions, charges)
    pos = _as_f64(positions)
    q = _as_f64(charges)
    if q.dim() != 1:
        q = q.reshape(-1)
    ext = _host_ext()
    if ext is not None and workspace is None:
        if box.device.type != "cpu" or box.dtype != torch.float64 or not box.is_contiguous():
            box = box.detach().to("cpu", torch.float64).contiguous()
        k0, k1, k2 = (int(k) for k in kmesh)
        e, f, g = ext.pme_fused(pos, q, box, float(rcut), float(kappa), k0, k1, k2,
                                _lib.PRECISION[precision], _lib.KCONV[convention],
                                float(dielectric), torch._C._cuda_getCurrentRawStream(pos.device.index))
        return e, f, g
    lib = _lib.load()
    n = pos.shape[0]
    dev = pos.device
    prec = _lib.PRECISION[precision]
    nbytes, boxh, kmh = _fused_shape(lib, n, box, rcut, kmesh, prec)
    stream = torch._C._cuda_getCurrentRawStream(dev.index)
    if workspace is None or workspace.numel() < nbytes:
        wkey = (dev.index, stream)
        ws = _FUSED_WS.get(wkey)
        if ws is None or ws.numel() < nbytes:
            if ws is None and len(_FUSED_WS) >= _FUSED_MAX_ENTRIES:
                _FUSED_WS.clear()
            ws = _FUSED_WS[wkey] = _lib.workspace(nbytes, dev)
        workspace = ws
    out = torch.empty(14 + 3 * n, dtype=torch.float64, device=dev)
    energy, dedbox, forces = out[:5], out[5:14].view(3, 3), out[14:].view(n, 3)
    _lib.check(lib.admp_pme_fused(pos.data_ptr(), q.data_ptr(), n, boxh, rcut, kappa, kmh, prec,
                                  _lib.KCONV[convention], dielectric, energy.data_ptr(),
                                  forces.data_ptr(), dedbox.data_ptr(), workspace.data_ptr(),
                                  nbytes, stream), "admp_pme_fused")
    return energy, forces, dedbox


def _host_array(values, dtype, shape):
    import numpy as np
    arr = np.ascontiguousarray(np.asarray(values, dtype=dtype).reshape(shape))
    return arr, ctypes.c_void_p(arr.ctypes.data)


def _per_system(value, n_sys: int, what: str):
    if isinstance(value, (int, float)):
        return [float(value)] * n_sys
    vals = [float(v) for v in value]
    if len(vals) != n_sys:
        raise ValueError(f"{what}: expected {n_sys} values, got {len(vals)}")
    return vals


class PMEBatchPlan:
    """Host description of a batch of periodic systems for admp_pme_batch.

    n_atoms, boxes ((n_sys, 3, 3) host values), cutoffs, kappas and kmesh
    ((n_sys, 3)) fix the workspace size; the same plan serves any positions
    and charges of that shape.
    """

    def __init__(self, n_atoms, boxes, cutoffs, kappas, kmesh, precision: str = "float64",
                 convention: str = "reference", dielectric: float = DIELECTRIC):
        import numpy as np
        self.n_atoms = [int(n) for n in n_atoms]
        self.n_sys = len(self.n_atoms)
        if self.n_sys < 1:
            raise ValueError("a batch needs at least one system")
        boxes = np.asarray([b.detach().cpu().numpy() if isinstance(b, torch.Tensor) else b
                            for b in boxes], dtype=np.float64).reshape(self.n_sys, 9)
        self._n, self._pn = _host_array(self.n_atoms, np.int64, (self.n_sys,))
        self._box, self._pbox = _host_array(boxes, np.float64, (self.n_sys, 9))
        self._rc, self._prc = _host_array(_per_system(cutoffs, self.n_sys, "cutoffs"),
                                          np.float64, (self.n_sys,))
        self._kap, self._pkap = _host_array(_per_system(kappas, self.n_sys, "kappas"),
                                            np.float64, (self.n_sys,))
        km = np.asarray(kmesh, dtype=np.int32)
        if km.ndim == 1:
            km = np.tile(km, (self.n_sys, 1))
        self._km, self._pkm = _host_array(km, np.int32, (self.n_sys, 3))
        self.precision = precision
        self.convention = convention
        self.dielectric = float(dielectric)
        self.offsets = np.concatenate([[0], np.cumsum(self.n_atoms)]).astype(np.int64)
        self.n_total = int(self.offsets[-1])
        lib = _lib.load()
        nbytes = ctypes.c_size_t()
        _lib.check(lib.admp_pme_batch_workspace_bytes(self.n_sys, self._pn, self._pbox, self._prc,
                                                      self._pkm, _lib.PRECISION[precision],
                                                      ctypes.byref(nbytes)),
                   "admp_pme_batch_workspace_bytes")
        self.workspace_bytes = int(nbytes.value)

    def run(self, pos: torch.Tensor, q: torch.Tensor, workspace: Optional[torch.Tensor] = None,
            want_box_grad: bool = True, overlap: bool = True):
        """Energies (n_sys, 5), forces (n_total, 3) and dE/dbox (n_sys, 3, 3)
        for positions (n_total, 3) and charges (n_total,) stored back to back.
        overlap=False keeps every kernel on the current stream (no real-space
        side stream): for callers with batches in flight on several streams."""
        _lib.require_gpu(pos, q)
        pos = _f64(pos)
        q = _f64(q).reshape(-1)
        if pos.shape != (self.n_total, 3) or q.shape[0] != self.n_total:
            raise ValueError(f"expected ({self.n_total}, 3) positions and {self.n_total} charges")
        dev = pos.device
        if workspace is None or workspace.numel() < self.workspace_bytes:
            workspace = _lib.workspace(self.workspace_bytes, dev)
        energy = torch.empty(self.n_sys, 5, dtype=torch.float64, device=dev)
        forces = torch.empty(self.n_total, 3, dtype=torch.float64, device=dev)
        dedbox = (torch.empty(self.n_sys, 3, 3, dtype=torch.float64, device=dev)
                  if want_box_grad else None)
        lib = _lib.load()
        _lib.check(lib.admp_pme_batch_ex(self.n_sys, self._pn, self._pbox, self._prc, self._pkap,
                                         self._pkm, _lib.PRECISION[self.precision],
                                         _lib.KCONV[self.convention], self.dielectric,
                                         _lib.ptr(pos), _lib.ptr(q), _lib.ptr(energy),
                                         _lib.ptr(forces), _lib.ptr(dedbox),
                                         0 if overlap else _lib.ADMP_BATCH_NO_OVERLAP,
                                         _lib.ptr(workspace), self.workspace_bytes,
                                         _lib.stream_of(pos)),
                   "admp_pme_batch_ex")
        return energy, forces, dedbox


def pme_batch(positions: Sequence[torch.Tensor], boxes, charges: Sequence[torch.Tensor], rcut,
              kappa, kmesh, precision: str = "float64", convention: str = "reference",
              dielectric: float = DIELECTRIC):
    """pme_fused over a list of independent periodic systems in one launch sequence.

    rcut and kappa are scalars or one value per system; kmesh is one (3,) mesh
    or one per system.  Returns (energy (n_sys, 5), list of (n_i, 3) forces,
    dE/dbox (n_sys, 3, 3)); each system's values equal pme_fused on it alone.
    """
    _lib.require_gpu(*positions, *charges)
    plan = PMEBatchPlan([p.shape[0] for p in positions], boxes, rcut, kappa, kmesh, precision,
                        convention, dielectric)
    pos = torch.cat([_f64(p).reshape(-1, 3) for p in positions])
    q = torch.cat([_f64(c).reshape(-1) for c in charges])
    energy, forces, dedbox = plan.run(pos, q)
    return energy, list(torch.split(forces, plan.n_atoms)), dedbox


class PMEGraph:
    """A batch captured in a HIP graph for repeated evaluation.

    Positions and charges live in the static device buffers ``positions``
    (n_total, 3) and ``charges`` (n_total,); ``replay()`` relaunches the whole
    kernel sequence (one graph launch instead of ~14 kernel launches) and the
    results land in ``energy`` (n_sys, 5), ``forces`` and ``dedbox``.  The
    systems' boxes, cutoffs and meshes are fixed at capture.  The system table
    is uploaded once before capture (admp_pme_batch_upload); the captured
    sequence (admp_pme_batch_run) contains kernels only.
    """

    def __init__(self, plan: PMEBatchPlan, device="cuda"):
        dev = torch.device(device)
        _lib.require_gpu(torch.empty(0, device=dev))
        self.plan = plan
        f64 = torch.float64
        self.positions = torch.zeros(plan.n_total, 3, dtype=f64, device=dev)
        self.charges = torch.zeros(plan.n_total, dtype=f64, device=dev)
        self.energy = torch.zeros(plan.n_sys, 5, dtype=f64, device=dev)
        self.forces = torch.zeros(plan.n_total, 3, dtype=f64, device=dev)
        self.dedbox = torch.zeros(plan.n_sys, 3, 3, dtype=f64, device=dev)
        self.workspace = _lib.workspace(plan.workspace_bytes, dev)
        lib = _lib.load()
        side = torch.cuda.Stream(dev)
        side.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(side):
            _lib.check(lib.admp_pme_batch_upload(
                plan.n_sys, plan._pn, plan._pbox, plan._prc, plan._pkap, plan._pkm,
                _lib.PRECISION[plan.precision], _lib.KCONV[plan.convention], plan.dielectric,
                _lib.ptr(self.workspace), plan.workspace_bytes, _lib.stream_of(self.workspace)),
                "admp_pme_batch_upload")
            self._launch(lib)  # eager run: loads the kernels before capture
        torch.cuda.current_stream(dev).wait_stream(side)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._launch(lib)

    def _launch(self, lib):
        p = self.plan
        _lib.check(lib.admp_pme_batch_run(
            p.n_sys, p._pn, p._pbox, p._prc, p._pkap, p._pkm, _lib.PRECISION[p.precision],
            _lib.KCONV[p.convention], p.dielectric, _lib.ptr(self.positions),
            _lib.ptr(self.charges), _lib.ptr(self.energy), _lib.ptr(self.forces),
            _lib.ptr(self.dedbox), _lib.ptr(self.workspace), p.workspace_bytes,
            _lib.stream_of(self.positions)), "admp_pme_batch_run")

    def replay(self):
        self.graph.replay()
        return self.energy, self.forces, self.dedbox

    def run(self, positions: torch.Tensor, charges: torch.Tensor):
        """Copy inputs into the static buffers and replay."""
        self.positions.copy_(positions.detach().reshape(-1, 3))
        self.charges.copy_(charges.detach().reshape(-1))
        return self.replay()
