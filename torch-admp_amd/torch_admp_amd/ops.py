"""torch.autograd.Function wrappers around the HIP kernels.

Each forward launches the library kernels on torch's current stream; each
backward is a hand-written adjoint kernel (no autograd through torch ops):

* PairDistance     -- pair distances and their transpose (scatter of dL/dds)
* PairEnergy       -- erfc / bare / Gaussian-damped pair sums, dE/dds and dE/dq
* PMEReciprocal    -- spread -> r2c FFT -> influence function -> c2r FFT; the
                      backward is the gather (the transpose of the spread) plus
                      the k-vector and volume terms of dE/dbox.
"""

from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import torch

from . import _lib
from .utils import DIELECTRIC


def _f64(t: torch.Tensor) -> torch.Tensor:
    return t.detach().to(torch.float64).contiguous()


class PairDistance(torch.autograd.Function):
    """ds = |pbc_shift(r_i - r_j)| (src/torch_admp/nblist.py:517-551); open
    boundaries when box is None."""

    @staticmethod
    def forward(ctx, positions: torch.Tensor, box: Optional[torch.Tensor], pairs: torch.Tensor):
        _lib.require_gpu(positions, box, pairs)
        lib = _lib.load()
        pos = _f64(positions)
        prs = pairs.detach().to(torch.int64).contiguous()
        npairs = prs.shape[0]
        ds = torch.empty(npairs, dtype=torch.float64, device=pos.device)
        dr = torch.empty(npairs, 3, dtype=torch.float64, device=pos.device)
        boxh = _lib.box_array(box)
        _lib.check(lib.admp_pair_ds(_lib.ptr(pos), _lib.ptr(prs), npairs, boxh, _lib.ptr(ds),
                                    _lib.ptr(dr), _lib.stream_of(pos)), "admp_pair_ds")
        ctx.save_for_backward(pos, prs, dr, ds)
        ctx.boxh = boxh
        ctx.dtype = positions.dtype
        ctx.box_dtype = None if box is None else box.dtype
        return ds.to(positions.dtype)

    @staticmethod
    def backward(ctx, grad_ds):
        pos, prs, dr, ds = ctx.saved_tensors
        lib = _lib.load()
        g = _f64(grad_ds)
        gpos = torch.zeros_like(pos)
        need_box = ctx.boxh is not None and ctx.needs_input_grad[1]
        gbox = torch.zeros(3, 3, dtype=torch.float64, device=pos.device) if need_box else None
        _lib.check(lib.admp_pair_ds_backward(_lib.ptr(g), _lib.ptr(prs), prs.shape[0],
                                             _lib.ptr(pos), ctx.boxh, _lib.ptr(dr), _lib.ptr(ds),
                                             _lib.ptr(gpos), _lib.ptr(gbox),
                                             _lib.stream_of(pos)), "admp_pair_ds_backward")
        return (gpos.to(ctx.dtype), None if gbox is None else gbox.to(ctx.box_dtype), None)


PAIR_EWALD, PAIR_BARE, PAIR_DAMPED = 0, 1, 2


class PairEnergy(torch.autograd.Function):
    """Per-frame pair energy over (nf, P) pair rows with float64 accumulation.

    mode 0 (src/torch_admp/pme.py:312-349), 1 (pme.py:529-561),
    2 (src/torch_admp/qeq.py:106-121).
    """

    @staticmethod
    def forward(ctx, q, ds, pairs, bs, eta, mode: int, kappa: float, eps: float):
        _lib.require_gpu(q, ds, pairs, bs, eta)
        lib = _lib.load()
        nf, n = q.shape
        npairs = pairs.shape[1]
        q64 = _f64(q)
        ds64 = _f64(ds)
        bs64 = _f64(bs)
        prs = pairs.detach().to(torch.int64).contiguous()
        eta64 = None if eta is None else _f64(eta)
        dev = q64.device
        energy = torch.empty(nf, dtype=torch.float64, device=dev)
        need = ctx.needs_input_grad
        gds = torch.empty(nf, npairs, dtype=torch.float64, device=dev) if need[1] else None
        gq = torch.empty(nf, n, dtype=torch.float64, device=dev) if need[0] else None
        geta = (torch.empty(nf, n, dtype=torch.float64, device=dev)
                if need[4] and mode == PAIR_DAMPED and eta is not None else None)
        nbytes = ctypes.c_size_t()
        _lib.check(lib.admp_pair_energy_workspace_bytes(nf, ctypes.byref(nbytes)), "workspace")
        ws = _lib.workspace(nbytes.value, dev)
        _lib.check(lib.admp_pair_energy(mode, nf, n, npairs, _lib.ptr(q64), _lib.ptr(eta64),
                                        _lib.ptr(prs), _lib.ptr(ds64), _lib.ptr(bs64), kappa, eps,
                                        _lib.ptr(energy), _lib.ptr(gds), _lib.ptr(gq),
                                        _lib.ptr(geta), _lib.ptr(ws), nbytes.value,
                                        _lib.stream_of(q64)),
                   "admp_pair_energy")
        ctx.save_for_backward(gds, gq, geta)
        ctx.dtypes = (q.dtype, ds.dtype, None if eta is None else eta.dtype)
        return energy.to(q.dtype)

    @staticmethod
    def backward(ctx, grad_e):
        gds, gq, geta = ctx.saved_tensors
        ge = grad_e.detach().to(torch.float64).reshape(-1, 1)
        out_q = None if gq is None else (gq * ge).to(ctx.dtypes[0])
        out_ds = None if gds is None else (gds * ge).to(ctx.dtypes[1])
        out_eta = None if geta is None else (geta * ge).to(ctx.dtypes[2])
        return out_q, out_ds, None, None, out_eta, None, None, None


def _inverse_transpose(boxh) -> torch.Tensor:
    b = torch.tensor(list(boxh), dtype=torch.float64).reshape(3, 3)
    return torch.linalg.inv(b).transpose(0, 1)


class PMEReciprocal(torch.autograd.Function):
    """One frame of the SPME reciprocal energy (src/torch_admp/pme.py:378-446).

    Differentiable in positions, box and charges; the backward launches the
    gather kernel on the potential mesh kept from the forward.
    """

    @staticmethod
    def forward(ctx, positions, box, q, kmesh: Sequence[int], kappa: float, eps: float,
                precision: int, convention: int):
        _lib.require_gpu(positions, box, q)
        lib = _lib.load()
        pos = _f64(positions)
        q64 = _f64(q).reshape(-1)
        n = pos.shape[0]
        dev = pos.device
        boxh = _lib.box_array(box)
        km = _lib.kmesh_array(kmesh)
        mesh_dtype = torch.float32 if precision == _lib.ADMP_F32 else torch.float64
        mesh = torch.empty(int(kmesh[0]) * int(kmesh[1]) * int(kmesh[2]), dtype=mesh_dtype,
                           device=dev)
        nbytes = ctypes.c_size_t()
        _lib.check(lib.admp_pme_workspace_bytes(km, precision, n, ctypes.byref(nbytes)),
                   "admp_pme_workspace_bytes")
        ws = _lib.workspace(nbytes.value, dev)
        stream = _lib.stream_of(pos)
        _lib.check(lib.admp_pme_spread(_lib.ptr(pos), _lib.ptr(q64), n, boxh, km, precision,
                                       _lib.ptr(mesh), _lib.ptr(ws), nbytes.value, stream),
                   "admp_pme_spread")
        out = torch.empty(10, dtype=torch.float64, device=dev)
        need_grad = any(ctx.needs_input_grad[:3])
        _lib.check(lib.admp_pme_convolve(_lib.ptr(mesh), boxh, km, precision, convention, kappa,
                                         eps, 1.0, _lib.ptr(mesh) if need_grad else None,
                                         _lib.ptr(out), _lib.ptr(out[1:]), _lib.ptr(ws),
                                         nbytes.value, stream), "admp_pme_convolve")
        energy = out[0].clone()
        if need_grad:
            ctx.save_for_backward(pos, q64, mesh, out)
        ctx.meta = (boxh, km, precision, positions.dtype, None if box is None else box.dtype)
        return energy.to(positions.dtype)

    @staticmethod
    def backward(ctx, grad_e):
        pos, q64, phi, out = ctx.saved_tensors
        boxh, km, precision, pdtype, bdtype = ctx.meta
        lib = _lib.load()
        n = pos.shape[0]
        dev = pos.device
        gpos = torch.empty(n, 3, dtype=torch.float64, device=dev)
        gq = torch.empty(n, dtype=torch.float64, device=dev)
        dEdA = torch.empty(9, dtype=torch.float64, device=dev)
        nbytes = ctypes.c_size_t()
        _lib.check(lib.admp_pme_workspace_bytes(km, precision, n, ctypes.byref(nbytes)),
                   "admp_pme_workspace_bytes")
        ws = _lib.workspace(nbytes.value, dev)
        _lib.check(lib.admp_pme_gather(_lib.ptr(pos), _lib.ptr(q64), n, boxh, km, precision,
                                       _lib.ptr(phi), _lib.ptr(gpos), _lib.ptr(gq),
                                       _lib.ptr(dEdA), 0, _lib.ptr(ws), nbytes.value,
                                       _lib.stream_of(pos)), "admp_pme_gather")
        ge = grad_e.detach().to(torch.float64)
        grad_box = None
        if ctx.needs_input_grad[1]:
            at = _inverse_transpose(boxh).to(dev)                     # A^T
            dA = (out[1:] + dEdA).reshape(3, 3)
            grad_box = (-(at @ dA @ at) - out[0] * at) * ge
            grad_box = grad_box.to(bdtype)
        return ((gpos * ge).to(pdtype), grad_box, (gq * ge), None, None, None, None, None)


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
            want_box_grad: bool = True):
        """Energies (n_sys, 5), forces (n_total, 3) and dE/dbox (n_sys, 3, 3)
        for positions (n_total, 3) and charges (n_total,) stored back to back."""
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
        _lib.check(lib.admp_pme_batch(self.n_sys, self._pn, self._pbox, self._prc, self._pkap,
                                      self._pkm, _lib.PRECISION[self.precision],
                                      _lib.KCONV[self.convention], self.dielectric,
                                      _lib.ptr(pos), _lib.ptr(q), _lib.ptr(energy),
                                      _lib.ptr(forces), _lib.ptr(dedbox), _lib.ptr(workspace),
                                      self.workspace_bytes, _lib.stream_of(pos)),
                   "admp_pme_batch")
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
