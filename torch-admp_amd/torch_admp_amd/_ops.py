"""The HIP entry points as PyTorch operators, namespace ``admp``.

Registering them with the dispatcher (torch.library) makes the force modules
scriptable: TorchScript compiles ``torch.ops.admp.*`` calls, and a scripted
module saved with ``torch.jit.save`` loads again once this package is
imported.  Each operator's implementation calls the C ABI (include/admp_amd.h)
on the caller's current stream; none has a CPU implementation.

Derivatives.  Every operator that returns an energy has an autograd formula
made of other ``admp`` operators, so gradients of gradients work where the
reference's callers take them:

* ``pair_energy`` -> ``pair_energy_grads`` (dE/dq, dE/dds, dE/deta in one
  kernel).  Its own backward, for an upstream v on dE/dq: d/dq = the pair
  operator applied to v (the kernel's dE/dq at charges v, the operator being
  symmetric); d/dds and d/deta from the polarisation identity
  [f(q + v) - f(q - v)] / 2, the kernel's dE/dds and dE/deta being quadratic
  in the charges.
* ``pme_reciprocal`` -> ``pme_reciprocal_grads`` (gather and box terms).  Its
  own backward for v on dE/dq: d/dq = the reciprocal potential of charges v at
  the atoms (one spread / FFT / gather at charges v); d/dr and d/dbox by the
  same polarisation identity (two evaluations at q +- v).

That covers the second derivative in the charges that the reference's
QEqForceModule.calc_hessian takes (src/torch_admp/qeq.py:507-514), mixed with
positions, box, ds and eta.  Differentiating forces, dE/dds or dE/deta again
(force-matching losses) raises NotImplementedError rather than returning a
partial gradient.

Auxiliary outputs (the potential mesh and k-space box derivative of
``pme_reciprocal``) feed the matching backward operator only; gradients sent
into them are ignored.
"""

from __future__ import annotations

import ctypes
import math
from typing import List, Optional, Tuple

import torch
from torch import Tensor

from . import _lib

_F64 = torch.float64
_SECOND_ORDER = ("second derivatives through {} (differentiating forces, dE/dds or dE/deta "
                 "again) are not implemented; derivatives of dE/dq are")


def _f64(t: Tensor) -> Tensor:
    return t.detach().to(_F64).contiguous()


def _ws(nbytes: int, dev, kind: str) -> Tensor:
    return _lib.stream_workspace(kind, nbytes, dev)


_SIZES: "dict" = {}


def _ws_bytes(fn_name: str, *args) -> int:
    """Workspace size from the library's size query, remembered per argument
    values (the queries are pure functions of them)."""
    key = (fn_name,) + tuple(tuple(a) if isinstance(a, ctypes.Array) else a for a in args)
    nb = _SIZES.get(key)
    if nb is None:
        out = ctypes.c_size_t()
        _lib.check(getattr(_lib.load(), fn_name)(*args, ctypes.byref(out)), fn_name)
        nb = out.value
        if len(_SIZES) > 4096:
            _SIZES.clear()
        _SIZES[key] = nb
    return nb


# ------------------------------------------------------------------ neighbour pairs
def _search(positions: Tensor, box: Optional[Tensor], cutoff: float,
            box_host: Optional[Tensor], rows: bool) -> Tuple[Tensor, Tensor, Tensor]:
    _lib.require_gpu(positions)
    lib = _lib.load()
    pos = _f64(positions)
    n = pos.shape[0]
    boxh = _lib.box_array(box, box_host)
    nbytes = _ws_bytes("admp_nblist_workspace_bytes", n, boxh, float(cutoff))
    ws = _ws(nbytes, pos.device, "nblist")
    stream = _lib.stream_of(pos)
    npairs = ctypes.c_int64()
    _lib.check(lib.admp_nblist_count(_lib.ptr(pos), n, boxh, float(cutoff), _lib.ptr(ws),
                                     nbytes, ctypes.byref(npairs), stream),
               "admp_nblist_count")
    p = npairs.value
    pairs = torch.empty(p, 2, dtype=torch.int64, device=pos.device)
    if rows:
        row_ptr = torch.empty(n + 1, dtype=torch.int64, device=pos.device)
        reverse = torch.empty(p, dtype=torch.int32, device=pos.device)
        scratch = torch.empty(p, dtype=torch.int32, device=pos.device)
        if n == 0:
            row_ptr.zero_()
        ptrs = (_lib.ptr(row_ptr), _lib.ptr(reverse), _lib.ptr(scratch))
    else:
        row_ptr = reverse = torch.empty(0, dtype=torch.int64, device=pos.device)
        ptrs = (None, None, None)
    _lib.check(lib.admp_nblist_fill(_lib.ptr(pos), n, boxh, float(cutoff), _lib.ptr(ws),
                                    nbytes, _lib.ptr(pairs), p, *ptrs, stream),
               "admp_nblist_fill")
    return pairs, row_ptr, reverse


@torch.library.custom_op("admp::neighbor_pairs", mutates_args=())
def neighbor_pairs(positions: Tensor, box: Optional[Tensor], cutoff: float,
                   box_host: Optional[Tensor]) -> Tensor:
    """Ordered pairs (P, 2) int64 with 0.01 < d < cutoff (box) or d < cutoff
    (box None), both orderings, grouped by the first atom in ascending index
    (src/torch_admp/nblist.py:335-490).  One stream synchronisation returns P.
    box_host: the box's values on the host when the caller has them (saves
    the device-to-host copy), else None; the same holds for every operator."""
    return _search(positions, box, cutoff, box_host, False)[0]


@torch.library.custom_op("admp::neighbor_list", mutates_args=())
def neighbor_list(positions: Tensor, box: Optional[Tensor], cutoff: float,
                  box_host: Optional[Tensor]) -> Tuple[Tensor, Tensor, Tensor]:
    """The pairs of neighbor_pairs with their row form: row_ptr (n + 1,
    int64; n rows, one per atom in cell order, row r = pairs[row_ptr[r]:
    row_ptr[r + 1]] sharing one first atom, some empty) and reverse (P, int32;
    reverse[o] = index of the pair (j, i) of o = (i, j)).  pair_distance
    given both runs its backward without per-pair atomics."""
    return _search(positions, box, cutoff, box_host, True)


@torch.library.custom_op("admp::max_displacement", mutates_args=())
def max_displacement(positions: Tensor, reference: Tensor, box: Optional[Tensor],
                     box_host: Optional[Tensor]) -> Tensor:
    """max_i |positions_i - reference_i| (minimum image when box is given),
    a 0-dim float64 device tensor."""
    _lib.require_gpu(positions, reference)
    lib = _lib.load()
    pos, ref = _f64(positions), _f64(reference)
    out = torch.empty((), dtype=_F64, device=pos.device)
    _lib.check(lib.admp_max_displacement(_lib.ptr(pos), _lib.ptr(ref), pos.shape[0],
                                         _lib.box_array(box, box_host), _lib.ptr(out),
                                         _lib.stream_of(pos)), "admp_max_displacement")
    return out


# ------------------------------------------------------------------ pair distances
@torch.library.custom_op("admp::pair_distance", mutates_args=())
def pair_distance(positions: Tensor, box: Optional[Tensor], pairs: Tensor,
                  box_host: Optional[Tensor], row_ptr: Optional[Tensor] = None,
                  reverse: Optional[Tensor] = None) -> Tensor:
    """ds = |pbc_shift(r_j - r_i)| (src/torch_admp/nblist.py:517-551,
    src/torch_admp/spatial.py:15-36), |r_j - r_i| when box is None; (P,) in the
    positions' dtype.  row_ptr and reverse (from neighbor_list, for these
    pairs) only change how the backward sums."""
    _lib.require_gpu(positions, pairs)
    lib = _lib.load()
    pos = _f64(positions)
    prs = pairs.detach().to(torch.int64).contiguous()
    npairs = prs.shape[0]
    ds = torch.empty(npairs, dtype=_F64, device=pos.device)
    _lib.check(lib.admp_pair_ds(_lib.ptr(pos), _lib.ptr(prs), npairs, _lib.box_array(box, box_host),
                                _lib.ptr(ds), None, _lib.stream_of(pos)), "admp_pair_ds")
    return ds.to(positions.dtype)


@torch.library.custom_op("admp::pair_distance_backward", mutates_args=())
def pair_distance_backward(grad_ds: Tensor, positions: Tensor, box: Optional[Tensor],
                           pairs: Tensor, ds: Tensor, box_host: Optional[Tensor],
                           row_ptr: Optional[Tensor] = None,
                           reverse: Optional[Tensor] = None) -> Tuple[Tensor, Tensor]:
    """(dL/dpositions, dL/dbox) of L(ds); dL/dbox is zero when box is None.
    The displacement is recomputed from the positions.  With row_ptr and
    reverse each atom's row is summed in one pass (no per-pair atomics)."""
    _lib.require_gpu(grad_ds, positions, pairs)
    rows = row_ptr is not None and reverse is not None and row_ptr.numel() == positions.shape[0] + 1
    lib = _lib.load()
    pos = _f64(positions)
    prs = pairs.detach().to(torch.int64).contiguous()
    g = _f64(grad_ds)
    gpos = torch.zeros_like(pos)
    gbox = torch.zeros(3, 3, dtype=_F64, device=pos.device)
    boxh = _lib.box_array(box, box_host)
    nbytes = _ws_bytes("admp_pair_ds_backward_workspace_bytes", prs.shape[0])
    ws = _ws(nbytes, pos.device, "pair_ds")
    rp = rv = None
    if rows:
        assert row_ptr is not None and reverse is not None
        rp = _lib.ptr(row_ptr.to(torch.int64).contiguous())
        rv = _lib.ptr(reverse.to(torch.int32).contiguous())
    _lib.check(lib.admp_pair_ds_backward(_lib.ptr(g), _lib.ptr(prs), prs.shape[0], _lib.ptr(pos),
                                         boxh, None, _lib.ptr(_f64(ds)), _lib.ptr(gpos),
                                         _lib.ptr(gbox) if boxh is not None else None,
                                         rp, rv, pos.shape[0], _lib.ptr(ws), nbytes,
                                         _lib.stream_of(pos)),
               "admp_pair_ds_backward")
    if box is not None:
        gbox = gbox.to(device=box.device, dtype=box.dtype)
    return gpos.to(positions.dtype), gbox


def _pd_setup(ctx, inputs, output):
    positions, box, pairs, box_host, row_ptr, reverse = inputs
    ctx.save_for_backward(positions, box, pairs, output, box_host, row_ptr, reverse)
    ctx.set_materialize_grads(False)


def _pd_backward(ctx, g_ds):
    positions, box, pairs, ds, box_host, row_ptr, reverse = ctx.saved_tensors
    if g_ds is None:
        return None, None, None, None, None, None
    gpos, gbox = torch.ops.admp.pair_distance_backward(g_ds, positions, box, pairs, ds, box_host,
                                                       row_ptr, reverse)
    return (gpos if ctx.needs_input_grad[0] else None,
            gbox if box is not None and ctx.needs_input_grad[1] else None,
            None, None, None, None)


def _pdb_setup(ctx, inputs, output):
    ctx.set_materialize_grads(False)


def _pdb_backward(ctx, g_gpos, g_gbox):
    if g_gpos is not None or g_gbox is not None:
        raise NotImplementedError(_SECOND_ORDER.format("the pair distances"))
    return (None,) * 8


torch.library.register_autograd("admp::pair_distance", _pd_backward, setup_context=_pd_setup)
torch.library.register_autograd("admp::pair_distance_backward", _pdb_backward,
                                setup_context=_pdb_setup)


# ------------------------------------------------------------------ pair energies
PAIR_EWALD, PAIR_BARE, PAIR_DAMPED = 0, 1, 2


def _pair_call(q: Tensor, ds: Tensor, pairs: Tensor, bs: Tensor, eta: Optional[Tensor],
               mode: int, kappa: float, eps: float, want_q: bool, want_ds: bool,
               want_eta: bool):
    _lib.require_gpu(q, ds, pairs, bs, eta)
    lib = _lib.load()
    nf, n = q.shape
    npairs = pairs.shape[1]
    q64, ds64, bs64 = _f64(q), _f64(ds), _f64(bs)
    prs = pairs.detach().to(torch.int64).contiguous()
    eta64 = None if eta is None else _f64(eta)
    dev = q64.device
    energy = torch.empty(nf, dtype=_F64, device=dev)
    gq = torch.empty(nf, n, dtype=_F64, device=dev) if want_q else None
    gds = torch.empty(nf, npairs, dtype=_F64, device=dev) if want_ds else None
    geta = (torch.empty(nf, n, dtype=_F64, device=dev)
            if want_eta and mode == PAIR_DAMPED and eta is not None else None)
    nbytes = _ws_bytes("admp_pair_energy_workspace_bytes", nf)
    ws = _ws(nbytes, dev, "pair_energy")
    _lib.check(lib.admp_pair_energy(mode, nf, n, npairs, _lib.ptr(q64), _lib.ptr(eta64),
                                    _lib.ptr(prs), _lib.ptr(ds64), _lib.ptr(bs64), kappa, eps,
                                    _lib.ptr(energy), _lib.ptr(gds), _lib.ptr(gq), _lib.ptr(geta),
                                    _lib.ptr(ws), nbytes, _lib.stream_of(q64)),
               "admp_pair_energy")
    return energy, gq, gds, geta


@torch.library.custom_op("admp::pair_energy", mutates_args=())
def pair_energy(q: Tensor, ds: Tensor, pairs: Tensor, bs: Tensor, eta: Optional[Tensor],
                mode: int, kappa: float, eps: float) -> Tensor:
    """Per-frame pair energy over (nf, P) pair rows, float64 accumulation:
    mode 0 eps erfc(kappa d) qi qj / d bs (src/torch_admp/pme.py:312-349),
    1 eps qi qj / d bs (pme.py:529-561), 2 the Gaussian damping
    -eps erfc(d / (2 eta_ij)) qi qj / d bs (src/torch_admp/qeq.py:106-121)."""
    energy, _, _, _ = _pair_call(q, ds, pairs, bs, eta, mode, kappa, eps, False, False, False)
    return energy.to(q.dtype)


@torch.library.custom_op("admp::pair_energy_grads", mutates_args=())
def pair_energy_grads(q: Tensor, ds: Tensor, pairs: Tensor, bs: Tensor, eta: Optional[Tensor],
                      mode: int, kappa: float, eps: float, want_q: bool, want_ds: bool,
                      want_eta: bool) -> Tuple[Tensor, Tensor, Tensor]:
    """(dE/dq (nf, n), dE/dds (nf, P), dE/deta (nf, n)) of pair_energy; the
    ones not wanted (and dE/deta outside mode 2) are zero."""
    _, gq, gds, geta = _pair_call(q, ds, pairs, bs, eta, mode, kappa, eps, want_q, want_ds,
                                  want_eta)
    nf, n = q.shape
    gq = torch.zeros(nf, n, dtype=q.dtype, device=q.device) if gq is None else gq.to(q.dtype)
    gds = (torch.zeros(ds.shape, dtype=ds.dtype, device=ds.device) if gds is None
           else gds.reshape(ds.shape).to(ds.dtype))
    if geta is None:
        geta = torch.zeros(nf, n, dtype=q.dtype, device=q.device)
    elif eta is not None:
        geta = geta.reshape(eta.shape).to(eta.dtype)
    return gq, gds, geta


def _pe_setup(ctx, inputs, output):
    q, ds, pairs, bs, eta, mode, kappa, eps = inputs
    ctx.save_for_backward(q, ds, pairs, bs, eta)
    ctx.args = (mode, kappa, eps)
    ctx.set_materialize_grads(False)


def _pe_backward(ctx, g_e):
    q, ds, pairs, bs, eta = ctx.saved_tensors
    mode, kappa, eps = ctx.args
    if g_e is None:
        return (None,) * 8
    need = ctx.needs_input_grad
    want_eta = bool(need[4]) and eta is not None and mode == PAIR_DAMPED
    gq, gds, geta = torch.ops.admp.pair_energy_grads(q, ds, pairs, bs, eta, mode, kappa, eps,
                                                     bool(need[0]), bool(need[1]), want_eta)
    ge = g_e.reshape(-1, 1)
    return (ge * gq if need[0] else None, ge * gds if need[1] else None, None, None,
            ge * geta if want_eta else None, None, None, None)


def _peg_setup(ctx, inputs, output):
    q, ds, pairs, bs, eta, mode, kappa, eps = inputs[:8]
    ctx.save_for_backward(q, ds, pairs, bs, eta)
    ctx.args = (mode, kappa, eps)
    ctx.set_materialize_grads(False)


def _peg_backward(ctx, g_gq, g_gds, g_geta):
    if g_gds is not None or g_geta is not None:
        raise NotImplementedError(_SECOND_ORDER.format("dE/dds or dE/deta of the pair energy"))
    nothing = (None,) * 11
    if g_gq is None:
        return nothing
    q, ds, pairs, bs, eta = ctx.saved_tensors
    mode, kappa, eps = ctx.args
    need = ctx.needs_input_grad
    d_q = d_ds = d_eta = None
    if need[0]:
        # the pair operator is symmetric: v^T (C q) differentiates to C v
        d_q = torch.ops.admp.pair_energy_grads(g_gq, ds, pairs, bs, eta, mode, kappa, eps,
                                               True, False, False)[0]
    want_eta = bool(need[4]) and eta is not None and mode == PAIR_DAMPED
    if need[1] or want_eta:
        v = g_gq.to(q.dtype)
        plus = torch.ops.admp.pair_energy_grads(q + v, ds, pairs, bs, eta, mode, kappa, eps,
                                                False, bool(need[1]), want_eta)
        minus = torch.ops.admp.pair_energy_grads(q - v, ds, pairs, bs, eta, mode, kappa, eps,
                                                 False, bool(need[1]), want_eta)
        if need[1]:
            d_ds = 0.5 * (plus[1] - minus[1])
        if want_eta:
            d_eta = 0.5 * (plus[2] - minus[2])
    return (d_q, d_ds, None, None, d_eta) + (None,) * 6


torch.library.register_autograd("admp::pair_energy", _pe_backward, setup_context=_pe_setup)
torch.library.register_autograd("admp::pair_energy_grads", _peg_backward,
                                setup_context=_peg_setup)


# ------------------------------------------------------------------ self and non-neutral terms
@torch.library.custom_op("admp::ewald_self_terms", mutates_args=())
def ewald_self_terms(q: Tensor, box: Tensor, kappa: float, eps: float) -> Tuple[Tensor, Tensor]:
    """Per frame (q (nf, n), box (nf, 3, 3) on the device): the self energy
    -eps kappa / sqrt(pi) sum q^2 and the non-neutral energy
    -eps pi Q^2 / (2 V kappa^2) (src/torch_admp/pme.py:448-493), in one kernel
    with no host copy of the box."""
    _lib.require_gpu(q, box)
    q64, b64 = _f64(q), _f64(box)
    nf, n = q64.shape
    e_self = torch.empty(nf, dtype=_F64, device=q64.device)
    e_nn = torch.empty(nf, dtype=_F64, device=q64.device)
    _lib.check(_lib.load().admp_ewald_self_terms(_lib.ptr(q64), nf, n, _lib.ptr(b64), kappa, eps,
                                                 _lib.ptr(e_self), _lib.ptr(e_nn),
                                                 _lib.stream_of(q64)), "admp_ewald_self_terms")
    return e_self.to(q.dtype), e_nn.to(q.dtype)


@torch.library.custom_op("admp::ewald_self_terms_backward", mutates_args=())
def ewald_self_terms_backward(q: Tensor, box: Tensor, kappa: float, eps: float,
                              g_self: Optional[Tensor], g_nn: Optional[Tensor]
                              ) -> Tuple[Tensor, Tensor]:
    """(dL/dq (nf, n), dL/dbox (nf, 3, 3)) of ewald_self_terms for upstream
    gradients g_self, g_nn (nf each; None for zero)."""
    _lib.require_gpu(q, box)
    q64, b64 = _f64(q), _f64(box)
    nf, n = q64.shape
    dq = torch.empty(nf, n, dtype=_F64, device=q64.device)
    dbox = torch.empty(nf, 3, 3, dtype=_F64, device=q64.device)
    gs = None if g_self is None else _f64(g_self).reshape(nf)
    gn = None if g_nn is None else _f64(g_nn).reshape(nf)
    _lib.check(_lib.load().admp_ewald_self_terms_backward(
        _lib.ptr(q64), nf, n, _lib.ptr(b64), kappa, eps, _lib.ptr(gs), _lib.ptr(gn),
        _lib.ptr(dq), _lib.ptr(dbox), _lib.stream_of(q64)), "admp_ewald_self_terms_backward")
    return dq.to(q.dtype), dbox.to(box.dtype)


def _st_setup(ctx, inputs, output):
    q, box, kappa, eps = inputs
    ctx.save_for_backward(q, box)
    ctx.args = (kappa, eps)
    ctx.set_materialize_grads(False)


def _st_backward(ctx, g_self, g_nn):
    q, box = ctx.saved_tensors
    kappa, eps = ctx.args
    if g_self is None and g_nn is None:
        return None, None, None, None
    if torch.is_grad_enabled():
        # a graph of the gradient is being built (second derivatives, as the
        # reference's Hessian takes them): the same formulas as torch
        # expressions, differentiable in q, box and the upstream gradients
        cs = -eps * kappa / math.sqrt(math.pi)
        ra, rb, rc = box[:, 0, :], box[:, 1, :], box[:, 2, :]
        # d det / d box row by row: b x c, c x a, a x b (the cofactor matrix)
        dvol = torch.stack([torch.cross(rb, rc, dim=-1), torch.cross(rc, ra, dim=-1),
                            torch.cross(ra, rb, dim=-1)], dim=1)
        vol = torch.sum(ra * dvol[:, 0, :], dim=-1)
        qt = torch.sum(q, dim=-1)
        cn = -eps * math.pi / (2.0 * vol * kappa * kappa)
        dq = torch.zeros_like(q)
        dbox = torch.zeros_like(box)
        if g_self is not None:
            dq = dq + (2.0 * cs) * g_self.unsqueeze(-1) * q
        if g_nn is not None:
            dq = dq + (2.0 * g_nn * cn * qt).unsqueeze(-1)
            dbox = dbox - (g_nn * cn * qt * qt / vol).reshape(-1, 1, 1) * dvol
    else:
        dq, dbox = torch.ops.admp.ewald_self_terms_backward(q, box, kappa, eps, g_self, g_nn)
    need = ctx.needs_input_grad
    return (dq if need[0] else None, dbox if need[1] else None, None, None)


torch.library.register_autograd("admp::ewald_self_terms", _st_backward, setup_context=_st_setup)


# ------------------------------------------------------------------ PME reciprocal space
def _pme_ws(lib, km, precision: int, n: int, dev) -> Tuple[Tensor, int]:
    nbytes = _ws_bytes("admp_pme_workspace_bytes", km, precision, n)
    return _ws(nbytes, dev, "pme"), nbytes


@torch.library.custom_op("admp::pme_reciprocal", mutates_args=())
def pme_reciprocal(positions: Tensor, box: Tensor, q: Tensor, kmesh: List[int], kappa: float,
                   eps: float, precision: int, convention: int, need_potential: bool,
                   box_host: Optional[Tensor]) -> Tuple[Tensor, Tensor, Tensor]:
    """One frame of the SPME reciprocal energy (src/torch_admp/pme.py:351-446):
    spread, r2c FFT, influence function, c2r FFT.  Returns (energy (0-dim, the
    positions' dtype), the potential mesh dE/dQ (empty unless
    need_potential), dE/dA through the k vectors (9,) float64)."""
    _lib.require_gpu(positions, q)
    lib = _lib.load()
    pos = _f64(positions)
    q64 = _f64(q).reshape(-1)
    n = pos.shape[0]
    dev = pos.device
    boxh = _lib.box_array(box, box_host)
    km = _lib.kmesh_array(kmesh)
    mesh_dtype = torch.float32 if precision == _lib.ADMP_F32 else _F64
    mesh = torch.empty(int(kmesh[0]) * int(kmesh[1]) * int(kmesh[2]), dtype=mesh_dtype,
                       device=dev)
    ws, nbytes = _pme_ws(lib, km, precision, n, dev)
    stream = _lib.stream_of(pos)
    _lib.check(lib.admp_pme_spread(_lib.ptr(pos), _lib.ptr(q64), n, boxh, km, precision,
                                   _lib.ptr(mesh), _lib.ptr(ws), nbytes, stream),
               "admp_pme_spread")
    energy = torch.empty((), dtype=_F64, device=dev)
    dEdA = torch.empty(9, dtype=_F64, device=dev)
    _lib.check(lib.admp_pme_convolve(_lib.ptr(mesh), boxh, km, precision, convention, kappa,
                                     eps, 1.0, _lib.ptr(mesh) if need_potential else None,
                                     _lib.ptr(energy), _lib.ptr(dEdA), _lib.ptr(ws), nbytes,
                                     stream), "admp_pme_convolve")
    phi = mesh if need_potential else mesh.new_empty(0)
    return energy.to(positions.dtype), phi, dEdA


@torch.library.custom_op("admp::pme_reciprocal_grads", mutates_args=())
def pme_reciprocal_grads(positions: Tensor, box: Tensor, q: Tensor, potential: Tensor,
                         dEdA_k: Tensor, energy: Tensor, kmesh: List[int], kappa: float,
                         eps: float, precision: int, convention: int,
                         box_host: Optional[Tensor]) -> Tuple[Tensor, Tensor, Tensor]:
    """(dE/dpositions (n, 3), dE/dq (n,), dE/dbox (3, 3)) of pme_reciprocal
    from its potential mesh: the gather (the transpose of the spread) plus the
    k-vector and volume terms of dE/dbox."""
    _lib.require_gpu(positions, q, potential)
    lib = _lib.load()
    pos = _f64(positions)
    q64 = _f64(q).reshape(-1)
    n = pos.shape[0]
    dev = pos.device
    boxh = _lib.box_array(box, box_host)
    km = _lib.kmesh_array(kmesh)
    gpos = torch.empty(n, 3, dtype=_F64, device=dev)
    gq = torch.empty(n, dtype=_F64, device=dev)
    dEdA = torch.empty(9, dtype=_F64, device=dev)
    ws, nbytes = _pme_ws(lib, km, precision, n, dev)
    _lib.check(lib.admp_pme_gather(_lib.ptr(pos), _lib.ptr(q64), n, boxh, km, precision,
                                   _lib.ptr(potential.contiguous()), _lib.ptr(gpos), _lib.ptr(gq),
                                   _lib.ptr(dEdA), 0, _lib.ptr(ws), nbytes, _lib.stream_of(pos)),
               "admp_pme_gather")
    gbox = torch.empty(3, 3, dtype=_F64, device=dev)
    _lib.check(lib.admp_pme_box_grad(boxh, _lib.ptr(_f64(dEdA_k).contiguous()), _lib.ptr(dEdA),
                                     _lib.ptr(_f64(energy).contiguous()), _lib.ptr(gbox),
                                     _lib.stream_of(pos)), "admp_pme_box_grad")
    return (gpos.to(positions.dtype), gq.to(q.dtype).reshape(q.shape),
            gbox.to(device=box.device, dtype=box.dtype))


def _pr_setup(ctx, inputs, output):
    positions, box, q, kmesh, kappa, eps, precision, convention, need_pot, box_host = inputs
    energy, phi, dEdA_k = output
    ctx.save_for_backward(positions, box, q, phi, dEdA_k, energy, box_host)
    ctx.args = (list(kmesh), kappa, eps, precision, convention, need_pot)
    ctx.set_materialize_grads(False)


def _pr_backward(ctx, g_e, g_phi, g_dA):
    positions, box, q, phi, dEdA_k, energy, box_host = ctx.saved_tensors
    kmesh, kappa, eps, precision, convention, need_pot = ctx.args
    if g_e is None:
        return (None,) * 10
    if not need_pot:
        raise RuntimeError("pme_reciprocal was called with need_potential=False; its energy "
                           "cannot be differentiated")
    gpos, gq, gbox = torch.ops.admp.pme_reciprocal_grads(positions, box, q, phi, dEdA_k,
                                                         energy.detach(), kmesh, kappa, eps,
                                                         precision, convention, box_host)
    need = ctx.needs_input_grad
    return (g_e * gpos if need[0] else None, g_e * gbox if need[1] else None,
            g_e * gq if need[2] else None) + (None,) * 7


def _pme_eval_grads(positions, box, q, kmesh, kappa, eps, precision, convention, box_host):
    e, phi, dA = torch.ops.admp.pme_reciprocal(positions, box, q, kmesh, kappa, eps, precision,
                                               convention, True, box_host)
    return torch.ops.admp.pme_reciprocal_grads(positions, box, q, phi, dA, e, kmesh, kappa, eps,
                                               precision, convention, box_host)


def _prg_setup(ctx, inputs, output):
    positions, box, q = inputs[:3]
    ctx.save_for_backward(positions, box, q, inputs[11])
    ctx.args = tuple(inputs[6:11])
    ctx.set_materialize_grads(False)


def _prg_backward(ctx, g_gpos, g_gq, g_gbox):
    if g_gpos is not None or g_gbox is not None:
        raise NotImplementedError(_SECOND_ORDER.format("the reciprocal-space forces"))
    nothing = (None,) * 12
    if g_gq is None:
        return nothing
    positions, box, q, box_host = ctx.saved_tensors
    kmesh, kappa, eps, precision, convention = ctx.args
    kmesh = list(kmesh)
    need = ctx.needs_input_grad
    v = g_gq.to(q.dtype).reshape(q.shape)
    d_pos = d_box = d_q = None
    if need[2]:
        # E is quadratic in q: v^T (dE/dq) differentiates to the potential of charges v
        d_q = _pme_eval_grads(positions, box, v, kmesh, kappa, eps, precision, convention,
                              box_host)[1]
    if need[0] or need[1]:
        gp = _pme_eval_grads(positions, box, q + v, kmesh, kappa, eps, precision, convention,
                             box_host)
        gm = _pme_eval_grads(positions, box, q - v, kmesh, kappa, eps, precision, convention,
                             box_host)
        if need[0]:
            d_pos = 0.5 * (gp[0] - gm[0])
        if need[1]:
            d_box = 0.5 * (gp[2] - gm[2])
    return (d_pos, d_box, d_q) + (None,) * 9


torch.library.register_autograd("admp::pme_reciprocal", _pr_backward, setup_context=_pr_setup)
torch.library.register_autograd("admp::pme_reciprocal_grads", _prg_backward,
                                setup_context=_prg_setup)
