"""GPU parity of the HIP path against the float64 CPU oracle.

Tolerances: float64 kernels agree with the oracle to 1e-10 relative energy and
1e-8 eV/A force RMSE (rounding only); float32-bulk kernels to 1e-6 relative
energy and 1e-4 eV/A force RMSE.  Pair sets are compared exactly.
"""

import math

import numpy as np
import pytest
import torch

from oracle import nblist_ref, pme_ref
from torch_admp_amd import ops
from torch_admp_amd.nblist import TorchNeighborList
from torch_admp_amd.pme import CoulombForceModule
from torch_admp_amd.utils import calc_grads

pytestmark = pytest.mark.gpu

F64 = torch.float64


def _system(n, box, seed, neutral=True):
    rng = np.random.default_rng(seed)
    box = torch.as_tensor(box, dtype=F64)
    frac = torch.tensor(rng.random((n, 3)))
    pos = frac @ box
    q = torch.tensor(rng.uniform(-1, 1, n))
    if neutral:
        q -= q.mean()
    return pos, q, box


TRICLINIC = [[12.0, 0.0, 0.0], [2.5, 13.0, 0.0], [-1.5, 1.0, 14.0]]
BIG_TRICLINIC_NB = [[36.0, 0.0, 0.0], [5.0, 34.0, 0.0], [-3.0, 4.0, 38.0]]
SYMMETRIC = [[12.0, 0.4, 0.3], [0.4, 13.0, 0.2], [0.3, 0.2, 14.0]]


def _oracle_ef(pos, box, q, rc, kmesh, **kw):
    pairs = nblist_ref.nblist_pbc(pos, box, rc)
    bs = nblist_ref.buffer_scales(pairs)
    orc = pme_ref.CoulombOracle(rc, 1e-5, kmesh=kmesh, **kw)
    e, f, gb = pme_ref.energy_forces(
        pos, box, q, pairs, lambda p, b: nblist_ref.pairs_ds(p, pairs, b), bs, orc)
    return e, f, gb, orc


@pytest.mark.parametrize("box", [np.eye(3) * 13.0, TRICLINIC, SYMMETRIC])
def test_nblist_pair_set_exact(dev, box):
    pos, _, box = _system(300, box, 1)
    # positions partly outside the box exercise the wrapping
    pos = pos + torch.tensor([[30.0, -25.0, 41.0]], dtype=F64) * (torch.arange(300) % 3 == 0)[:, None]
    rc = 5.0
    nbl = TorchNeighborList(rc)
    pairs = nbl(pos.to(dev), box.to(dev)).cpu()
    got = set(map(tuple, pairs.tolist()))
    ref = set(map(tuple, nblist_ref.nblist_pbc(pos, box, rc).tolist()))
    brute, amb = nblist_ref.brute_force_pairs(pos.numpy(), box.numpy(), rc)
    assert not amb
    assert got == ref == brute
    assert len(got) == pairs.shape[0]  # no duplicates
    _assert_grouped(pairs)
    assert torch.equal(nbl.get_buffer_scales().cpu(), (pairs[:, 0] < pairs[:, 1]).long())


def _assert_grouped(pairs):
    """Each first atom's pairs are contiguous."""
    first = pairs[:, 0]
    starts = torch.cat([torch.ones(1, dtype=torch.bool), first[1:] != first[:-1]])
    assert int(starts.sum()) == len(torch.unique(first))


@pytest.mark.parametrize("box,rc", [(np.eye(3) * 20.0, 6.0), (np.diag([33.0, 41.0, 29.0]), 7.5),
                                    (BIG_TRICLINIC_NB, 6.0)])
def test_nblist_many_cells(dev, box, rc):
    """Boxes of 27-200 cells (the dense 20 A box stages more candidates than
    one LDS chunk holds per cell), against brute force; grouped by first atom."""
    pos, _, box = _system(3000, box, 21)
    pairs = TorchNeighborList(rc)(pos.to(dev), box.to(dev)).cpu()
    got = set(map(tuple, pairs.tolist()))
    brute, amb = nblist_ref.brute_force_pairs(pos.numpy(), box.numpy(), rc)
    assert not amb
    assert got == brute and len(got) == pairs.shape[0]
    _assert_grouped(pairs)


def test_nblist_obc(dev):
    pos, _, _ = _system(200, np.eye(3) * 10.0, 2)
    rc = 3.0
    pairs = TorchNeighborList(rc)(pos.to(dev), None).cpu()
    ref = set(map(tuple, nblist_ref.nblist_obc(pos, rc).tolist()))
    assert set(map(tuple, pairs.tolist())) == ref


@pytest.mark.parametrize("order", ["grouped", "shuffled"])
@pytest.mark.parametrize("box", [np.eye(3) * 13.0, TRICLINIC])
def test_pair_ds_and_backward(dev, box, order):
    """ds and its transpose, for pair lists grouped by first atom and in
    random order, with a third of the upstream gradients zero (as the i > j
    half of a reference list is, through its zero buffer scales)."""
    pos, _, box = _system(150, box, 3)
    pairs = nblist_ref.nblist_pbc(pos, box, 5.0)
    rng = np.random.default_rng(4)
    if order == "shuffled":
        pairs = pairs[torch.tensor(rng.permutation(pairs.shape[0]))]
    p = pos.clone().requires_grad_(True)
    b = box.clone().requires_grad_(True)
    ds_ref = nblist_ref.pairs_ds(p, pairs, b)
    w = torch.tensor(rng.normal(size=ds_ref.shape[0]) * (rng.random(ds_ref.shape[0]) < 0.67))
    gp_ref, gb_ref = torch.autograd.grad((ds_ref * w).sum(), [p, b])
    pg = pos.to(dev).requires_grad_(True)
    bg = box.to(dev).requires_grad_(True)
    ds = ops.pair_distance(pg, bg, pairs.to(dev))
    gp, gb = torch.autograd.grad((ds * w.to(dev)).sum(), [pg, bg])
    assert torch.allclose(ds.detach().cpu(), ds_ref.detach(), rtol=0, atol=1e-12)
    assert torch.allclose(gp.cpu(), gp_ref, rtol=0, atol=1e-10)
    assert torch.allclose(gb.cpu(), gb_ref, rtol=0, atol=1e-9)


@pytest.mark.parametrize("box,rc,n", [(np.eye(3) * 20.0, 6.0, 3000), (BIG_TRICLINIC_NB, 6.0, 3000),
                                      (np.eye(3) * 13.0, 5.0, 200), (TRICLINIC, 5.0, 300),
                                      (None, 4.0, 400)])
def test_row_form_and_reverse_index(dev, box, rc, n):
    """The search's row form: one row per atom (cell order, empty rows
    allowed) with row_ptr their starts, each row one first atom, and
    reverse[o] the index of (j, i) for o = (i, j)
    (cell grids of 3+ cells per side, one cell with images, triclinic, open
    boundaries).  The distance backward through it (no per-pair atomics)
    against the oracle's autograd, with the buffer-scale zeros and random
    zeros in the upstream gradient; then with part of the reverse index
    removed (pairs without a reverse add their second-atom term by atomics)."""
    pos, _, boxt = _system(n, np.eye(3) * 12.0 if box is None else box, 31)
    if box is None:
        boxt = None
    nbl = TorchNeighborList(rc)
    pairs = nbl(pos.to(dev), None if boxt is None else boxt.to(dev))
    row_ptr, rev = nbl.row_ptr.cpu(), nbl.reverse.cpu().long()
    pc = pairs.cpu()
    assert row_ptr.shape[0] == n + 1 and int(row_ptr[0]) == 0 and int(row_ptr[-1]) == pc.shape[0]
    lens = row_ptr[1:] - row_ptr[:-1]
    assert bool((lens >= 0).all())
    row_of = torch.repeat_interleave(torch.arange(n), lens)
    first = torch.full((n,), -1, dtype=torch.long)
    first[row_of] = pc[:, 0]
    assert torch.equal(pc[:, 0], first[row_of])  # one first atom per row
    counts = torch.bincount(pc[:, 0], minlength=n)
    assert torch.equal(torch.sort(lens[lens > 0]).values, torch.sort(counts[counts > 0]).values)
    assert len(torch.unique(first[lens > 0])) == int((lens > 0).sum())  # one row per atom
    assert bool((rev >= 0).all())
    assert torch.equal(pc[rev], pc[:, [1, 0]])
    assert torch.equal(rev[rev], torch.arange(pc.shape[0]))
    # backward through the row form against the oracle
    rng = np.random.default_rng(5)
    p = pos.clone().requires_grad_(True)
    b = None if boxt is None else boxt.clone().requires_grad_(True)
    ds_ref = nblist_ref.pairs_ds(p, pc, b) if b is not None else (p[pc[:, 1]] - p[pc[:, 0]]).norm(dim=1)
    bs = (pc[:, 0] < pc[:, 1]).to(F64)
    w = torch.tensor(rng.normal(size=pc.shape[0]) * (rng.random(pc.shape[0]) < 0.8)) * bs
    wrt = [p] if b is None else [p, b]
    ref = torch.autograd.grad((ds_ref * w).sum(), wrt)
    pg = pos.to(dev).requires_grad_(True)
    bg = None if boxt is None else boxt.to(dev).requires_grad_(True)
    pairs2 = nbl(pg, bg)
    ds = nbl.get_ds()
    assert torch.equal(pairs2.cpu(), pc)
    got = torch.autograd.grad((ds * w.to(dev)).sum(), [pg] if bg is None else [pg, bg])
    for g, r in zip(got, ref):
        assert torch.allclose(g.cpu(), r, rtol=0, atol=1e-10 * max(1.0, float(r.abs().max())))
    # the same with a third of the reverse index removed (both pairs of each
    # removed couple), through the operator
    cut = torch.tensor(rng.random(pc.shape[0]) < 0.33)
    cut = cut | cut[rev]
    rev_part = torch.where(cut, torch.full_like(rev, -1), rev).to(torch.int32).to(dev)
    pg2 = pos.to(dev).requires_grad_(True)
    ds2 = torch.ops.admp.pair_distance(pg2, bg, pairs2, None, nbl.row_ptr, rev_part)
    got2 = torch.autograd.grad((ds2 * w.to(dev)).sum(), [pg2])[0]
    assert torch.allclose(got2.cpu(), ref[0], rtol=0, atol=1e-10 * max(1.0, float(ref[0].abs().max())))


@pytest.mark.parametrize("n", [1, 2, 65, 1025, 32768, 32769, 50000])
def test_pair_count_scan_and_buffer_scales(dev, n):
    """Sizes either side of the single-work-group scan limit (32768 per-atom
    counts) and the trivial ones: the total pair count against scipy's
    periodic k-d tree, the row starts against the rows' first atoms, the
    buffer scales the search writes against i < j, admp::pair_scales with
    distance bounds against the same mask in torch, and the pair energy with
    int64 scales bit-equal to the same with float64 scales."""
    from scipy.spatial import cKDTree

    rc = 4.0
    side = max((n / 0.05) ** (1.0 / 3.0), 2.0 * rc + 0.5)
    rng = np.random.default_rng(n)
    pos_np = rng.random((n, 3)) * side
    box = torch.eye(3, dtype=F64) * side
    nbl = TorchNeighborList(rc)
    pairs = nbl(torch.tensor(pos_np).to(dev), box.to(dev))
    pc = pairs.cpu()
    tree = cKDTree(pos_np, boxsize=side)
    close = tree.query_pairs(0.01)
    assert pc.shape[0] == 2 * (len(tree.query_pairs(rc)) - len(close))
    row_ptr = nbl.row_ptr.cpu()
    lens = row_ptr[1:] - row_ptr[:-1]
    assert int(row_ptr[0]) == 0 and int(row_ptr[-1]) == pc.shape[0] and bool((lens >= 0).all())
    row_of = torch.repeat_interleave(torch.arange(n), lens)
    if pc.shape[0]:
        first = torch.full((n,), -1, dtype=torch.long)
        first[row_of] = pc[:, 0]
        assert torch.equal(pc[:, 0], first[row_of])
    bs = nbl.get_buffer_scales()
    assert bs.dtype == torch.long
    assert torch.equal(bs.cpu(), (pc[:, 0] < pc[:, 1]).long())
    ds = nbl.get_ds().detach()
    for lo, hi in [(0.01, rc), (1.5, 3.0), (-1.0, 2.5)]:
        got = torch.ops.admp.pair_scales(pairs, ds, lo, hi).cpu()
        d = ds.cpu()
        assert torch.equal(got, ((pc[:, 0] < pc[:, 1]) & (d > lo) & (d < hi)).long())
    assert torch.equal(torch.ops.admp.pair_scales(pairs, None, 0.0, 0.0).cpu(),
                       (pc[:, 0] < pc[:, 1]).long())
    q = torch.tensor(rng.uniform(-1, 1, n)).to(dev)[None]
    e_i64 = torch.ops.admp.pair_energy(q, ds[None], pairs[None], bs[None], None, 0, 0.4, 14.4)
    e_f64 = torch.ops.admp.pair_energy(q, ds[None], pairs[None], bs.to(F64)[None], None, 0,
                                       0.4, 14.4)
    assert torch.equal(e_i64, e_f64)


def test_real_space_pair_energy(dev):
    pos, q, box = _system(300, np.eye(3) * 14.0, 5)
    rc, kappa = 6.0, 0.4
    pairs = nblist_ref.nblist_pbc(pos, box, rc)
    bs = nblist_ref.buffer_scales(pairs).to(F64)
    ds = nblist_ref.pairs_ds(pos, pairs, box)
    qr = q[None].clone().requires_grad_(True)
    dr = ds[None].clone().requires_grad_(True)
    e_ref = pme_ref.real_energy(qr, pairs[None], dr, bs[None], kappa)
    gq_ref, gd_ref = torch.autograd.grad(e_ref.sum(), [qr, dr])
    qg = q[None].to(dev).requires_grad_(True)
    dg = ds[None].to(dev).requires_grad_(True)
    e = ops.pair_energy(qg, dg, pairs[None].to(dev), bs[None].to(dev), None,
                             ops.PAIR_EWALD, kappa, pme_ref.DIELECTRIC)
    gq, gd = torch.autograd.grad(e.sum(), [qg, dg])
    assert abs(float(e) - float(e_ref)) <= 1e-11 * abs(float(e_ref))
    assert torch.allclose(gq.cpu(), gq_ref, atol=1e-11, rtol=1e-10)
    assert torch.allclose(gd.cpu(), gd_ref, atol=1e-11, rtol=1e-10)


@pytest.mark.parametrize("tiled", [False, True])
@pytest.mark.parametrize("precision,tol", [("float64", 1e-12), ("float32", 2e-6)])
@pytest.mark.parametrize("kmesh", [[16, 18, 20], [25, 12, 27], [74, 30, 22], [7, 9, 64]])
def test_spread_mesh(dev, precision, tol, kmesh, tiled):
    """admp_pme_spread without a workspace (per-atom stencils) and with one
    (atoms binned into cells of about 8 mesh points, one LDS tile per cell)."""
    import ctypes

    from torch_admp_amd import _lib

    pos, q, box = _system(120, TRICLINIC, 6)
    pos = pos - 7.0  # negative fractional coordinates
    ref = pme_ref.spread_charges(pos, torch.linalg.inv(box), q, kmesh)
    lib = _lib.load()
    dt = torch.float32 if precision == "float32" else F64
    mesh = torch.empty(int(np.prod(kmesh)), dtype=dt, device=dev)
    pd, qd = pos.to(dev).contiguous(), q.to(dev).contiguous()
    ws, nb = None, 0
    if tiled:
        nbytes = ctypes.c_size_t()
        _lib.check(lib.admp_pme_workspace_bytes(_lib.kmesh_array(kmesh), _lib.PRECISION[precision],
                                                pos.shape[0], ctypes.byref(nbytes)), "ws")
        ws, nb = _lib.workspace(nbytes.value, dev), nbytes.value
    _lib.check(lib.admp_pme_spread(_lib.ptr(pd), _lib.ptr(qd), pos.shape[0], _lib.box_array(box),
                                   _lib.kmesh_array(kmesh), _lib.PRECISION[precision],
                                   _lib.ptr(mesh), _lib.ptr(ws), nb, _lib.stream_of(pd)), "spread")
    got = mesh.double().cpu().reshape(kmesh)
    assert (got - ref).abs().max() <= tol * ref.abs().max()
    assert abs(float(got.sum()) - float(q.sum())) < 1e-9 + tol * len(q)


@pytest.mark.parametrize("conv", ["reference", "columns"])
@pytest.mark.parametrize("box,kmesh", [(np.eye(3) * 13.0, [24, 24, 24]),
                                       (TRICLINIC, [30, 32, 36]),
                                       (np.diag([11.0, 12.5, 13.0]), [25, 27, 33]),
                                       (np.eye(3) * 14.0, [74, 37, 22])])
def test_reciprocal_energy_grads_fp64(dev, box, kmesh, conv):
    pos, q, box = _system(160, box, 8, neutral=False)
    kappa = 0.45
    p = pos.clone().requires_grad_(True)
    b = box.clone().requires_grad_(True)
    qq = q.clone().requires_grad_(True)
    e_ref = pme_ref.reciprocal_energy_frame(p, b, qq, kappa, kmesh, conv)
    gp_ref, gb_ref, gq_ref = torch.autograd.grad(e_ref, [p, b, qq])
    pg = pos.to(dev).requires_grad_(True)
    bg = box.to(dev).requires_grad_(True)
    qg = q.to(dev).requires_grad_(True)
    from torch_admp_amd import _lib

    e = ops.pme_reciprocal(pg, bg, qg, kmesh, kappa, pme_ref.DIELECTRIC, _lib.ADMP_F64,
                                _lib.KCONV[conv])
    gp, gb, gq = torch.autograd.grad(e, [pg, bg, qg])
    assert abs(float(e) - float(e_ref)) <= 1e-10 * abs(float(e_ref))
    assert float((gp.cpu() - gp_ref).abs().max()) <= 1e-9 * float(gp_ref.abs().max())
    assert float((gq.cpu() - gq_ref).abs().max()) <= 1e-9 * float(gq_ref.abs().max())
    assert float((gb.cpu() - gb_ref).abs().max()) <= 1e-8 * float(gb_ref.abs().max())


@pytest.mark.parametrize("box,kmesh,conv", [(np.eye(3) * 13.0, 24, "reference"),
                                            (TRICLINIC, [30, 32, 36], "reference"),
                                            (TRICLINIC, [30, 32, 36], "columns")])
def test_module_energy_forces(dev, box, kmesh, conv):
    pos, q, box = _system(250, box, 9)
    rc = 5.5
    km = [kmesh] * 3 if isinstance(kmesh, int) else kmesh
    e_ref, f_ref, gb_ref, orc = _oracle_ef(pos, box, q, rc, km, convention=conv)
    p = pos.to(dev).requires_grad_(True)
    b = box.to(dev).requires_grad_(True)
    nbl = TorchNeighborList(rc)
    pairs = nbl(p, b)
    mod = CoulombForceModule(rc, 1e-5, kmesh=kmesh, convention=conv)
    e = mod(p, b, pairs, nbl.get_ds(), nbl.get_buffer_scales(), {"charge": q.to(dev)})
    gp, gb = torch.autograd.grad(e.sum(), [p, b])
    assert abs(float(e) - float(e_ref)) <= 1e-10 * abs(float(e_ref))
    assert float(((-gp.cpu() - f_ref) ** 2).mean().sqrt()) < 1e-8
    assert float((gb.cpu() - gb_ref).abs().max()) <= 1e-8 * float(gb_ref.abs().max())
    assert abs(float(mod.real_energy) - float(orc.real)) < 1e-9 * abs(float(orc.real))
    assert abs(float(mod.reciprocal_energy) - float(orc.recip)) < 1e-9 * abs(float(orc.recip))


BIG_TRICLINIC = [[24.0, 0.0, 0.0], [3.0, 23.0, 0.0], [-2.0, 1.5, 25.0]]


@pytest.mark.parametrize("precision,etol,ftol", [("float64", 1e-10, 1e-8),
                                                 ("float32", 1e-6, 1e-4)])
@pytest.mark.parametrize("box,kmesh,n", [(np.eye(3) * 16.0, [32, 32, 32], 400),
                                         (TRICLINIC, [30, 32, 36], 400),
                                         (np.eye(3) * 24.0, [32, 30, 27], 1200),
                                         (BIG_TRICLINIC, [25, 24, 40], 1200)])
def test_fused_energy_forces(dev, precision, etol, ftol, box, kmesh, n):
    """Small boxes use one cell with periodic images; the 24 A boxes use the
    cell grid (nc >= 4) and the LDS-tiled spread."""
    pos, q, box = _system(n, box, 10)
    rc = 5.5
    e_ref, f_ref, gb_ref, orc = _oracle_ef(pos, box, q, rc, kmesh)
    mod = CoulombForceModule(rc, 1e-5, kmesh=kmesh, precision=precision)
    e, f, gb = mod.energy_forces(pos.to(dev), box.to(dev), q.to(dev))
    assert abs(float(e) - float(e_ref)) <= etol * abs(float(e_ref))
    assert float(((f.cpu() - f_ref) ** 2).mean().sqrt()) < ftol
    assert float((gb.cpu() - gb_ref).abs().max()) <= 100 * etol * float(gb_ref.abs().max())


def test_fused_two_streams_in_flight(dev):
    """Without a workspace argument the fused call keeps one workspace per
    stream: calls in flight on two streams at once each give their own result
    (float64; 1e-12, the order of the gradient atomics aside)."""
    pos_a, q_a, box = _system(1200, np.eye(3) * 24.0, 21)
    pos_b, q_b, _ = _system(1200, np.eye(3) * 24.0, 22)
    args = (5.5, 0.35, [32, 30, 27], "float64")
    pa, qa, pb, qb = (t.to(dev) for t in (pos_a, q_a, pos_b, q_b))
    ref_a = ops.pme_fused(pa, box, qa, *args)
    ref_b = ops.pme_fused(pb, box, qb, *args)
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(dev), torch.cuda.Stream(dev)
    outs = []
    for _ in range(3):
        with torch.cuda.stream(s1):
            r1 = ops.pme_fused(pa, box, qa, *args)
        with torch.cuda.stream(s2):
            r2 = ops.pme_fused(pb, box, qb, *args)
        outs.append((r1, r2))
    torch.cuda.synchronize()
    for r1, r2 in outs:
        for got, want in list(zip(r1, ref_a)) + list(zip(r2, ref_b)):
            assert torch.allclose(got, want, rtol=1e-12, atol=1e-12)


def test_module_batched_frames(dev):
    frames = [_system(120, np.eye(3) * 12.0, s) for s in (11, 12, 13)]
    pos = torch.stack([f[0] for f in frames])
    q = torch.stack([f[1] for f in frames])
    box = torch.stack([f[2] for f in frames])
    rc = 5.0
    # same pair count per frame is required by the batched layout: pad with zero-scale rows
    plist = [nblist_ref.nblist_pbc(pos[i], box[i], rc) for i in range(3)]
    npmax = max(p.shape[0] for p in plist)
    pairs = torch.zeros(3, npmax, 2, dtype=torch.long)
    bs = torch.zeros(3, npmax, dtype=F64)
    for i, p in enumerate(plist):
        pairs[i, :p.shape[0]] = p
        bs[i, :p.shape[0]] = nblist_ref.buffer_scales(p).to(F64)
    ds = torch.stack([nblist_ref.pairs_ds(pos[i], pairs[i], box[i]) for i in range(3)])
    orc = pme_ref.CoulombOracle(rc, 1e-5, kmesh=20)
    e_ref = orc.energy(pos, box, pairs, ds, bs, q)
    mod = CoulombForceModule(rc, 1e-5, kmesh=20)
    e = mod(pos.to(dev), box.to(dev), pairs.to(dev), ds.to(dev), bs.to(dev), {"charge": q.to(dev)})
    assert torch.allclose(e.cpu(), e_ref, rtol=1e-10, atol=0)
    assert mod._kmesh.tolist() == [20, 20, 20]


def test_kmesh_rule_and_obc(dev):
    pos, q, box = _system(100, np.eye(3) * 10.0, 14)
    rc = 4.0
    nbl = TorchNeighborList(rc)
    p = pos.to(dev).requires_grad_(True)
    pairs = nbl(p, None)
    mod = CoulombForceModule(rc + 1.0, 1e-5)
    e = mod(p, None, pairs, nbl.get_ds(), nbl.get_buffer_scales(), {"charge": q.to(dev)})
    f = -calc_grads(e, p)
    pr = nblist_ref.nblist_obc(pos, rc)
    pp = pos.clone().requires_grad_(True)
    ds = nblist_ref.pairs_ds(pp, pr, None)
    e_ref = pme_ref.obc_energy(q[None], pr[None], ds[None], nblist_ref.buffer_scales(pr)[None].to(F64))
    f_ref = -torch.autograd.grad(e_ref.sum(), pp)[0]
    assert abs(float(e) - float(e_ref)) < 1e-10 * abs(float(e_ref))
    assert torch.allclose(f.cpu(), f_ref, atol=1e-9)


def test_ewald_self_terms_operator(dev):
    """The fused self and non-neutral terms (two frames, non-neutral charges,
    triclinic boxes) against the reference's torch expressions
    (src/torch_admp/pme.py:448-493): energies, the backward kernel (no graph)
    and the graph-building backward (second derivatives in q and the box)."""
    rng = np.random.default_rng(9)
    q = torch.tensor(rng.uniform(-1, 1, (2, 50)) + 0.3, device=dev)
    box = torch.tensor(np.stack([np.array(TRICLINIC), np.array(SYMMETRIC) * 1.1]), device=dev)
    kappa, eps = 0.37, 14.399645
    w = torch.tensor([0.7, -1.3], device=dev, dtype=F64)

    def ref(qq, bb):
        e_self = -torch.sum(kappa / math.sqrt(math.pi) * qq ** 2, dim=-1) * eps
        e_nn = -math.pi / (2 * torch.det(bb) * kappa ** 2) * eps * torch.sum(qq, dim=-1) ** 2
        return e_self, e_nn

    qa, ba = q.clone().requires_grad_(True), box.clone().requires_grad_(True)
    qb, bb = q.clone().requires_grad_(True), box.clone().requires_grad_(True)
    es, en = torch.ops.admp.ewald_self_terms(qa, ba, kappa, eps)
    rs, rn = ref(qb, bb)
    assert torch.allclose(es, rs, rtol=1e-13, atol=0) and torch.allclose(en, rn, rtol=1e-13, atol=0)
    ga = torch.autograd.grad((es * w).sum() + (en * w * 2).sum(), [qa, ba])
    gr = torch.autograd.grad((rs * w).sum() + (rn * w * 2).sum(), [qb, bb])
    for a, b in zip(ga, gr):
        assert torch.allclose(a, b, rtol=1e-12, atol=1e-12 * float(b.abs().max()))
    # second derivatives through the graph-building backward
    es, en = torch.ops.admp.ewald_self_terms(qa, ba, kappa, eps)
    rs, rn = ref(qb, bb)
    (gqa,) = torch.autograd.grad((es + en).sum(), qa, create_graph=True)
    (gqb,) = torch.autograd.grad((rs + rn).sum(), qb, create_graph=True)
    v = torch.tensor(rng.normal(size=(2, 50)), device=dev)
    ha = torch.autograd.grad((gqa * v).sum(), [qa, ba])
    hb = torch.autograd.grad((gqb * v).sum(), [qb, bb])
    for a, b in zip(ha, hb):
        assert torch.allclose(a, b, rtol=1e-12, atol=1e-12 * float(b.abs().max()))
