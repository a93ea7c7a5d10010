"""The drop-in boundary's autograd and TorchScript obligations on the GPU.

* The Hessian in the charges exactly as the reference's
  QEqForceModule.calc_hessian takes it (src/torch_admp/qeq.py:499-514: one
  gradient with create_graph=True, then one backward per component) through
  the drop-in func_energy, against the oracle's Hessian to 1e-10 relative
  (periodic, open and slab-corrected systems).
* Mixed second derivatives: L = w . dE/dq differentiated in positions, box,
  eta, chi and hardness (through ds and the reciprocal term), against the
  float64 oracle's autograd to 1e-9 relative.
* TorchScript: every force module scripted gives the eager energy and dE/dq
  (the pattern of tests/test_jit.py:29-114) before and after torch.jit.save /
  torch.jit.load.
* Differentiating forces again (a force-matching loss) raises
  NotImplementedError instead of returning a partial gradient.
"""

import numpy as np
import pytest
import torch

from oracle import nblist_ref, qeq_ref
from test_oracle_qeq import qeq_system
from torch_admp_amd.nblist import TorchNeighborList
from torch_admp_amd.pme import CoulombForceModule
from torch_admp_amd.qeq import GaussianDampingForceModule, QEqForceModule, SiteForceModule
from torch_admp_amd.utils import calc_grads

pytestmark = pytest.mark.gpu
F64 = torch.float64


def _hessian_by_double_backward(func_energy, positions, box, chi, hardness, eta, pairs, ds, bs):
    """d2E/dq2 the way the reference builds it: the gradient at q = 0 with a
    graph, then the gradient of each of its components."""
    n = positions.shape[0]
    q0 = torch.zeros(n, device=positions.device, dtype=positions.dtype, requires_grad=True)
    y = func_energy(q0, positions, box, chi, hardness, eta, pairs, ds, bs)
    (g,) = torch.autograd.grad(y, q0, retain_graph=True, create_graph=True)
    rows = [torch.autograd.grad(gi, q0, retain_graph=True)[0] for gi in g]
    return torch.stack(rows).reshape(n, n)


def _to(dev, *xs):
    return [None if x is None else x.to(dev) for x in xs]


@pytest.mark.parametrize("box,slab", [(True, False), (False, False), (True, True)])
def test_reference_hessian_loop_matches_oracle(dev, box, slab):
    pos, b, chi, hard, eta, pairs, ds, bs = qeq_system(n=48, seed=11, box=box)
    orc = qeq_ref.QEqOracle(4.5, kmesh=20, slab_corr=slab)
    H_ref = orc.hessian(pos, b, chi, hard, eta, pairs, ds, bs)
    mod = QEqForceModule(4.5, kmesh=20, slab_corr=slab)
    H = _hessian_by_double_backward(mod.func_energy, *_to(dev, pos, b, chi, hard, eta, pairs, ds,
                                                          bs)).cpu()
    assert float((H - H_ref).abs().max()) <= 1e-10 * float(H_ref.abs().max())
    # the matrix-free calc_hessian agrees too
    H2 = mod.calc_hessian(*_to(dev, pos, b, chi, hard, eta, pairs, ds, bs)).cpu()
    assert float((H2 - H_ref).abs().max()) <= 1e-10 * float(H_ref.abs().max())


def test_mixed_second_derivatives_match_oracle(dev):
    """L = w . dE/dq at random charges, differentiated in positions (through
    ds and the reciprocal term), box, eta, chi and hardness."""
    pos, b, chi, hard, eta, pairs, _, bs = qeq_system(n=40, seed=12)
    rng = np.random.default_rng(3)
    q = torch.tensor(rng.normal(size=pos.shape[0]))
    w = torch.tensor(rng.normal(size=pos.shape[0]))

    def mixed(func_energy, make_ds, tensors):
        p, bx, c, h, e = (t.clone().requires_grad_(True) for t in tensors)
        qq = q.to(p.device).clone().requires_grad_(True)
        ds = make_ds(p, bx)
        y = func_energy(qq, p, bx, c, h, e, pairs.to(p.device), ds, bs.to(p.device))
        (g,) = torch.autograd.grad(y.sum(), qq, create_graph=True)
        return torch.autograd.grad((g * w.to(p.device)).sum(), (p, bx, c, h, e))

    orc = qeq_ref.QEqOracle(4.5, kmesh=20)
    ref = mixed(orc.func_energy, lambda p, bx: nblist_ref.pairs_ds(p, pairs, bx),
                (pos, b, chi, hard, eta))
    mod = QEqForceModule(4.5, kmesh=20)
    got = mixed(mod.func_energy,
                lambda p, bx: TorchNeighborList.pairs_ds(p, pairs.to(dev), bx),
                _to(dev, pos, b, chi, hard, eta))
    for name, a, r in zip(("positions", "box", "chi", "hardness", "eta"), got, ref):
        assert float((a.cpu() - r).abs().max()) <= 1e-9 * float(r.abs().max()), name


def _jit_inputs(dev, periodic, seed=1024, n=100, L=10.0, rc=4.0):
    """The inputs of the reference's JIT test (tests/test_jit.py:40-79)."""
    rng = np.random.default_rng(seed)
    pos = torch.tensor(rng.random((n, 3)) * L, device=dev).requires_grad_(True)
    box = torch.eye(3, dtype=F64, device=dev) * L if periodic else None
    q = torch.tensor(rng.uniform(-1.0, 1.0, n), device=dev)
    q = (q - q.mean()).requires_grad_(True)
    nbl = TorchNeighborList(cutoff=rc)
    pairs = nbl(pos, box)
    params = {"charge": q, "eta": torch.ones(n, dtype=F64, device=dev),
              "chi": torch.ones(n, dtype=F64, device=dev),
              "hardness": torch.zeros(n, dtype=F64, device=dev)}
    return pos, box, pairs, nbl.get_ds(), nbl.get_buffer_scales(), params


JIT_CASES = {
    "coulomb_obc": (False, lambda: CoulombForceModule(rcut=4.0, ethresh=1e-5)),
    "coulomb_pbc": (True, lambda: CoulombForceModule(rcut=4.0, ethresh=1e-5)),
    "slab_x": (True, lambda: CoulombForceModule(rcut=4.0, ethresh=1e-5, slab_corr=True,
                                                slab_axis=0)),
    "slab_y": (True, lambda: CoulombForceModule(rcut=4.0, ethresh=1e-5, slab_corr=True,
                                                slab_axis=1)),
    "slab_z": (True, lambda: CoulombForceModule(rcut=4.0, ethresh=1e-5, slab_corr=True,
                                                slab_axis=2)),
    "damping": (True, GaussianDampingForceModule),
    "site": (True, SiteForceModule),
    "qeq": (True, lambda: QEqForceModule(rcut=4.0, ethresh=1e-5)),
}


@pytest.mark.parametrize("case", list(JIT_CASES))
def test_scripted_module_equals_eager(dev, case, tmp_path):
    periodic, make = JIT_CASES[case]
    pos, box, pairs, ds, bs, params = _jit_inputs(dev, periodic)
    module = make()
    scripted = torch.jit.script(make())
    path = tmp_path / "frozen_model.pth"
    torch.jit.save(scripted, str(path))
    loaded = torch.jit.load(str(path))
    e = module(pos, box, pairs, ds, bs, params)
    g = calc_grads(e, params["charge"])
    for m in (scripted, loaded):
        ej = m(pos, box, pairs, ds, bs, params)
        gj = calc_grads(ej, params["charge"])
        assert abs(float(ej) - float(e)) <= 1e-12 * max(1.0, abs(float(e))), case
        assert float((gj - g).abs().max()) <= 1e-12 * max(1.0, float(g.abs().max())), case


def test_force_matching_second_derivative_raises(dev):
    pos, box, pairs, ds, bs, params = _jit_inputs(dev, True)
    e = CoulombForceModule(rcut=4.0, ethresh=1e-5)(pos, box, pairs, ds, bs, params)
    (f,) = torch.autograd.grad(e.sum(), pos, create_graph=True)
    with pytest.raises(NotImplementedError, match="second derivatives"):
        torch.autograd.grad((f ** 2).sum(), params["charge"])
