"""Implicit-function backward through the QEq charge solve.

The loss L = w . q*(positions, box, chi, hardness, eta) of the equilibrated
charges is differentiated (1) by QEqForceModule.solve_implicit (adjoint CG
solve + the library's Coulomb/damping backward kernels) and (2) by the oracle,
autograd through the dense KKT solve with a differentiable Hessian (the
reference's route, src/torch_admp/qeq.py:460-607).  Float64, 1e-7 relative to
the largest gradient entry; the oracle itself is pinned by finite differences
in test_oracle_qeq.py.  A central finite difference of the GPU solve checks the
position gradient end to end.
"""

import numpy as np
import pytest
import torch

from oracle import nblist_ref, qeq_ref
from test_oracle_qeq import qeq_system
from torch_admp_amd.nblist import TorchNeighborList
from torch_admp_amd.qeq import QEqForceModule

pytestmark = pytest.mark.gpu
F64 = torch.float64
RC, KM = 4.5, 20


def _oracle_grads(pos, b, chi, hard, eta, pairs, bs, A, bval, w, box_grad):
    orc = qeq_ref.QEqOracle(RC, kmesh=KM)
    p = pos.clone().requires_grad_(True)
    c = chi.clone().requires_grad_(True)
    h = hard.clone().requires_grad_(True)
    e = eta.clone().requires_grad_(True)
    bx = b.clone().requires_grad_(box_grad)
    ds = nblist_ref.pairs_ds(p, pairs, bx)
    H = qeq_ref.hessian_graph(orc, p, bx, c, h, e, pairs, ds, bs)
    q, _, _ = qeq_ref.kkt_solve(H, c, A, bval)
    inputs = [p, c, h, e] + ([bx] if box_grad else [])
    return q.detach(), torch.autograd.grad((w * q).sum(), inputs)


def _gpu(dev, pos, b, chi, hard, eta, A, bval, w, box_grad):
    mod = QEqForceModule(RC, kmesh=KM)
    p = pos.to(dev).requires_grad_(True)
    c = chi.to(dev).requires_grad_(True)
    h = hard.to(dev).requires_grad_(True)
    e = eta.to(dev).requires_grad_(True)
    bx = b.to(dev).requires_grad_(box_grad)
    nbl = TorchNeighborList(RC)
    pairs = nbl(p, bx)
    ds = nbl.get_ds()
    q = mod.solve_implicit(p, bx, c, h, e, pairs, ds, nbl.get_buffer_scales(), A.to(dev),
                           bval.to(dev), tol=1e-11)
    inputs = [p, c, h, e] + ([bx] if box_grad else [])
    return q.detach().cpu(), [g.cpu() for g in torch.autograd.grad((w.to(dev) * q).sum(), inputs)]


@pytest.mark.parametrize("box_grad", [False, True])
def test_implicit_gradients_match_oracle(dev, box_grad):
    pos, b, chi, hard, eta, pairs, _, bs = qeq_system(n=40, L=10.0, seed=6, rc=RC)
    n = pos.shape[0]
    A = torch.ones(1, n, dtype=F64)
    bval = torch.zeros(1, dtype=F64)
    w = torch.tensor(np.random.default_rng(2).normal(size=n))
    q_ref, g_ref = _oracle_grads(pos, b, chi, hard, eta, pairs, bs, A, bval, w, box_grad)
    q, g = _gpu(dev, pos, b, chi, hard, eta, A, bval, w, box_grad)
    assert float((q - q_ref).abs().max()) < 1e-8
    for name, a, r in zip(["positions", "chi", "hardness", "eta", "box"], g, g_ref):
        scale = float(r.abs().max())
        assert float((a - r).abs().max()) <= 1e-7 * scale, (name, float((a - r).abs().max()), scale)


def test_implicit_constraint_gradient_and_fd(dev):
    """dL/db (the total-charge constraint value) equals the adjoint multiplier,
    checked by a finite difference of the GPU solve; so is one position entry."""
    pos, b, chi, hard, eta, _, _, _ = qeq_system(n=30, L=9.0, seed=9, rc=RC)
    n = pos.shape[0]
    A = torch.ones(1, n, dtype=F64, device=dev)
    w = torch.tensor(np.random.default_rng(4).normal(size=n), device=dev)
    mod = QEqForceModule(RC, kmesh=KM)
    nbl = TorchNeighborList(RC)

    def loss(p, bv):
        pairs = nbl(p, b.to(dev))
        q = mod.solve_implicit(p, b.to(dev), chi.to(dev), hard.to(dev), eta.to(dev), pairs,
                               nbl.get_ds(), nbl.get_buffer_scales(), A, bv, tol=1e-12)
        return (w * q).sum()

    p0 = pos.to(dev).requires_grad_(True)
    bv0 = torch.zeros(1, dtype=F64, device=dev).requires_grad_(True)
    gp, gb = torch.autograd.grad(loss(p0, bv0), (p0, bv0))
    h = 1e-4
    with torch.no_grad():
        fd_b = (loss(pos.to(dev), bv0 + h) - loss(pos.to(dev), bv0 - h)) / (2 * h)
        assert abs(float(fd_b) - float(gb)) < 1e-6 * max(1.0, abs(float(fd_b)))
        dp = torch.zeros_like(p0)
        dp[3, 1] = h
        fd_p = (loss(pos.to(dev) + dp, bv0) - loss(pos.to(dev) - dp, bv0)) / (2 * h)
        assert abs(float(fd_p) - float(gp[3, 1])) < 1e-5 * max(1.0, abs(float(fd_p)))
