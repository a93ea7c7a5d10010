"""Pins the QEq oracle: exact quadratic form, Hessian symmetry and finite
differences of dE/dq, KKT optimality on the constraint set, the invariant
q.J.q / 2 = E_coul + E_damp of the reference test (tests/test_qeq_numerical.py:331-369),
and the closed-form damping energy of two charges."""

import math

import numpy as np
import pytest
import torch
from scipy.special import erfc

from oracle import nblist_ref, pme_ref, qeq_ref

F64 = torch.float64


def qeq_system(n=40, L=10.0, seed=0, rc=4.5, box=True):
    rng = np.random.default_rng(seed)
    pos = torch.tensor(rng.random((n, 3)) * L)
    boxt = torch.eye(3, dtype=F64) * L if box else None
    chi = torch.tensor(rng.normal(5.0, 1.0, n))
    hardness = torch.tensor(rng.uniform(6.0, 10.0, n)) / 2
    eta = torch.tensor(rng.uniform(0.8, 1.2, n))
    if box:
        pairs = nblist_ref.nblist_pbc(pos, boxt, rc)
        ds = nblist_ref.pairs_ds(pos, pairs, boxt)
    else:
        pairs = nblist_ref.nblist_obc(pos, rc)
        ds = nblist_ref.pairs_ds(pos, pairs, None)
    bs = nblist_ref.buffer_scales(pairs)
    return pos, boxt, chi, hardness, eta, pairs, ds, bs


@pytest.mark.parametrize("box", [True, False])
def test_energy_is_exact_quadratic_and_hessian_symmetric(box):
    pos, b, chi, hard, eta, pairs, ds, bs = qeq_system(box=box)
    orc = qeq_ref.QEqOracle(4.5, kmesh=20)
    H = orc.hessian(pos, b, chi, hard, eta, pairs, ds, bs)
    assert torch.allclose(H, H.T, atol=1e-10)
    q = torch.tensor(np.random.default_rng(3).normal(size=pos.shape[0]))
    e = float(orc.func_energy(q, pos, b, chi, hard, eta, pairs, ds, bs))
    assert abs(e - float(chi @ q + 0.5 * q @ H @ q)) < 1e-10 * abs(e)


def test_hessian_matches_finite_difference_of_gradient():
    pos, b, chi, hard, eta, pairs, ds, bs = qeq_system(n=24)
    orc = qeq_ref.QEqOracle(4.5, kmesh=18)
    H = orc.hessian(pos, b, chi, hard, eta, pairs, ds, bs)

    def grad(q):
        q = q.clone().requires_grad_(True)
        e = orc.func_energy(q, pos, b, chi, hard, eta, pairs, ds, bs)
        return torch.autograd.grad(e.sum(), q)[0]

    q0 = torch.zeros(pos.shape[0], dtype=F64)
    for j in (0, 7, 23):
        dq = torch.zeros_like(q0)
        dq[j] = 1e-4
        fd = (grad(q0 + dq) - grad(q0 - dq)) / 2e-4
        assert torch.allclose(fd, H[:, j], atol=1e-8)


def test_kkt_solution_is_constrained_minimum():
    pos, b, chi, hard, eta, pairs, ds, bs = qeq_system()
    orc = qeq_ref.QEqOracle(4.5, kmesh=20)
    H = orc.hessian(pos, b, chi, hard, eta, pairs, ds, bs)
    n = pos.shape[0]
    A = torch.ones(1, n, dtype=F64)
    bv = torch.zeros(1, dtype=F64)
    q, lam, fermi = qeq_ref.kkt_solve(H, chi, A, bv)
    assert abs(float(q.sum())) < 1e-10
    g = chi + H @ q
    assert qeq_ref.pgrad_norm(g, A) < 1e-12
    # the gradient is constant (= -lambda = fermi level) on a total-charge constraint
    assert torch.allclose(g, torch.full_like(g, float(fermi)), atol=1e-9)
    # feasible perturbations raise the energy
    rng = np.random.default_rng(5)
    e0 = float(orc.func_energy(q, pos, b, chi, hard, eta, pairs, ds, bs))
    for _ in range(3):
        d = torch.tensor(rng.normal(size=n)) * 1e-2
        d -= d.mean()
        assert float(orc.func_energy(q + d, pos, b, chi, hard, eta, pairs, ds, bs)) > e0
    # q.J.q / 2 = E_coul + E_damp with J = H - 2 diag(hardness)
    J = H - 2 * torch.diag(hard)
    e_cd = float(orc.coulomb.energy(pos[None], b[None], pairs[None], ds[None], bs[None].to(F64),
                                    q[None]))
    e_cd += float(qeq_ref.damping_energy(q[None], eta[None], pairs[None], ds[None],
                                         bs[None].to(F64)))
    assert abs(0.5 * float(q @ J @ q) - e_cd) < 1e-9 * abs(e_cd)


def test_damping_closed_form_two_charges():
    q = torch.tensor([[0.7, -0.4]], dtype=F64)
    eta = torch.tensor([[0.9, 1.3]], dtype=F64)
    pairs = torch.tensor([[[0, 1], [1, 0]]])
    ds = torch.tensor([[2.5, 2.5]], dtype=F64)
    bs = torch.tensor([[1.0, 0.0]], dtype=F64)
    e = float(qeq_ref.damping_energy(q, eta, pairs, ds, bs))
    eij = math.sqrt((0.9 ** 2 + 1.3 ** 2) / 2)
    ref = 14.399645468667817 * (-erfc(2.5 / (2 * eij)) * 0.7 * -0.4 / 2.5
                                + 0.7 ** 2 / (2 * math.sqrt(math.pi) * 0.9)
                                + 0.4 ** 2 / (2 * math.sqrt(math.pi) * 1.3))
    assert abs(e - ref) < 1e-13 * abs(ref)


def test_kkt_gradients_through_differentiable_hessian_match_finite_differences():
    """d(w.q*)/d(chi, positions) through hessian_graph + kkt_solve against central
    differences of re-solved charges (pins the oracle for the implicit backward)."""
    pos, b, chi, hard, eta, pairs, _, bs = qeq_system(n=16, L=8.0, seed=5)
    orc = qeq_ref.QEqOracle(4.0, kmesh=16)
    A = torch.ones(1, pos.shape[0], dtype=F64)
    bval = torch.zeros(1, dtype=F64)
    w = torch.tensor(np.random.default_rng(8).normal(size=pos.shape[0]))

    def loss(p, c):
        ds = nblist_ref.pairs_ds(p, pairs, b)
        H = qeq_ref.hessian_graph(orc, p, b, c, hard, eta, pairs, ds, bs)
        q, _, _ = qeq_ref.kkt_solve(H, c, A, bval)
        return (w * q).sum()

    p0 = pos.clone().requires_grad_(True)
    c0 = chi.clone().requires_grad_(True)
    gp, gc = torch.autograd.grad(loss(p0, c0), (p0, c0))
    h = 1e-5
    for i, d in ((0, 0), (5, 2), (11, 1)):
        dp = torch.zeros_like(pos)
        dp[i, d] = h
        fd = (loss(pos + dp, chi) - loss(pos - dp, chi)) / (2 * h)
        assert abs(float(fd) - float(gp[i, d])) < 1e-6 * max(1.0, abs(float(fd)))
    for i in (1, 9):
        dc = torch.zeros_like(chi)
        dc[i] = h
        fd = (loss(pos, chi + dc) - loss(pos, chi - dc)) / (2 * h)
        assert abs(float(fd) - float(gc[i])) < 1e-7 * max(1.0, abs(float(fd)))
