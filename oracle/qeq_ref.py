"""Float64 CPU restatement of the reference QEq path (test oracle).

TEST INFRASTRUCTURE ONLY -- never imported by the product package.

=====================  ===========================================  ==============================
function               reference                                    pinned by (tests/)
=====================  ===========================================  ==============================
site_energy            src/torch_admp/qeq.py:132-209                closed form (quadratic in q)
damping_energy         src/torch_admp/qeq.py:33-129                 closed form per pair; the
                                                                    erfc(d/2eta_ij) - 1 limit
func_energy            src/torch_admp/qeq.py:408-458                sum of the three modules
calc_hessian           src/torch_admp/qeq.py:460-515, 1620-1670     symmetry; finite differences
                                                                    of dE/dq; E exactly quadratic
kkt_solve              src/torch_admp/qeq.py:517-607, 1530-1617     KKT residual; constrained
                                                                    optimality; invariant
                                                                    q.Jq/2 = E_coul + E_damp
hessian_graph          src/torch_admp/qeq.py:460-607                central finite differences of
                                                                    q*(chi, positions)
pgrad_norm             src/torch_admp/qeq.py:714-765,               projection P A^T = 0
                       src/torch_admp/utils.py:115-210
=====================  ===========================================  ==============================

The reference's iterative solvers (LBFGS, Polak-Ribiere with parabolic line
search, torchopt implicit differentiation) are pinned only through their fixed
point, the KKT solution above: their iteration trajectories are parity
unpinned (torch.optim.LBFGS internals, torchopt 0.7.3 absent here).
"""

from __future__ import annotations

import math

import torch

from . import pme_ref

SQRT_PI = math.sqrt(math.pi)


def site_energy(q, chi, hardness):
    """sum(chi q + hardness q^2), no 1/2 (src/torch_admp/qeq.py:200-209)."""
    return (chi * q + hardness * q ** 2).sum(-1)


def damping_energy(q, eta, pairs, ds, bs):
    """eps [sum_P -erfc(d / (2 eta_ij)) qi qj safe_inv(d, 1e-4) bs
             + sum q^2 safe_inv(eta, 1e-4) / (2 sqrt(pi))]
    with eta_ij = sqrt((eta_i^2 + eta_j^2) / 2), eta_ij == 0 -> 1e-10
    (src/torch_admp/qeq.py:106-129).  q, eta: (nf, N); pairs (nf, P, 2)."""
    ei = torch.gather(eta, 1, pairs[:, :, 0])
    ej = torch.gather(eta, 1, pairs[:, :, 1])
    eij = torch.sqrt((ei ** 2 + ej ** 2) / 2)
    eij = torch.where(eij == 0, torch.full_like(eij, 1e-10), eij)
    qi = torch.gather(q, 1, pairs[:, :, 0])
    qj = torch.gather(q, 1, pairs[:, :, 1])
    e_pair = (-torch.erfc(ds / (2 * eij)) * qi * qj * pme_ref.safe_inverse(ds, 1e-4) * bs).sum(-1)
    e_self = (pme_ref.safe_inverse(eta, 1e-4) / (2 * SQRT_PI) * q * q).sum(-1)
    return (e_pair + e_self) * pme_ref.DIELECTRIC


class QEqOracle:
    """site + coulomb (CoulombOracle) + optional damping (src/torch_admp/qeq.py:242-458)."""

    def __init__(self, rcut, ethresh=1e-5, kspace=True, rspace=True, slab_corr=False,
                 slab_axis=2, damping=True, kappa=None, kmesh=None, spacing=None,
                 convention="reference"):
        self.coulomb = pme_ref.CoulombOracle(rcut, ethresh, kspace, rspace, slab_corr, slab_axis,
                                             kappa=kappa, kmesh=kmesh, spacing=spacing,
                                             convention=convention)
        self.damping = damping

    def func_energy(self, q, pos, box, chi, hardness, eta, pairs, ds, bs):
        """Single frame: q, chi, hardness, eta (N,); pos (N, 3); pairs (P, 2)."""
        bsf = bs.to(pos.dtype)[None]
        e = site_energy(q[None], chi[None], hardness[None])
        e = e + self.coulomb.energy(pos[None], None if box is None else box[None], pairs[None],
                                    ds[None], bsf, q[None])
        if self.damping:
            e = e + damping_energy(q[None], eta[None], pairs[None], ds[None], bsf)
        return e

    def hessian(self, pos, box, chi, hardness, eta, pairs, ds, bs):
        """d2E/dq2 by double backward, row by row (src/torch_admp/qeq.py:460-515)."""
        n = pos.shape[0]
        q = torch.zeros(n, dtype=pos.dtype, requires_grad=True)
        y = self.func_energy(q, pos.detach(), box, chi, hardness, eta, pairs, ds.detach(), bs)
        (g,) = torch.autograd.grad(y.sum(), q, create_graph=True)
        rows = [torch.autograd.grad(g[i], q, retain_graph=True)[0] for i in range(n)]
        return torch.stack(rows)


def hessian_graph(oracle: "QEqOracle", pos, box, chi, hardness, eta, pairs, ds, bs):
    """d2E/dq2 kept differentiable in pos, box, hardness, eta and ds (double
    backward with create_graph), for gradients through the dense KKT solve as
    the reference takes them (src/torch_admp/qeq.py:460-607)."""
    n = pos.shape[0]
    q = torch.zeros(n, dtype=pos.dtype, requires_grad=True)
    y = oracle.func_energy(q, pos, box, chi, hardness, eta, pairs, ds, bs)
    (g,) = torch.autograd.grad(y.sum(), q, create_graph=True)
    rows = [torch.autograd.grad(g[i], q, create_graph=True)[0] for i in range(n)]
    return torch.stack(rows)


def kkt_solve(hessian, chi, constraint_matrix=None, constraint_vals=None):
    """[[H, A^T], [A, 0]] [q; lambda] = [-chi; b] (src/torch_admp/qeq.py:572-607).

    Returns q*, the multipliers and fermi = mean(chi + H q*)."""
    n = hessian.shape[0]
    if constraint_matrix is None:
        q = torch.linalg.solve(hessian, -chi)
        lam = torch.zeros(0, dtype=hessian.dtype)
    else:
        nc = constraint_matrix.shape[0]
        kkt = torch.cat([torch.cat([hessian, constraint_matrix.T], 1),
                         torch.cat([constraint_matrix,
                                    torch.zeros(nc, nc, dtype=hessian.dtype)], 1)], 0)
        sol = torch.linalg.solve(kkt, torch.cat([-chi, constraint_vals]))
        q, lam = sol[:n], sol[n:]
    fermi = torch.mean(chi + hessian @ q)
    return q, lam, fermi


def pgrad_norm(grad, constraint_matrix):
    """|P g| / N with P = I - A^T (A A^T)^-1 A (src/torch_admp/qeq.py:714-765)."""
    A = constraint_matrix
    coeff = A.T @ torch.inverse(A @ A.T)
    pg = grad - coeff @ (A @ grad)
    return float(pg.norm() / grad.shape[0])
