"""GPU parity of the QEq Hessian product and the projected-CG solve against
the float64 oracle (KKT solution).  Tolerances: matvec 1e-10 relative in
float64; solve to |P grad E|/N < 1e-9 giving charges within 1e-7 and energy
within 1e-9 relative (tighter than the reference's own matinv-vs-pgrad test,
tests/test_qeq_numerical.py:209-329, which uses 1e-6 / 5e-6)."""

import numpy as np
import pytest
import torch

from oracle import qeq_ref
from test_oracle_qeq import qeq_system
from torch_admp_amd.qeq import QEqForceModule, matinv_optimize, pgrad_optimize
from torch_admp_amd.utils import calc_grads

pytestmark = pytest.mark.gpu
F64 = torch.float64


def _to(dev, *xs):
    return [None if x is None else x.to(dev) for x in xs]


@pytest.mark.parametrize("box,slab", [(True, False), (False, False), (True, True)])
def test_hessian_matvec(dev, box, slab):
    pos, b, chi, hard, eta, pairs, ds, bs = qeq_system(n=60, box=box)
    orc = qeq_ref.QEqOracle(4.5, kmesh=20, slab_corr=slab)
    H = orc.hessian(pos, b, chi, hard, eta, pairs, ds, bs)
    mod = QEqForceModule(4.5, kmesh=20, slab_corr=slab)
    v = torch.tensor(np.random.default_rng(1).normal(size=pos.shape[0]))
    hv = mod.hessian_matvec(*_to(dev, v, pos, b, chi, hard, eta, pairs, ds, bs)).cpu()
    ref = H @ v
    assert float((hv - ref).abs().max()) <= 1e-10 * float(ref.abs().max())


@pytest.mark.parametrize("precision", ["float64", "float32"])
def test_pgrad_solve_matches_kkt(dev, precision):
    pos, b, chi, hard, eta, pairs, ds, bs = qeq_system(n=80, L=11.0, seed=2)
    n = pos.shape[0]
    orc = qeq_ref.QEqOracle(4.5, kmesh=24)
    H = orc.hessian(pos, b, chi, hard, eta, pairs, ds, bs)
    A = torch.ones(1, n, dtype=F64)
    bv = torch.zeros(1, dtype=F64)
    q_ref, _, _ = qeq_ref.kkt_solve(H, chi, A, bv)
    e_ref = float(orc.func_energy(q_ref, pos, b, chi, hard, eta, pairs, ds, bs))
    tol = 1e-9 if precision == "float64" else 1e-6
    mod = QEqForceModule(4.5, kmesh=24, eps=tol, precision=precision)
    p = pos.to(dev).requires_grad_(True)
    args = _to(dev, b, chi, hard, eta, pairs, ds, bs, A, bv)
    e, q = pgrad_optimize(mod, torch.zeros(n, dtype=F64, device=dev), p, *args, reinit_q=True)
    assert mod.converge_iter >= 0 and mod.residual < tol
    qtol, etol = (1e-7, 1e-9) if precision == "float64" else (2e-5, 1e-6)
    assert float((q.detach().cpu() - q_ref).abs().max()) < qtol
    assert abs(float(e.detach()) - e_ref) < etol * abs(e_ref)
    f = -calc_grads(e, p)
    assert torch.isfinite(f).all()


@pytest.mark.parametrize("kmesh", [[7, 11, 13], [6, 9, 21], [17, 18, 33], [12, 18, 30]])
def test_pcg_solve_odd_meshes(dev, kmesh):
    """The solver spreads each iteration by mesh point over atoms binned by
    stencil base (4-point bins, work-groups of 16 x 8 x 16 points): meshes
    with partial bins and blocks, wrap-around within one block and the
    smallest mesh it accepts (6) against the oracle's KKT solution, with a
    nonzero total charge.  Even K1 and K2 ([12, 18, 30]: radices 4 3, 2 3 3,
    2 3 5) take the one-work-group-per-x-slab z and y transforms; odd ones
    the line passes."""
    pos, b, chi, hard, eta, pairs, ds, bs = qeq_system(n=70, L=10.5, seed=4)
    n = pos.shape[0]
    orc = qeq_ref.QEqOracle(4.5, kmesh=kmesh)
    H = orc.hessian(pos, b, chi, hard, eta, pairs, ds, bs)
    A = torch.ones(1, n, dtype=F64)
    bv = torch.full((1,), 1.5, dtype=F64)
    q_ref, _, _ = qeq_ref.kkt_solve(H, chi, A, bv)
    mod = QEqForceModule(4.5, kmesh=kmesh)
    with torch.no_grad():
        q = mod.solve_implicit(*_to(dev, pos, b, chi, hard, eta, pairs, ds, bs, A, bv), tol=1e-11)
    assert mod.converge_iter > 0 and mod.residual < 1e-11
    assert float((q.cpu() - q_ref).abs().max()) < 1e-8
    # the single Hessian product (atomic spread) agrees with the oracle on the same mesh
    v = torch.tensor(np.random.default_rng(2).normal(size=n))
    hv = mod.hessian_matvec(*_to(dev, v, pos, b, chi, hard, eta, pairs, ds, bs)).cpu()
    ref = H @ v
    assert float((hv - ref).abs().max()) <= 1e-10 * float(ref.abs().max())


def test_full_size_solve_is_optimal_under_the_other_product(dev):
    """3000 atoms of 3 species, 30 A cube, rc 6 A, mesh 120^3 (the timed QEq
    workload): the solver's charges (row-list pair product, point spread,
    x-slab transforms, one-work-group vector step) satisfy the optimality
    condition evaluated by the single Hessian product, which takes the other
    code paths (pair atomics, atomic spread): |P (H q + chi)| / n within 10x
    the solve tolerance, and sum q = 0."""
    from torch_admp_amd.nblist import TorchNeighborList
    from torch_admp_amd.workloads import qeq_species_box

    pos, box, chi, hard, eta = [torch.tensor(a, device=dev)
                                for a in qeq_species_box(3000, 30.0, seed=3)]
    n = pos.shape[0]
    nbl = TorchNeighborList(6.0)
    mod = QEqForceModule(6.0, 1e-5, kmesh=[120, 120, 120])
    A = torch.ones(1, n, dtype=F64, device=dev)
    b = torch.zeros(1, dtype=F64, device=dev)
    tol = 1e-10 / n
    with torch.no_grad():
        pairs = nbl(pos, box)
        ds, bs = nbl.get_ds(), nbl.get_buffer_scales()
        q = mod.solve_implicit(pos, box, chi, hard, eta, pairs, ds, bs, A, b, tol=tol)
        g = mod.hessian_matvec(q, pos, box, chi, hard, eta, pairs, ds, bs) + chi
    assert mod.converge_iter > 0
    pg = g - g.mean()
    assert float(pg.norm()) / n < 10 * tol
    assert abs(float(q.sum())) < 1e-10


def test_matrix_inversion_and_forces_match_oracle(dev):
    pos, b, chi, hard, eta, pairs, ds, bs = qeq_system(n=50, seed=4)
    n = pos.shape[0]
    orc = qeq_ref.QEqOracle(4.5, kmesh=20)
    H = orc.hessian(pos, b, chi, hard, eta, pairs, ds, bs)
    A = torch.ones(1, n, dtype=F64)
    bv = torch.zeros(1, dtype=F64)
    q_ref, _, fermi_ref = qeq_ref.kkt_solve(H, chi, A, bv)
    # oracle forces at fixed q*
    pp = pos.clone().requires_grad_(True)
    from oracle import nblist_ref
    ds_p = nblist_ref.pairs_ds(pp, pairs, b)
    e_ref = orc.func_energy(q_ref, pp, b, chi, hard, eta, pairs, ds_p, bs)
    f_ref = -torch.autograd.grad(e_ref.sum(), pp)[0]

    from torch_admp_amd.nblist import TorchNeighborList
    p = pos.to(dev).requires_grad_(True)
    nbl = TorchNeighborList(4.5)
    pr = nbl(p, b.to(dev))
    mod = QEqForceModule(4.5, kmesh=20)
    args = (p, *_to(dev, b, chi, hard, eta), pr, nbl.get_ds(), nbl.get_buffer_scales(),
            *_to(dev, A, bv))
    _, q4, hdiag, fermi = mod.solve_matrix_inversion(*args)
    assert torch.allclose(hdiag.cpu(), torch.diag(H), rtol=1e-9, atol=1e-9)
    assert float((q4.detach().cpu() - q_ref).abs().max()) < 1e-9
    assert abs(float(fermi) - float(fermi_ref)) < 1e-8
    # the module-level form returns (energy, q_opt) like the reference's
    e, q = matinv_optimize(mod, *args)
    assert q.requires_grad and q.is_leaf
    assert float((q.detach().cpu() - q_ref).abs().max()) < 1e-9
    assert abs(float(e.detach()) - float(e_ref)) < 1e-10 * abs(float(e_ref))
    f = -calc_grads(e, p)
    assert float((f.cpu() - f_ref).abs().max()) < 1e-8


def test_damping_gradients_match_oracle(dev):
    """dE_damp/d(q, ds, eta) from the pair kernel's backward against autograd of
    the oracle damping energy (1e-10 relative)."""
    from torch_admp_amd.qeq import GaussianDampingForceModule

    pos, b, chi, hard, eta, pairs, ds, bs = qeq_system(n=50, seed=7)
    q = torch.tensor(np.random.default_rng(5).normal(size=pos.shape[0]))
    qo, dso, etao = (t.clone().requires_grad_(True) for t in (q, ds, eta))
    e_ref = qeq_ref.damping_energy(qo[None], etao[None], pairs[None], dso[None],
                                   bs.to(F64)[None]).sum()
    g_ref = torch.autograd.grad(e_ref, (qo, dso, etao))
    qg, dsg, etag = (t.to(dev).requires_grad_(True) for t in (q, ds, eta))
    mod = GaussianDampingForceModule()
    e = mod(pos.to(dev), b.to(dev), pairs.to(dev), dsg, bs.to(dev), {"charge": qg, "eta": etag})
    g = torch.autograd.grad(e.sum(), (qg, dsg, etag))
    assert abs(float(e.sum()) - float(e_ref)) <= 1e-12 * abs(float(e_ref))
    for a, r in zip(g, g_ref):
        assert float((a.cpu() - r).abs().max()) <= 1e-10 * float(r.abs().max())


def test_dependent_constraint_rows_raise(dev):
    """A duplicated total-charge row makes A A^T singular; the reference's
    vector_projection_coeff_matrix (torch.inverse, src/torch_admp/utils.py:151-171)
    raises there, and so does the solve (ADMP_ERR_INVALID -> AdmpError)."""
    from torch_admp_amd._lib import AdmpError

    pos, b, chi, hard, eta, pairs, ds, bs = qeq_system(n=40, seed=3)
    n = pos.shape[0]
    A = torch.ones(2, n, dtype=F64)
    bv = torch.zeros(2, dtype=F64)
    mod = QEqForceModule(4.5, kmesh=16)
    with pytest.raises(AdmpError, match="invalid"):
        with torch.no_grad():
            mod.solve_implicit(*_to(dev, pos, b, chi, hard, eta, pairs, ds, bs, A, bv), tol=1e-9)
    # independent rows still solve: a species-group row beside the total charge
    A2 = torch.ones(2, n, dtype=F64)
    A2[1, n // 2:] = 0.0
    bv2 = torch.tensor([0.0, 0.5], dtype=F64)
    with torch.no_grad():
        q = mod.solve_implicit(*_to(dev, pos, b, chi, hard, eta, pairs, ds, bs, A2, bv2),
                               tol=1e-11)
    assert float((A2 @ q.cpu() - bv2).abs().max()) < 1e-10


def test_large_solve_multi_kernel_path_is_optimal(dev):
    """20000 atoms (above the 16384-atom limit of the one-work-group vector
    step): the solve takes the multi-kernel iteration (dot products with
    last-block sums, separate step, preconditioner, projection and direction
    kernels, blocking residual reads).  Its charges satisfy the optimality
    condition under the single Hessian product, |P (H q + chi)| / n within 10x
    the tolerance, and sum q = 0."""
    from torch_admp_amd.nblist import TorchNeighborList
    from torch_admp_amd.workloads import qeq_species_box

    n = 20000
    length = (n / 0.1) ** (1.0 / 3.0)
    pos, box, chi, hard, eta = [torch.tensor(a, device=dev)
                                for a in qeq_species_box(n, length, seed=6)]
    nbl = TorchNeighborList(6.0)
    mod = QEqForceModule(6.0, 1e-5, kmesh=[60, 60, 60])
    A = torch.ones(1, n, dtype=F64, device=dev)
    b = torch.zeros(1, dtype=F64, device=dev)
    tol = 1e-13
    with torch.no_grad():
        pairs = nbl(pos, box)
        ds, bs = nbl.get_ds(), nbl.get_buffer_scales()
        q = mod.solve_implicit(pos, box, chi, hard, eta, pairs, ds, bs, A, b, tol=tol)
        g = mod.hessian_matvec(q, pos, box, chi, hard, eta, pairs, ds, bs) + chi
    assert mod.converge_iter > 0
    pg = g - g.mean()
    assert float(pg.norm()) / n < 10 * tol
    assert abs(float(q.sum())) < 1e-9
