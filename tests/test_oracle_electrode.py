"""Pins the electrode oracle by the physics of CONP and CONQ.

At the solved charges the derivative of the whole-system energy with respect
to each electrode charge must be the same on every atom of a group (the
Lagrange multiplier of that group's constraint): for CONP that means the
electrostatic potential minus the applied potential is one constant over all
electrode atoms (with `symm`), for CONQ the group charges equal their targets
and the potential is constant within each group.  Tolerances 1e-8 relative.
"""

import numpy as np
import pytest
import torch

from oracle import electrode_ref, nblist_ref, qeq_ref

F64 = torch.float64


def electrode_system(seed=0, n_ions=24, edge=9.0, lz=30.0):
    """Two 3x3x2 square electrodes at the ends of a 9 x 9 x 30 A cell and
    random +-1 ions between them (min separation 2 A)."""
    rng = np.random.default_rng(seed)
    a = edge / 3
    g = np.stack(np.meshgrid(np.arange(3), np.arange(3), np.arange(2), indexing="ij"), -1)
    layer = g.reshape(-1, 3) * np.array([a, a, 2.0]) + np.array([a / 2, a / 2, 1.0])
    left = layer.copy()
    right = layer.copy()
    right[:, 2] += lz - 5.0
    ions = []
    while len(ions) < n_ions:
        p = rng.random(3) * np.array([edge, edge, lz - 14.0]) + np.array([0, 0, 7.0])
        if all(np.linalg.norm(p - o) > 2.0 for o in ions):
            ions.append(p)
    pos = np.concatenate([left, right, np.array(ions)])
    q = np.r_[np.zeros(36), np.tile([1.0, -1.0], n_ions // 2)]
    box = np.diag([edge, edge, lz])
    return torch.tensor(pos), torch.tensor(box), torch.tensor(q)


def _potential(orc, q, pos, box, eta):
    pairs = nblist_ref.nblist_pbc(pos, box, orc.coulomb.rcut)
    ds = nblist_ref.pairs_ds(pos, pairs, box)
    bs = nblist_ref.buffer_scales(pairs)
    return electrode_ref.full_potential(orc, q, pos, box, eta, pairs, ds, bs)


def test_setup_from_lammps_arrays_and_errors():
    groups = [dict(indices=range(4), mode="conp", value=1.5, eta=1.6, ffield=True),
              dict(indices=range(4, 8), mode="conp", value=0.0, eta=2.0, ffield=True)]
    mask, eta, chi, hard, A, b, ffm, ffp = electrode_ref.setup_from_lammps(10, groups, symm=True)
    assert mask.tolist() == [True] * 8 + [False] * 2
    assert torch.allclose(eta[:4], torch.full((4,), np.sqrt(2) / 2 / 1.6, dtype=F64))
    assert torch.allclose(eta[4:8], torch.full((4,), np.sqrt(2) / 2 / 2.0, dtype=F64))
    assert chi[:4].tolist() == [-1.5] * 4 and chi[4:].tolist() == [0.0] * 6
    assert A.tolist() == [[1.0] * 8] and b.tolist() == [0.0]
    assert ffm.shape == (2, 10) and ffp.tolist() == [1.5, 0.0]
    conq = [dict(indices=range(3), mode="conq", value=-2.0, eta=1.6),
            dict(indices=range(3, 6), mode="conq", value=2.0, eta=1.6)]
    _, _, chi, _, A, b, ffm, _ = electrode_ref.setup_from_lammps(8, conq)
    assert A.tolist() == [[1, 1, 1, 0, 0, 0], [0, 0, 0, 1, 1, 1]] and b.tolist() == [-2.0, 2.0]
    assert ffm is None and chi.abs().max() == 0
    with pytest.raises(AttributeError, match="symm should be False"):
        electrode_ref.setup_from_lammps(8, conq, symm=True)
    with pytest.raises(AttributeError, match="ffield with conq"):
        electrode_ref.setup_from_lammps(8, [dict(conq[0], ffield=True)])
    with pytest.raises(AttributeError, match="should be 0 or 2"):
        electrode_ref.setup_from_lammps(8, [groups[0]])


@pytest.mark.parametrize("slab", [True, False])
def test_conp_equal_potential(slab):
    pos, box, q = electrode_system(seed=1)
    groups = [dict(indices=range(18), mode="conp", value=2.0, eta=1.6),
              dict(indices=range(18, 36), mode="conp", value=0.0, eta=1.6)]
    mask, eta, chi, hard, A, b, _, _ = electrode_ref.setup_from_lammps(len(q), groups, symm=True)
    orc = qeq_ref.QEqOracle(4.0, kappa=0.5, kmesh=[12, 12, 40], slab_corr=slab)
    q_e, _ = electrode_ref.charge_optimization(orc, pos, box, q, mask, eta, chi, hard, A, b)
    assert abs(float(q_e.sum())) < 1e-10
    q_full = q.clone()
    q_full[mask] = q_e
    psi = _potential(orc, q_full, pos, box, eta)
    shifted = (psi + chi)[mask]  # psi - V on the electrode atoms
    spread = float(shifted.max() - shifted.min())
    assert spread < 1e-8 * float(psi[mask].abs().max())
    # the two electrodes carry opposite, nonzero charge under a 2 V drop
    assert float(q_e[:18].sum()) * float(q_e[18:].sum()) < 0


def test_conq_group_charges_and_potential():
    pos, box, q = electrode_system(seed=2)
    groups = [dict(indices=range(18), mode="conq", value=-1.0, eta=1.6),
              dict(indices=range(18, 36), mode="conq", value=1.0, eta=1.6)]
    mask, eta, chi, hard, A, b, _, _ = electrode_ref.setup_from_lammps(len(q), groups)
    orc = qeq_ref.QEqOracle(4.0, kappa=0.5, kmesh=[12, 12, 40], slab_corr=True)
    q_e, _ = electrode_ref.charge_optimization(orc, pos, box, q, mask, eta, chi, hard, A, b)
    assert abs(float(q_e[:18].sum()) + 1.0) < 1e-10 and abs(float(q_e[18:].sum()) - 1.0) < 1e-10
    q_full = q.clone()
    q_full[mask] = q_e
    psi = _potential(orc, q_full, pos, box, eta)[mask]
    for g in (psi[:18], psi[18:]):
        assert float(g.max() - g.min()) < 1e-8 * float(psi.abs().max())


def test_conp_finite_field_equal_potential():
    pos, box, q = electrode_system(seed=3)
    groups = [dict(indices=range(18), mode="conp", value=1.0, eta=1.6, ffield=True),
              dict(indices=range(18, 36), mode="conp", value=0.0, eta=1.6, ffield=True)]
    mask, eta, chi, hard, A, b, ffm, ffp = electrode_ref.setup_from_lammps(len(q), groups,
                                                                          symm=True)
    orc = qeq_ref.QEqOracle(4.0, kappa=0.5, kmesh=[12, 12, 40])
    q_e, field = electrode_ref.charge_optimization(orc, pos, box, q, mask, eta, chi, hard, A, b,
                                                   ffm, ffp)
    assert float(field) == pytest.approx(1.0 / 30.0)  # right electrode is higher along z
    q_full = q.clone()
    q_full[mask] = q_e
    psi = _potential(orc, q_full, pos, box, eta)
    ramp, _ = electrode_ref.finite_field_add_chi(pos, box, ffm, ffp)
    shifted = (psi + chi)[mask] + ramp
    assert float(shifted.max() - shifted.min()) < 1e-8 * float(psi[mask].abs().max())


def test_conq_finite_field_self_consistent_drop():
    """CONQ with a finite field (LAMMPS ffield on): the drop dV is solved with the
    charges; on each group psi + dV r is one constant V_g and dV = V_0 - V_1."""
    pos, box, q = electrode_system(seed=6)
    groups = [dict(indices=range(18), mode="conq", value=-1.0, eta=1.6, ffield=True),
              dict(indices=range(18, 36), mode="conq", value=1.0, eta=1.6, ffield=True)]
    mask, eta, chi, hard, A, b, ffm, ffp = electrode_ref.setup_from_lammps(len(q), groups,
                                                                          conq_ffield=True)
    assert ffp is None
    orc = qeq_ref.QEqOracle(4.0, kappa=0.5, kmesh=[12, 12, 40])
    q_e, field = electrode_ref.charge_optimization(orc, pos, box, q, mask, eta, chi, hard, A, b,
                                                   ffm, ffp)
    assert abs(float(q_e[:18].sum()) + 1.0) < 1e-10 and abs(float(q_e[18:].sum()) - 1.0) < 1e-10
    q_full = q.clone()
    q_full[mask] = q_e
    psi = _potential(orc, q_full, pos, box, eta)
    ramp, unit = electrode_ref.finite_field_add_chi(pos, box, ffm,
                                                    torch.tensor([1.0, 0.0], dtype=F64))
    dv = float(field / unit)
    v = psi[mask] + dv * ramp
    scale = float(psi[mask].abs().max())
    for g in (v[:18], v[18:]):
        assert float(g.max() - g.min()) < 1e-8 * scale
    assert abs(float(v[:18].mean() - v[18:].mean()) - dv) < 1e-8 * scale
    with pytest.raises(AttributeError, match="ffield with conq"):
        electrode_ref.setup_from_lammps(len(q), groups)


@pytest.mark.parametrize("deck,qtol,ftol", [("lmp_conp_slab_3d", 5e-4, 5e-5),
                                            ("lmp_conq_interface_3d_bias", 2e-5, 5e-4)])
def test_oracle_matches_lammps_charges_and_forces(deck, qtol, ftol):
    """The oracle's solved electrode charges and forces against LAMMPS `fix
    electrode` (the deck settings of the reference's tests,
    tests/test_electrode_conp.py:108-323, tests/test_electrode_conq.py:29-223).
    The per-atom charges of the dump are a pin the reference does not assert."""
    from pathlib import Path

    from test_gpu_electrode import DECKS

    slab, _, groups, symm = DECKS[deck]
    d = np.load(Path(__file__).resolve().parent / "golden" / f"{deck}.npz")
    cell = np.diag(d["edges"]).astype(np.float64)
    if slab:
        cell[2, 2] *= 3.0
    orc = qeq_ref.QEqOracle(5.0, ethresh=1e-6, kappa=0.5, slab_corr=slab)
    ref = electrode_ref.setup_from_lammps(len(d["q"]), [dict(g, eta=1.6) for g in groups], symm,
                                          conq_ffield=True)
    _, f, q = electrode_ref.infer(orc, torch.tensor(d["positions"]), torch.tensor(cell),
                                  torch.tensor(d["q"]), *ref)
    mask = ref[0].numpy()
    np.testing.assert_allclose(q.numpy()[mask], d["q"][mask], rtol=0, atol=qtol)
    np.testing.assert_allclose(f.detach().numpy(), d["forces"], rtol=ftol, atol=ftol)
