"""Host-side electrode setup (no GPU): the product's setup_from_lammps and
finite_field_add_chi against the oracle restatement, and the reference's errors."""

import numpy as np
import pytest
import torch

from oracle import electrode_ref
from test_oracle_electrode import electrode_system
from torch_admp_amd.electrode import (LAMMPSElectrodeConstraint, finite_field_add_chi,
                                      setup_from_lammps)

GROUPS = [dict(indices=np.arange(18), mode="conp", value=1.5, eta=1.6, ffield=True),
          dict(indices=np.arange(18, 36), mode="conp", value=0.25, eta=2.0, ffield=True)]


def test_setup_matches_oracle():
    prod = setup_from_lammps(60, [LAMMPSElectrodeConstraint(**g) for g in GROUPS], True,
                             device="cpu")
    ref = electrode_ref.setup_from_lammps(60, GROUPS, True)
    for a, r in zip(prod, ref):
        assert torch.equal(a.to(r.dtype), r)
    conq = [dict(indices=np.arange(10), mode="conq", value=-3.0, eta=1.6)]
    prod = setup_from_lammps(20, [LAMMPSElectrodeConstraint(**g) for g in conq], device="cpu")
    ref = electrode_ref.setup_from_lammps(20, conq)
    for a, r in zip(prod[:6], ref[:6]):
        assert torch.equal(a.to(r.dtype), r)


def test_finite_field_matches_oracle():
    pos, box, _ = electrode_system(seed=4)
    mask = setup_from_lammps(len(pos), [LAMMPSElectrodeConstraint(**g) for g in GROUPS], True,
                             device="cpu")
    ffm, ffp = mask[6], mask[7]
    pot, field = finite_field_add_chi(pos, box, ffm, ffp)
    pot_r, field_r = electrode_ref.finite_field_add_chi(pos, box, ffm, ffp)
    assert torch.allclose(pot, pot_r, rtol=0, atol=0) and float(field) == float(field_r)


def test_reference_errors():
    conq_ff = [LAMMPSElectrodeConstraint(indices=np.arange(3), mode="conq", value=1.0, eta=1.6,
                                         ffield=True)]
    with pytest.raises(AttributeError, match="ffield with conq has not been implemented yet"):
        setup_from_lammps(6, conq_ff, device="cpu")
    # opt-in: two constant-charge ffield groups, the drop left to the solve
    two = conq_ff + [LAMMPSElectrodeConstraint(indices=np.arange(3, 6), mode="conq",
                                               value=-1.0, eta=1.6, ffield=True)]
    out = setup_from_lammps(8, two, device="cpu", conq_ffield=True)
    assert out[6].shape == (2, 8) and out[7] is None and out[4].shape == (2, 6)
    mixed = [two[0], LAMMPSElectrodeConstraint(indices=np.arange(3, 6), mode="conp", value=0.0,
                                               eta=1.6, ffield=True)]
    with pytest.raises(AttributeError, match="both be conp or both be conq"):
        setup_from_lammps(8, mixed, device="cpu", conq_ffield=True)
    conq = [LAMMPSElectrodeConstraint(indices=np.arange(3), mode="conq", value=1.0, eta=1.6)]
    with pytest.raises(AttributeError, match="symm should be False for conq"):
        setup_from_lammps(6, conq, True, device="cpu")
    with pytest.raises(AttributeError, match="number of ffield group should be 0 or 2"):
        setup_from_lammps(60, [LAMMPSElectrodeConstraint(**GROUPS[0])], device="cpu")
    with pytest.raises(AssertionError):
        LAMMPSElectrodeConstraint(indices=np.arange(3), mode="cong", value=1.0, eta=1.6)
