"""The drop-in boundary on the host (no GPU): shape and value validation,
units, Ewald parameter rules, the cutoff check, TorchScript compilation and
the absence of a CPU fallback.

Restates the reference's own boundary tests against torch_admp_amd:
* tests/test_base_force.py:41-607 -- every standardize_input_tensor case
  (valid single / batched inputs and each invalid shape), forward with a
  trivial _forward_impl, and units_dict handling (nm -> 10);
* tests/test_pme_numerical.py:352-507 -- CoulombForceModule's ValueError texts,
  get_rcut / get_sel and every setup_ewald_parameters branch;
* nblist.py:620-660 -- the check_cutoff assertion;
* tests/test_jit.py:29-114 -- that each force module scripts and saves (the
  numerical comparison of scripted and eager modules is a GPU test).
The mesh sizes of the benchmark boxes are the values worked out independently
of the rule's code (74, 110, 171 and 147 x 147 x 439).
"""

import io
import math
from typing import Dict, Optional

import numpy as np
import pytest
import torch
from scipy.special import erfc, erfcinv

from oracle import pme_ref
from torch_admp_amd.base_force import BaseForceModule
from torch_admp_amd.nblist import TorchNeighborList, check_cutoff
from torch_admp_amd.pme import CoulombForceModule, setup_ewald_parameters
from torch_admp_amd.qeq import GaussianDampingForceModule, QEqForceModule, SiteForceModule
from torch_admp_amd.utils import TorchConstants

F64 = torch.float64


class _Tester(BaseForceModule):
    def _forward_impl(self, positions: torch.Tensor, box: Optional[torch.Tensor],
                      pairs: torch.Tensor, ds: torch.Tensor, buffer_scales: torch.Tensor,
                      params: Dict[str, torch.Tensor]) -> torch.Tensor:
        return torch.tensor(0.0, device=positions.device, dtype=positions.dtype)


def _inputs(nf=None, n=10, npairs=20, seed=0):
    """Valid inputs: single system when nf is None, else nf frames."""
    rng = np.random.default_rng(seed)
    lead = () if nf is None else (nf,)
    pos = torch.tensor(rng.random(lead + (n, 3)))
    box = torch.eye(3, dtype=F64) if nf is None else torch.eye(3, dtype=F64).repeat(nf, 1, 1)
    pairs = torch.tensor(rng.integers(0, n, lead + (npairs, 2)))
    ds = torch.tensor(rng.normal(size=lead + (npairs,)))
    bs = torch.ones(lead + (npairs,), dtype=F64)
    return pos, box, pairs, ds, bs


def _bad(nf, **change):
    pos, box, pairs, ds, bs = _inputs(nf)
    args = dict(positions=pos, box=box, pairs=pairs, ds=ds, buffer_scales=bs)
    args.update(change)
    return args


def _rand(*shape):
    return torch.tensor(np.random.default_rng(1).random(shape))


# (case, frames of the valid inputs, replaced argument, message)
INVALID = [
    ("positions_2d_last", None, dict(positions=_rand(10, 2)), "positions must have shape (natoms, 3)"),
    ("positions_3d_last", 5, dict(positions=_rand(5, 10, 2)),
     "positions must have shape (nframes, natoms, 3)"),
    ("positions_ndim", None, dict(positions=_rand(10)), "positions must be 2D or 3D tensor, got 1D"),
    ("box_2d_shape", None, dict(box=torch.eye(2, dtype=F64)), "box must have shape (3, 3)"),
    ("box_3d_shape", 5, dict(box=torch.eye(2, dtype=F64).repeat(5, 1, 1)),
     "box must have shape (nframes, 3, 3)"),
    ("box_frames", 5, dict(box=torch.eye(3, dtype=F64).repeat(3, 1, 1)),
     "box is expected to have 5 frame(s), got 3"),
    ("box_no_frame_axis", 5, dict(box=torch.eye(3, dtype=F64)),
     "box must include a frame dimension when positions has 5 frames"),
    ("box_ndim", None, dict(box=torch.ones(3, dtype=F64)), "box must be 2D or 3D tensor, got 1D"),
    ("pairs_2d_last", None, dict(pairs=torch.zeros(20, 3, dtype=torch.long)),
     "pairs must have shape (n_pairs, 2)"),
    ("pairs_3d_last", 5, dict(pairs=torch.zeros(5, 20, 3, dtype=torch.long)),
     "pairs must have shape (nframes, n_pairs, 2)"),
    ("pairs_frames", 5, dict(pairs=torch.zeros(3, 20, 2, dtype=torch.long)),
     "pairs is expected to have 5 frame(s), got 3"),
    ("pairs_no_frame_axis", 5, dict(pairs=torch.zeros(20, 2, dtype=torch.long)),
     "pairs must include a frame dimension when positions has 5 frames"),
    ("pairs_ndim", None, dict(pairs=torch.zeros(20, dtype=torch.long)),
     "pairs must be 2D or 3D tensor, got 1D"),
    ("ds_frames", 5, dict(ds=_rand(3, 20)), "ds is expected to have 5 frame(s), got 3"),
    ("ds_pairs_2d", 5, dict(ds=_rand(5, 15)), "ds is expected to have 20 pairs(s), got 15"),
    ("ds_pairs_1d", None, dict(ds=_rand(15)), "ds is expected to have 20 pairs(s), got 15"),
    ("ds_no_frame_axis", 5, dict(ds=_rand(20)),
     "ds must include a frame dimension when positions has 5 frames"),
    ("ds_ndim", None, dict(ds=_rand(1, 20, 1)), "ds must be 1D or 2D tensor, got 3D"),
    ("bs_frames", 5, dict(buffer_scales=_rand(3, 20)),
     "buffer_scales is expected to have 5 frame(s), got 3"),
    ("bs_pairs_2d", 5, dict(buffer_scales=_rand(5, 15)),
     "buffer_scales is expected to have 20 pairs(s), got 15"),
    ("bs_pairs_1d", None, dict(buffer_scales=_rand(15)),
     "buffer_scales is expected to have 20 pairs(s), got 15"),
    ("bs_no_frame_axis", 5, dict(buffer_scales=_rand(20)),
     "buffer_scales must include a frame dimension when positions has 5 frames"),
    ("bs_ndim", None, dict(buffer_scales=_rand(1, 20, 1)),
     "buffer_scales must be 1D or 2D tensor, got 3D"),
]


@pytest.mark.parametrize("case,nf,change,msg", INVALID, ids=[c[0] for c in INVALID])
def test_standardize_rejects_with_reference_text(case, nf, change, msg):
    with pytest.raises(ValueError) as err:
        _Tester().standardize_input_tensor(**_bad(nf, **change))
    assert msg in str(err.value)


@pytest.mark.parametrize("nf", [None, 5])
def test_standardize_valid_and_forward(nf):
    tester = _Tester()
    pos, box, pairs, ds, bs = _inputs(nf)
    p, b, pr, d, s = tester.standardize_input_tensor(pos, box, pairs, ds, bs)
    f = 1 if nf is None else nf
    assert p.shape == (f, 10, 3) and b.shape == (f, 3, 3) and pr.shape == (f, 20, 2)
    assert d.shape == (f, 20) and s.shape == (f, 20)
    params = {"charges": torch.tensor(np.random.default_rng(2).normal(size=(f, 10)))}
    assert tester(pos, box, pairs, ds, bs, params) == torch.tensor(0.0, dtype=F64)
    # open boundaries: box None passes through
    p, b, *_ = tester.standardize_input_tensor(pos, None, pairs, ds, bs)
    assert b is None
    assert tester(pos, None, pairs, ds, bs, params) == torch.tensor(0.0, dtype=F64)


def test_units():
    """units_dict (src/torch_admp/utils.py:213-282; tests/test_base_force.py:592-607)."""
    assert _Tester(units_dict={"length": "nm"}).const_lib.length_coeff == 10.0
    assert _Tester().const_lib.length_coeff == 1.0 and _Tester().const_lib.energy_coeff == 1.0
    # 1 eV = 96.4853 kJ/mol (CODATA 2014, the ase default)
    c = TorchConstants({"energy": "kJ/mol"})
    assert abs(1.0 / c.energy_coeff - 96.485332) < 1e-5
    assert abs(1.0 / TorchConstants({"energy": "kcal/mol"}).energy_coeff - 23.060548) < 1e-5
    assert TorchConstants({"length": "Bohr"}).length_coeff == pytest.approx(0.52917721067)
    # the Coulomb constant 1/(4 pi eps0) in eV A / e^2 from scipy's SI constants,
    # as the reference derives it (src/torch_admp/utils.py:271-277)
    from scipy import constants

    eps = constants.epsilon_0 / constants.elementary_charge * constants.angstrom
    assert c.dielectric == pytest.approx(1.0 / (4.0 * math.pi * eps), rel=1e-14)
    with pytest.raises(ValueError, match="Unknown length unit"):
        TorchConstants({"length": "furlong"})


def test_coulomb_module_validation_and_getters():
    """src/torch_admp/pme.py:98-145 texts; tests/test_pme_numerical.py:382-389,468-480."""
    m = CoulombForceModule(rcut=5.0, ethresh=1e-5, sel=[10, 20])
    assert m.get_rcut() == 5.0 and m.get_sel() == [10, 20]
    for kw, msg in ((dict(rcut=-1.0), "rcut must be positive, got -1.0"),
                    (dict(rcut=5.0, ethresh=-1.0), "ethresh must be positive, got -1.0"),
                    (dict(rcut=5.0, slab_axis=3), "slab_axis must be 0/1/2, got 3"),
                    (dict(rcut=5.0, kappa=-0.1), "kappa must be positive, got -0.1"),
                    (dict(rcut=5.0, kmesh=[16, 0, 16]),
                     "kmesh values must be positive, got kmesh[1] = 0"),
                    (dict(rcut=5.0, spacing=[1.0, -1.0, 1.0]),
                     "spacing values must be positive, got spacing[1] = -1.0")):
        with pytest.raises(ValueError) as err:
            CoulombForceModule(**kw)
        assert msg in str(err.value)
    # kappa from ethresh: exp(-(kappa rc)^2) = 2 ethresh
    k = CoulombForceModule(rcut=6.0, ethresh=1e-5).kappa
    assert math.exp(-(k * 6.0) ** 2) == pytest.approx(2e-5, rel=1e-12)
    assert CoulombForceModule(rcut=6.0, kspace=False).kappa == 0.0


# Mesh sizes of the benchmark boxes from the reference's accuracy rule
# (src/torch_admp/pme.py:390-392), worked by hand: 74 = 2 x 37, 110, 171 and
# 147 x 147 x 439 (439 is prime).
WORKED = [(6.0, [20.0, 20.0, 20.0], (74, 74, 74)), (6.0, [30.0, 30.0, 30.0], (110, 110, 110)),
          (8.0, [62.0856] * 3, (171, 171, 171)), (6.0, [40.0, 40.0, 120.0], (147, 147, 439))]


@pytest.mark.parametrize("rc,edges,mesh", WORKED)
def test_mesh_rule_worked_values(rc, edges, mesh):
    box = np.diag(edges)
    kappa, kx, ky, kz = setup_ewald_parameters(rc, box, 1e-5)
    assert (kx, ky, kz) == mesh
    assert CoulombForceModule(rc, 1e-5).mesh_for_box(torch.tensor(box)) == list(mesh)
    assert pme_ref.setup_ewald_parameters(rc, box, 1e-5)[1:] == mesh
    assert pme_ref.kmesh_rule(torch.tensor(box), kappa, 1e-5) == list(mesh)


def test_setup_ewald_parameters_branches():
    """Every branch of src/torch_admp/pme.py:564-658 (tests/test_pme_numerical.py:404-507)."""
    box = np.diag([10.0, 10.0, 10.0])
    kappa, kx, ky, kz = setup_ewald_parameters(rcut=5.0, box=box, threshold=1e-5, method="openmm")
    assert kappa > 0 and min(kx, ky, kz) >= 1
    # gromacs: kappa solves erfc(kappa rc) = threshold, mesh = ceil(L / spacing)
    kappa, kx, ky, kz = setup_ewald_parameters(rcut=5.0, box=box, threshold=1e-5, spacing=1.0,
                                               method="gromacs")
    assert kappa == pytest.approx(erfcinv(1e-5) / 5.0, rel=1e-12)
    assert erfc(kappa * 5.0) == pytest.approx(1e-5, rel=1e-9)
    assert (kx, ky, kz) == (10, 10, 10)
    assert setup_ewald_parameters(rcut=5.0, box=None) == (0.1, 1, 1, 1)
    with pytest.raises(ValueError) as err:
        setup_ewald_parameters(rcut=5.0, box=box, threshold=1e-5, method="invalid")
    assert "Invalid method: invalid" in str(err.value)
    assert "Valid methods: 'openmm', 'gromacs'" in str(err.value)
    with pytest.raises(AssertionError):
        setup_ewald_parameters(rcut=5.0, box=box, threshold=1e-5, method="gromacs")
    with pytest.raises(AssertionError):
        setup_ewald_parameters(rcut=5.0, box=np.array([[10.0, 1.0, 0.0], [0.0, 10.0, 0.0],
                                                        [0.0, 0.0, 10.0]]), threshold=1e-5)
    for rc in (-1.0, 0.0):
        with pytest.raises(ValueError) as err:
            setup_ewald_parameters(rcut=rc, box=box)
        assert f"rcut must be positive, got {rc}" in str(err.value)


def test_check_cutoff():
    """rc <= half the smallest face height (nblist.py:620-660); heights of a
    triclinic box from V / |a x b|."""
    check_cutoff(torch.eye(3, dtype=F64) * 10.0, 5.0)
    with pytest.raises(AssertionError, match="larger than half the minimum height"):
        check_cutoff(torch.eye(3, dtype=F64) * 10.0, 5.1)
    box = torch.tensor([[12.0, 0.0, 0.0], [4.0, 11.0, 0.0], [-3.0, 2.0, 13.0]], dtype=F64)
    vol = float(torch.det(box))
    h = min(vol / float(torch.linalg.cross(box[a], box[b]).norm())
            for a, b in ((1, 2), (2, 0), (0, 1)))
    check_cutoff(box, h / 2 - 1e-9)
    with pytest.raises(AssertionError):
        check_cutoff(box, h / 2 + 1e-9)
    with pytest.raises(AssertionError):
        TorchNeighborList(5.1)(torch.zeros(4, 3, dtype=F64), torch.eye(3, dtype=F64) * 10.0)


@pytest.mark.parametrize("make", [
    lambda: CoulombForceModule(rcut=4.0, ethresh=1e-5),
    lambda: CoulombForceModule(rcut=4.0, ethresh=1e-5, slab_corr=True, slab_axis=0),
    lambda: CoulombForceModule(rcut=4.0, ethresh=1e-5, slab_corr=True, slab_axis=1),
    lambda: CoulombForceModule(rcut=4.0, ethresh=1e-5, slab_corr=True, slab_axis=2),
    lambda: CoulombForceModule(rcut=4.0, kmesh=20, precision="float32", convention="columns"),
    GaussianDampingForceModule, SiteForceModule,
    lambda: QEqForceModule(rcut=4.0, ethresh=1e-5),
    lambda: TorchNeighborList(4.0)])
def test_modules_script_and_save(make):
    """torch.jit.script and torch.jit.save / load of every module (the pattern
    of tests/test_jit.py:29-114)."""
    scripted = torch.jit.script(make())
    buf = io.BytesIO()
    torch.jit.save(scripted, buf)
    buf.seek(0)
    loaded = torch.jit.load(buf)
    assert hasattr(loaded, "forward")


def test_no_cpu_fallback():
    """The operators have no CPU implementation: CPU tensors raise."""
    pos, box, pairs, ds, bs = _inputs()
    q = torch.zeros(1, 10, dtype=F64)
    with pytest.raises(RuntimeError, match="ROCm GPU"):
        torch.ops.admp.pair_energy(q, ds[None], pairs[None], bs[None], None, 0, 0.3, 14.4)
    with pytest.raises(RuntimeError, match="ROCm GPU"):
        CoulombForceModule(rcut=0.4, kmesh=8)(pos, box, pairs, ds, bs, {"charge": q})
    with pytest.raises(RuntimeError, match="ROCm GPU"):
        TorchNeighborList(0.4)(pos, box)
