"""Pins the float64 CPU oracle before it is trusted as the parity reference.

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a wrong sign or index, a transposed operand) fails it:
  * B-spline: knot values, partition of unity, symmetry, Cox-de Boor recursion,
    derivative by finite differences;
  * spread: one-atom brute-force stencil, charge conservation;
  * k points: numpy.fft.fftfreq ordering;
  * reciprocal energy: direct (mesh-free) Ewald k-sum in both k conventions;
  * Ewald split: kappa-independence of the total;
  * real space / open boundary: LAMMPS coul/cut fixture (energy and forces);
  * slab correction: LAMMPS kspace_modify slab difference fixture;
  * full periodic sum: direct Ewald (the LAMMPS pppm 3D fixture uses a 16-point
    mesh over 50 A and is not converged, so only its slab difference is used);
  * forces and dE/dbox: central finite differences;
  * neighbour lists: brute-force enumeration (set equality).
"""

import itertools
import math
from pathlib import Path

import numpy as np
import pytest
import torch

from oracle import ewald_ref, nblist_ref, pme_ref

F64 = torch.float64
GOLDEN = Path(__file__).resolve().parent / "golden"


def _system(n, box, seed, neutral=True):
    rng = np.random.default_rng(seed)
    box = torch.as_tensor(box, dtype=F64)
    pos = torch.tensor(rng.random((n, 3))) @ box
    q = torch.tensor(rng.uniform(-1, 1, n))
    if neutral:
        q -= q.mean()
    return pos, q, box


TRICLINIC = [[10.0, 0.0, 0.0], [2.0, 11.0, 0.0], [1.5, -1.0, 12.0]]


# ---------------------------------------------------------------- B-spline
def test_bspline_knot_values():
    vals = pme_ref.bspline6(torch.arange(0, 7, dtype=F64)) * 120
    assert torch.allclose(vals, torch.tensor([0, 1, 26, 66, 26, 1, 0], dtype=F64), atol=1e-9)


def test_bspline_partition_symmetry_recursion():
    f = torch.linspace(0, 0.999, 37, dtype=F64)
    tot = sum(pme_ref.bspline6(f + j) for j in range(6))
    assert torch.allclose(tot, torch.ones_like(f), atol=1e-12)
    u = torch.linspace(0.01, 5.99, 101, dtype=F64)
    assert torch.allclose(pme_ref.bspline6(u), pme_ref.bspline6(6 - u), atol=1e-11)

    assert np.allclose(pme_ref.bspline6(u).numpy(), _cox_de_boor(6, u.numpy()), atol=1e-11)


def test_bspline_derivative_fd():
    u = torch.linspace(0.1, 5.9, 59, dtype=F64).requires_grad_(True)
    g = torch.autograd.grad(pme_ref.bspline6(u).sum(), u)[0]
    h = 1e-6
    fd = (pme_ref.bspline6(u.detach() + h) - pme_ref.bspline6(u.detach() - h)) / (2 * h)
    assert torch.allclose(g, fd, atol=1e-7)


# ---------------------------------------------------------------- spread
def _cox_de_boor(n, x):
    """Cardinal B-spline of order n on the knots 0..n (zero outside [0, n))."""
    if n == 2:
        return np.where((x >= 0) & (x < 1), x, np.where((x >= 1) & (x < 2), 2 - x, 0.0))
    return (x * _cox_de_boor(n - 1, x) + (n - x) * _cox_de_boor(n - 1, x - 1)) / (n - 1)


def test_spread_single_atom_brute_force():
    """Q(g) = q prod_d M6(g_d - R_d + 3) summed over every periodic image of g."""
    box = torch.tensor(TRICLINIC, dtype=F64)
    K = [12, 14, 16]
    r = torch.tensor([[3.3, -2.1, 17.7]], dtype=F64)
    mesh = pme_ref.spread_charges(r, torch.linalg.inv(box), torch.tensor([0.7], dtype=F64), K)
    R = ((r @ torch.linalg.inv(box))[0] * torch.tensor(K, dtype=F64)).numpy()
    ref = np.zeros(K)
    for g in itertools.product(*[range(int(np.floor(R[d])) - 8, int(np.floor(R[d])) + 9)
                                 for d in range(3)]):
        w = 0.7 * np.prod([_cox_de_boor(6, np.array(g[d] - R[d] + 3.0)) for d in range(3)])
        ref[g[0] % K[0], g[1] % K[1], g[2] % K[2]] += w
    assert np.allclose(mesh.numpy(), ref, atol=1e-14)
    assert abs(float(mesh.sum()) - 0.7) < 1e-12


def test_kpts_integer_fft_order():
    for K in ([4, 5, 6], [7, 8, 9]):
        kint = pme_ref.kpts_integer(K)
        for d in range(3):
            f = np.fft.fftfreq(K[d]) * K[d]
            if K[d] % 2 == 0:
                f[K[d] // 2] = -K[d] // 2
            axis = kint[:, d].reshape(K)
            line = axis.movedim(d, 0).reshape(K[d], -1)[:, 0].numpy()
            assert np.array_equal(line, f.astype(int))


# ---------------------------------------------------------------- reciprocal
@pytest.mark.parametrize("box", [np.diag([10.0, 11.0, 12.0]), TRICLINIC,
                                 [[10.0, 0.4, 0.3], [0.4, 11.0, 0.2], [0.3, 0.2, 12.0]]])
@pytest.mark.parametrize("conv", ["reference", "columns"])
def test_spme_matches_direct_ewald(box, conv):
    pos, q, box = _system(24, box, 3)
    kappa = 0.4
    e_spme = float(pme_ref.reciprocal_energy_frame(pos, box, q, kappa, [40, 44, 48], conv))
    e_dir = ewald_ref.ewald_reciprocal(pos.numpy(), box.numpy(), q.numpy(), kappa, [14] * 3, conv)
    assert abs(e_spme - e_dir) < 3e-8 * abs(e_dir)


def test_ewald_total_kappa_independent_and_rows_defect():
    pos, q, box = _system(20, TRICLINIC, 4)
    tot = {c: [ewald_ref.ewald_total(pos.numpy(), box.numpy(), q.numpy(), k, 14.0, [16] * 3, c)
               for k in (0.35, 0.45, 0.55)] for c in ("columns", "reference")}
    assert max(tot["columns"]) - min(tot["columns"]) < 1e-9 * abs(tot["columns"][0])
    # the rows convention of the reference is not kappa independent for this box
    assert max(tot["reference"]) - min(tot["reference"]) > 1e-4 * abs(tot["reference"][0])


def test_oracle_total_matches_ewald_total():
    pos, q, box = _system(40, np.diag([11.0, 12.0, 13.0]), 5)
    rc, kappa = 5.4, 0.6
    pairs = nblist_ref.nblist_pbc(pos, box, rc)
    ds = nblist_ref.pairs_ds(pos, pairs, box)
    bs = nblist_ref.buffer_scales(pairs).to(F64)
    orc = pme_ref.CoulombOracle(rc, kappa=kappa, kmesh=[48, 48, 48])
    e = float(orc.energy(pos[None], box[None], pairs[None], ds[None], bs[None], q[None]))
    e_dir = ewald_ref.ewald_total(pos.numpy(), box.numpy(), q.numpy(), kappa, rc, [20] * 3)
    assert abs(e - e_dir) < 1e-7 * abs(e_dir)


def test_non_neutral_term_matches_ewald():
    pos, q, box = _system(30, np.diag([11.0, 12.0, 13.0]), 6, neutral=False)
    rc, kappa = 5.4, 0.6
    pairs = nblist_ref.nblist_pbc(pos, box, rc)
    ds = nblist_ref.pairs_ds(pos, pairs, box)
    bs = nblist_ref.buffer_scales(pairs).to(F64)
    orc = pme_ref.CoulombOracle(rc, kappa=kappa, kmesh=[48, 48, 48])
    e = float(orc.energy(pos[None], box[None], pairs[None], ds[None], bs[None], q[None]))
    e_dir = ewald_ref.ewald_total(pos.numpy(), box.numpy(), q.numpy(), kappa, rc, [20] * 3)
    assert abs(float(q.sum())) > 0.1
    assert abs(e - e_dir) < 1e-7 * abs(e_dir)


# ---------------------------------------------------------------- LAMMPS fixtures
def _golden(name):
    g = np.load(GOLDEN / f"{name}.npz")
    return {k: torch.tensor(g[k]) for k in g.files}


def test_obc_matches_lammps_coul_cut():
    g = _golden("lmp_coul_obc")
    pos = g["positions"].clone().requires_grad_(True)
    pairs = nblist_ref.nblist_obc(pos, 4.0)
    ds = nblist_ref.pairs_ds(pos, pairs, None)
    e = pme_ref.obc_energy(g["charges"][None], pairs[None], ds[None],
                           nblist_ref.buffer_scales(pairs)[None].to(F64))
    f = -torch.autograd.grad(e.sum(), pos)[0]
    assert abs(float(e) - float(g["energy"])) < 1e-6
    assert torch.allclose(f, g["forces"], atol=1e-6, rtol=1e-6)


def test_slab_correction_matches_lammps_difference():
    a, b = _golden("lmp_coul_pbc"), _golden("lmp_coul_pbc_slab_corr")
    pos = a["positions"].clone().requires_grad_(True)
    box = a["box"]
    e = pme_ref.slab_corr_energy(a["charges"][None], pos[None], box[None], 2)
    f = -torch.autograd.grad(e.sum(), pos)[0]
    assert abs(float(e) - float(b["energy"] - a["energy"])) < 1e-6
    assert torch.allclose(f, b["forces"] - a["forces"], atol=1e-6, rtol=1e-6)


# ---------------------------------------------------------------- finite differences
@pytest.mark.parametrize("box", [np.diag([9.0, 10.0, 11.0]), TRICLINIC])
def test_forces_and_box_gradient_fd(box):
    pos, q, box = _system(30, box, 7)
    rc = 4.4
    pairs = nblist_ref.nblist_pbc(pos, box, rc)
    bs = nblist_ref.buffer_scales(pairs)
    orc = pme_ref.CoulombOracle(rc, kappa=0.5, kmesh=[20, 22, 24])

    def energy(p, b):
        ds = nblist_ref.pairs_ds(p, pairs, b)
        return float(orc.energy(p[None], b[None], pairs[None], ds[None], bs[None].to(F64), q[None]))

    e, f, gb = pme_ref.energy_forces(pos, box, q, pairs,
                                     lambda p, b: nblist_ref.pairs_ds(p, pairs, b), bs, orc)
    h = 1e-5
    for (i, d) in [(0, 0), (3, 1), (11, 2), (29, 0)]:
        pp, pm = pos.clone(), pos.clone()
        pp[i, d] += h
        pm[i, d] -= h
        assert abs(-(energy(pp, box) - energy(pm, box)) / (2 * h) - float(f[i, d])) < 1e-6
    for (a, b) in [(0, 0), (1, 0), (2, 1), (2, 2)]:
        bp, bm = box.clone(), box.clone()
        bp[a, b] += h
        bm[a, b] -= h
        assert abs((energy(pos, bp) - energy(pos, bm)) / (2 * h) - float(gb[a, b])) < 1e-5


# ---------------------------------------------------------------- neighbour lists
@pytest.mark.parametrize("box", [np.eye(3) * 12.0, TRICLINIC])
def test_nblist_oracles_match_brute_force(box):
    pos, _, box = _system(150, box, 8)
    pos = pos + 25.0
    rc = 4.5
    a = set(map(tuple, nblist_ref.nblist_pbc(pos, box, rc).tolist()))
    b = set(map(tuple, nblist_ref.nblist_pbc_cells(pos, box, rc).tolist()))
    c, amb = nblist_ref.brute_force_pairs(pos.numpy(), box.numpy(), rc)
    assert not amb and a == b == c
    # symmetric, no self pairs
    assert all((j, i) in a for (i, j) in a) and all(i != j for (i, j) in a)


def test_check_cutoff_assert():
    with pytest.raises(AssertionError, match="larger than half the minimum height"):
        nblist_ref.check_cutoff(torch.eye(3, dtype=F64) * 10.0, 5.1)
    nblist_ref.check_cutoff(torch.eye(3, dtype=F64) * 10.0, 5.0)


def test_multi_image_pairs_energy_matches_ewald():
    """A box smaller than 2 rc: the multi-image pair list in the reference's
    dp_nblist convention gives the direct Ewald real-space sum."""
    pos, q, box = _system(6, np.diag([5.0, 5.5, 6.0]), 9)
    rc, kappa = 7.0, 0.45
    pairs, ds, bs = nblist_ref.multi_image_pairs(pos, box, rc)
    e = float(pme_ref.real_energy(q[None], pairs[None], ds[None], bs[None], kappa))
    e_dir = ewald_ref.ewald_real(pos.numpy(), box.numpy(), q.numpy(), kappa, rc)
    assert abs(e - e_dir) < 1e-12 * abs(e_dir) + 1e-12


def test_setup_ewald_parameters_rules():
    k, kx, ky, kz = pme_ref.setup_ewald_parameters(6.0, np.diag([20.0, 30.0, 40.0]), 1e-5)
    assert abs(k - math.sqrt(-math.log(2e-5)) / 6.0) < 1e-14
    assert (kx, ky, kz) == tuple(int(math.ceil(2 * k * L / (3 * 1e-5 ** 0.2))) for L in (20, 30, 40))
    k2, *_ = pme_ref.setup_ewald_parameters(6.0, np.diag([20.0, 30.0, 40.0]), 1e-5, 1.0, "gromacs")
    from scipy.special import erfc

    assert abs(erfc(k2 * 6.0) - 1e-5) < 1e-9
    with pytest.raises(ValueError):
        pme_ref.setup_ewald_parameters(6.0, np.eye(3) * 20, 1e-5, method="bad")
