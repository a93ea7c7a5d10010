"""The benchmark workloads at their full sizes on the GPU, against the float64 oracle.

* Ionic box: 1000 +-1 charges (net zero) in a 20 A cube, rc 6 A, ethresh 1e-5
  (kappa 0.548224), order 6, at the reference's rule mesh (74^3) and at an
  explicit 32^3, through the fused, the batched and the drop-in module paths.
* Water box: 8000 rigid SPC-like waters (24,000 atoms) in a 62.09 A cube, rc 8 A,
  kappa 0.411168 (ethresh 1e-5), mesh 64^3: energy, forces and dE/dbox.
* Electrode cell: 2 x 800 fcc(111) electrode atoms and 2000 ions, 40 x 40 x 120
  A, rc 6 A, mesh 150 x 150 x 450, slab correction, 1 V with symm: the CONP
  optimality conditions under an independent evaluation of the potential.
* Serving stream: 36 requests of 64-8192 atoms (orthorhombic and triclinic,
  boxes below 2 rc among them) through the float64 and float32 engines.

Bars: float64 energy within 1e-10 relative of the oracle, forces within 1e-8
eV/A RMSE, dE/dbox within 1e-8 of its largest element; float32 bulk energy
within 1e-6 relative, forces within 1e-4 eV/A RMSE, dE/dbox within 1e-4.

For random +-1 ions the total energy is the small difference of large
reciprocal and self terms (a stream request with E = -7.9 eV has E_recip =
+7769 eV and E_self = -8149 eV).  float32 arithmetic controls the error
relative to those terms, so the float32 serving bars are stated per term and
against the energy scale |E_real| + |E_recip| + |E_self| + |E_nn| (1e-6), beside
the total-relative error divided by that condition number.
"""

import numpy as np
import pytest
import torch

from oracle import nblist_ref, pme_ref
from torch_admp_amd import ops, serving
from torch_admp_amd.nblist import TorchNeighborList
from torch_admp_amd.pme import CoulombForceModule
from torch_admp_amd.workloads import ionic_box, spc_water

pytestmark = pytest.mark.gpu

BARS = {"float64": (1e-10, 1e-8, 1e-8), "float32": (1e-6, 1e-4, 1e-4)}


def _check(e, f, g, ref, precision, what):
    e_ref, f_ref, g_ref = ref
    etol, ftol, gtol = BARS[precision]
    e, f = float(e), f.detach().cpu().to(torch.float64)
    rel = abs(e - float(e_ref)) / abs(float(e_ref))
    rmse = float(((f - f_ref) ** 2).mean().sqrt())
    print(f"{what} {precision}: E rel {rel:.2e}, F rmse {rmse:.2e}", flush=True)
    assert rel <= etol, what
    assert rmse <= ftol, what
    if g is not None:
        g = g.detach().cpu().reshape(3, 3)
        assert float((g - g_ref).abs().max()) <= gtol * float(g_ref.abs().max()), what


# ---------------------------------------------------------------- ionic box
RC1, ETH = 6.0, 1e-5


@pytest.fixture(scope="module")
def ionic():
    pos, q, box = (torch.tensor(a) for a in ionic_box(1000, 20.0, seed=0))
    pairs = nblist_ref.nblist_pbc(pos, box, RC1)
    bs = nblist_ref.buffer_scales(pairs)
    refs = {}
    for km in (74, 32):
        orc = pme_ref.CoulombOracle(RC1, ETH, kmesh=None if km == 74 else km)
        # the reference's mesh rule gives 74 here (= ceil(2 kappa 20 / (3 ethresh^0.2)))
        assert orc.mesh_for(box) == [km] * 3
        refs[km] = pme_ref.energy_forces(pos, box, q, pairs,
                                         lambda p, b: nblist_ref.pairs_ds(p, pairs, b), bs, orc)
    return pos, q, box, refs


@pytest.mark.parametrize("precision", ["float64", "float32"])
@pytest.mark.parametrize("km", [74, 32])
def test_ionic_box_fused(dev, ionic, precision, km):
    pos, q, box, refs = ionic
    kappa = pme_ref.kappa_from_ethresh(RC1, ETH)
    assert abs(kappa - 0.548224) < 1e-6
    e, f, g = ops.pme_fused(pos.to(dev), box, q.to(dev), RC1, kappa, [km] * 3,
                            precision=precision)
    _check(e[4], f, g, refs[km], precision, f"fused K={km}")


@pytest.mark.parametrize("precision", ["float64", "float32"])
def test_ionic_box_batched_both_meshes(dev, ionic, precision):
    """Both meshes in one batch: the rule mesh 74 = 2 x 37 takes the generic
    prime pass of the batched FFT, 32 the radix-2/4/8 passes."""
    pos, q, box, refs = ionic
    kappa = pme_ref.kappa_from_ethresh(RC1, ETH)
    p, c = pos.to(dev), q.to(dev)
    e, f, g = ops.pme_batch([p, p], [box, box], [c, c], RC1, kappa, [[74] * 3, [32] * 3],
                            precision=precision)
    for i, km in enumerate((74, 32)):
        _check(e[i, 4], f[i], g[i], refs[km], precision, f"batched K={km}")


@pytest.mark.parametrize("precision", ["float64", "float32"])
@pytest.mark.parametrize("km", [74, 32])
def test_ionic_box_dropin(dev, ionic, precision, km):
    """TorchNeighborList + CoulombForceModule (mesh from the module's rule for
    74, explicit kmesh=32) with forces and dE/dbox through autograd."""
    pos, q, box, refs = ionic
    p = pos.to(dev).requires_grad_(True)
    b = box.to(dev).requires_grad_(True)
    nbl = TorchNeighborList(RC1)
    pairs = nbl(p, b)
    mod = CoulombForceModule(RC1, ETH, kmesh=None if km == 74 else km, precision=precision)
    e = mod(p, b, pairs, nbl.get_ds(), nbl.get_buffer_scales(), {"charge": q.to(dev)})
    assert mod._kmesh.tolist() == [km] * 3
    gp, gb = torch.autograd.grad(e.sum(), (p, b))
    _check(e.detach()[0], -gp, gb, refs[km], precision, f"drop-in K={km}")


# ---------------------------------------------------------------- water box
@pytest.fixture(scope="module")
def water():
    pos, q, box = (torch.tensor(a) for a in spc_water(8000, seed=0))
    rc = 8.0
    orc = pme_ref.CoulombOracle(rc, ETH, kmesh=64)
    pairs = nblist_ref.nblist_pbc_cells(pos, box, rc)
    bs = nblist_ref.buffer_scales(pairs)
    ref = pme_ref.energy_forces(pos, box, q, pairs,
                                lambda p, b: nblist_ref.pairs_ds(p, pairs, b), bs, orc)
    return pos, q, box, rc, orc.kappa, ref


@pytest.mark.parametrize("precision", ["float64", "float32"])
def test_water_box_full_size(dev, water, precision):
    pos, q, box, rc, kappa, ref = water
    assert pos.shape == (24000, 3) and abs(kappa - 0.411168) < 1e-6
    e, f, g = ops.pme_fused(pos.to(dev), box, q.to(dev), rc, kappa, [64] * 3,
                            precision=precision)
    _check(e[4], f, g, ref, precision, "water fused")
    # the batched path on the same box
    eb, fb, gb = ops.pme_batch([pos.to(dev)], [box], [q.to(dev)], rc, kappa, [[64] * 3],
                               precision=precision)
    _check(eb[0, 4], fb[0], gb[0], ref, precision, "water batched")


# ---------------------------------------------------------------- electrode cell
@pytest.mark.parametrize("precision,eps", [("float64", 1e-10), ("float32", 1e-8)])
def test_electrode_cell_conp_optimality(dev, precision, eps):
    """The solved electrode charges satisfy CONP optimality: the whole-system
    potential minus the applied potential is one constant over all 1600
    electrode atoms, and the electrode charge is zero in total (symm).  The
    potential is evaluated by the drop-in module's forward and its charge
    gradient (pair kernel + PME reciprocal adjoint), not by the solver's
    Hessian product."""
    from torch_admp_amd.electrode import (LAMMPSElectrodeConstraint, PolarizableElectrode,
                                          infer, setup_from_lammps)
    from torch_admp_amd.workloads import electrode_cell

    pos_np, q_np, box_np = electrode_cell(seed=0)
    n = pos_np.shape[0]
    eta_lmp = 2.0 ** 0.5 / (2.0 * 0.5)
    groups = [LAMMPSElectrodeConstraint(range(800), "conp", 1.0, eta_lmp),
              LAMMPSElectrodeConstraint(range(800, 1600), "conp", 0.0, eta_lmp)]
    mask, eta, chi, hard, A, b, ffm, ffp = setup_from_lammps(n, groups, True, device=dev)
    pos = torch.tensor(pos_np, device=dev).requires_grad_(True)
    box = torch.tensor(box_np, device=dev)
    q = torch.tensor(q_np, device=dev)
    calc = PolarizableElectrode(rcut=6.0, ethresh=1e-5, kmesh=[150, 150, 450], slab_corr=True,
                                eps=eps, precision=precision)
    nbl = TorchNeighborList(6.0)
    pairs = nbl(pos, box)
    energy, forces, q_opt = infer(calc, pos, box, q, pairs, nbl.get_ds(),
                                  nbl.get_buffer_scales(), mask, eta, chi, hard, A, b, ffm, ffp)
    assert calc.converge_iter > 0 and torch.isfinite(forces).all()
    # psi_i = dE/dq_i of the whole system at the solved charges
    psi, _ = calc.calc_coulomb_potential(None, pos.detach(), box, eta, q_opt.detach(), pairs,
                                         nbl.get_ds().detach(), nbl.get_buffer_scales())
    dev_pot = (psi + chi)[mask]
    spread = float((dev_pot - dev_pot.mean()).abs().max())
    tol = 1e-6 if precision == "float64" else 1e-4
    print(f"electrode {precision}: potential spread {spread:.2e} V, "
          f"sum q_e {float(q_opt[mask].sum()):.2e}", flush=True)
    assert spread <= tol
    assert abs(float(q_opt[mask].sum())) <= 1e-8
    # 1 V across the cell pulls charge: left electrode positive, right negative
    assert float(q_opt[:800].sum()) > 1.0 and float(q_opt[800:1600].sum()) < -1.0


# ---------------------------------------------------------------- serving stream
def _stream():
    reqs = serving.synthetic_stream(33, 0.0, seed=7)
    # the largest request of the range, triclinic
    reqs += serving.synthetic_stream(1, 0.0, seed=21, n_min=8192, n_max=8192,
                                     triclinic_fraction=1.0)
    # two boxes far below 2 rc (64 atoms: an 8.6 A box at rc 6-8)
    reqs += serving.synthetic_stream(2, 0.0, seed=22, n_min=64, n_max=64)
    return reqs


def _oracle(r):
    pos, box, q = (torch.tensor(a) for a in (r.positions, r.box, r.charges))
    vol = float(torch.linalg.det(box).abs())
    heights = [vol / float(torch.linalg.cross(box[a], box[b]).norm())
               for a, b in ((1, 2), (2, 0), (0, 1))]
    if min(heights) < 2 * r.rcut:
        pairs, ds, bs = nblist_ref.multi_image_pairs(pos, box, r.rcut)
    else:
        pairs = nblist_ref.nblist_pbc_cells(pos, box, r.rcut)
        bs = nblist_ref.buffer_scales(pairs)
        ds = nblist_ref.pairs_ds(pos, pairs, box)
    orc = pme_ref.CoulombOracle(r.rcut, r.ethresh, kmesh=r.kmesh, kappa=r.kappa)
    e = orc.energy(pos[None], box[None], pairs[None], ds[None], bs[None].to(torch.float64),
                   q[None])[0]
    return float(e), float(orc.real[0]), float(orc.recip[0]), min(heights) < 2 * r.rcut


def test_serving_stream_full_range(dev):
    reqs = _stream()
    n = np.array([r.n_atoms for r in reqs])
    tric = np.array([r.box[1, 0] != 0.0 or r.box[2, 0] != 0.0 for r in reqs])
    assert len(reqs) >= 32 and n.max() == 8192 and tric.sum() >= 8 and (~tric).sum() >= 8
    served = {}
    for prec in ("float64", "float32"):
        server = serving.PMEServer(dev, hbm_budget=4 << 30, max_systems=64, streams=2,
                                   precision=prec)
        rs = [serving.PMERequest(r.positions, r.box, r.charges, r.rcut, r.ethresh)
              for r in reqs]
        stats = server.serve(rs)
        assert stats["requests"] == len(rs)
        served[prec] = rs
    r64, r32 = served["float64"], served["float32"]
    # the float64 engine against the oracle: the 8192-atom triclinic box, two
    # boxes below 2 rc and a mid-size orthorhombic one
    small = [i for i, r in enumerate(reqs) if r.n_atoms == 64]
    ortho = [i for i, r in enumerate(reqs) if not tric[i] and 1000 < r.n_atoms < 3000][:1]
    for i in [int(np.argmax(n))] + small + ortho:
        e_ref, real_ref, recip_ref, below = _oracle(reqs[i])
        e = r64[i].energy
        scale = abs(real_ref) + abs(recip_ref) + abs(e[2]) + abs(e[3])
        assert abs(e[4] - e_ref) <= 1e-10 * abs(e_ref) + 1e-12 * scale, (i, reqs[i].n_atoms)
        assert abs(e[4] - e_ref) <= r64[i].energy_error, i  # the stated bound holds
        assert abs(e[0] - real_ref) <= 1e-10 * scale and abs(e[1] - recip_ref) <= 1e-10 * scale
    # the float32 engine against the float64 engine on every request
    worst = 0.0
    for i, (a, b) in enumerate(zip(r32, r64)):
        scale = float(np.abs(b.energy[:4]).sum())
        assert abs(a.energy[1] - b.energy[1]) <= 1e-6 * abs(b.energy[1]), i   # reciprocal
        assert abs(a.energy[0] - b.energy[0]) <= 1e-6 * scale, i               # real space
        d_total = abs(a.energy[4] - b.energy[4])
        assert d_total <= 1e-6 * scale, i
        # every float32 result states a bound on its total's error, and the
        # error stays within it (the float64 engine's own error is ~1e-10 of it)
        assert d_total <= a.energy_error, (i, d_total, a.energy_error)
        worst = max(worst, d_total / scale)
        rmse = float(np.sqrt(((a.forces - b.forces) ** 2).mean()))
        assert rmse <= 1e-4, i
    print(f"serving float32: worst |dE| / scale {worst:.2e}", flush=True)
