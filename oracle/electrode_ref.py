"""Float64 CPU restatement of the reference electrode layer (test oracle).

TEST INFRASTRUCTURE ONLY -- never imported by the product package.

=====================  ===========================================  ==============================
function               reference                                    pinned by (tests/)
=====================  ===========================================  ==============================
setup_from_lammps      src/torch_admp/electrode.py:203-285          closed form of the arrays
finite_field_add_chi   src/torch_admp/electrode.py:288-336          closed form (linear ramp)
charge_optimization    src/torch_admp/electrode.py:456-560          CONP: equal electrostatic
                                                                    potential minus applied
                                                                    potential on every electrode
                                                                    atom; CONQ: group charges and
                                                                    equal potential per group;
                                                                    the full-system KKT residual
conq + ffield          LAMMPS fix electrode/conq ffield (the        the self-consistent drop:
                       reference raises, electrode.py:235-236)      (psi + dV r) constant per
                                                                    group, equal to V_g, with
                                                                    dV = V_0 - V_1; LAMMPS deck
infer                  src/torch_admp/electrode.py:339-453          forces = -dE/dr at q*, field
                                                                    work; LAMMPS fix electrode
                                                                    forces through the GPU path
=====================  ===========================================  ==============================

The electrode problem is the QEq of the electrode atoms alone with the
electrolyte's potential added to chi; the energy being quadratic, this split is
exact, and charge_optimization is checked against the whole-system optimality
conditions instead of a literal copy of that split.
"""

from __future__ import annotations

import numpy as np
import torch

from . import nblist_ref, qeq_ref


def setup_from_lammps(n_atoms, groups, symm=False, conq_ffield=False):
    """groups: list of dicts {indices, mode ('conp'|'conq'), value, eta, chi=0,
    hardness=0, ffield=False}.  Same outputs as the reference function on the CPU;
    conq_ffield=True admits two constant-charge ffield groups (ffield potential
    None: the drop is solved)."""
    mask = np.zeros(n_atoms, dtype=bool)
    eta = np.zeros(n_atoms)
    chi = np.zeros(n_atoms)
    hard = np.zeros(n_atoms)
    rows, vals, ff_rows, ff_pot = [], [], [], []
    for g in groups:
        idx = np.asarray(g["indices"], dtype=int)
        mask[idx] = True
        eta[idx] = 1.0 / g["eta"] * np.sqrt(2.0) / 2.0
        chi[idx] = g.get("chi", 0.0)
        hard[idx] = g.get("hardness", 0.0)
        if g["mode"] == "conq":
            if symm:
                raise AttributeError("symm should be False for conq, user can implement symm "
                                     "by conq")
            if g.get("ffield", False) and not conq_ffield:
                raise AttributeError("ffield with conq has not been implemented yet")
            r = np.zeros(n_atoms)
            r[idx] = 1.0
            rows.append(r)
            vals.append(g["value"])
        if g["mode"] == "conp":
            chi[idx] -= g["value"]
        if g.get("ffield", False):
            r = np.zeros(n_atoms, dtype=bool)
            r[idx] = True
            ff_rows.append(r)
            ff_pot.append(g["value"])
    if len(ff_rows) not in (0, 2):
        raise AttributeError("number of ffield group should be 0 or 2")
    if symm:
        rows.append(np.ones(n_atoms))
        vals.append(0.0)
    A = torch.tensor(np.stack(rows)[:, mask]) if rows else torch.zeros((0, int(mask.sum())),
                                                                       dtype=torch.float64)
    b = torch.tensor(vals, dtype=torch.float64)
    ffm = torch.tensor(np.stack(ff_rows)) if ff_rows else None
    conq_ff = any(g["mode"] == "conq" and g.get("ffield", False) for g in groups)
    ffp = torch.tensor(ff_pot, dtype=torch.float64) if ff_rows and not conq_ff else None
    return (torch.tensor(mask), torch.tensor(eta), torch.tensor(chi), torch.tensor(hard), A, b,
            ffm, ffp)


def finite_field_add_chi(pos, box, ff_mask, ff_pot, axis=2):
    """Potential of the uniform field on the two ffield groups and the field
    (electrode.py:288-336): ramp drop * (-+ z / L) with the sign set by which
    group sits higher along the axis."""
    drop = ff_pot[0] - ff_pot[1]
    lz = box[axis, axis]
    z = pos[:, axis] / lz
    if pos[ff_mask[0], axis].max() > pos[ff_mask[1], axis].max():
        pot, field = drop * z, -drop / lz
    else:
        pot, field = -drop * z, drop / lz
    return pot[ff_mask[0] | ff_mask[1]], field


def full_potential(orc: qeq_ref.QEqOracle, q, pos, box, eta, pairs, ds, bs):
    """psi = dE_coul+damp/dq of the whole system at charges q."""
    qq = q.clone().requires_grad_(True)
    zeros = torch.zeros_like(q)
    e = orc.func_energy(qq, pos, box, zeros, zeros, eta, pairs, ds, bs)
    return torch.autograd.grad(e.sum(), qq)[0]


def charge_optimization(orc: qeq_ref.QEqOracle, pos, box, q, mask, eta, chi, hard, A, b,
                        ff_mask=None, ff_pot=None, axis=2):
    """Electrode charges (electrode.py:456-560) by the dense KKT solve of the
    electrode subsystem with chi + the electrolyte potential."""
    pairs = nblist_ref.nblist_pbc(pos, box, orc.coulomb.rcut)
    ds = nblist_ref.pairs_ds(pos, pairs, box)
    bs = nblist_ref.buffer_scales(pairs)
    q0 = torch.where(mask, torch.zeros_like(q), q)
    phi = full_potential(orc, q0, pos, box, eta, pairs, ds, bs)
    chi_e = (chi + phi)[mask]
    field = None
    ramp = None
    if ff_mask is not None:
        if ff_pot is not None:
            add, field = finite_field_add_chi(pos, box, ff_mask, ff_pot, axis)
            chi_e = chi_e + add
        else:
            ramp, unit = finite_field_add_chi(pos, box, ff_mask,
                                              torch.tensor([1.0, 0.0], dtype=pos.dtype), axis)
    pe = pos[mask]
    sub_pairs = nblist_ref.nblist_pbc(pe, box, orc.coulomb.rcut)
    sub_ds = nblist_ref.pairs_ds(pe, sub_pairs, box)
    sub_bs = nblist_ref.buffer_scales(sub_pairs)
    H = orc.hessian(pe, box, chi_e, hard[mask], eta[mask], sub_pairs, sub_ds, sub_bs)
    if ramp is not None:
        # unknowns [q, mu_0, mu_1, dV]: H q + A^T mu + dV r = -chi, A q = b,
        # dV = V_0 - V_1 with the electrode potentials V_g = -mu_g
        n, nc = H.shape[0], A.shape[0]
        rows = [int(torch.argmax((A.bool() == g[mask][None]).all(1).to(torch.int)))
                for g in ff_mask]
        M = torch.zeros(n + nc + 1, n + nc + 1, dtype=H.dtype)
        M[:n, :n] = H
        M[:n, n:n + nc] = A.T
        M[:n, -1] = ramp
        M[n:n + nc, :n] = A
        M[-1, n + rows[0]] = 1.0
        M[-1, n + rows[1]] = -1.0
        M[-1, -1] = 1.0
        rhs = torch.cat([-chi_e, b, torch.zeros(1, dtype=H.dtype)])
        sol = torch.linalg.solve(M, rhs)
        return sol[:n], sol[-1] * unit
    if A.shape[0] == 0:
        q_e, _, _ = qeq_ref.kkt_solve(H, chi_e)
    else:
        q_e, _, _ = qeq_ref.kkt_solve(H, chi_e, A, b)
    return q_e, field


def infer(orc: qeq_ref.QEqOracle, pos, box, q, mask, eta, chi, hard, A, b, ff_mask=None,
          ff_pot=None, axis=2):
    """(energy, forces, q_opt) as electrode.infer (electrode.py:339-453, 86-133)."""
    q_e, field = charge_optimization(orc, pos, box, q, mask, eta, chi, hard, A, b, ff_mask,
                                     ff_pot, axis)
    q_opt = q.clone()
    q_opt[mask] = q_e
    p = pos.clone().requires_grad_(True)
    pairs = nblist_ref.nblist_pbc(pos, box, orc.coulomb.rcut)
    ds = nblist_ref.pairs_ds(p, pairs, box)
    bs = nblist_ref.buffer_scales(pairs)
    zeros = torch.zeros_like(q)
    e = orc.func_energy(q_opt, p, box, zeros, zeros, eta, pairs, ds, bs).sum()
    f = -torch.autograd.grad(e, p)[0]
    if field is not None:
        fv = torch.zeros(3, dtype=pos.dtype)
        fv[axis] = field
        f = f + q_opt[:, None] * fv
        e = e + (fv[None] * q_opt[:, None] * pos).sum()
    return e.detach(), f, q_opt
