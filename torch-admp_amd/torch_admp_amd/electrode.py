"""Constant-potential (CONP) and constant-charge (CONQ) electrodes with the
reference call surface (src/torch_admp/electrode.py:21-560), on the HIP QEq
solver.

An electrode is a QEq subsystem: the electrolyte charges are fixed, the
electrode charges minimise the total energy under the electrode constraints.
The electrolyte enters the electrode problem through the electrostatic
potential it creates on the electrode atoms (dE/dq at zero electrode charge),
added to the electronegativity; CONP shifts chi by the applied potential,
CONQ adds a total-charge row to the constraint matrix, `symm` adds global
neutrality, and the finite field (CONP only) adds the potential ramp of a
uniform field along the slab axis.
"""

from __future__ import annotations

from typing import List, Optional, Tuple, Union

import numpy as np
import torch

from .qeq import QEqForceModule, matinv_optimize, pgrad_optimize
from .utils import calc_grads


class PolarizableElectrode(QEqForceModule):
    """QEq module with the electrode helpers (src/torch_admp/electrode.py:21-133)."""

    def __init__(self, rcut: float, ethresh: float = 1e-5, **kwargs) -> None:
        super().__init__(rcut, ethresh, **kwargs)

    def calc_coulomb_potential(self, electrode_mask: Optional[torch.Tensor],
                               positions: torch.Tensor, box: torch.Tensor, eta: torch.Tensor,
                               charges: torch.Tensor, pairs: torch.Tensor, ds: torch.Tensor,
                               buffer_scales: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
        """Potential dE/dq on every atom with the electrode charges set to zero,
        and that energy (src/torch_admp/electrode.py:48-84)."""
        if electrode_mask is None:
            q = charges.detach().clone()
        else:
            q = torch.where(electrode_mask == 0, charges.detach(), torch.zeros_like(charges))
        q.requires_grad_(True)
        zeros = torch.zeros_like(eta)
        energy = self.forward(positions, box, pairs, ds, buffer_scales,
                              {"charge": q, "eta": eta, "hardness": zeros, "chi": zeros})
        assert energy.size(0) == 1
        return calc_grads(energy[0], q), energy

    def coulomb_calculator(self, positions: torch.Tensor, box: torch.Tensor,
                           charges: torch.Tensor, eta: torch.Tensor, pairs: torch.Tensor,
                           ds: torch.Tensor, buffer_scales: torch.Tensor,
                           efield: Optional[torch.Tensor] = None
                           ) -> Tuple[torch.Tensor, torch.Tensor]:
        """Energy and forces at fixed charges, plus the q E work of a uniform
        field along the slab axis (src/torch_admp/electrode.py:86-133)."""
        zeros = torch.zeros_like(eta)
        energy = self.forward(positions, box, pairs, ds, buffer_scales,
                              {"charge": charges, "eta": eta, "hardness": zeros, "chi": zeros})
        assert energy.size(0) == 1
        energy = energy[0]
        if not positions.requires_grad:
            raise ValueError("positions must require grad to compute forces; call "
                             "positions.requires_grad_(True)")
        forces = -calc_grads(energy, positions)
        if efield is not None:
            field = torch.zeros(3, dtype=positions.dtype, device=positions.device)
            field[self.slab_axis] = efield[self.slab_axis]
            forces = forces + charges.unsqueeze(1) * field
            energy = energy + torch.sum(field.reshape(1, 3) * charges.unsqueeze(1) * positions)
        return energy, forces


class LAMMPSElectrodeConstraint:
    """One electrode group as LAMMPS `fix electrode` declares it
    (src/torch_admp/electrode.py:136-200): mode "conp" (value = potential, V)
    or "conq" (value = total charge, e); eta in 1/length as LAMMPS takes it."""

    def __init__(self, indices: Union[List[int], np.ndarray], mode: str, value: float,
                 eta: float, chi: float = 0.0, hardness: float = 0.0,
                 ffield: bool = False) -> None:
        self.indices = np.array(indices, dtype=int)
        assert self.indices.ndim == 1
        assert mode in ["conp", "conq"], f"mode {mode} not supported"
        self.mode = mode
        self.value = value
        self.eta = eta
        self.chi = chi
        self.hardness = hardness
        self.ffield = ffield


def setup_from_lammps(n_atoms: int, constraint_list: List[LAMMPSElectrodeConstraint],
                      symm: bool = False, device=None, dtype=torch.float64,
                      conq_ffield: bool = False):
    """Per-atom electrode arrays from LAMMPS-style groups (src/torch_admp/electrode.py:203-285).

    Returns (mask, eta, chi, hardness, constraint_matrix over electrode atoms,
    constraint_vals, ffield_electrode_mask or None, ffield_potential or None).
    The Gaussian width is eta_i = sqrt(2) / (2 eta_lammps).

    ``conq_ffield=True`` enables the finite field with two constant-charge
    groups (LAMMPS `fix electrode/conq ... ffield on`), which the reference
    rejects: the potential drop is then an unknown, returned ffield_potential
    is None and charge_optimization solves it self-consistently.
    """
    if device is None:
        device = torch.device("cuda") if torch.cuda.is_available() else torch.device("cpu")
    mask = np.zeros(n_atoms, dtype=bool)
    eta = np.zeros(n_atoms)
    chi = np.zeros(n_atoms)
    hardness = np.zeros(n_atoms)
    rows, vals, ff_rows, ff_pot = [], [], [], []
    for c in constraint_list:
        mask[c.indices] = True
        eta[c.indices] = np.sqrt(2.0) / 2.0 / c.eta
        chi[c.indices] = c.chi
        hardness[c.indices] = c.hardness
        if c.mode == "conq":
            if symm:
                raise AttributeError("symm should be False for conq, user can implement symm "
                                     "by conq")
            if c.ffield and not conq_ffield:
                raise AttributeError("ffield with conq has not been implemented yet")
            row = np.zeros(n_atoms)
            row[c.indices] = 1.0
            rows.append(row)
            vals.append(c.value)
        else:
            chi[c.indices] -= c.value
        if c.ffield:
            row = np.zeros(n_atoms, dtype=bool)
            row[c.indices] = True
            ff_rows.append(row)
            ff_pot.append(c.value)
    if len(ff_rows) == 0:
        ff_mask, ff_potential = None, None
    elif len(ff_rows) == 2:
        ff_mask = torch.tensor(np.stack(ff_rows), dtype=torch.bool, device=device)
        modes = [c.mode for c in constraint_list if c.ffield]
        if "conq" in modes:
            if modes != ["conq", "conq"]:
                raise AttributeError("ffield groups must both be conp or both be conq")
            ff_potential = None  # the drop is solved with the charges
        else:
            ff_potential = torch.tensor(ff_pot, dtype=dtype, device=device)
    else:
        raise AttributeError("number of ffield group should be 0 or 2")
    if symm:
        rows.append(np.ones(n_atoms))
        vals.append(0.0)
    n_e = int(mask.sum())
    if rows:
        cmat = torch.tensor(np.stack(rows)[:, mask], dtype=dtype, device=device)
        cvals = torch.tensor(vals, dtype=dtype, device=device)
    else:
        cmat = torch.zeros((0, n_e), dtype=dtype, device=device)
        cvals = torch.zeros(0, dtype=dtype, device=device)

    def t(a, dt=dtype):
        return torch.tensor(a, dtype=dt, device=device)

    return (t(mask, torch.bool), t(eta), t(chi), t(hardness), cmat, cvals, ff_mask, ff_potential)


def finite_field_add_chi(positions: torch.Tensor, box: torch.Tensor,
                         ffield_electrode_mask: torch.Tensor, ffield_potential: torch.Tensor,
                         slab_axis: int = 2):
    """Electrode-atom potential of a uniform field that drops potential[0] -
    potential[1] across the cell, and that field (orthogonal cells only;
    src/torch_admp/electrode.py:288-336, LAMMPS fix electrode convention)."""
    assert positions.dim() == 2 and positions.shape[1] == 3
    assert box.dim() == 2
    assert ffield_potential.dim() == 1 and ffield_potential.shape[0] == 2
    assert ffield_electrode_mask.dim() == 2 and ffield_electrode_mask.shape[0] == 2
    assert ffield_electrode_mask.shape[1] == positions.shape[0]
    first, second = ffield_electrode_mask[0], ffield_electrode_mask[1]
    drop = ffield_potential[0] - ffield_potential[1]
    lz = box[slab_axis][slab_axis]
    z = positions[:, slab_axis] / lz
    if torch.max(positions[first, slab_axis]) > torch.max(positions[second, slab_axis]):
        offset, efield = z, -drop / lz
    else:
        offset, efield = -z, drop / lz
    potential = drop * offset
    return potential[first | second], efield


def charge_optimization(calculator: PolarizableElectrode, positions, box, charges, pairs, ds,
                        buffer_scales, electrode_mask, eta, chi, hardness, constraint_matrix,
                        constraint_vals, ffield_electrode_mask, ffield_potential,
                        method: str = "lbfgs"):
    """Electrode charges for fixed electrolyte charges (src/torch_admp/electrode.py:456-560).
    Single frame, (N, 3) positions.  Returns (electrode charges, field or None)."""
    device, dtype = positions.device, positions.dtype
    if int(electrode_mask.sum()) == 0:
        return charges[electrode_mask], None
    if ffield_electrode_mask is not None and calculator.slab_corr:
        raise ValueError("Slab correction and finite field cannot be used together.")
    phi, _ = calculator.calc_coulomb_potential(electrode_mask, positions, box, eta, charges,
                                               pairs, ds, buffer_scales)
    chi_e = (chi + phi)[electrode_mask]
    efield = None
    ramp = None
    if ffield_electrode_mask is not None:
        drop = ffield_potential if ffield_potential is not None else torch.tensor(
            [1.0, 0.0], dtype=dtype, device=device)
        chi_ff, ef = finite_field_add_chi(positions, box, ffield_electrode_mask, drop,
                                          calculator.slab_axis)
        if ffield_potential is not None:
            chi_e = chi_e + chi_ff
            efield = torch.zeros(3, dtype=dtype, device=device)
            efield[calculator.slab_axis] = ef
        else:
            ramp, unit_field = chi_ff, ef  # per unit drop; the drop is solved below
    # the electrode-electrode pairs, renumbered within the electrode subsystem
    pair_mask = electrode_mask[pairs[:, 0]] & electrode_mask[pairs[:, 1]]
    remap = torch.full((electrode_mask.size(0),), -1, dtype=torch.long, device=device)
    remap[electrode_mask] = torch.arange(int(electrode_mask.sum()), device=device)
    sub_pairs = remap[pairs[pair_mask]]
    kwargs = dict(module=calculator, positions=positions[electrode_mask], box=box,
                  chi=chi_e.detach(), hardness=hardness[electrode_mask], eta=eta[electrode_mask],
                  pairs=sub_pairs, ds=ds[pair_mask], buffer_scales=buffer_scales[pair_mask],
                  constraint_matrix=constraint_matrix, constraint_vals=constraint_vals)
    if constraint_matrix is not None and constraint_matrix.shape[0] == 0:
        kwargs["constraint_matrix"] = None
        kwargs["constraint_vals"] = None
    if ramp is not None:
        q_e, drop = _conq_ffield_solve(calculator, kwargs, ffield_electrode_mask[:, electrode_mask],
                                       ramp, method)
        efield = torch.zeros(3, dtype=dtype, device=device)
        efield[calculator.slab_axis] = drop * unit_field
        return q_e, efield
    if method == "matinv":
        _, q_e = matinv_optimize(**kwargs)
    else:
        _, q_e = pgrad_optimize(q0=charges[electrode_mask].reshape(-1, 1), method=method,
                                reinit_q=True, **kwargs)
    return q_e, efield


def _solve(kwargs, chi, vals, method):
    kw = dict(kwargs, chi=chi, constraint_vals=vals)
    if method == "matinv":
        return matinv_optimize(**kw)[1].detach()
    return pgrad_optimize(q0=None, method=method, reinit_q=True, **kw)[1].detach()


def _conq_ffield_solve(calculator, kwargs, ff_groups, ramp, method):
    """Constant group charges under the finite field whose potential drop dV is
    itself unknown.  Stationarity H q + c + dV r + A^T mu = 0 with A q = b and
    the electrode potentials -mu_g giving dV = mu_1 - mu_0.  Being linear,
    q = q_a + dV q_b with q_a the solve at dV = 0 and q_b the solve of the
    unit ramp with zero group charges; their multipliers fix dV."""
    A, b = kwargs["constraint_matrix"], kwargs["constraint_vals"]
    chi = kwargs["chi"]
    q_a = _solve(kwargs, chi, b, method)
    q_b = _solve(kwargs, ramp, torch.zeros_like(b), method)
    args = [kwargs[k] for k in ("positions", "box", "chi", "hardness", "eta", "pairs", "ds",
                                "buffer_scales")]
    module = kwargs["module"]

    def multipliers(q, c):
        hq = module.hessian_matvec(q, *args).to(q.dtype)
        return -torch.linalg.solve(A @ A.T, A @ (hq + c))

    # the two ffield groups as rows of the constraint matrix
    rows = [int(torch.argmax((A.bool() == g[None]).all(1).to(torch.int))) for g in ff_groups]
    mu_a = multipliers(q_a, chi)[rows]
    mu_b = multipliers(q_b, ramp)[rows]
    drop = (mu_a[1] - mu_a[0]) / (1.0 - (mu_b[1] - mu_b[0]))
    return q_a + drop * q_b, drop


def infer(calculator: PolarizableElectrode, positions: torch.Tensor, box: torch.Tensor,
          charges: torch.Tensor, pairs: torch.Tensor, ds: torch.Tensor,
          buffer_scales: torch.Tensor, electrode_mask: torch.Tensor, eta: torch.Tensor,
          chi: torch.Tensor, hardness: torch.Tensor, constraint_matrix: Optional[torch.Tensor],
          constraint_vals: Optional[torch.Tensor], ffield_electrode_mask: Optional[torch.Tensor],
          ffield_potential: Optional[torch.Tensor], method: str = "lbfgs"):
    """Solve the electrode charges, then energy and forces at those charges
    (src/torch_admp/electrode.py:339-453).  Returns (energy, forces, q_opt)."""
    pos_s, box_s, pairs_s, ds_s, bs_s = calculator.standardize_input_tensor(
        positions, box, pairs, ds, buffer_scales)
    assert pos_s.shape[0] == 1
    assert box_s is not None
    q_e, efield = charge_optimization(calculator, pos_s[0], box_s[0], charges, pairs_s[0],
                                      ds_s[0], bs_s[0], electrode_mask, eta, chi, hardness,
                                      constraint_matrix, constraint_vals, ffield_electrode_mask,
                                      ffield_potential, method)
    q_opt = charges.clone()
    q_opt[electrode_mask] = q_e.detach().reshape(-1).to(q_opt.dtype)
    energy, forces = calculator.coulomb_calculator(positions, box, q_opt, eta, pairs, ds,
                                                   buffer_scales, efield)
    return energy, forces, q_opt
