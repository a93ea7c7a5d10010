"""Particle-mesh Ewald Coulomb module with the reference call surface
(src/torch_admp/pme.py:24-683), computed by the HIP library.

Extra keyword arguments beyond the reference:
  precision   "float64" (default) or "float32": float32 mesh / FFT / spline
              arithmetic with float64 accumulation of every energy and virial.
  convention  "reference" (default) reproduces k = 2 pi m box^-1 (rows,
              src/torch_admp/recip.py:304); "columns" uses the reciprocal
              lattice vectors proper.  They differ only for non-symmetric
              triclinic boxes.
"""

import math
from typing import Dict, List, Optional, Tuple, Union

import numpy as np
import torch

from . import _lib
from .base_force import BaseForceModule
from . import _ops  # noqa: F401  (registers torch.ops.admp)
from .ops import pme_fused
from .utils import safe_inverse  # noqa: F401  (re-export, reference API)


def _volume(box: torch.Tensor) -> torch.Tensor:
    """det of each (3, 3) box as the triple product a . (b x c): the same
    value and derivative as torch.det, in elementwise kernels rather than a
    device LU factorisation."""
    return torch.sum(box[..., 0, :] * torch.cross(box[..., 1, :], box[..., 2, :], dim=-1), dim=-1)


class CoulombForceModule(BaseForceModule):
    """Coulomb energy with PME (src/torch_admp/pme.py:24-561).

    TorchScript-compatible: forward, func-style helpers and mesh_for_box
    script (torch.jit.script) and save; the pair and reciprocal terms are the
    torch.ops.admp operators.  energy_forces (the fused single-call path) is
    Python only."""

    kmesh: Optional[torch.Tensor]
    spacing: Optional[torch.Tensor]
    sel: Optional[List[int]]

    def __init__(self, rcut: float, ethresh: float = 1e-5, kspace: bool = True,
                 rspace: bool = True, slab_corr: bool = False, slab_axis: int = 2,
                 units_dict: Optional[Dict] = None, sel: Optional[List[int]] = None,
                 kappa: Optional[float] = None, spacing: Union[List[float], float, None] = None,
                 kmesh: Union[List[int], int, None] = None, precision: str = "float64",
                 convention: str = "reference") -> None:
        BaseForceModule.__init__(self, units_dict)
        # validation order and messages of src/torch_admp/pme.py:98-145
        if rcut <= 0.0:
            raise ValueError(f"rcut must be positive, got {rcut}")
        if ethresh <= 0.0:
            raise ValueError(f"ethresh must be positive, got {ethresh}")
        if slab_axis not in (0, 1, 2):
            raise ValueError(f"slab_axis must be 0/1/2, got {slab_axis}")
        if precision not in _lib.PRECISION:
            raise ValueError(f"precision must be 'float64' or 'float32', got {precision!r}")
        if convention not in _lib.KCONV:
            raise ValueError(f"convention must be 'reference' or 'columns', got {convention!r}")
        self.kspace_flag = kspace
        if kappa is not None:
            if kappa <= 0.0:
                raise ValueError(f"kappa must be positive, got {kappa}")
            self.kappa = float(kappa)
        elif kspace:
            self.kappa = math.sqrt(-math.log(2 * ethresh)) / rcut / self.const_lib.length_coeff
        else:
            self.kappa = 0.0
        self.ethresh = float(ethresh)
        if kmesh is not None:
            if isinstance(kmesh, int):
                kmesh = [kmesh, kmesh, kmesh]
            for i, k in enumerate(kmesh):
                if k <= 0:
                    raise ValueError(f"kmesh values must be positive, got kmesh[{i}] = {k}")
            self.kmesh = torch.tensor(list(kmesh), dtype=torch.long)
        else:
            self.kmesh = None
        self._kmesh = torch.zeros(3, dtype=torch.long)
        if spacing is not None:
            if isinstance(spacing, float):
                spacing = [spacing, spacing, spacing]
            for i, sp in enumerate(spacing):
                if sp <= 0:
                    raise ValueError(f"spacing values must be positive, got spacing[{i}] = {sp}")
            self.spacing = torch.tensor(list(spacing), dtype=torch.float64)
        else:
            self.spacing = None
        self.rspace_flag = rspace
        self.slab_corr_flag = slab_corr
        self.slab_axis = slab_axis
        self.precision = precision
        self.convention = convention
        self._precision_code: int = _lib.PRECISION[precision]
        self._convention_code: int = _lib.KCONV[convention]
        zero = torch.zeros(1, dtype=torch.float64)
        self.real_energy = zero
        self.reciprocal_energy = zero
        self.self_energy = zero
        self.non_neutral_energy = zero
        self.slab_corr_energy = zero
        self.pme_order: int = 6
        self.rcut = float(rcut)
        self.sel = sel

    def get_rcut(self) -> float:
        return self.rcut

    def get_sel(self) -> Optional[List[int]]:
        return self.sel

    @torch.jit.export
    def mesh_for_box(self, box_frame: torch.Tensor) -> List[int]:
        """Mesh size: explicit kmesh, else ceil(diag / spacing), else the
        accuracy rule ceil(2 kappa L / (3 ethresh^(1/5))) on the box diagonal
        (src/torch_admp/pme.py:380-392)."""
        kmesh = self.kmesh
        if kmesh is not None:
            return torch.jit.annotate(List[int], kmesh.tolist())
        diag = torch.diagonal(box_frame.detach().reshape(3, 3)).to(torch.device("cpu"),
                                                                  torch.float64)
        spacing = self.spacing
        if spacing is not None:
            k = torch.ceil(diag / spacing)
        else:
            k = torch.ceil(2 * self.kappa * diag / (3.0 * self.ethresh ** (1.0 / 5.0)))
        return torch.jit.annotate(List[int], k.to(torch.long).tolist())

    def _forward_impl(self, positions: torch.Tensor, box: Optional[torch.Tensor],
                      pairs: torch.Tensor, ds: torch.Tensor, buffer_scales: torch.Tensor,
                      params: Dict[str, torch.Tensor]) -> torch.Tensor:
        nf = positions.size(0)
        na = positions.size(1)
        charges = params["charge"].reshape(nf, na)
        if box is not None:
            return self._forward_pbc(charges, positions, box, pairs, ds, buffer_scales)
        return self._forward_obc(charges, pairs, ds, buffer_scales)

    def _forward_pbc(self, charges: torch.Tensor, positions: torch.Tensor, box: torch.Tensor,
                     pairs: torch.Tensor, ds: torch.Tensor,
                     buffer_scales: torch.Tensor) -> torch.Tensor:
        """Sum of the five terms (src/torch_admp/pme.py:261-310); terms that are
        switched off contribute zero and keep their previous attribute value."""
        terms: List[torch.Tensor] = []
        if self.rspace_flag:
            self.real_energy = self._forward_pbc_real(charges, pairs, ds, buffer_scales)
            terms.append(self.real_energy)
        if self.kspace_flag:
            self.reciprocal_energy = self._forward_pbc_reciprocal(charges, positions, box)
            # the terms of _forward_pbc_self and _forward_pbc_non_neutral in one
            # kernel (and one backward kernel)
            e_self, e_nn = torch.ops.admp.ewald_self_terms(charges, box, self.kappa,
                                                           self.const_lib.dielectric)
            self.self_energy = e_self
            self.non_neutral_energy = e_nn
            terms.append(self.reciprocal_energy)
            terms.append(self.self_energy)
            terms.append(self.non_neutral_energy)
        if self.slab_corr_flag:
            self.slab_corr_energy = self._forward_slab_corr(charges, positions, box, ds)
            terms.append(self.slab_corr_energy)
        if len(terms) == 0:
            return torch.zeros(charges.size(0), device=positions.device, dtype=positions.dtype)
        energy = terms[0]
        for t in terms[1:]:
            energy = energy + t
        return energy

    def _forward_pbc_real(self, charges: torch.Tensor, pairs: torch.Tensor, ds: torch.Tensor,
                          buffer_scales: torch.Tensor) -> torch.Tensor:
        """eps erfc(kappa d) qi qj / d over the pair rows (src/torch_admp/pme.py:312-349)."""
        # pair mode 0: erfc-screened pairs (the kernel reads int64 scales as
        # they are)
        return torch.ops.admp.pair_energy(charges, ds, pairs, buffer_scales, None, 0,
                                          self.kappa, self.const_lib.dielectric)

    def _forward_pbc_reciprocal(self, charges: torch.Tensor, positions: torch.Tensor,
                                box: torch.Tensor) -> torch.Tensor:
        """One reciprocal launch sequence per frame (src/torch_admp/pme.py:351-446)."""
        out: List[torch.Tensor] = []
        for ii in range(positions.size(0)):
            # one host copy of the frame's box for the mesh rule and the kernels
            # (the neighbour list's copy of the same box when it has one)
            if torch.jit.is_scripting():
                box_host = box[ii].detach().to(torch.device("cpu"), torch.float64)
            else:
                box_host = _lib.host_box(box[ii])
            km = self.mesh_for_box(box_host)
            need = torch.is_grad_enabled() and (positions.requires_grad or box.requires_grad
                                                or charges.requires_grad)
            e, _, _ = torch.ops.admp.pme_reciprocal(positions[ii], box[ii], charges[ii], km,
                                                    self.kappa, self.const_lib.dielectric,
                                                    self._precision_code, self._convention_code,
                                                    need, box_host)
            out.append(e)
            if torch.jit.annotate(List[int], self._kmesh.tolist()) != km:
                self._kmesh = torch.tensor(km, dtype=torch.long)
        if len(out) == 1:
            return out[0].reshape(1)
        return torch.stack(out)

    def _forward_pbc_self(self, charges: torch.Tensor) -> torch.Tensor:
        """-eps kappa / sqrt(pi) sum q^2 (src/torch_admp/pme.py:448-461)."""
        if self.kspace_flag:
            coeff = self.kappa / self.const_lib.sqrt_pi
            return -torch.sum(coeff * charges ** 2, dim=-1) * self.const_lib.dielectric
        return torch.zeros(charges.size(0), device=charges.device, dtype=charges.dtype)

    def _forward_pbc_non_neutral(self, charges: torch.Tensor, box: torch.Tensor) -> torch.Tensor:
        """-eps pi Q^2 / (2 V kappa^2) (src/torch_admp/pme.py:463-493)."""
        volume = _volume(box)
        q_tot = torch.sum(charges, dim=-1)
        coeff = -self.const_lib.pi / (2 * volume * self.kappa ** 2) * self.const_lib.dielectric
        return coeff * q_tot ** 2

    def _forward_slab_corr(self, charges: torch.Tensor, positions: torch.Tensor,
                           box: torch.Tensor, ds: torch.Tensor) -> torch.Tensor:
        """Yeh-Berkowitz slab term (src/torch_admp/pme.py:495-527)."""
        volume = _volume(box)
        pre = 2 * self.const_lib.pi / volume * self.const_lib.dielectric
        z = positions[:, :, self.slab_axis]
        mz = torch.sum(charges * z, dim=-1)
        q_tot = torch.sum(charges, dim=-1)
        lz = torch.norm(box[:, self.slab_axis], dim=-1)
        return pre * (mz ** 2 - q_tot * torch.sum(charges * z ** 2, dim=-1)
                      - q_tot ** 2 * lz ** 2 / 12)

    def _forward_obc(self, charges: torch.Tensor, pairs: torch.Tensor, ds: torch.Tensor,
                     buffer_scales: torch.Tensor) -> torch.Tensor:
        """Bare Coulomb over the pair list (src/torch_admp/pme.py:529-561)."""
        # pair mode 1: bare pairs
        return torch.ops.admp.pair_energy(charges, ds, pairs, buffer_scales, None, 1,
                                          0.0, self.const_lib.dielectric)

    # -- fused single-call path (no pair list crosses the boundary) -------------
    @torch.jit.unused
    def energy_forces(self, positions: torch.Tensor, box: torch.Tensor, charges: torch.Tensor,
                      workspace: Optional[torch.Tensor] = None
                      ) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
        """Energy, forces and dE/dbox of one periodic frame through the fused
        cell-list + PME launch sequence.  Same result as forward() on the
        TorchNeighborList(rcut) pair list followed by autograd when every box
        height is at least 2 rcut; below that every periodic image inside the
        cutoff is summed (the lattice sum).  Real and reciprocal space only
        (kspace and rspace on, no slab term)."""
        if not (self.kspace_flag and self.rspace_flag) or self.slab_corr_flag:
            raise ValueError("energy_forces requires kspace=True, rspace=True, slab_corr=False")
        lc = self.const_lib.length_coeff
        ec = self.const_lib.energy_coeff
        if lc != 1.0:
            positions, box = positions * lc, box * lc
        km = self.mesh_for_box(box)
        d = self.__dict__
        if km != d.get("_kmesh_list"):
            self._kmesh = torch.tensor(km, dtype=torch.long)
            d["_kmesh_list"] = km
        e, f, gbox = pme_fused(positions, box, charges, self.rcut * lc, float(self.kappa), km,
                               self.precision, self.convention, self.const_lib.dielectric,
                               workspace)
        if ec != 1.0 or lc != 1.0:
            e, f, gbox = e / ec, f * (lc / ec), gbox * (lc / ec)
        # plain attributes: written into __dict__ directly, skipping nn.Module's
        # parameter/buffer checks on this per-step path
        (d["real_energy"], d["reciprocal_energy"], d["self_energy"], d["non_neutral_energy"],
         total) = e.split(1)
        return total.view(()), f, gbox


def setup_ewald_parameters(rcut: float, box: Union[torch.Tensor, np.ndarray, None] = None,
                           threshold: float = 1e-5, spacing: Optional[float] = None,
                           method: str = "openmm") -> Tuple[float, int, int, int]:
    """kappa and mesh from cutoff and accuracy (src/torch_admp/pme.py:564-658)."""
    from scipy import special

    if rcut <= 0.0:
        raise ValueError(f"rcut must be positive, got {rcut}")
    if box is None:
        return 0.1, 1, 1, 1
    if isinstance(box, torch.Tensor):
        box = box.detach().cpu().numpy()
    box = np.asarray(box)
    for a, b in ((0, 1), (0, 2), (1, 2)):
        assert np.inner(box[a], box[b]) == 0.0, "Only orthogonal box is supported currently."
    if method == "openmm":
        kappa = np.sqrt(-np.log(2 * threshold)) / rcut
        ks = [np.ceil(2 * kappa * box[i, i] / (3.0 * threshold ** (1.0 / 5.0))).astype(int)
              for i in range(3)]
    elif method == "gromacs":
        assert spacing is not None, "Spacing must be provided for gromacs method."
        kappa = 5.0
        i = 0
        while special.erfc(kappa * rcut) > threshold:
            i += 1
            kappa *= 2
        low, high = 0.0, kappa
        for _ in range(i + 60):
            kappa = (low + high) / 2
            if special.erfc(kappa * rcut) > threshold:
                low = kappa
            else:
                high = kappa
        ks = [np.ceil(box[i, i] / spacing).astype(int) for i in range(3)]
    else:
        raise ValueError(f"Invalid method: {method}.Valid methods: 'openmm', 'gromacs'")
    return kappa, ks[0], ks[1], ks[2]
