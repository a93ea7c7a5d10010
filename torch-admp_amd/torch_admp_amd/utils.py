"""Constants, unit handling and the small tensor utilities of the reference
interface (src/torch_admp/utils.py).  Host-side only: plain torch ops."""

import math
from typing import Dict, List, Optional

import torch

# ASE-compatible unit values (CODATA 2014, the default of ase.units) for the
# names the reference accepts through units_dict (src/torch_admp/utils.py:227-258).
_E_CHARGE = 1.6021766208e-19
_AVOGADRO = 6.022140857e23
_UNITS = {
    "eV": 1.0,
    "Ang": 1.0,
    "Angstrom": 1.0,
    "nm": 10.0,
    "m": 1e10,
    "cm": 1e8,
    "Bohr": 0.52917721067,
    "Hartree": 27.21138602,
    "Ry": 13.605693009,
    "J": 1.0 / _E_CHARGE,
    "kJ": 1000.0 / _E_CHARGE,
    "kcal": 4184.0 / _E_CHARGE,
    "mol": _AVOGADRO,
}

# 1/(4 pi eps0) in eV Angstrom / e^2 from scipy CODATA 2018 (src/torch_admp/utils.py:268-278)
DIELECTRIC = 14.399645468667817


class TorchConstants(torch.nn.Module):
    """pi, sqrt(pi), the Coulomb constant and unit coefficients
    (src/torch_admp/utils.py:213-282)."""

    def __init__(self, units_dict: Optional[Dict] = None):
        super().__init__()
        self.pi = math.pi
        self.sqrt_pi = math.sqrt(math.pi)
        units_dict = units_dict or {}
        user_energy = units_dict.get("energy", "eV")
        user_length = units_dict.get("length", "Ang")
        if user_length not in _UNITS:
            raise ValueError(f"Unknown length unit: {user_length}")
        if user_energy == "kJ/mol":
            energy_coeff = _UNITS["kJ"] / _UNITS["mol"]
        elif user_energy == "kcal/mol":
            energy_coeff = _UNITS["kcal"] / _UNITS["mol"]
        elif user_energy in _UNITS:
            energy_coeff = _UNITS[user_energy]
        else:
            raise ValueError(f"Unknown energy unit: {user_energy}")
        self.length_coeff = _UNITS[user_length]
        self.energy_coeff = energy_coeff
        self.dielectric = DIELECTRIC


def safe_inverse(x: torch.Tensor, threshold: float = 1e-8) -> torch.Tensor:
    """1/x where |x| > threshold else 0 (src/torch_admp/utils.py:66-81)."""
    return torch.where(x.abs() > threshold, 1 / x, torch.zeros_like(x))


def calc_grads(t_out: torch.Tensor, t_in: torch.Tensor) -> torch.Tensor:
    """d(sum t_out)/d t_in with the graph retained (src/torch_admp/utils.py:84-112)."""
    assert t_in.requires_grad, "Input tensor requires grad"
    grad = torch.autograd.grad([t_out], [t_in], grad_outputs=[torch.ones_like(t_out)],
                               retain_graph=True)[0]
    assert grad is not None
    return grad


def calc_pgrads(t_out: torch.Tensor, t_in: torch.Tensor, constraint_matrix: torch.Tensor,
                coeff_matrix: torch.Tensor) -> torch.Tensor:
    """Projected gradient g - C (A g) (src/torch_admp/utils.py:115-148)."""
    raw = calc_grads(t_out, t_in).reshape(-1, 1)
    residual = -torch.matmul(constraint_matrix, raw)
    return (raw + torch.matmul(coeff_matrix, residual)).reshape(-1)


def vector_projection_coeff_matrix(constraint_matrix: torch.Tensor) -> torch.Tensor:
    """A^T (A A^T)^-1 (src/torch_admp/utils.py:151-171)."""
    at = constraint_matrix.transpose(0, 1)
    return torch.matmul(at, torch.inverse(torch.matmul(constraint_matrix, at)))


def vector_projection(vector_in: torch.Tensor, constraint_matrix: torch.Tensor,
                      constraint_vector: Optional[torch.Tensor] = None) -> torch.Tensor:
    """x + A^T (A A^T)^-1 (b - A x) (src/torch_admp/utils.py:174-210)."""
    if constraint_vector is None:
        constraint_vector = torch.zeros([constraint_matrix.shape[0]], device=vector_in.device,
                                        dtype=vector_in.dtype)
    coeff = vector_projection_coeff_matrix(constraint_matrix)
    residual = constraint_vector.reshape(-1, 1) - torch.matmul(constraint_matrix,
                                                              vector_in.reshape(-1, 1))
    return (vector_in.reshape(-1, 1) + torch.matmul(coeff, residual)).reshape(-1)


def to_numpy_array(x):
    """Detach to a numpy array (host conversion helper, src/torch_admp/utils.py:285-360)."""
    if x is None:
        return None
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy()
    import numpy as np

    return np.asarray(x)


def to_torch_tensor(x, device=None, dtype=None):
    """numpy / list -> tensor (src/torch_admp/utils.py:363-438)."""
    if x is None:
        return None
    t = x if isinstance(x, torch.Tensor) else torch.as_tensor(x)
    return t.to(device=device if device is not None else t.device,
                dtype=dtype if dtype is not None else t.dtype)


__all__: List[str] = [
    "TorchConstants", "safe_inverse", "calc_grads", "calc_pgrads",
    "vector_projection_coeff_matrix", "vector_projection", "to_numpy_array", "to_torch_tensor",
    "DIELECTRIC",
]
