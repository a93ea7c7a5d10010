"""torch_admp_amd: MI355X (gfx950) HIP implementation of the torch-admp
electrostatics hot path behind the reference call surface.

    from torch_admp_amd.nblist import TorchNeighborList
    from torch_admp_amd.pme import CoulombForceModule

The compute path is the HIP library in ``_lib/libadmp_amd.so``; calls on CPU
tensors or without the library raise.
"""

from . import _lib, base_force, electrode, lammps_io, nblist, ops, pme, qeq, serving, utils  # noqa: F401
from .base_force import BaseForceModule  # noqa: F401
from .electrode import PolarizableElectrode, infer, setup_from_lammps  # noqa: F401
from .nblist import TorchNeighborList, VerletNeighborList  # noqa: F401
from .pme import CoulombForceModule, setup_ewald_parameters  # noqa: F401
from .qeq import QEqForceModule, matinv_optimize, pgrad_optimize  # noqa: F401
from .utils import calc_grads, safe_inverse  # noqa: F401

__version__ = "0.1.0"
