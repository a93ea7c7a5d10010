"""The HIP entry points as PyTorch operators, namespace ``admp``.

The operators are registered with the dispatcher by the native library
``_lib/libadmp_torch_ops.so`` (``csrc/torch_ops.cpp``): schemas, argument
checks, per-stream workspaces and autograd formulas in C++, each body one or
two calls of the C ABI (include/admp_amd.h) on the caller's current stream.
Importing this module loads it; the force modules then script
(``torch.jit.script``), and a scripted module saved with ``torch.jit.save``
loads again in any process that has loaded that library.  No operator has a
CPU implementation: CPU tensors raise.

Operators (``torch.ops.admp.*``):

* ``neighbor_pairs(positions, box?, cutoff, box_host?) -> pairs (P, 2)``:
  ordered pairs with 0.01 < d < cutoff (box) or d < cutoff (box None), both
  orderings (src/torch_admp/nblist.py:335-490).  One stream synchronisation
  returns P.  ``box_host``: the box's values on the host when the caller has
  them (saves the device-to-host copy); the same holds for every operator.
* ``neighbor_list(...) -> (pairs, row_ptr, reverse)``: the same pairs with
  their row form: row_ptr (n + 1, int64; n rows, one per atom in cell order,
  row r = pairs[row_ptr[r]:row_ptr[r + 1]] sharing one first atom, some
  empty) and reverse (P, int32; the index of the pair (j, i) of each (i, j)).
* ``max_displacement(positions, reference, box?, box_host?)``: max_i
  |positions_i - reference_i| (minimum image with a box), 0-dim float64.
* ``pair_distance(positions, box?, pairs, box_host?, row_ptr?, reverse?)``:
  ds = |pbc_shift(r_j - r_i)| (src/torch_admp/nblist.py:517-551,
  src/torch_admp/spatial.py:15-36); with the row form the backward sums each
  atom's row without per-pair atomics.  ``pair_distance_backward`` is its
  gradient.
* ``pair_energy(q (nf, n), ds, pairs (nf, P, 2), bs, eta?, mode, kappa,
  eps)``: per-frame pair energy, mode 0 eps erfc(kappa d) qi qj / d bs
  (src/torch_admp/pme.py:312-349), 1 eps qi qj / d bs (pme.py:529-561), 2 the
  Gaussian damping -eps erfc(d / (2 eta_ij)) qi qj / d bs
  (src/torch_admp/qeq.py:106-121).  ``pair_energy_grads`` gives dE/dq, dE/dds
  and dE/deta in one kernel.
* ``ewald_self_terms(q (nf, n), box (nf, 3, 3), kappa, eps)``: the self and
  non-neutral energies (src/torch_admp/pme.py:448-493), no host copy of the
  box; ``ewald_self_terms_backward`` is their gradient.
* ``pme_reciprocal(positions, box, q, kmesh, kappa, eps, precision,
  convention, need_potential, box_host?) -> (energy, potential mesh, dE/dA)``:
  one frame of the SPME reciprocal energy (src/torch_admp/pme.py:351-446);
  ``pme_reciprocal_grads`` gathers dE/dr, dE/dq and dE/dbox from the mesh.

Derivatives.  Every operator that returns an energy has an autograd formula
made of other ``admp`` operators, so gradients of gradients work where the
reference's callers take them (the second derivative in the charges of
QEqForceModule.calc_hessian, src/torch_admp/qeq.py:507-514, mixed with
positions, box, ds and eta): d/dq of v . dE/dq is the same operator applied to
v; d/dds, d/deta, d/dr and d/dbox follow from the polarisation identity
[f(q + v) - f(q - v)] / 2, those gradients being quadratic in the charges.
Differentiating forces, dE/dds or dE/deta again raises NotImplementedError
rather than returning a partial gradient.  The potential mesh and dE/dA of
``pme_reciprocal`` are not differentiable outputs.
"""

from pathlib import Path

import torch

PAIR_EWALD, PAIR_BARE, PAIR_DAMPED = 0, 1, 2

OPS_LIBRARY = Path(__file__).resolve().parent / "_lib" / "libadmp_torch_ops.so"


def _register() -> None:
    if not OPS_LIBRARY.exists():
        raise ImportError(
            f"torch_admp_amd: operator library not found at {OPS_LIBRARY}; build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'`")
    torch.ops.load_library(str(OPS_LIBRARY))


_register()
