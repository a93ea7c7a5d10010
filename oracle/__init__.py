"""CPU oracle for the electrostatics hot path -- TEST INFRASTRUCTURE ONLY.

This package is a float64 torch-CPU restatement of the reference
(torch-admp 1.1.5) algorithms on the PME / neighbour-list / QEq path.  It is the
checker: only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` may import it.  The product package ``torch_admp_amd`` never
imports, links or executes anything from here, and must fail loudly when its
HIP library is missing rather than fall back to this code.

Modules
-------
pme_ref     CoulombForceModule restatement (pme.py, recip.py, base_force.py)
nblist_ref  TorchNeighborList restatement (nblist.py, spatial.py) plus an
            O(N^2) brute-force pair enumeration that pins it
ewald_ref   Direct (non-mesh) Ewald summation, numpy, used to pin the SPME
            restatement on small systems
qeq_ref     QEq energy terms and solvers (qeq.py, utils.py, optimizer.py)

Pinning status of every function is listed in each module header.
"""
