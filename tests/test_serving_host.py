"""Host-side logic of the serving loop (no GPU needed)."""

import math

import numpy as np
import pytest
import torch

from oracle import pme_ref
from torch_admp_amd import serving


def test_ewald_parameters_follow_module_rule():
    box = np.diag([31.0, 27.5, 40.2])
    for rc, eth in ((6.0, 1e-5), (8.0, 1e-4), (7.3, 5e-6)):
        kappa, kmesh = serving.ewald_parameters(box, rc, eth)
        assert kappa == pytest.approx(pme_ref.kappa_from_ethresh(rc, eth), rel=1e-15)
        assert kmesh == pme_ref.kmesh_rule(torch.tensor(box), kappa, eth)


def test_synthetic_stream_shape():
    reqs = serving.synthetic_stream(300, 500.0, seed=1)
    n = np.array([r.n_atoms for r in reqs])
    assert n.min() >= 64 and n.max() <= 8194 and np.all(n % 2 == 0)
    # log-uniform: the median sits near the geometric mean of the range
    assert 400 < np.median(n) < 1400
    arr = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(arr) >= 0)
    assert 0.3 < arr[-1] < 1.0  # 300 arrivals at 500 Hz
    for r in reqs[:50]:
        assert abs(r.charges.sum()) == 0.0
        vol = abs(np.linalg.det(r.box))
        assert vol == pytest.approx(r.n_atoms / 0.1, rel=1e-9)
        assert 6.0 <= r.rcut <= 8.0
        assert r.kappa == pytest.approx(math.sqrt(-math.log(2e-5)) / r.rcut)


def test_server_requires_gpu():
    with pytest.raises(RuntimeError, match="ROCm GPU"):
        serving.PMEServer("cpu", hbm_budget=1 << 20)


def test_qeq_request_checks_and_cost():
    from torch_admp_amd.workloads import qeq_species_box

    pos, box, chi, hard, eta = qeq_species_box(300, 15.0, seed=2)
    r = serving.QEqRequest(pos, box, chi, hard, eta, rcut=6.0)
    assert r.n_atoms == 300 and r.kmesh == serving.ewald_parameters(box, 6.0, 1e-5)[1]
    # pair list (half sphere of rc at this density), per-atom arrays and mesh
    pairs = 300 * (4.0 / 3.0 * math.pi * 6.0 ** 3 * 300 / 15.0 ** 3) / 2.0
    mesh = int(np.prod(r.kmesh))
    assert serving.PMEServer.qeq_cost(r) == int(64 * pairs + 512 * 300 + 48 * mesh + (1 << 20))
    with pytest.raises(ValueError, match="eta"):
        serving.QEqRequest(pos, box, chi, hard, eta[:-1], rcut=6.0)


def test_energy_error_bound_forms():
    """The stated bound of a result: float32 scales with the sum of the term
    magnitudes plus the reciprocal term, float64 with the term magnitudes."""
    from torch_admp_amd import serving

    e = np.array([-100.0, 2000.0, -2500.0, -1.0, -601.0])
    b32 = serving.energy_error_bound(e, "float32")
    assert b32 == pytest.approx(serving.F32_SCALE_ERROR * 4601.0 + serving.F32_RECIP_ERROR * 2000.0)
    assert serving.energy_error_bound(e, "float64") == pytest.approx(1e-10 * 4601.0)
    # the total-relative bound grows as the terms cancel
    e2 = np.array([-100.0, 2000.0, -1899.0, 0.0, 1.0])
    assert serving.energy_error_bound(e2, "float32") / 1.0 > 1e-4
