"""Serving loop: continuous batching under an HBM budget.

Each served result equals the same request evaluated alone (a batch of one)
to 1e-6 relative in the float32 engine and 1e-12 in float64, dE/dbox to 10x
that in float64 and to the float32 engine's dE/dbox accuracy bar (1e-4
relative, as against the oracle in the batch tests) in float32: float32
atomics and the in-cell atom order, both fixed by timing, change float32
partial sums, and dE/dbox amplifies those differences through the cancelling
pair and atomic virials.  Results also match the oracle within the float32
bars (energy 1e-6 relative, forces 1e-4 eV/A).
"""

import sys

import numpy as np
import pytest
import torch

from oracle import nblist_ref, pme_ref
from torch_admp_amd import ops, serving

pytestmark = pytest.mark.gpu


def _alone(req, dev, precision):
    plan = ops.PMEBatchPlan([req.n_atoms], [req.box], req.rcut, req.kappa, [req.kmesh],
                            precision)
    pos = torch.tensor(req.positions, device=dev)
    q = torch.tensor(req.charges, device=dev)
    e, f, g = plan.run(pos, q)
    return e[0].cpu().numpy(), f.cpu().numpy(), g[0].cpu().numpy()


@pytest.mark.parametrize("precision,tol", [("float32", 1e-6), ("float64", 1e-12)])
def test_served_results_equal_alone(dev, precision, tol):
    reqs = serving.synthetic_stream(48, 4000.0, seed=3, n_min=64, n_max=2048)
    server = serving.PMEServer(dev, hbm_budget=64 << 20, max_systems=16, streams=2,
                               precision=precision)
    stats = server.serve(reqs)
    assert stats["requests"] == len(reqs)
    assert stats["batches"] > 1
    for r in reqs:
        e1, f1, g1 = _alone(r, dev, precision)
        assert np.abs(r.energy - e1).max() <= tol * max(1.0, np.abs(e1).max()), r.rid
        assert np.abs(r.forces - f1).max() <= tol * max(1.0, np.abs(f1).max()), r.rid
        gtol = 1e-4 if precision == "float32" else 10 * tol
        assert np.abs(r.dedbox - g1).max() <= gtol * max(1.0, np.abs(g1).max()), r.rid
        assert r.t_admit >= r.arrival and r.t_done >= r.t_admit


def test_served_results_match_oracle(dev):
    reqs = serving.synthetic_stream(6, 0.0, seed=9, n_min=100, n_max=600)
    server = serving.PMEServer(dev, hbm_budget=256 << 20, streams=2, precision="float32")
    server.serve(reqs)
    for r in reqs:
        pos, box, q = (torch.tensor(r.positions), torch.tensor(r.box), torch.tensor(r.charges))
        vol = float(torch.linalg.det(box).abs())
        heights = [vol / float(torch.linalg.cross(box[a], box[b]).norm())
                   for a, b in ((1, 2), (2, 0), (0, 1))]
        orc = pme_ref.CoulombOracle(r.rcut, r.ethresh, kmesh=r.kmesh, kappa=r.kappa)
        if min(heights) < 2 * r.rcut:
            pairs, _, bs = nblist_ref.multi_image_pairs(pos, box, r.rcut)
            ds_fn = lambda p, b: nblist_ref.multi_image_pairs(p, b, r.rcut)[1]  # noqa: E731
        else:
            pairs = nblist_ref.nblist_pbc(pos, box, r.rcut)
            bs = nblist_ref.buffer_scales(pairs)
            ds_fn = lambda p, b: nblist_ref.pairs_ds(p, pairs, b)  # noqa: E731
        e, f, _ = pme_ref.energy_forces(pos, box, q, pairs, ds_fn, bs, orc)
        assert abs(r.energy[4] - float(e)) <= 1e-6 * abs(float(e)), r.rid
        assert float(np.sqrt(((r.forces - f.numpy()) ** 2).mean())) < 1e-4, r.rid


def test_admission_respects_budget(dev):
    reqs = serving.synthetic_stream(40, 0.0, seed=4, n_min=500, n_max=4000)
    server = serving.PMEServer(dev, hbm_budget=48 << 20, max_systems=64, streams=2)
    peak = 0
    for r in reqs:
        server.submit(r)
    while not server.idle():
        server.pump()
        peak = max(peak, server.bytes_in_flight)
    assert 0 < peak <= server.peak_bytes_in_flight <= server.hbm_budget
    assert len(server.completed) == len(reqs)
    assert server.batches_launched >= 2
    big = serving.synthetic_stream(1, 0.0, seed=5, n_min=8000, n_max=8192)[0]
    with pytest.raises(ValueError, match="HBM budget"):
        serving.PMEServer(dev, hbm_budget=1 << 20).submit(big)


def test_qeq_requests_served_beside_pme(dev):
    """QEq solves run on the server's worker stream while PME batches are in
    flight, queued solves together (qeq.solve_batch): charges equal the same
    solve made directly with QEqForceModule within 1e-7 (both stop below the
    same residual, the batched solve at the first iteration below it, the
    direct one at its check every 4 iterations), the constraint holds, and the
    PME results still equal the request evaluated alone."""
    from torch_admp_amd.nblist import TorchNeighborList
    from torch_admp_amd.qeq import QEqForceModule
    from torch_admp_amd.workloads import qeq_species_box

    pme = serving.synthetic_stream(24, 0.0, seed=11, n_min=200, n_max=1500)
    qeq = []
    for k, (n, length, qtot) in enumerate(((300, 15.0, 0.0), (600, 18.0, 1.0), (450, 16.0, -2.0))):
        pos, box, chi, hard, eta = qeq_species_box(n, length, seed=20 + k)
        qeq.append(serving.QEqRequest(pos, box, chi, hard, eta, rcut=6.0, total_charge=qtot,
                                      tol=1e-10, arrival=0.0))
    switch = sys.getswitchinterval()
    with serving.PMEServer(dev, hbm_budget=256 << 20, max_systems=8, streams=2,
                           precision="float64") as server:
        stats = server.serve(pme + qeq)
    assert sys.getswitchinterval() == switch  # restored when the worker stops
    assert stats["requests"] == len(pme) + len(qeq)
    assert stats["qeq"]["requests"] == len(qeq) and stats["pme"]["requests"] == len(pme)
    assert server.bytes_in_flight == 0
    assert 1 <= server.qeq_batches <= len(qeq)
    # PME batches and QEq solves reserve from one budget under one lock
    assert 0 < server.peak_bytes_in_flight <= server.hbm_budget
    for r in qeq:
        t = lambda a: torch.tensor(a, device=dev)  # noqa: E731
        pos, box = t(r.positions), t(r.box)
        mod = QEqForceModule(r.rcut, r.ethresh, kappa=r.kappa, kmesh=r.kmesh)
        nbl = TorchNeighborList(r.rcut)
        with torch.no_grad():
            pairs = nbl(pos, box)
            A = torch.ones(1, r.n_atoms, dtype=torch.float64, device=dev)
            b = torch.full((1,), r.total_charge, dtype=torch.float64, device=dev)
            q = mod.solve_implicit(pos, box, t(r.chi), t(r.hardness), t(r.eta), pairs,
                                   nbl.get_ds(), nbl.get_buffer_scales(), A, b, tol=r.tol)
        q = q.cpu().numpy()
        assert np.abs(r.charges - q).max() <= 1e-7, r.rid
        assert abs(r.charges.sum() - r.total_charge) <= 1e-9, r.rid
        assert r.iterations > 0 and r.residual <= r.tol, r.rid
        assert r.t_done >= r.t_admit >= r.arrival
    for r in pme:
        e1, f1, _ = _alone(r, dev, "float64")
        assert np.abs(r.energy - e1).max() <= 1e-12 * max(1.0, np.abs(e1).max()), r.rid
        assert np.abs(r.forces - f1).max() <= 1e-12 * max(1.0, np.abs(f1).max()), r.rid


def test_close_cancels_queued_qeq_and_refuses_new(dev):
    """close() with QEq requests still queued: they move to ``cancelled`` with
    an error text, idle() becomes true, a later QEq submission raises, and PME
    requests are still served."""
    from torch_admp_amd.workloads import qeq_species_box

    pos, box, chi, hard, eta = qeq_species_box(300, 15.0, seed=2)
    # a budget that admits no QEq solve keeps the request queued
    server = serving.PMEServer(dev, hbm_budget=64 << 20, precision="float64")
    req = serving.QEqRequest(pos, box, chi, hard, eta, rcut=6.0)
    with server._cv:
        server.bytes_in_flight += server.hbm_budget  # hold the budget: nothing is admitted
    server.submit(req)
    assert not server.idle()
    server.close()
    with server._cv:
        server.bytes_in_flight -= server.hbm_budget
    assert server.cancelled == [req] and req.error and req.charges is None
    assert server.idle()
    with pytest.raises(RuntimeError, match="closed"):
        server.submit(serving.QEqRequest(pos, box, chi, hard, eta, rcut=6.0))
    pme = serving.synthetic_stream(3, 0.0, seed=2, n_min=100, n_max=300)
    stats = server.serve(pme)
    assert stats["requests"] == 3 and all(r.energy is not None for r in pme)


@pytest.mark.parametrize("precision,tol", [("float64", 1e-12), ("float32", 1e-6)])
def test_replica_server_replays_equal_eager(dev, precision, tol):
    """Resident replicas stepped by graph replay: after each in-place position
    update, every replay equals the eager batch of the same positions (eager
    work between the replays), and a replica equals its system alone."""
    reqs = serving.synthetic_stream(12, 0.0, seed=17, n_max=2000)
    rs = serving.ReplicaServer(reqs, dev, precision=precision)
    gen = torch.Generator(device=dev).manual_seed(5)
    for step in range(3):
        rs.positions.add_(0.02 * torch.randn(rs.positions.shape, device=dev, dtype=torch.float64,
                                             generator=gen))
        e, f, g = (t.clone() for t in rs.step())
        pos = rs.positions.clone()
        q = rs.graph.charges.clone()
        e1, f1, g1 = rs.plan.run(pos, q)
        torch.cuda.synchronize()
        scale = max(1.0, float(e1.abs().max()))
        assert float((e - e1).abs().max()) <= tol * scale, step
        assert float((f - f1).abs().max()) <= tol * max(1.0, float(f1.abs().max())), step
        assert float((g - g1).abs().max()) <= 10 * tol * max(1.0, float(g1.abs().max())), step
    r0 = reqs[0]
    n0 = r0.n_atoms
    e0, f0, _ = ops.pme_fused(rs.positions[:n0].clone(), torch.tensor(r0.box),
                              rs.graph.charges[:n0].clone(), r0.rcut, r0.kappa, r0.kmesh,
                              precision=precision)
    e, f, _ = rs.step()
    assert abs(float(e[0, 4] - e0[4])) <= tol * max(1.0, abs(float(e0[4])))
    assert float((f[:n0] - f0).abs().max()) <= tol * max(1.0, float(f0.abs().max()))
