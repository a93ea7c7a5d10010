"""Continuous-batching request loop for PME energy + force evaluations and
QEq charge solves.

Requests of any atom count, box and mesh arrive over time.  The server keeps
a FIFO queue and up to ``streams`` batches in flight, one per HIP stream.
Whenever a stream is free it forms a batch from the head of the queue, as
many requests as fit the HBM budget that in-flight batches leave over (and
at most ``max_systems``), copies their positions and charges to the device
once, and launches the whole batch as one admp_pme_batch launch sequence.
Requests that arrive while other batches run are admitted as soon as a
stream and memory free up, so their work overlaps on the device with the
batches already there.  Each request's result is the same whether it runs
alone or in a mix (the batched kernels treat every system independently).

QEq requests (charge equilibration) share the HBM budget.  A worker thread
admits the queued QEq requests in arrival order as the budget allows (up to
``max_qeq_batch`` at a time) and solves them together on its own stream,
concurrently with the PME batches: admp_qeq_solve_batch forms all their
Hessian products in one launch sequence per iteration and each solve stops at
its own tolerance (a device-side flag, read every few iterations).

The reference evaluates one frame per CoulombForceModule call
(src/torch_admp/pme.py:261-310); this module has no counterpart there.
"""

from __future__ import annotations

import math
import sys
import threading
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Deque, List, Optional, Union

import numpy as np
import torch

from . import _lib
from .ops import PMEBatchPlan
from .utils import DIELECTRIC


# Stated error of a float32 result's energies against the float64 engine (which
# equals the oracle to 1e-10), per result:
#   |dE_total| <= F32_SCALE_ERROR (|E_real| + |E_recip| + |E_self| + |E_nn|)
#                 + F32_RECIP_ERROR |E_recip|.
# Measured over 512 requests of the serving stream (64-8192 atoms, orthorhombic
# and triclinic): the real-space error at most 5.2e-8 of that scale (float32
# pair terms), the reciprocal error at most 1.9e-7 of E_recip (float32 FFT
# rounding: an exact-twiddle float32 FFT of the same float64 mesh gives up to
# 1.3e-7); the constants keep 2x and 1.6x above those.  The total itself is
# often a small difference of the terms (random +-1 ions: scale / |E_total| has
# median 6.5 and reaches ~1000), so the relative error of the total is the
# stated bound over |E_total|.
F32_SCALE_ERROR = 1.0e-7
F32_RECIP_ERROR = 3.0e-7
# the float64 engine's stated error (its tested bound against the oracle)
F64_SCALE_ERROR = 1.0e-10


def energy_error_bound(energy: np.ndarray, precision: str) -> float:
    """Stated bound on |E_total - E_total(float64 oracle)| for one result's
    energies (real, reciprocal, self, non-neutral, total)."""
    scale = float(np.abs(np.asarray(energy[:4], dtype=np.float64)).sum())
    if precision == "float32":
        return F32_SCALE_ERROR * scale + F32_RECIP_ERROR * abs(float(energy[1]))
    return F64_SCALE_ERROR * scale


def ewald_parameters(box: np.ndarray, rcut: float, ethresh: float):
    """kappa and mesh of CoulombForceModule for a box (src/torch_admp/pme.py:113-115,383-392):
    kappa = sqrt(-ln(2 ethresh)) / rcut, K_d = ceil(2 kappa h_dd / (3 ethresh^(1/5)))."""
    kappa = math.sqrt(-math.log(2.0 * ethresh)) / rcut
    diag = np.diagonal(np.asarray(box, dtype=np.float64))
    kmesh = [int(math.ceil(2.0 * kappa * d / (3.0 * ethresh ** 0.2))) for d in diag]
    return kappa, kmesh


@dataclass
class PMERequest:
    """One energy + force evaluation: positions (n, 3), box (3, 3) rows,
    charges (n,), cutoff; kappa and mesh follow CoulombForceModule's rule for
    ``ethresh`` unless given.  ``arrival`` is in seconds from the stream start."""

    positions: np.ndarray
    box: np.ndarray
    charges: np.ndarray
    rcut: float
    ethresh: float = 1e-5
    kappa: Optional[float] = None
    kmesh: Optional[List[int]] = None
    arrival: float = 0.0
    rid: int = -1
    # results (host float64)
    energy: Optional[np.ndarray] = None  # real, reciprocal, self, non-neutral, total
    forces: Optional[np.ndarray] = None
    dedbox: Optional[np.ndarray] = None
    energy_error: float = math.nan  # stated bound on the total's error (energy_error_bound)
    t_admit: float = math.nan
    t_done: float = math.nan
    cost: int = 0

    def __post_init__(self):
        self.positions = np.ascontiguousarray(self.positions, dtype=np.float64).reshape(-1, 3)
        self.charges = np.ascontiguousarray(self.charges, dtype=np.float64).reshape(-1)
        self.box = np.asarray(self.box, dtype=np.float64).reshape(3, 3)
        if self.kappa is None or self.kmesh is None:
            kappa, kmesh = ewald_parameters(self.box, self.rcut, self.ethresh)
            self.kappa = kappa if self.kappa is None else self.kappa
            self.kmesh = kmesh if self.kmesh is None else list(self.kmesh)

    @property
    def n_atoms(self) -> int:
        return self.positions.shape[0]

    @property
    def latency(self) -> float:
        return self.t_done - self.arrival


@dataclass
class QEqRequest:
    """One charge equilibration: q* = argmin E(q) on sum q = total_charge for
    positions (n, 3), box (3, 3) rows, chi, hardness (the q^2 coefficient) and
    Gaussian widths eta (n,), to |P grad E| / n <= tol (float64 solve).  Mesh
    and kappa follow CoulombForceModule's rule for ``ethresh``."""

    positions: np.ndarray
    box: np.ndarray
    chi: np.ndarray
    hardness: np.ndarray
    eta: np.ndarray
    rcut: float
    ethresh: float = 1e-5
    total_charge: float = 0.0
    tol: float = 1e-10
    arrival: float = 0.0
    rid: int = -1
    # results
    charges: Optional[np.ndarray] = None
    iterations: int = -1
    residual: float = math.nan
    error: Optional[str] = None
    t_admit: float = math.nan
    t_done: float = math.nan
    cost: int = 0

    def __post_init__(self):
        self.positions = np.ascontiguousarray(self.positions, dtype=np.float64).reshape(-1, 3)
        n = self.positions.shape[0]
        for name in ("chi", "hardness", "eta"):
            v = np.ascontiguousarray(getattr(self, name), dtype=np.float64).reshape(-1)
            if v.shape[0] != n:
                raise ValueError(f"{name}: expected {n} values, got {v.shape[0]}")
            setattr(self, name, v)
        self.box = np.asarray(self.box, dtype=np.float64).reshape(3, 3)
        self.kappa, self.kmesh = ewald_parameters(self.box, self.rcut, self.ethresh)

    @property
    def n_atoms(self) -> int:
        return self.positions.shape[0]

    @property
    def latency(self) -> float:
        return self.t_done - self.arrival


def _stats(reqs: List, wall: float) -> dict:
    lat = np.array([r.latency for r in reqs]) * 1e3
    if len(lat) == 0:
        return {"requests": 0}
    return {"requests": len(reqs), "throughput_req_s": len(reqs) / wall,
            "atoms_per_s": sum(r.n_atoms for r in reqs) / wall,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean())}


@dataclass
class _Batch:
    requests: List[PMERequest]
    plan: PMEBatchPlan
    stream: torch.cuda.Stream
    done: torch.cuda.Event
    host_energy: torch.Tensor
    host_forces: torch.Tensor
    host_dedbox: torch.Tensor
    keep: list = field(default_factory=list)  # device buffers alive until the batch ends
    cost: int = 0


class PMEServer:
    """Continuous batching of PME requests, and QEq solves beside them, under an
    HBM budget (bytes)."""

    def __init__(self, device="cuda", hbm_budget: Optional[int] = None, max_systems: int = 1024,
                 streams: int = 2, precision: str = "float32", convention: str = "reference",
                 dielectric: float = DIELECTRIC, max_qeq_batch: int = 64):
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise RuntimeError("PMEServer runs only on a ROCm GPU (no CPU fallback)")
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            hbm_budget = int(free * 0.5)
        self.hbm_budget = int(hbm_budget)
        self.max_systems = int(max_systems)
        self.max_qeq_batch = max(1, int(max_qeq_batch))
        self.qeq_batches = 0
        self.precision = precision
        self.convention = convention
        self.dielectric = float(dielectric)
        self.streams = [torch.cuda.Stream(self.device) for _ in range(max(1, streams))]
        self.queue: Deque[PMERequest] = deque()
        self.inflight: List[Optional[_Batch]] = [None] * len(self.streams)
        self.bytes_in_flight = 0
        self.peak_bytes_in_flight = 0  # highest reservation seen, PME and QEq together
        self.completed: List[PMERequest] = []
        self.batches_launched = 0
        self._next_id = 0
        self._t0 = time.perf_counter()
        # QEq: queue, worker thread and the lock that also guards bytes_in_flight
        self.qeq_queue: Deque[QEqRequest] = deque()
        self._cv = threading.Condition()
        self._qeq_done: List[QEqRequest] = []
        self._qeq_busy = False
        self._stop = False
        self._closed = False
        self.cancelled: List[QEqRequest] = []  # QEq requests still queued at close()
        self._worker: Optional[threading.Thread] = None
        self._switch: Optional[float] = None

    # ------------------------------------------------------------------ admission
    def request_cost(self, req: PMERequest) -> int:
        """HBM bytes a request holds while in flight: its share of the batch
        workspace plus positions, charges and results."""
        return self.request_costs([req])[0]

    def request_costs(self, reqs: List[PMERequest]) -> List[int]:
        """request_cost of many requests from one library call."""
        n = len(reqs)
        if n == 0:
            return []
        n_atoms = np.fromiter((r.n_atoms for r in reqs), dtype=np.int64, count=n)
        boxes = np.empty((n, 9), dtype=np.float64)
        for k, r in enumerate(reqs):
            boxes[k] = np.asarray(r.box, dtype=np.float64).reshape(9)
        cut = np.fromiter((r.rcut for r in reqs), dtype=np.float64, count=n)
        km = np.asarray([r.kmesh for r in reqs], dtype=np.int32).reshape(n, 3)
        out = np.empty(n, dtype=np.uint64)
        lib = _lib.load()
        _lib.check(lib.admp_pme_batch_system_bytes(
            n, n_atoms.ctypes.data, boxes.ctypes.data, cut.ctypes.data, km.ctypes.data,
            _lib.PRECISION[self.precision], out.ctypes.data), "admp_pme_batch_system_bytes")
        return [int(b) + 56 * int(a) + 8 * 14 + 4096 for b, a in zip(out, n_atoms)]

    def now(self) -> float:
        return time.perf_counter() - self._t0

    @staticmethod
    def qeq_cost(req: QEqRequest) -> int:
        """HBM bytes a QEq solve holds: its pair list (cutoff sphere at the
        request's density), per-atom arrays and the reciprocal mesh buffers."""
        n = req.n_atoms
        vol = abs(float(np.linalg.det(req.box)))
        pairs = n * (4.0 / 3.0 * math.pi * req.rcut ** 3 * n / max(vol, 1e-30)) / 2.0
        mesh = int(np.prod(req.kmesh))
        return int(64 * pairs + 512 * n + 48 * mesh + (1 << 20))

    def submit_many(self, reqs: List[Union[PMERequest, QEqRequest]]) -> List[int]:
        """submit() of several requests, the PME costs from one library call."""
        pme = [r for r in reqs if not isinstance(r, QEqRequest)]
        for r, c in zip(pme, self.request_costs(pme)):
            r.cost = c
        return [self.submit(r, _costed=not isinstance(r, QEqRequest)) for r in reqs]

    def submit(self, req: Union[PMERequest, QEqRequest], _costed: bool = False) -> int:
        if not _costed:
            req.cost = (self.qeq_cost(req) if isinstance(req, QEqRequest)
                        else self.request_cost(req))
        if req.cost > self.hbm_budget:
            raise ValueError(f"request of {req.n_atoms} atoms needs {req.cost} bytes, more than "
                             f"the HBM budget of {self.hbm_budget}")
        req.rid = self._next_id
        self._next_id += 1
        if isinstance(req, QEqRequest):
            with self._cv:
                if self._closed:
                    raise RuntimeError("PMEServer is closed: QEq requests are no longer served")
                if self._worker is None:
                    # The serving loop polls and the worker returns from a library
                    # call every few CG iterations: with Python's default 5 ms
                    # thread switch interval each side waits up to 5 ms for the
                    # other at every hand-over.  0.1 ms while the worker lives.
                    self._switch = sys.getswitchinterval()
                    sys.setswitchinterval(min(self._switch, 1e-4))
                    self._worker = threading.Thread(target=self._qeq_worker, daemon=True)
                    self._worker.start()
                self.qeq_queue.append(req)
                self._cv.notify_all()
        else:
            self.queue.append(req)
        return req.rid

    # ------------------------------------------------------------------ QEq
    def _qeq_worker(self) -> None:
        torch.cuda.set_device(self.device)
        stream = torch.cuda.Stream(self.device)
        while True:
            with self._cv:
                while not self._stop and not (
                        self.qeq_queue and
                        self.bytes_in_flight + self.qeq_queue[0].cost <= self.hbm_budget):
                    self._cv.wait(0.01)
                if self._stop:
                    return
                # every queued solve that fits the budget, in arrival order
                take: List[QEqRequest] = []
                cost = 0
                while (self.qeq_queue and len(take) < self.max_qeq_batch and
                       self.bytes_in_flight + cost + self.qeq_queue[0].cost <= self.hbm_budget):
                    r = self.qeq_queue.popleft()
                    take.append(r)
                    cost += r.cost
                self._reserve(cost)
                self._qeq_busy = True
            t = self.now()
            for r in take:
                r.t_admit = t
            try:
                with torch.cuda.stream(stream):
                    self._solve_qeq_batch(take)
            except Exception as exc:  # reported by pump() on the serving thread
                for r in take:
                    r.error = repr(exc)
            t = self.now()
            for r in take:
                r.t_done = t
            with self._cv:
                self.bytes_in_flight -= cost
                self._qeq_busy = False
                self._qeq_done.extend(take)
                self.qeq_batches += 1
                self._cv.notify_all()

    def _solve_qeq_batch(self, reqs: List[QEqRequest]) -> None:
        """The solves of `reqs` together (qeq.solve_batch): one Hessian product
        for all of them per iteration, each stopping at its own tolerance."""
        from .qeq import solve_batch

        def dev(a):
            return torch.tensor(a, dtype=torch.float64, device=self.device)

        with torch.no_grad():
            qs, iters, resid = solve_batch(
                [dev(r.positions) for r in reqs], [dev(r.box) for r in reqs],
                [dev(r.chi) for r in reqs], [dev(r.hardness) for r in reqs],
                [dev(r.eta) for r in reqs], [r.rcut for r in reqs],
                total_charge=[float(r.total_charge) for r in reqs], tol=[r.tol for r in reqs],
                kappa=[r.kappa for r in reqs], kmesh=[r.kmesh for r in reqs],
                convention=self.convention)
        for r, q, it, res in zip(reqs, qs, iters, resid):
            r.charges = q.cpu().numpy()  # waits for this stream only
            r.iterations = it
            r.residual = res
            if it < 0:
                r.error = f"QEq solve did not converge (|P grad E| / n = {res:.3e})"

    def __enter__(self) -> "PMEServer":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def close(self) -> None:
        """Stop the QEq worker and restore the thread switch interval it set.
        QEq requests still queued are not served: they get an error text and
        move to ``cancelled``, so that idle() and drain() see an empty queue.
        Later QEq submissions raise; PME requests are still served."""
        with self._cv:
            self._stop = True
            self._closed = True
            self._cv.notify_all()
        if self._worker is not None:
            self._worker.join()
            self._worker = None
            sys.setswitchinterval(self._switch)
        with self._cv:
            while self.qeq_queue:
                r = self.qeq_queue.popleft()
                r.error = "PMEServer closed before the request was served"
                self.cancelled.append(r)

    def _reserve(self, cost: int) -> None:
        """Add `cost` to the bytes in flight; the caller holds self._cv."""
        self.bytes_in_flight += cost
        self.peak_bytes_in_flight = max(self.peak_bytes_in_flight, self.bytes_in_flight)

    # ------------------------------------------------------------------ batches
    def _launch(self, slot: int) -> bool:
        take: List[PMERequest] = []
        cost = 0
        # pick the batch and reserve its bytes in one hold of the lock, so that
        # the QEq worker's admission sees them
        with self._cv:
            room = self.hbm_budget - self.bytes_in_flight
            while self.queue and len(take) < self.max_systems:
                r = self.queue[0]
                if cost + r.cost > room:
                    break
                take.append(self.queue.popleft())
                cost += r.cost
            self._reserve(cost)
        if not take:
            return False
        try:
            self._start_batch(slot, take, cost)
        except BaseException:
            with self._cv:
                self.bytes_in_flight -= cost
                self._cv.notify_all()
            self.queue.extendleft(reversed(take))
            raise
        self.batches_launched += 1
        return True

    def _start_batch(self, slot: int, take: List[PMERequest], cost: int) -> None:
        plan = PMEBatchPlan([r.n_atoms for r in take], [r.box for r in take],
                            [r.rcut for r in take], [r.kappa for r in take],
                            [r.kmesh for r in take], self.precision, self.convention,
                            self.dielectric)
        n = plan.n_total
        stream = self.streams[slot]
        t = self.now()
        for r in take:
            r.t_admit = t
        host_in = torch.empty(n, 4, dtype=torch.float64, pin_memory=True)
        if n:
            host_in.numpy()[:, :3] = np.concatenate([r.positions for r in take])
            host_in.numpy()[:, 3] = np.concatenate([r.charges for r in take])
        with torch.cuda.stream(stream):
            dev_in = host_in.to(self.device, non_blocking=True)
            pos = dev_in[:, :3].contiguous()
            q = dev_in[:, 3].contiguous()
            ws = _lib.workspace(plan.workspace_bytes, self.device)
            # every kernel on this batch's stream: the library's real-space side
            # stream is one per device, and two batches in flight sharing it
            # wait on each other (p50 15-20 ms instead of 1.5-2 ms at 70 % load)
            energy, forces, dedbox = plan.run(pos, q, workspace=ws, overlap=False)
            h_e = torch.empty(energy.shape, dtype=torch.float64, pin_memory=True)
            h_f = torch.empty(forces.shape, dtype=torch.float64, pin_memory=True)
            h_g = torch.empty(dedbox.shape, dtype=torch.float64, pin_memory=True)
            h_e.copy_(energy, non_blocking=True)
            h_f.copy_(forces, non_blocking=True)
            h_g.copy_(dedbox, non_blocking=True)
            done = torch.cuda.Event()
            done.record(stream)
        self.inflight[slot] = _Batch(take, plan, stream, done, h_e, h_f, h_g,
                                     [host_in, dev_in, pos, q, ws, energy, forces, dedbox], cost)

    def _finish(self, slot: int) -> List[PMERequest]:
        b = self.inflight[slot]
        t = self.now()
        e = b.host_energy.numpy()
        f = b.host_forces.numpy()
        g = b.host_dedbox.numpy()
        off = b.plan.offsets
        for k, r in enumerate(b.requests):
            r.energy = e[k].copy()
            r.forces = f[off[k]:off[k + 1]].copy()
            r.dedbox = g[k].copy()
            r.energy_error = energy_error_bound(r.energy, self.precision)
            r.t_done = t
        self.inflight[slot] = None
        with self._cv:
            self.bytes_in_flight -= b.cost
            self._cv.notify_all()
        self.completed.extend(b.requests)
        return b.requests

    def pump(self) -> List[PMERequest]:
        """Collect finished batches, then fill free streams from the queue.
        Returns the requests completed by this call."""
        out: List[PMERequest] = []
        for slot, b in enumerate(self.inflight):
            if b is not None and b.done.query():
                out.extend(self._finish(slot))
        for slot, b in enumerate(self.inflight):
            if b is None and self.queue:
                if not self._launch(slot):
                    break
        with self._cv:
            done, self._qeq_done = self._qeq_done, []
        for r in done:
            if r.error is not None:
                raise RuntimeError(f"QEq request {r.rid} failed: {r.error}")
        self.completed.extend(done)
        out.extend(done)
        return out

    def idle(self) -> bool:
        with self._cv:
            qeq_idle = not self.qeq_queue and not self._qeq_busy and not self._qeq_done
        return qeq_idle and not self.queue and all(b is None for b in self.inflight)

    def drain(self) -> None:
        while not self.idle():
            self.pump()

    # ------------------------------------------------------------------ streams
    def serve(self, requests: List[PMERequest]) -> dict:
        """Replay requests at their arrival times (seconds from now) and return
        latency and throughput statistics."""
        order = sorted(requests, key=lambda r: r.arrival)
        self.completed = []
        self._t0 = time.perf_counter()
        i = 0
        while i < len(order) or not self.idle():
            now = self.now()
            i0 = i
            # arrivals due by now, at most one batch's worth per pass so that
            # launching overlaps the admission of a large backlog
            while (i < len(order) and order[i].arrival <= now and
                   i - i0 < self.max_systems):
                i += 1
            if i > i0:
                self.submit_many(order[i0:i])
            if not self.pump() and i == i0 and self._worker is not None:
                time.sleep(5e-5)  # nothing moved: let the QEq worker have the interpreter
        wall = self.now()
        out = dict(_stats(self.completed, wall), wall_s=wall, batches=self.batches_launched)
        qeq = [r for r in self.completed if isinstance(r, QEqRequest)]
        if qeq:
            out["pme"] = _stats([r for r in self.completed if isinstance(r, PMERequest)], wall)
            out["qeq"] = _stats(qeq, wall)
        return out


class ReplicaServer:
    """Resident systems evaluated again and again with new positions: MD
    replicas, or repeated queries of the same systems.

    Boxes, cutoffs, meshes and charges are fixed at construction and stay on
    the device, with the batch plan's system table; the whole batch is
    captured once in a HIP graph (ops.PMEGraph).  A step copies the new
    positions into the graph's position buffer on the device and replays the
    graph: no host staging, no planning and one graph launch instead of the
    ~14 kernel launches of a batch.  Results are device tensors of the graph
    (valid until the next step).  ``replicas`` are PMERequest objects giving
    the systems (their positions are the initial ones)."""

    def __init__(self, replicas: List[PMERequest], device="cuda", precision: str = "float32",
                 convention: str = "reference", dielectric: float = DIELECTRIC):
        from .ops import PMEGraph

        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise RuntimeError("ReplicaServer runs only on a ROCm GPU (no CPU fallback)")
        self.replicas = list(replicas)
        self.plan = PMEBatchPlan([r.n_atoms for r in replicas], [r.box for r in replicas],
                                 [r.rcut for r in replicas], [r.kappa for r in replicas],
                                 [r.kmesh for r in replicas], precision, convention, dielectric)
        self.offsets = self.plan.offsets
        self.graph = PMEGraph(self.plan, self.device)
        pos = np.concatenate([r.positions for r in replicas]) if replicas else np.zeros((0, 3))
        q = np.concatenate([r.charges for r in replicas]) if replicas else np.zeros(0)
        self.graph.positions.copy_(torch.as_tensor(pos, dtype=torch.float64))
        self.graph.charges.copy_(torch.as_tensor(q, dtype=torch.float64))

    @property
    def positions(self) -> torch.Tensor:
        """The resident positions (n_total, 3), replicas back to back; update
        them in place (or pass new ones to step)."""
        return self.graph.positions

    def step(self, positions: Optional[torch.Tensor] = None):
        """(energies (n_rep, 5), forces (n_total, 3), dE/dbox (n_rep, 3, 3)) for
        the resident positions, after copying ``positions`` (a device tensor of
        the same layout) into them when given."""
        if positions is not None:
            self.graph.positions.copy_(positions.detach().reshape(-1, 3))
        return self.graph.replay()


def synthetic_stream(n_requests: int, rate_hz: float, seed: int = 0, n_min: int = 64,
                     n_max: int = 8192, density: float = 0.1, rcut_range=(6.0, 8.0),
                     triclinic_fraction: float = 0.5, ethresh: float = 1e-5) -> List[PMERequest]:
    """Poisson arrivals of PME requests: atom counts log-uniform in [n_min, n_max],
    orthorhombic or triclinic boxes at ``density`` atoms per A^3 (liquid water is
    0.1), uniform positions and random neutral +-1 charges."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate_hz, n_requests)) if rate_hz > 0 else np.zeros(n_requests)
    reqs = []
    for k in range(n_requests):
        n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        n += n % 2
        vol = n / density
        aspect = rng.uniform(0.8, 1.25, 3)
        edges = aspect * (vol / aspect.prod()) ** (1.0 / 3.0)
        box = np.diag(edges)
        if rng.random() < triclinic_fraction:
            # lower-triangular tilt keeps the volume (and the diagonal used by the mesh rule)
            box[1, 0] = rng.uniform(-0.3, 0.3) * edges[0]
            box[2, 0] = rng.uniform(-0.3, 0.3) * edges[0]
            box[2, 1] = rng.uniform(-0.3, 0.3) * edges[1]
        pos = rng.random((n, 3)) @ box
        q = np.r_[np.ones(n // 2), -np.ones(n - n // 2)]
        rng.shuffle(q)
        reqs.append(PMERequest(pos, box, q, float(rng.uniform(*rcut_range)), ethresh,
                               arrival=float(t[k])))
    return reqs
