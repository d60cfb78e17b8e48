"""Serving steady solves as they arrive, on one GPU, with continuous batching.

A `SolveServer` takes solve requests (mesh, operator, physics, initial
state) at any time and returns a `Ticket` per request.  Admission is by HBM
footprint: every request is priced with `pdes_estimate_bytes` (the device
counterpart of the reference's estimateMem, src/estimateMem.jl:1-74) and the
batches on the device never hold more than the budget.

Requests of one kind (operator and physics kind) are served by one lane: a
worker thread that owns a packed batch (`pdes_batch_*`, one launch per kernel
for all of its instances) and drives it a few passes of the batched Newton
loop at a time (`pdes_batch_advance`).  Between those calls the lane

  * hands each instance that has converged (or failed) back to its client at
    once, with its state and report;
  * takes in same-kind requests that have arrived, and gives back the memory
    of its finished instances when some request is waiting for memory, by
    repacking: the unfinished instances are exported at a GMRES restart
    (`pdes_batch_export`), the old batch is freed, and a new batch holding
    them and the newcomers continues their solves (`pdes_batch_import`).

Lanes of different kinds run concurrently on their own streams, so requests
admitted while others are on the GPU overlap with them on the device.  A
result does not depend on what it was batched with, on repacks, or on when it
arrived: the packed batch keeps every reduction, Krylov scalar and
convergence test per instance, and a moved solve continues from the exact
state it had (its Jacobian and preconditioner are rebuilt from that state).
"""
from __future__ import annotations

import ctypes as C
import threading
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import abi
from ._lib import check, lib
from .batch import BatchSolver
from .linsolve import NewtonOptions


def estimate_bytes(mesh, op, phys, opts: NewtonOptions, batched: bool = True) -> int:
    """HBM bytes one solve of this problem allocates (pdes_estimate_bytes)."""
    b = abi.make_descriptors(mesh, op, phys)
    out = C.c_int64(0)
    check(lib().pdes_estimate_bytes(C.byref(b.mesh), C.byref(b.sbp), int(opts.fp64_blocks), int(opts.pc_kind),
                                    int(opts.gmres_restart), int(batched), C.byref(out)))
    return int(out.value)


def device_memory():
    """(free, total) bytes of device memory."""
    f, t = C.c_int64(0), C.c_int64(0)
    check(lib().pdes_device_memory(C.byref(f), C.byref(t)))
    return int(f.value), int(t.value)


def _kind(mesh, op, phys):
    """Requests of one kind can share a packed group (same operator and
    physics kind)."""
    return (mesh.dim, op.kind, op.degree, phys.volume_integral_type, phys.face_flux, phys.gamma, phys.R_gas,
            phys.check_positivity)


@dataclass
class Ticket:
    mesh: object
    op: object
    phys: object
    q0: np.ndarray
    nbytes: int
    t_submit: float
    t_admit: float = 0.0
    t_done: float = 0.0
    q: np.ndarray | None = None
    report: object = None
    error: str | None = None
    batch_size: int = 0
    _event: threading.Event = field(default_factory=threading.Event)

    def done(self) -> bool:
        return self._event.is_set()

    def wait(self, timeout: float | None = None) -> bool:
        return self._event.wait(timeout)

    @property
    def latency(self) -> float:
        return self.t_done - self.t_submit


class _Lane:
    """The running batch of one request kind."""

    def __init__(self, kind):
        self.kind = kind
        self.batch: BatchSolver | None = None
        self.members: list[Ticket] = []     # instance i of the batch is members[i]
        self.live: list[bool] = []          # still being solved
        self.nbytes = 0                     # bytes admitted for the current batch (finished members included)
        self.thread: threading.Thread | None = None

    def finished_bytes(self):
        return sum(t.nbytes for t, a in zip(self.members, self.live) if not a)

    def live_members(self):
        return [t for t, a in zip(self.members, self.live) if a]


class SolveServer:
    def __init__(self, hbm_budget_bytes: int, lanes: int = 4, max_batch: int = 32,
                 opts: NewtonOptions | None = None, passes: int = 2):
        self.budget = int(hbm_budget_bytes)
        self.opts = opts or NewtonOptions(res_abstol=1e-10)
        self.max_batch = max_batch
        self.max_lanes = lanes
        self.passes = passes
        self._queue: deque[Ticket] = deque()
        self._cv = threading.Condition()
        self._lanes: dict = {}
        self._in_use = 0
        self._stop = False
        self.peak_bytes = 0
        self.batches = 0          # batches created
        self.repacks = 0          # of which repacks of a running batch

    # ---- client side
    def submit(self, mesh, op, phys, q0) -> Ticket:
        nbytes = estimate_bytes(mesh, op, phys, self.opts)
        if nbytes > self.budget:
            raise ValueError(f"request needs {nbytes} bytes of HBM, above the {self.budget}-byte budget")
        t = Ticket(mesh, op, phys, np.ascontiguousarray(q0, dtype=np.float64), nbytes, time.perf_counter())
        with self._cv:
            if self._stop:
                raise RuntimeError("server is shut down")
            self._queue.append(t)
            self._start_lanes()
            self._cv.notify_all()
        return t

    def shutdown(self):
        """Serves what has been submitted, then stops the lanes."""
        with self._cv:
            self._stop = True
            self._cv.notify_all()
            while self._lanes or self._queue:
                self._start_lanes()
                self._cv.wait(0.05)

    # ---- admission (under the lock)
    def _start_lanes(self):
        for t in list(self._queue):
            k = _kind(t.mesh, t.op, t.phys)
            if k in self._lanes or len(self._lanes) >= self.max_lanes:
                continue
            lane = _Lane(k)
            lane.thread = threading.Thread(target=self._run_lane, args=(lane,), name="pdes-lane", daemon=True)
            self._lanes[k] = lane
            lane.thread.start()

    def _take(self, lane, reclaim):
        """Waiting requests of the lane's kind, oldest first, that fit the
        budget with the lane's batch replaced by its live members plus them
        (reclaim) or added to it."""
        keep = sum(t.nbytes for t in lane.live_members()) if reclaim else lane.nbytes
        room = self.budget - (self._in_use - lane.nbytes + keep)
        picked = []
        nlive = len(lane.live_members())
        for t in list(self._queue):
            if nlive + len(picked) >= self.max_batch:
                break
            if _kind(t.mesh, t.op, t.phys) == lane.kind and t.nbytes <= room:
                picked.append(t)
                room -= t.nbytes
        for t in picked:
            self._queue.remove(t)
        return picked, keep

    def _reserve(self, lane, new_bytes):
        self._in_use += new_bytes - lane.nbytes
        lane.nbytes = new_bytes
        self.peak_bytes = max(self.peak_bytes, self._in_use)

    # ---- lanes
    def _run_lane(self, lane):
        try:
            while True:
                with self._cv:
                    finished = lane.finished_bytes()
                    if lane.batch is None:
                        joiners, keep = self._take(lane, reclaim=True)
                        if not joiners:
                            if any(_kind(t.mesh, t.op, t.phys) == lane.kind for t in self._queue):
                                # a request of this kind waits for memory another lane holds
                                self._cv.wait(0.05)
                                continue
                            del self._lanes[lane.kind]
                            self._start_lanes()
                            self._cv.notify_all()
                            return
                        plan = "new"
                    else:
                        joiners, keep = self._take(lane, reclaim=True)
                        # give finished instances' memory back when a request waits for memory
                        waiting = bool(self._queue)
                        plan = "repack" if joiners or (finished and waiting) else "run"
                    if plan != "run":
                        self._reserve(lane, keep + sum(t.nbytes for t in joiners))
                if plan != "run":
                    self._repack(lane, joiners)
                    with self._cv:
                        self._cv.notify_all()
                reps, active, done = lane.batch.advance(self.opts, self.passes)
                self._harvest(lane, reps, active)
                if done or not any(lane.live):
                    lane.batch.free()
                    with self._cv:
                        lane.batch = None
                        lane.members, lane.live = [], []
                        self._in_use -= lane.nbytes
                        lane.nbytes = 0
                        self._cv.notify_all()
        except Exception as exc:  # the lane's requests fail loudly to their clients
            with self._cv:
                for t in lane.live_members():
                    t.error = str(exc)
                    t.t_done = time.perf_counter()
                    t._event.set()
                if lane.batch is not None:
                    lane.batch.free()
                self._in_use -= lane.nbytes
                self._lanes.pop(lane.kind, None)
                self._cv.notify_all()

    def _repack(self, lane, joiners):
        """New batch = the lane's live instances (their solves moved over) +
        the joiners (fresh solves); the old batch is freed first."""
        carried = []
        if lane.batch is not None:
            for i, (t, a) in enumerate(zip(lane.members, lane.live)):
                if a:
                    carried.append((t, lane.batch.export(i)))
            lane.batch.free()
            lane.batch = None
            self.repacks += 1
        members = [t for t, _ in carried] + list(joiners)
        bs = BatchSolver([(t.mesh, t.op, t.phys) for t in members])
        for i, (t, (st, q, dq, rhs)) in enumerate(carried):
            bs.import_(i, st, q, dq, rhs)
        now = time.perf_counter()
        for i, t in enumerate(joiners):
            bs.set_state(len(carried) + i, t.q0)
            t.t_admit = now
        for t in members:
            t.batch_size = max(t.batch_size, len(members))
        lane.batch, lane.members, lane.live = bs, members, [True] * len(members)
        self.batches += 1

    def _harvest(self, lane, reps, active):
        """Hands back, at once, every instance that finished in the last
        advance."""
        for i, t in enumerate(lane.members):
            if lane.live[i] and not active[i]:
                t.q = lane.batch.get_state(i)
                t.report = reps[i]
                t.t_done = time.perf_counter()
                lane.live[i] = False
                t._event.set()
