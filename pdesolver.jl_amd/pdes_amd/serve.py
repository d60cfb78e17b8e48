"""Serving steady solves as they arrive, on one GPU.

A `SolveServer` takes solve requests (mesh, operator, physics, initial
state) at any time and returns a `Ticket` per request.  Admission is by HBM
footprint: every request is priced with `pdes_estimate_bytes` (the device
counterpart of the reference's estimateMem, src/estimateMem.jl:1-74) and a
request is admitted only while the admitted bytes stay within the budget.

Work is done by `lanes` worker threads, each owning its own device stream
through its own packed batch.  A free lane takes the oldest waiting request
and every other waiting request with the same operator and physics kind that
still fits the budget (up to `max_batch`), packs them into one batch
(`pdes_batch_*`, one launch per kernel for all of them), runs Newton with
converged instances retired between Newton steps, and hands each result back.
Lanes run concurrently, so requests admitted while others are on the GPU
overlap with them on the device; memory is released as each batch finishes
and waiting requests are admitted as slots and memory free up.

Each result is independent of what it was batched with: the packed batch keeps
every reduction, Krylov scalar and convergence test per instance.
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


class SolveServer:
    def __init__(self, hbm_budget_bytes: int, lanes: int = 4, max_batch: int = 32,
                 opts: NewtonOptions | None = None):
        self.budget = int(hbm_budget_bytes)
        self.opts = opts or NewtonOptions(res_abstol=1e-10)
        self.max_batch = max_batch
        self._queue: deque[Ticket] = deque()
        self._cv = threading.Condition()
        self._in_use = 0
        self._stop = False
        self.peak_bytes = 0
        self.batches = 0
        self._lanes = [threading.Thread(target=self._lane, name=f"pdes-lane-{i}", daemon=True) for i in range(lanes)]
        for t in self._lanes:
            t.start()

    # ---- client side
    def submit(self, mesh, op, phys, q0) -> Ticket:
        nbytes = estimate_bytes(mesh, op, phys, self.opts)
        if nbytes > self.budget:
            raise ValueError(f"request needs {nbytes} bytes of HBM, above the {self.budget}-byte budget")
        t = Ticket(mesh, op, phys, np.ascontiguousarray(q0, dtype=np.float64), nbytes, time.perf_counter())
        with self._cv:
            self._queue.append(t)
            self._cv.notify_all()
        return t

    def shutdown(self):
        with self._cv:
            self._stop = True
            self._cv.notify_all()
        for t in self._lanes:
            t.join()

    # ---- lanes
    def _take(self):
        """Under the lock: the oldest request that fits, plus same-kind
        requests behind it that also fit.  None if nothing fits yet."""
        if not self._queue:
            return None
        head = self._queue[0]
        if self._in_use + head.nbytes > self.budget:
            return None
        kind = _kind(head.mesh, head.op, head.phys)
        picked = [head]
        used = self._in_use + head.nbytes
        for t in list(self._queue)[1:]:
            if len(picked) >= self.max_batch:
                break
            if _kind(t.mesh, t.op, t.phys) == kind and used + t.nbytes <= self.budget:
                picked.append(t)
                used += t.nbytes
        for t in picked:
            self._queue.remove(t)
        self._in_use = used
        self.peak_bytes = max(self.peak_bytes, used)
        return picked

    def _lane(self):
        while True:
            with self._cv:
                while True:
                    if self._stop and not self._queue:
                        return
                    picked = self._take()
                    if picked is not None:
                        break
                    self._cv.wait(0.05)
            now = time.perf_counter()
            for t in picked:
                t.t_admit = now
                t.batch_size = len(picked)
            bs = None
            try:
                bs = BatchSolver([(t.mesh, t.op, t.phys) for t in picked])
                for i, t in enumerate(picked):
                    bs.set_state(i, t.q0)
                reps = bs.solve(self.opts)
                for i, t in enumerate(picked):
                    t.q = bs.get_state(i)
                    t.report = reps[i]
            except Exception as exc:  # the request fails loudly to its client
                for t in picked:
                    t.error = str(exc)
            finally:
                if bs is not None:
                    bs.free()
                done = time.perf_counter()
                with self._cv:
                    self._in_use -= sum(t.nbytes for t in picked)
                    self.batches += 1
                    self._cv.notify_all()
                for t in picked:
                    t.t_done = done
                    t._event.set()
