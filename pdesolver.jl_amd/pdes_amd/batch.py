"""Batched and served steady solves on one GPU.

`BatchSolver` wraps the packed-batch C ABI (pdes_batch_*): many independent
problems (meshes of different sizes, operators of different degree/family,
different freestream / inflow Mach) solved together; instances sharing an
operator are packed into one group whose kernels cover all its unconverged
instances per launch, and groups run concurrently on separate streams.
Per-instance reductions make each result independent of the batch mix.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from ._lib import check, lib
from .linsolve import NewtonOptions


class BatchSolver:
    def __init__(self, problems):
        """problems: list of (mesh, sbp, phys)."""
        self.problems = list(problems)
        n = len(self.problems)
        self._bundles = [abi.make_descriptors(m, op, ph) for (m, op, ph) in self.problems]
        self._meshes = (abi.MeshDesc * n)(*[b.mesh for b in self._bundles])
        self._sbps = (abi.SBPDesc * n)(*[b.sbp for b in self._bundles])
        self._phys = (abi.PhysDesc * n)(*[b.phys for b in self._bundles])
        h = C.c_void_p()
        check(lib().pdes_batch_create(n, C.cast(self._meshes, C.c_void_p), C.cast(self._sbps, C.c_void_p),
                                      C.cast(self._phys, C.c_void_p), C.byref(h)))
        self.handle = h
        self.n = n

    def num_groups(self):
        g = C.c_int()
        check(lib().pdes_batch_num_groups(self.handle, C.byref(g)))
        return g.value

    def graph_stats(self):
        """(Arnoldi blocks captured into HIP graphs, graph replays) so far."""
        a, b = C.c_int64(), C.c_int64()
        check(lib().pdes_batch_graph_stats(self.handle, C.byref(a), C.byref(b)))
        return a.value, b.value

    def shape(self, i):
        m, op, _ = self.problems[i]
        return (m.numEl, op.numnodes, m.numDofPerNode)

    def set_state(self, i, q):
        qc = np.ascontiguousarray(q, dtype=np.float64)
        assert qc.shape == self.shape(i)
        check(lib().pdes_batch_set_state(self.handle, int(i), qc.ctypes.data))

    def get_state(self, i):
        out = np.empty(self.shape(i))
        check(lib().pdes_batch_get_state(self.handle, int(i), out.ctypes.data))
        return out

    def eval_residual(self, i):
        out = np.empty(self.shape(i))
        check(lib().pdes_batch_eval_residual(self.handle, int(i), out.ctypes.data))
        return out

    def solve(self, opts: NewtonOptions | None = None):
        """Newton on every instance; returns the list of NewtonReport
        (converged: 1 ok, 0 itermax reached, -1 checkDensity/checkPressure)."""
        opts = opts or NewtonOptions()
        o = opts.desc()
        reps = (abi.NewtonReport * self.n)()
        check(lib().pdes_batch_newton_solve(self.handle, C.byref(o), C.cast(reps, C.c_void_p)))
        return list(reps)

    def advance(self, opts: NewtonOptions, max_passes: int = 1):
        """pdes_batch_advance: up to max_passes passes of every group's Newton
        loop (the first call starts the solve).  Returns (reports, active,
        done): every instance's NewtonReport so far, a bool per instance that
        is still being solved, and whether no instance has work left."""
        o = opts.desc()
        reps = (abi.NewtonReport * self.n)()
        active = np.zeros(self.n, dtype=np.int32)
        done = C.c_int(0)
        check(lib().pdes_batch_advance(self.handle, C.byref(o), int(max_passes), C.cast(reps, C.c_void_p),
                                       active.ctypes.data, C.byref(done)))
        return list(reps), active.astype(bool), bool(done.value)

    def export(self, i):
        """State of an unfinished instance between advance calls:
        (BatchInstState, q, dq, rhs); dq and rhs are None unless it is in the
        middle of a linear solve."""
        st = abi.BatchInstState()
        q = np.empty(self.shape(i))
        dq = np.empty(self.shape(i))
        rhs = np.empty(self.shape(i))
        check(lib().pdes_batch_export(self.handle, int(i), C.byref(st), q.ctypes.data, dq.ctypes.data,
                                      rhs.ctypes.data))
        if not st.in_solve:
            dq = rhs = None
        return st, q, dq, rhs

    def import_(self, i, st, q, dq=None, rhs=None):
        """Continue an exported solve as instance i (before the first advance)."""
        qc = np.ascontiguousarray(q, dtype=np.float64)
        assert qc.shape == self.shape(i)
        dqc = None if dq is None else np.ascontiguousarray(dq, dtype=np.float64)
        rc = None if rhs is None else np.ascontiguousarray(rhs, dtype=np.float64)
        check(lib().pdes_batch_import(self.handle, int(i), C.byref(st), qc.ctypes.data,
                                      None if dqc is None else dqc.ctypes.data,
                                      None if rc is None else rc.ctypes.data))

    def free(self):
        if getattr(self, "handle", None):
            lib().pdes_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
