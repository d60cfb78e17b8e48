"""Host-side mirror of the reference's Euler physics interface, backed by the
gfx950 library.

Reference names are kept so the calls read as in PDESolver.jl:
  EulerData            src/solver/euler/types.jl:401-723 (eqn object: q, res)
  evalResidual         src/solver/euler/euler.jl:111-175
  calcNorm             src/Utils/Utils.jl:427-449 (strongres=True)
  evaldRdqProduct      src/interface2.jl:454-498
  rk4                  src/NonlinearSolvers/rk4.jl:144-344
Arrays use the reference layout with axes reversed: q[numEl, Np, ndof].
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from ._lib import check, lib
from .mesh import DGMesh
from .sbp import SBPOperator


class EulerData:
    """The eqn object: owns the device context (state, residual, tables).

    `q` and `res` are host views; `set_q`/`get_res` move data to/from HBM.
    """

    def __init__(self, mesh: DGMesh, sbp: SBPOperator, phys: abi.Physics):
        self.mesh = mesh
        self.sbp = sbp
        self.phys = phys
        self._desc = abi.make_descriptors(mesh, sbp, phys)
        handle = C.c_void_p()
        check(lib().pdes_ctx_create(C.byref(self._desc.mesh), C.byref(self._desc.sbp),
                                    C.byref(self._desc.phys), C.byref(handle)))
        self.ctx = handle
        self.shape = (mesh.numEl, sbp.numnodes, mesh.numDofPerNode)
        self.q = np.zeros(self.shape)
        self.res = np.zeros(self.shape)
        # mass matrix inverse per dof (calcMassMatrixInverse, src/Utils/mass_matrix.jl:20-55)
        self.Minv = np.repeat((1.0 / (sbp.w[None, :] / mesh.jac))[:, :, None],
                              mesh.numDofPerNode, axis=2)

    @property
    def numDof(self):
        return int(np.prod(self.shape))

    def set_q(self, q=None):
        if q is not None:
            self.q[...] = q
        qc = np.ascontiguousarray(self.q, dtype=np.float64)
        check(lib().pdes_set_state(self.ctx, qc.ctypes.data))

    def get_q(self):
        out = np.empty(self.shape)
        check(lib().pdes_get_state(self.ctx, out.ctypes.data))
        self.q[...] = out
        return out

    def get_res(self):
        check(lib().pdes_get_residual(self.ctx, self.res.ctypes.data))
        return self.res

    def device_ptrs(self):
        q = C.c_void_p()
        r = C.c_void_p()
        check(lib().pdes_state_ptr(self.ctx, C.byref(q), C.byref(r)))
        return q.value, r.value

    def free(self):
        if getattr(self, "ctx", None):
            lib().pdes_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def evalResidual(mesh, sbp, eqn: EulerData, opts=None, t: float = 0.0):
    """evalResidual(mesh, sbp, eqn, opts, t): device q -> device res, copied
    back to eqn.res (the reference writes eqn.res in place)."""
    check(lib().pdes_eval_residual(eqn.ctx, float(t)))
    eqn.get_res()
    return None


def evalResidual_device(eqn: EulerData, t: float = 0.0):
    """Residual on the device only (no host copy), stream-ordered."""
    check(lib().pdes_eval_residual(eqn.ctx, float(t)))


def calcNorm(eqn: EulerData, res=None, strongres: bool = True):
    """calcNorm(eqn, res, strongres=true) of the device residual (res=None)
    or of a host vector."""
    if res is None:
        out = C.c_double()
        check(lib().pdes_residual_norm(eqn.ctx, C.byref(out)))
        return out.value
    w = eqn.Minv if strongres else 1.0 / eqn.Minv
    return float(np.sqrt((np.asarray(res) ** 2 * w).sum()))


def evaldRdqProduct(eqn: EulerData, x: np.ndarray) -> np.ndarray:
    """b = dR/dq x at the device state (complex-step product of the
    reference, evaluated with forward-mode derivatives on the device)."""
    xc = np.ascontiguousarray(x, dtype=np.float64).reshape(eqn.shape)
    b = np.empty(eqn.shape)
    check(lib().pdes_jv_complex_step(eqn.ctx, xc.ctypes.data, b.ctypes.data))
    return b


def rk4(eqn: EulerData, nsteps: int, h: float, timed: bool = False):
    """nsteps classical RK4 steps on the device state; returns elapsed ms when
    timed (HIP events on the context stream)."""
    ms = C.c_double(0.0)
    check(lib().pdes_rk4(eqn.ctx, int(nsteps), float(h), C.byref(ms) if timed else None))
    return ms.value if timed else None


def autotune(eqn: EulerData, iters: int = 10):
    """Pick the fastest residual launch variant for this mesh/operator
    (plain and RK4-stage kernels); returns the chosen variant indices."""
    out = (C.c_int32 * 2)()
    check(lib().pdes_autotune(eqn.ctx, int(iters), out))
    return int(out[0]), int(out[1])


def copy_bandwidth(nbytes: int = 1 << 30, iters: int = 20) -> float:
    """Average duration (ms) of a plain device copy of nbytes (read nbytes,
    write nbytes) over `iters` launches."""
    ms = C.c_double(0.0)
    check(lib().pdes_copy_bandwidth(int(nbytes), int(iters), C.byref(ms)))
    return ms.value


def time_residual(eqn: EulerData, iters: int) -> float:
    """Average duration (ms) of the residual kernel over `iters` launches."""
    ms = C.c_double(0.0)
    check(lib().pdes_time_residual(eqn.ctx, int(iters), C.byref(ms)))
    return ms.value
