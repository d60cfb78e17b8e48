"""Host-side mirror of the reference's Jacobian / linear-solver / Newton
interfaces, backed by the gfx950 library.

  evalJacobian            src/solver/euler/evalJacobian.jl:6-72 (explicit, BSR on the device)
  getSparsityPattern      src/interface2.jl:544-555 / src/jacobian/sparse.jl:250-495
  AbstractLO / AbstractPC src/linearsolvers/LinearSolvers.jl:185-645
    HIPBSRLO        ~ PetscMatLO (assembled matrix)          lo_petscmat.jl
    HIPMatFreeLO    ~ PetscMatFreeLO (complex-step products) newton_setup.jl:632-662
    HIPBJacobiPC    ~ NewtonBDiagPC (element block Jacobi)   preconditioning_types.jl:41
  StandardLinearSolver    LinearSolvers.jl:73-154 ; ls_standard.jl:28-491
  newton                  src/NonlinearSolvers/newton.jl:54-304
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import abi
from ._lib import check, lib
from .euler import EulerData


def evalJacobian(mesh, sbp, eqn: EulerData, opts=None, fp64_blocks: bool = False):
    """Assemble the explicit Jacobian at the device state into the device BSR
    matrix (block size ndof, FP32 blocks unless fp64_blocks)."""
    check(lib().pdes_jacobian_setup(eqn.ctx, int(fp64_blocks)))
    check(lib().pdes_assemble_jacobian(eqn.ctx))


def calcJacobianSparse(mesh, sbp, eqn: EulerData, opts=None, fp64_blocks: bool = False):
    """The Jacobian from directional derivatives of the residual over a
    distance-2 element colouring (calcJacobianSparse,
    src/jacobian/jacobian.jl:365-446), into the same device BSR matrix."""
    check(lib().pdes_jacobian_setup(eqn.ctx, int(fp64_blocks)))
    check(lib().pdes_assemble_jacobian_coloring(eqn.ctx))


def getColoring(eqn: EulerData, fp64_blocks: bool = False):
    """The mesh colouring fields of the reference (numColors, color,
    pertNeighborEls; docs/src/interfaces.md:535-570) as the device colouring
    assembly uses them: (color[numEl], numColors, pertNeighborEls[numEl,
    numColors]), 0-based, -1 where no element of that colour is adjacent."""
    check(lib().pdes_jacobian_setup(eqn.ctx, int(fp64_blocks)))
    nc = C.c_int32()
    check(lib().pdes_get_coloring(eqn.ctx, None, C.byref(nc), None))
    color = np.zeros(eqn.mesh.numEl, dtype=np.int32)
    pert = np.zeros((eqn.mesh.numEl, nc.value), dtype=np.int32)
    check(lib().pdes_get_coloring(eqn.ctx, color.ctypes.data, C.byref(nc), pert.ctypes.data))
    return color, nc.value, pert


def getSparsityPattern(eqn: EulerData, fp64_blocks: bool = False):
    """(rowptr, colind) of the node-block BSR matrix (0-based, sorted)."""
    check(lib().pdes_jacobian_setup(eqn.ctx, int(fp64_blocks)))
    nr = C.c_int64()
    nz = C.c_int64()
    check(lib().pdes_jacobian_nnzb(eqn.ctx, C.byref(nr), C.byref(nz)))
    rowptr = np.zeros(nr.value + 1, dtype=np.int64)
    colind = np.zeros(nz.value, dtype=np.int32)
    check(lib().pdes_get_bsr_pattern(eqn.ctx, rowptr.ctypes.data, colind.ctypes.data))
    return rowptr, colind


def getJacobianValues(eqn: EulerData, fp64_blocks: bool = False):
    nr = C.c_int64()
    nz = C.c_int64()
    check(lib().pdes_jacobian_nnzb(eqn.ctx, C.byref(nr), C.byref(nz)))
    nd = eqn.mesh.numDofPerNode
    vals = np.zeros((nz.value, nd, nd), dtype=np.float64 if fp64_blocks else np.float32)
    check(lib().pdes_get_bsr_values(eqn.ctx, vals.ctypes.data))
    return vals


class AbstractLO:
    def calcLinearOperator(self, mesh, sbp, eqn, opts=None, ctx_residual=None, t=0.0):
        raise NotImplementedError

    def applyLinearOperator(self, mesh, sbp, eqn, opts, ctx_residual, t, x, b):
        raise NotImplementedError

    def free(self):
        pass


class HIPBSRLO(AbstractLO):
    """Assembled BSR Jacobian on the device; apply = BSR SpMV."""
    jv_mode = 0

    def __init__(self, eqn: EulerData, fp64_blocks: bool = False):
        self.eqn = eqn
        self.fp64_blocks = fp64_blocks

    def calcLinearOperator(self, mesh=None, sbp=None, eqn=None, opts=None, ctx_residual=None, t=0.0):
        evalJacobian(mesh, sbp, self.eqn, opts, self.fp64_blocks)

    def applyLinearOperator(self, mesh, sbp, eqn, opts, ctx_residual, t, x, b):
        xc = np.ascontiguousarray(x, dtype=np.float64)
        out = np.empty_like(xc)
        check(lib().pdes_bsr_matvec(self.eqn.ctx, xc.ctypes.data, out.ctypes.data))
        b[...] = out.reshape(b.shape)


def applyLinearOperatorTranspose(eqn: EulerData, x, b):
    """applyLinearOperatorTranspose: b = J^T x with the assembled BSR Jacobian."""
    xc = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(xc)
    check(lib().pdes_bsr_matvec_transpose(eqn.ctx, xc.ctypes.data, out.ctypes.data))
    b[...] = out.reshape(b.shape)


def evalResidual_revq(mesh, sbp, eqn: EulerData, opts, input_array, output_array, t: float = 0.0,
                      zero_output: bool = True):
    """evalResidual_revq (interface2.jl:392-428): output += psi^T dR/dq at the
    device state, psi = input_array (overwrite when zero_output).  The product
    is (dR/dq)^T psi with the FP64-block Jacobian assembled here at that state,
    which replaces the context's assembled Jacobian."""
    evalJacobian(mesh, sbp, eqn, opts, fp64_blocks=True)
    out = np.empty(eqn.q.shape, dtype=np.float64)
    applyLinearOperatorTranspose(eqn, np.reshape(input_array, eqn.q.shape), out)
    dst = np.reshape(output_array, eqn.q.shape)
    if zero_output:
        dst[...] = out
    else:
        dst += out


def evalResidual_revm(mesh, sbp, eqn: EulerData, opts, input_array, t: float = 0.0):
    """evalResidual_revm (interface2.jl:329-372, evaldRdm.jl:19-70): adds
    psi^T dR/dm, psi = input_array, into the mesh's `_bar` arrays
    (dxidx_bar, nrm_face_bar, nrm_bndry_bar, coords_bndry_bar, created as
    zeros on first use; zeroBarArrays resets them), at the device state."""
    psi = np.ascontiguousarray(input_array, dtype=np.float64).reshape(-1)
    assert psi.size == eqn.q.size
    outs = {}
    for name, ref in (("dxidx_bar", mesh.dxidx), ("nrm_face_bar", mesh.nrm_face),
                      ("nrm_bndry_bar", mesh.nrm_bndry), ("coords_bndry_bar", mesh.coords_bndry)):
        outs[name] = np.zeros(np.shape(ref), dtype=np.float64)
    check(lib().pdes_eval_residual_revm(eqn.ctx, psi.ctypes.data, outs["dxidx_bar"].ctypes.data,
                                        outs["nrm_face_bar"].ctypes.data, outs["nrm_bndry_bar"].ctypes.data,
                                        outs["coords_bndry_bar"].ctypes.data))
    for name, val in outs.items():
        if getattr(mesh, name, None) is None:
            setattr(mesh, name, np.zeros_like(val))
        getattr(mesh, name)[...] += val


def zeroBarArrays(mesh):
    """zeroBarArrays: the mesh's `_bar` arrays set to zero."""
    for name in ("dxidx_bar", "nrm_face_bar", "nrm_bndry_bar", "coords_bndry_bar"):
        if getattr(mesh, name, None) is not None:
            getattr(mesh, name)[...] = 0.0


def setHomotopyBoundaryIntegrals(eqn: EulerData, on: bool):
    """opts["homotopy_addBoundaryIntegrals"] for this context's homotopy
    dissipation (default off)."""
    check(lib().pdes_set_homotopy_boundary(eqn.ctx, int(bool(on))))


def evalHomotopy(mesh, sbp, eqn: EulerData, opts=None):
    """evalHomotopy with "FirstOrderDissipation" (src/solver/euler/homotopy.jl:
    11-31, calcHomotopyDiss :51-249): D(q) at the device state, in the layout
    of q (boundary term per setHomotopyBoundaryIntegrals)."""
    out = np.empty(eqn.q.shape, dtype=np.float64)
    check(lib().pdes_eval_homotopy(eqn.ctx, out.ctypes.data))
    return out


def evalHomotopy_jv(eqn: EulerData, v):
    """(dD/dq) v at the device state."""
    vc = np.ascontiguousarray(v, dtype=np.float64)
    out = np.empty(eqn.q.shape, dtype=np.float64)
    check(lib().pdes_homotopy_jv(eqn.ctx, vc.ctypes.data, out.ctypes.data))
    return out


def calcHomotopyJacobian(mesh, sbp, eqn: EulerData, lam: float, fp64_blocks: bool = False):
    """Jacobian of (1 - lam) R + lam D into the device BSR matrix
    (HomotopyMatLO calcLinearOperator, predictor_corrector.jl)."""
    check(lib().pdes_jacobian_setup(eqn.ctx, int(fp64_blocks)))
    check(lib().pdes_assemble_homotopy_jacobian(eqn.ctx, float(lam)))


FUNCTIONALS = {"lift": 1, "drag": 2, "entropyflux": 3, "massflow": 4, "liftCoefficient": 5, "dragCoefficient": 6}


def _bcs(bcnums):
    return np.ascontiguousarray(bcnums, dtype=np.int32)


def evalFunctionalDeriv_q(eqn: EulerData, functional: str, bcnums):
    """evalFunctionalDeriv_q (interface2.jl:259-283): dJ/dq in the layout of q
    for a scalar boundary functional over the (0-based) BC groups bcnums."""
    bc = _bcs(bcnums)
    out = np.empty(eqn.q.shape, dtype=np.float64)
    check(lib().pdes_functional_deriv(eqn.ctx, FUNCTIONALS[functional], bc.ctypes.data, len(bc), out.ctypes.data))
    return out


def evalFunctionalDeriv_m(mesh, sbp, eqn: EulerData, opts, functional: str, bcnums, val_bar: float = 1.0):
    """evalFunctionalDeriv_m (interface2.jl:205-227): adds val_bar dJ/dm into
    the mesh's nrm_bndry_bar (the only metric the boundary functionals
    depend on), at the device state."""
    bc = _bcs(bcnums)
    out = np.zeros(np.shape(mesh.nrm_bndry), dtype=np.float64)
    check(lib().pdes_functional_deriv_m(eqn.ctx, FUNCTIONALS[functional], bc.ctypes.data, len(bc),
                                        out.ctypes.data))
    if getattr(mesh, "nrm_bndry_bar", None) is None:
        mesh.nrm_bndry_bar = np.zeros_like(out)
    mesh.nrm_bndry_bar += val_bar * out


def calcAdjoint(mesh, sbp, eqn: EulerData, functional: str, bcnums, opts: "NewtonOptions | None" = None,
                rtol: float = 1e-10):
    """calcAdjoint (steady_adjoint.jl:30-66): psi with J^T psi = -dJ/dq at the
    current state; the preconditioner (opts.pc_kind) is built from J^T.
    Returns (psi, GMRES iterations, final residual norm)."""
    opts = opts or NewtonOptions()
    bc = _bcs(bcnums)
    psi = np.empty(eqn.q.shape, dtype=np.float64)
    its = C.c_int32()
    res = C.c_double()
    check(lib().pdes_adjoint_solve(eqn.ctx, FUNCTIONALS[functional], bc.ctypes.data, len(bc), opts.pc_kind,
                                   int(opts.fp64_blocks), rtol, 0.0, opts.gmres_restart, opts.krylov_itermax,
                                   psi.ctypes.data, C.byref(its), C.byref(res)))
    return psi, its.value, res.value


class HIPMatFreeLO(AbstractLO):
    """Matrix-free operator: the directional derivative of the residual
    (complex-step product of the reference) evaluated on the device."""
    jv_mode = 1

    def __init__(self, eqn: EulerData):
        self.eqn = eqn

    def calcLinearOperator(self, *a, **k):
        pass

    def applyLinearOperator(self, mesh, sbp, eqn, opts, ctx_residual, t, x, b):
        xc = np.ascontiguousarray(x, dtype=np.float64)
        out = np.empty_like(xc)
        check(lib().pdes_jv_complex_step(self.eqn.ctx, xc.ctypes.data, out.ctypes.data))
        b[...] = out.reshape(b.shape)


class AbstractPC:
    def calcPC(self, mesh, sbp, eqn, opts=None, ctx_residual=None, t=0.0):
        raise NotImplementedError

    def applyPC(self, mesh, sbp, eqn, opts, t, b, x):
        raise NotImplementedError

    def free(self):
        pass


class PCNone(AbstractPC):
    kind = abi.PDES_PC_NONE

    def __init__(self, eqn: EulerData):
        self.eqn = eqn

    def calcPC(self, *a, **k):
        check(lib().pdes_pc_setup(self.eqn.ctx, self.kind))

    def applyPC(self, mesh, sbp, eqn, opts, t, b, x):
        x[...] = b


class HIPBJacobiPC(AbstractPC):
    """Element block-Jacobi (NewtonBDiagPC): the explicit inverse of each
    element's (Np*ndof)^2 diagonal block by Gauss-Jordan elimination with
    partial pivoting, in FP32, applied as one product per element."""
    kind = abi.PDES_PC_BJACOBI

    def __init__(self, eqn: EulerData):
        self.eqn = eqn

    def calcPC(self, mesh=None, sbp=None, eqn=None, opts=None, ctx_residual=None, t=0.0):
        check(lib().pdes_pc_setup(self.eqn.ctx, self.kind))

    def applyPC(self, mesh, sbp, eqn, opts, t, b, x):
        bc = np.ascontiguousarray(b, dtype=np.float64)
        out = np.empty_like(bc)
        check(lib().pdes_pc_apply(self.eqn.ctx, bc.ctypes.data, out.ctypes.data))
        x[...] = out.reshape(x.shape)


class HIPILU0PC(HIPBJacobiPC):
    """Element-block ILU(0) in the natural element order, FP32 factors:
    D'_e = A_ee - sum_{k<e neighbour} A_ek D'_k^-1 A_ke and
    M = (D' + L) D'^-1 (D' + U); the device stand-in for the PETSc
    bjacobi/ILU sub-preconditioner of the reference (read_input.jl:560-581)."""
    kind = abi.PDES_PC_ILU0


class StandardLinearSolver:
    """LinearSolver with a PC and an LO; linearSolve runs GMRES(restart)
    with right preconditioning on the device."""

    def __init__(self, pc: AbstractPC, lo: AbstractLO, reltol=1e-2, abstol=1e-50, itermax=1000, restart=30):
        self.pc = pc
        self.lo = lo
        self.reltol = reltol
        self.abstol = abstol
        self.itermax = itermax
        self.restart = restart
        self.last_iters = 0
        self.last_resnorm = 0.0

    def setTolerances(self, reltol=-1, abstol=-1, dtol=-1, itermax=-1):
        if reltol > 0:
            self.reltol = reltol
        if abstol > 0:
            self.abstol = abstol
        if itermax > 0:
            self.itermax = itermax

    def calcPCandLO(self, mesh=None, sbp=None, eqn=None, opts=None, ctx_residual=None, t=0.0):
        self.lo.calcLinearOperator(mesh, sbp, eqn, opts, ctx_residual, t)
        self.pc.calcPC(mesh, sbp, eqn, opts, ctx_residual, t)

    def linearSolve(self, b, x, verbose=0):
        eqn = self.lo.eqn
        bc = np.ascontiguousarray(b, dtype=np.float64)
        xc = np.ascontiguousarray(x, dtype=np.float64).copy()
        its = C.c_int32()
        rn = C.c_double()
        check(lib().pdes_gmres(eqn.ctx, bc.ctypes.data, xc.ctypes.data, float(self.reltol), float(self.abstol),
                               int(self.restart), int(self.itermax), int(self.lo.jv_mode), C.byref(its),
                               C.byref(rn)))
        x[...] = xc.reshape(x.shape)
        self.last_iters = its.value
        self.last_resnorm = rn.value
        return its.value


@dataclass
class NewtonOptions:
    """Option keys of the reference that fix the Newton solve
    (src/input/read_input.jl:405-500)."""
    res_abstol: float = 1e-10
    res_reltol: float = 1e-300
    step_tol: float = 0.0
    itermax: int = 50
    krylov_reltol: float = 1e-2
    krylov_abstol: float = 1e-50
    krylov_itermax: int = 1000
    gmres_restart: int = 30
    pc_kind: int = abi.PDES_PC_BJACOBI
    fp64_blocks: bool = False
    jv_mode: int = 0
    # implicit-Euler globalization (newton_globalize_euler, euler_tau,
    # newton_scale_euler; read_input.jl:421-422)
    globalize_euler: bool = False
    euler_tau: float = 1.0
    scale_euler: bool = False

    def desc(self):
        return abi.NewtonOpts(self.res_abstol, self.res_reltol, self.step_tol, self.itermax,
                              self.krylov_reltol, self.krylov_abstol, self.krylov_itermax,
                              self.gmres_restart, self.pc_kind, int(self.fp64_blocks), self.jv_mode,
                              int(self.globalize_euler), self.euler_tau, int(self.scale_euler))


def newton(mesh, sbp, eqn: EulerData, opts: NewtonOptions | None = None):
    """Steady Newton solve of R(q) = 0 on the device state (newtonInner):
    explicit Jacobian + preconditioner recalculated every iteration, inexact
    GMRES solve of J dq = -R, q += dq, until ||R||_{M^-1} < res_abstol.
    Returns the report (iterations, converged, residual norms, time)."""
    opts = opts or NewtonOptions()
    o = opts.desc()
    rep = abi.NewtonReport()
    check(lib().pdes_newton_solve(eqn.ctx, C.byref(o), C.byref(rep)))
    return rep


def predictorCorrectorHomotopy(mesh, sbp, eqn: EulerData, opts: NewtonOptions | None = None,
                               psi_max: float = 10.0 * np.pi / 180.0, addBoundaryIntegrals: bool = False):
    """predictorCorrectorHomotopy (src/NonlinearSolvers/predictor_corrector.jl:
    226-326) with the first-order dissipation as the easy problem, on the
    device state.  opts gives the final tolerances, itermax (predictor
    iterations), the preconditioner and the Krylov settings; psi_max is
    opts["homotopy_psi_max"].  Returns the report (iterations = predictor
    iterations).  addBoundaryIntegrals is opts["homotopy_addBoundaryIntegrals"]."""
    opts = opts or NewtonOptions()
    o = opts.desc()
    rep = abi.NewtonReport()
    setHomotopyBoundaryIntegrals(eqn, addBoundaryIntegrals)
    check(lib().pdes_homotopy_solve(eqn.ctx, C.byref(o), float(psi_max), C.byref(rep)))
    return rep
