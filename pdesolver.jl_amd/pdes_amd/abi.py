"""ctypes mirror of include/pdes.h descriptor structs.

`make_descriptors(mesh, op, phys)` returns the three descriptors together with
the contiguous numpy arrays they point into (kept alive by the returned
`DescBundle`).  Both the HIP library binding and the test-side oracle binding
consume the same structs.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from .mesh import DGMesh
from .sbp import SBPOperator

PDES_BC_ISENTROPIC_VORTEX = 0
PDES_BC_FREESTREAM = 1
PDES_FLUX_ROE = 0
PDES_FLUX_IRSLF = 1
PDES_FLUX_ESLF = 2
PDES_PC_NONE = 0
PDES_PC_BJACOBI = 1
PDES_PC_ILU0 = 2

ERRORS = {0: "ok", -1: "bad argument", -2: "HIP error", -3: "out of device memory",
          -4: "non-positive density or pressure", -5: "unsupported", -6: "no convergence"}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class MeshDesc(C.Structure):
    _fields_ = [("dim", C.c_int32), ("numEl", C.c_int32), ("numInterfaces", C.c_int32),
                ("numBoundaryFaces", C.c_int32), ("numBC", C.c_int32),
                ("dxidx", _dp), ("jac", _dp), ("interfaces", _ip), ("bndryfaces", _ip),
                ("bndry_offsets", _ip), ("bc_type", _ip), ("nrm_face", _dp),
                ("nrm_bndry", _dp), ("coords_bndry", _dp)]


class SBPDesc(C.Structure):
    _fields_ = [("Np", C.c_int32), ("Nf", C.c_int32), ("stencil", C.c_int32),
                ("norient", C.c_int32), ("diag_e", C.c_int32), ("w", _dp), ("Q", _dp),
                ("interp", _dp), ("perm", _ip), ("nbrperm", _ip), ("wface", _dp)]


class PhysDesc(C.Structure):
    _fields_ = [("gamma", C.c_double), ("R_gas", C.c_double), ("Ma", C.c_double),
                ("aoa", C.c_double), ("rho_free", C.c_double), ("p_free", C.c_double),
                ("vortex_Min", C.c_double), ("volume_integral_type", C.c_int32),
                ("face_flux", C.c_int32), ("check_positivity", C.c_int32)]


class NewtonOpts(C.Structure):
    _fields_ = [("res_abstol", C.c_double), ("res_reltol", C.c_double), ("step_tol", C.c_double),
                ("itermax", C.c_int32), ("krylov_reltol", C.c_double), ("krylov_abstol", C.c_double),
                ("krylov_itermax", C.c_int32), ("gmres_restart", C.c_int32), ("pc_kind", C.c_int32),
                ("fp64_blocks", C.c_int32), ("jv_mode", C.c_int32), ("globalize_euler", C.c_int32),
                ("euler_tau", C.c_double), ("scale_euler", C.c_int32)]


class NewtonReport(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32),
                ("total_krylov_iters", C.c_int32), ("res_norm0", C.c_double),
                ("res_norm", C.c_double), ("step_norm", C.c_double), ("elapsed_s", C.c_double)]


class BatchInstState(C.Structure):
    _fields_ = [("active", C.c_int32), ("failed", C.c_int32), ("in_solve", C.c_int32),
                ("newton_steps", C.c_int32), ("tau", C.c_double), ("tau_prev_norm", C.c_double),
                ("gmres_iters", C.c_int32), ("gmres_live", C.c_int32), ("gmres_beta0", C.c_double),
                ("report", NewtonReport)]


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


@dataclass
class Physics:
    """Subset of the reference options dictionary that fixes the residual
    (src/input/read_input.jl defaults: gamma 1.4 :139, R 287.058 :141,
    p_free 1/gamma :162)."""
    gamma: float = 1.4
    R_gas: float = 287.058
    Ma: float = 0.5
    aoa_deg: float = 0.0
    rho_free: float = 1.0
    p_free: float = None
    vortex_Min: float = 0.95
    volume_integral_type: int = 1
    face_flux: int = PDES_FLUX_ROE
    check_positivity: int = 1
    bc_type: list = field(default_factory=lambda: None)

    def desc(self) -> PhysDesc:
        p_free = self.p_free if self.p_free is not None else 1.0 / self.gamma
        return PhysDesc(self.gamma, self.R_gas, self.Ma, self.aoa_deg * math.pi / 180.0,
                        self.rho_free, p_free, self.vortex_Min, self.volume_integral_type,
                        self.face_flux, self.check_positivity)

    @property
    def E_free(self):
        p_free = self.p_free if self.p_free is not None else 1.0 / self.gamma
        return p_free / (self.gamma - 1) + 0.5 * self.rho_free * self.Ma ** 2


@dataclass
class DescBundle:
    mesh: MeshDesc
    sbp: SBPDesc
    phys: PhysDesc
    keep: list


def make_descriptors(mesh: DGMesh, op: SBPOperator, phys: Physics) -> DescBundle:
    keep = []

    def f64(a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        keep.append(a)
        return a

    def i32(a):
        a = np.ascontiguousarray(a, dtype=np.int32)
        keep.append(a)
        return a

    bc = phys.bc_type
    if bc is None:
        bc = [PDES_BC_ISENTROPIC_VORTEX if mesh.dim == 2 else PDES_BC_FREESTREAM] * mesh.numBC
    bc = i32(np.asarray(bc))
    assert bc.shape[0] == mesh.numBC
    md = MeshDesc(mesh.dim, mesh.numEl, mesh.numInterfaces, mesh.numBoundaryFaces, mesh.numBC,
                  dptr(f64(mesh.dxidx)), dptr(f64(mesh.jac)), iptr(i32(mesh.interfaces)),
                  iptr(i32(mesh.bndryfaces)), iptr(i32(mesh.bndry_offsets)), iptr(bc),
                  dptr(f64(mesh.nrm_face)), dptr(f64(mesh.nrm_bndry)), dptr(f64(mesh.coords_bndry)))
    face = op.face
    Qd = f64(np.transpose(op.Q, (2, 0, 1)))      # [d][i][j]
    sd = SBPDesc(op.numnodes, face.numnodes, face.stencilsize, face.nbrperm.shape[1],
                 0 if face.dense else 1, dptr(f64(op.w)), dptr(Qd), dptr(f64(face.interp)),
                 iptr(i32(face.perm)), iptr(i32(face.nbrperm)), dptr(f64(face.wface)))
    return DescBundle(md, sd, phys.desc(), keep)
