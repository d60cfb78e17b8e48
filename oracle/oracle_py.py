"""TEST INFRASTRUCTURE ONLY: ctypes binding of the CPU parity oracle
(oracle/build/liboracle.so).  Imported by tests/, the smoke check and the
CPU-baseline leg of bench.py; never by the product package."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "build", "liboracle.so")
_lib = None


def build(force: bool = False):
    if force or not os.path.exists(_LIB):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = C.CDLL(_LIB)
    return _lib


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _check(st, what):
    if st != 0:
        raise RuntimeError(f"oracle {what} failed with status {st}")


def residual(b, q):
    q = np.ascontiguousarray(q, dtype=np.float64)
    res = np.zeros_like(q)
    _check(lib().oracle_residual(C.byref(b.mesh), C.byref(b.sbp), C.byref(b.phys), _d(q), _d(res)),
           "residual")
    return res


def residual_status(b, q):
    q = np.ascontiguousarray(q, dtype=np.float64)
    res = np.zeros_like(q)
    return lib().oracle_residual(C.byref(b.mesh), C.byref(b.sbp), C.byref(b.phys), _d(q), _d(res))


def residual_scale(b, q):
    q = np.ascontiguousarray(q, dtype=np.float64)
    res = np.zeros_like(q)
    _check(lib().oracle_residual_scale(C.byref(b.mesh), C.byref(b.sbp), C.byref(b.phys), _d(q), _d(res)),
           "residual_scale")
    return res


def residual_complex(b, qc):
    qr = np.ascontiguousarray(qc.real, dtype=np.float64)
    qi = np.ascontiguousarray(qc.imag, dtype=np.float64)
    rr = np.zeros_like(qr)
    ri = np.zeros_like(qr)
    _check(lib().oracle_residual_complex(C.byref(b.mesh), C.byref(b.sbp), C.byref(b.phys),
                                         _d(qr), _d(qi), _d(rr), _d(ri)), "residual_complex")
    return rr + 1j * ri


def jv_complex_step(b, q, v, h=1e-20):
    q = np.ascontiguousarray(q, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.zeros_like(q)
    _check(lib().oracle_jv_complex_step(C.byref(b.mesh), C.byref(b.sbp), C.byref(b.phys),
                                        _d(q), _d(v), C.c_double(h), _d(out)), "jv")
    return out


def homotopy(b, q, bndry=False):
    """calcHomotopyDiss (src/solver/euler/homotopy.jl:51-249); bndry adds the
    boundary term of homotopy_addBoundaryIntegrals."""
    q = np.ascontiguousarray(q, dtype=np.float64)
    out = np.zeros_like(q)
    _check(lib().oracle_homotopy(C.byref(b.mesh), C.byref(b.sbp), C.byref(b.phys), _d(q), C.c_int(int(bndry)),
                                 _d(out)), "homotopy")
    return out


def homotopy_jv(b, q, v, h=1e-20, bndry=False):
    q = np.ascontiguousarray(q, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.zeros_like(q)
    _check(lib().oracle_homotopy_jv(C.byref(b.mesh), C.byref(b.sbp), C.byref(b.phys), _d(q), _d(v),
                                    C.c_double(h), C.c_int(int(bndry)), _d(out)), "homotopy jv")
    return out


def sparsity_bsr(b):
    nrows = b.mesh.numEl * b.sbp.Np
    rowptr = np.zeros(nrows + 1, dtype=np.int64)
    L = lib()
    L.oracle_sparsity_bsr.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    _check(L.oracle_sparsity_bsr(C.addressof(b.mesh), C.addressof(b.sbp),
                                 rowptr.ctypes.data, None), "sparsity")
    colind = np.zeros(rowptr[-1], dtype=np.int32)
    _check(L.oracle_sparsity_bsr(C.addressof(b.mesh), C.addressof(b.sbp),
                                 rowptr.ctypes.data, colind.ctypes.data), "sparsity")
    return rowptr, colind


def color_elements(b):
    color = np.zeros(b.mesh.numEl, dtype=np.int32)
    nc = C.c_int32(0)
    _check(lib().oracle_color_elements(C.byref(b.mesh), color.ctypes.data_as(C.POINTER(C.c_int32)),
                                       C.byref(nc)), "coloring")
    return color, nc.value


def pert_neighbor_els(b, color, ncolors):
    color = np.ascontiguousarray(color, dtype=np.int32)
    out = np.zeros((b.mesh.numEl, ncolors), dtype=np.int32)
    L = lib()
    L.oracle_pert_neighbor_els.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]
    _check(L.oracle_pert_neighbor_els(C.addressof(b.mesh), color.ctypes.data, ncolors, out.ctypes.data),
           "pertNeighborEls")
    return out


def jacobian_cs(b, q, rowptr, colind):
    nd = b.mesh.dim + 2
    q = np.ascontiguousarray(q, dtype=np.float64)
    val = np.zeros((colind.shape[0], nd, nd))
    L = lib()
    L.oracle_jacobian_cs.argtypes = [C.c_void_p] * 7
    _check(L.oracle_jacobian_cs(C.addressof(b.mesh), C.addressof(b.sbp), C.addressof(b.phys),
                                q.ctypes.data, rowptr.ctypes.data, colind.ctypes.data,
                                val.ctypes.data), "jacobian_cs")
    return val


def newton_direct(b, q0, Minv, res_abstol=1e-10, itermax=30, euler_tau=None, jac=None, scale_euler=False):
    """newtonInner (src/NonlinearSolvers/newton.jl:137-304) with the sparse
    direct solve of jac_type 2 (UMFPACK in the reference, SuperLU here):
    J dq = -R, q += dq until ||R||_{M^-1} < res_abstol.  Returns (q, norms).

    euler_tau (with jac, the mapping Jacobian per node) turns on the
    implicit-Euler globalization: calcTauVec
    (src/NonlinearSolvers/globalization.jl:181-201) gives
    tau_vec = tau/(1 + sqrt(jac)) per node, applyEuler (:284-320) adds
    -M/tau_vec to the Jacobian diagonal, updateEuler (:237-262) multiplies
    tau by ||R_prev||/||R|| at every Jacobian after the first (after the
    first too, with ||R_prev|| = 1, when scale_euler)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    rowptr, colind = sparsity_bsr(b)
    nd = b.mesh.dim + 2
    nrow = b.mesh.numEl * b.sbp.Np
    q = np.array(q0, dtype=np.float64, copy=True)
    norms = []
    tau = euler_tau
    prev = 1.0 if scale_euler else 0.0
    for it in range(itermax + 1):
        r = residual(b, q)
        nrm = float(np.sqrt((r * r * Minv).sum()))
        norms.append(nrm)
        if nrm < res_abstol or it == itermax:
            break
        val = jacobian_cs(b, q, rowptr, colind)
        J = sp.bsr_matrix((val, colind, rowptr), shape=(nrow * nd, nrow * nd)).tocsc()
        if euler_tau is not None:
            if prev != 0.0:
                tau = tau * prev / nrm
            prev = nrm
            tau_vec = np.repeat((tau / (1.0 + np.sqrt(np.asarray(jac, dtype=np.float64)))).ravel(), nd)
            M = 1.0 / np.asarray(Minv, dtype=np.float64).ravel()
            J = (J - sp.diags(M / tau_vec)).tocsc()
        dq = spla.spsolve(J, -r.ravel())
        q = q + dq.reshape(q.shape)
    return q, norms


def functional(m, op, ph, q, bcnums):
    """calcBndryFunctional (src/solver/euler/boundary_functional.jl:82-126) for
    the force, entropy-flux and mass-flow integrands (:340-436), summed over
    the faces of the listed BC groups.  Returns dict of values."""
    g = ph.gamma
    f = op.face
    dim = m.dim
    force = np.zeros(dim, dtype=np.result_type(q.dtype, np.float64))   # complex q: complex-step derivatives
    ent = 0.0
    mass = 0.0
    for bc in bcnums:
        for bi in range(m.bndry_offsets[bc], m.bndry_offsets[bc + 1]):
            e, fc = m.bndryfaces[bi]
            qb = f.interp.T @ q[e][f.perm[:, fc]]
            for k in range(f.numnodes):
                nrm = m.nrm_bndry[bi, k]
                w = f.wface[k]
                qk = qb[k]
                nh = nrm / np.sqrt((nrm * nrm).sum())
                nm = (nh * qk[1:dim + 1]).sum()
                qg = qk.copy()
                qg[1:dim + 1] -= nh * nm
                p = (g - 1) * (qg[-1] - 0.5 * (qg[1:dim + 1] ** 2).sum() / qg[0])
                U = (qg[1:dim + 1] * nrm).sum() / qg[0]
                F = np.concatenate([[qg[0] * U], qg[1:dim + 1] * U + nrm * p, [(qg[-1] + p) * U]])
                force += w * F[1:dim + 1]
                unet = (qk[1:dim + 1] * nrm).sum()
                pk = (g - 1) * (qk[-1] - 0.5 * (qk[1:dim + 1] ** 2).sum() / qk[0])
                s = np.log(pk) - g * np.log(qk[0])
                Uent = -qk[0] * s / (g - 1)
                ent += w * unet * Uent
                mass += w * unet
    aoa = ph.aoa_deg * np.pi / 180
    if dim == 2:
        lift = -np.sin(aoa) * force[0] + np.cos(aoa) * force[1]
        drag = np.cos(aoa) * force[0] + np.sin(aoa) * force[1]
    else:
        lift = -np.sin(aoa) * force[0] + np.cos(aoa) * force[2]
        drag = np.cos(aoa) * force[0] + np.sin(aoa) * force[2]
    # LiftCoefficient / DragCoefficient (boundary_functional.jl:180-191):
    # divided by the dynamic pressure 0.5 rho_free Ma^2
    qdyn = 0.5 * ph.rho_free * ph.Ma * ph.Ma
    return {"force": force, "lift": lift, "drag": drag, "entropy": ent, "mass": mass,
            "lift_coef": lift / qdyn, "drag_coef": drag / qdyn}
