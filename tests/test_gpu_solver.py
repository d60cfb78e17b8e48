"""GPU parity of the explicit Jacobian (BSR), preconditioner, GMRES, Newton
and functionals against the CPU oracle, through the C ABI.

Reference tests restated: explicit Jacobian vs complex-step coloring Jacobian
(test/euler/test_jac.jl:2284-2367), explicit matrix-vector product vs
complex-step Jv (test_jac.jl:2370-2457), exact sparsity (test_jac.jl:2453-2460),
Newton convergence to the absolute tolerance."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

from pdes_amd import abi, euler, linsolve, problems
from pdes_amd._lib import check, lib
from helpers import setup, state_for

pytestmark = pytest.mark.gpu

JCASES = [("vortex", 3, 1, "omega", "roe"), ("vortex", 3, 2, "omega", "roe"), ("vortex", 2, 3, "omega", "roe"),
          ("vortex", 3, 1, "diage", "roe"), ("vortex", 3, 2, "diage", "roe"), ("vortex", 2, 4, "diage", "roe"),
          ("vortex", 3, 1, "diage", "ir"), ("vortex", 3, 2, "diage", "ir"), ("vortex", 2, 4, "diage", "ir"),
          ("cube", 1, 1, "omega", "roe"), ("cube", 1, 2, "omega", "roe")]


def _eqn(kind, n, p, opk, flux, q=None):
    m, op, ph, b = setup(kind, n, p, opk, flux)
    if q is None:
        q = state_for(kind, m, ph)
    eqn = euler.EulerData(m, op, ph)
    eqn.set_q(q)
    return m, op, ph, b, q, eqn


@pytest.mark.parametrize("kind,n,p,opk,flux", JCASES)
def test_bsr_pattern_bit_exact(oracle, kind, n, p, opk, flux):
    print(f"case {kind} n={n} p={p} {opk} {flux}", flush=True)
    m, op, ph, b, q, eqn = _eqn(kind, n, p, opk, flux)
    rowptr, colind = linsolve.getSparsityPattern(eqn)
    r_ref, c_ref = oracle.sparsity_bsr(b)
    assert np.array_equal(rowptr, r_ref)
    assert np.array_equal(colind, c_ref)
    eqn.free()


@pytest.mark.parametrize("kind,n,p,opk,flux", JCASES)
def test_jacobian_fp64_matches_complex_step(oracle, kind, n, p, opk, flux):
    print(f"case {kind} n={n} p={p} {opk} {flux}", flush=True)
    m, op, ph, b, q, eqn = _eqn(kind, n, p, opk, flux)
    linsolve.evalJacobian(m, op, eqn, fp64_blocks=True)
    got = linsolve.getJacobianValues(eqn, fp64_blocks=True)
    rowptr, colind = oracle.sparsity_bsr(b)
    ref = oracle.jacobian_cs(b, q, rowptr, colind)
    scale = np.abs(ref).max()
    err = np.abs(got - ref).max()
    assert err <= 1e-12 * scale, f"max err {err:.3e} scale {scale:.3e}"
    # explicit matvec == complex-step Jv (test_jac.jl:2370-2457)
    v = np.random.default_rng(2).standard_normal(q.shape)
    jv_ref = oracle.jv_complex_step(b, q, v)
    lo = linsolve.HIPBSRLO(eqn, fp64_blocks=True)
    out = np.zeros_like(v)
    lo.applyLinearOperator(m, op, eqn, None, None, 0.0, v, out)
    assert np.abs(out - jv_ref).max() <= 1e-11 * np.abs(jv_ref).max()
    eqn.free()


@pytest.mark.parametrize("kind,n,p,opk,flux", [JCASES[1], JCASES[4], JCASES[7], JCASES[10]])
def test_jacobian_fp32_blocks(oracle, kind, n, p, opk, flux):
    """FP32 storage of FP64-computed blocks: each entry is the rounded FP64 value."""
    m, op, ph, b, q, eqn = _eqn(kind, n, p, opk, flux)
    linsolve.evalJacobian(m, op, eqn, fp64_blocks=False)
    got = linsolve.getJacobianValues(eqn, fp64_blocks=False)
    assert got.dtype == np.float32
    rowptr, colind = oracle.sparsity_bsr(b)
    ref = oracle.jacobian_cs(b, q, rowptr, colind)
    err = np.abs(got.astype(np.float64) - ref)
    assert np.all(err <= 6e-8 * np.abs(ref) + 1e-12 * np.abs(ref).max())
    eqn.free()


@pytest.mark.parametrize("kind,n,p,opk,flux", [JCASES[1], JCASES[5], JCASES[10]])
def test_block_jacobi_apply(kind, n, p, opk, flux):
    """NewtonBDiagPC: x = D_e^{-1} b per element, checked against numpy solves
    of the same FP32 element blocks (FP32 factor/apply tolerance)."""
    m, op, ph, b, q, eqn = _eqn(kind, n, p, opk, flux)
    linsolve.evalJacobian(m, op, eqn, fp64_blocks=False)
    vals = linsolve.getJacobianValues(eqn).astype(np.float64)
    rowptr, colind = linsolve.getSparsityPattern(eqn)
    pc = linsolve.HIPBJacobiPC(eqn)
    pc.calcPC()
    rhs = np.random.default_rng(3).standard_normal(q.shape)
    x = np.zeros_like(rhs)
    pc.applyPC(m, op, eqn, None, 0.0, rhs, x)
    Np, nd = op.numnodes, m.dim + 2
    nb = Np * nd
    worst = 0.0
    for e in range(m.numEl):
        D = np.zeros((nb, nb))
        for i in range(Np):
            row = e * Np + i
            for pos in range(rowptr[row], rowptr[row + 1]):
                j = colind[pos] - e * Np
                if 0 <= j < Np:
                    D[i * nd:(i + 1) * nd, j * nd:(j + 1) * nd] = vals[pos]
        xe = np.linalg.solve(D, rhs[e].ravel())
        worst = max(worst, np.abs(x[e].ravel() - xe).max() / np.abs(xe).max())
    assert worst < 1e-3, worst
    eqn.free()


@pytest.mark.parametrize("pc_kind", [abi.PDES_PC_NONE, abi.PDES_PC_BJACOBI])
def test_gmres_solves_to_tolerance(pc_kind):
    m, op, ph, b, q, eqn = _eqn("vortex", 4, 2, "diage", "roe")
    lo = linsolve.HIPBSRLO(eqn, fp64_blocks=True)
    pc = linsolve.HIPBJacobiPC(eqn) if pc_kind == abi.PDES_PC_BJACOBI else linsolve.PCNone(eqn)
    # without a preconditioner restarted GMRES(30) stagnates on this DG
    # Jacobian; a long restart keeps the unpreconditioned case convergent
    restart = 30 if pc_kind == abi.PDES_PC_BJACOBI else 400
    tol = 1e-10 if pc_kind == abi.PDES_PC_BJACOBI else 1e-8
    ls = linsolve.StandardLinearSolver(pc, lo, reltol=tol, restart=restart, itermax=3000)
    ls.calcPCandLO()
    rhs = np.random.default_rng(4).standard_normal(q.shape)
    x = np.zeros_like(rhs)
    its = ls.linearSolve(rhs, x)
    Ax = np.zeros_like(x)
    lo.applyLinearOperator(m, op, eqn, None, None, 0.0, x, Ax)
    rel = np.linalg.norm(Ax - rhs) / np.linalg.norm(rhs)
    print("gmres iterations", its, "rel residual", rel)
    assert its < 3000 and rel < 2 * tol
    # the Givens estimate tracks the true residual
    assert abs(ls.last_resnorm / np.linalg.norm(rhs) - rel) < 0.1 * rel + 1e-13
    eqn.free()


def _minv(m, op):
    return np.repeat((1.0 / (op.w[None, :] / m.jac))[:, :, None], m.dim + 2, axis=2)


def test_newton_config0_matches_direct_solve(oracle):
    """Config 0: 20x20 quarter annulus (800 triangles), Omega p=1, Roe,
    vortex BC, exact solution times (1 + U(-1e-3, 1e-3)).  The oracle runs
    Newton with a sparse direct solve; the device runs Newton with FP64
    blocks and GMRES to 1e-12: same iteration count within 1, both reach
    ||R|| <= 1e-10 and the converged states agree."""
    m, op, ph = problems.make_problem("vortex", 20, 1, "omega", "roe")
    b = abi.make_descriptors(m, op, ph)
    q0 = problems.vortex_noise_state(m, ph, seed=0)
    q_ref, norms = oracle.newton_direct(b, q0, _minv(m, op), res_abstol=1e-10)
    eqn = euler.EulerData(m, op, ph)
    eqn.set_q(q0)
    rep = linsolve.newton(m, op, eqn, linsolve.NewtonOptions(res_abstol=1e-10, krylov_reltol=1e-12,
                                                             fp64_blocks=True, krylov_itermax=5000))
    q = eqn.get_q()
    print("oracle norms", norms, "device iterations", rep.iterations, "norm", rep.res_norm)
    assert rep.converged and rep.res_norm <= 1e-10
    assert abs(rep.iterations - (len(norms) - 1)) <= 1
    assert np.abs(q - q_ref).max() <= 1e-9 * np.abs(q_ref).max()
    eqn.free()


def test_newton_mixed_precision_matches_fp64():
    """FP32 Jacobian blocks and FP32 block-Jacobi under the FP64 inexact Newton
    loop reach ||R|| <= 1e-10 and the same state as the FP64-everywhere run."""
    m, op, ph = problems.make_problem("vortex", 12, 2, "diage", "roe")
    q0 = problems.vortex_noise_state(m, ph, seed=0)
    out = {}
    for fp64 in (True, False):
        eqn = euler.EulerData(m, op, ph)
        eqn.set_q(q0)
        rep = linsolve.newton(m, op, eqn, linsolve.NewtonOptions(res_abstol=1e-10, krylov_reltol=1e-6,
                                                                 fp64_blocks=fp64))
        assert rep.converged, (fp64, rep.res_norm)
        out[fp64] = (eqn.get_q(), rep.iterations)
        eqn.free()
    dq = np.abs(out[True][0] - out[False][0]).max()
    print("iterations fp64", out[True][1], "mixed", out[False][1], "max |dq|", dq)
    assert dq <= 1e-10 * np.abs(out[True][0]).max()


def test_newton_matrix_free():
    """Matrix-free Jv (jv_mode 1) with the FP32 block-Jacobi preconditioner."""
    m, op, ph = problems.make_problem("vortex", 8, 1, "omega", "roe")
    eqn = euler.EulerData(m, op, ph)
    eqn.set_q(problems.vortex_noise_state(m, ph, seed=0))
    rep = linsolve.newton(m, op, eqn, linsolve.NewtonOptions(res_abstol=1e-10, krylov_reltol=1e-6, jv_mode=1))
    assert rep.converged and rep.res_norm <= 1e-10
    eqn.free()


@pytest.mark.parametrize("jv_mode", [0, 1])
def test_newton_implicit_euler_globalization(oracle, jv_mode):
    """A start 30% off the vortex (Omega p=2, 6x6 annulus): plain Newton
    drives the pressure negative in the oracle, the implicit-Euler
    globalization (euler_tau 1) converges.  The device follows the oracle's
    globalized residual history within 1e-6 relative per iteration (the only
    difference is the inexact GMRES solve) and reaches the same state; with
    jv_mode 1 the shift is applied to the matrix-free product."""
    m, op, ph = problems.make_problem("vortex", 6, 2, "omega", "roe", vortex_Min=0.9)
    b = abi.make_descriptors(m, op, ph)
    q0 = problems.vortex_noise_state(m, ph, seed=1, amp=0.3)
    with pytest.raises(RuntimeError):
        oracle.newton_direct(b, q0, _minv(m, op), res_abstol=1e-10)
    q_ref, norms = oracle.newton_direct(b, q0, _minv(m, op), res_abstol=1e-10, euler_tau=1.0, jac=m.jac)
    eqn = euler.EulerData(m, op, ph)
    eqn.set_q(q0)
    rep = linsolve.newton(m, op, eqn, linsolve.NewtonOptions(res_abstol=1e-10, krylov_reltol=1e-12, fp64_blocks=True,
                                                             krylov_itermax=5000, globalize_euler=True,
                                                             euler_tau=1.0, jv_mode=jv_mode))
    q = eqn.get_q()
    print("oracle norms", ["%.3e" % x for x in norms], "device iterations", rep.iterations, "norm", rep.res_norm)
    assert rep.converged and rep.res_norm <= 1e-10
    assert rep.iterations == len(norms) - 1
    assert np.abs(q - q_ref).max() <= 1e-9 * np.abs(q_ref).max()
    eqn.free()


def test_newton_3d_cube():
    m, op, ph = problems.make_problem("cube", 3, 2, "omega", "roe", Ma=0.5, aoa_deg=2.0)
    eqn = euler.EulerData(m, op, ph)
    eqn.set_q(problems.freestream_bump_state(m, ph, seed=2))
    rep = linsolve.newton(m, op, eqn, linsolve.NewtonOptions(res_abstol=1e-10, krylov_reltol=1e-4))
    print("3D newton iterations", rep.iterations, "krylov", rep.total_krylov_iters, "norm", rep.res_norm)
    assert rep.converged and rep.res_norm <= 1e-10
    eqn.free()


@pytest.mark.parametrize("kind,n,p,opk", [("vortex", 6, 2, "omega"), ("cube", 2, 2, "omega")])
def test_functionals(oracle, kind, n, p, opk):
    m, op, ph = problems.make_problem(kind, n, p, opk, "roe", Ma=0.5, aoa_deg=3.0)
    q = problems.vortex_noise_state(m, ph) if kind == "vortex" else problems.freestream_bump_state(m, ph)
    eqn = euler.EulerData(m, op, ph)
    eqn.set_q(q)
    bcs = list(range(m.numBC))
    ref = oracle.functional(m, op, ph, q, bcs)
    bc = np.asarray(bcs, dtype=np.int32)
    val = np.zeros(3)
    for kind_id, key in [(1, "lift"), (2, "drag"), (3, "entropy"), (4, "mass")]:
        check(lib().pdes_eval_functional(eqn.ctx, kind_id, bc.ctypes.data, len(bcs), val.ctypes.data_as(C.POINTER(C.c_double))))
        assert abs(val[0] - ref[key]) <= 1e-12 * (1 + abs(ref[key])), (key, val[0], ref[key])
    eqn.free()


def test_functionals_closed_form():
    """Uniform freestream: zero net force and zero entropy flux over the
    closed boundary (test/euler/test_functionals.jl:150-158)."""
    m, op, ph = problems.make_problem("cube", 2, 2, "omega", "roe", Ma=0.5, aoa_deg=3.0)
    eqn = euler.EulerData(m, op, ph)
    eqn.set_q(problems.freestream(m.coords.shape[:-1], 3, ph))
    bc = np.arange(m.numBC, dtype=np.int32)
    val = np.zeros(3)
    check(lib().pdes_eval_functional(eqn.ctx, 0, bc.ctypes.data, len(bc), val.ctypes.data_as(C.POINTER(C.c_double))))
    assert np.abs(val).max() < 1e-13
    check(lib().pdes_eval_functional(eqn.ctx, 3, bc.ctypes.data, len(bc), val.ctypes.data_as(C.POINTER(C.c_double))))
    assert abs(val[0]) < 1e-13
    eqn.free()


def _dilu_reference(vals, rowptr, colind, m, op, b):
    """Element-block ILU(0), natural element order, from the device's BSR
    blocks: D'_e = A_ee - sum_{k<e} A_ek D'_k^-1 A_ke, then the two sweeps of
    M = (D' + L) D'^-1 (D' + U)."""
    nd = m.dim + 2
    Np = op.numnodes
    bs = Np * nd
    nel = m.numEl
    N = nel * bs
    J = sp.bsr_matrix((vals.astype(np.float64), colind, rowptr), shape=(N, N)).tocsr()
    nb = [set() for _ in range(nel)]
    for (eL, eR, *_r) in m.interfaces:
        nb[eL].add(eR)
        nb[eR].add(eL)

    def blk(i, k):
        return J[i * bs:(i + 1) * bs, k * bs:(k + 1) * bs].toarray()

    Dp = {}
    for e in range(nel):
        D = blk(e, e)
        for k in nb[e]:
            if k < e:
                D -= blk(e, k) @ np.linalg.solve(Dp[k], blk(k, e))
        Dp[e] = D
    x = np.zeros(N)
    bb = b.ravel()
    for e in range(nel):
        s = bb[e * bs:(e + 1) * bs].copy()
        for k in nb[e]:
            if k < e:
                s -= blk(e, k) @ x[k * bs:(k + 1) * bs]
        x[e * bs:(e + 1) * bs] = np.linalg.solve(Dp[e], s)
    for e in reversed(range(nel)):
        t = np.zeros(bs)
        for k in nb[e]:
            if k > e:
                t += blk(e, k) @ x[k * bs:(k + 1) * bs]
        x[e * bs:(e + 1) * bs] -= np.linalg.solve(Dp[e], t)
    return x.reshape(b.shape)


@pytest.mark.parametrize("kind,n,p,opk,flux", [("vortex", 3, 2, "omega", "roe"), ("vortex", 3, 3, "diage", "ir"),
                                              ("vortex", 2, 4, "diage", "ir"), ("cube", 1, 2, "omega", "roe")])
def test_ilu0_apply_matches_numpy(kind, n, p, opk, flux):
    """The device ILU(0) (level-scheduled factor and sweeps, FP32 inverses of
    the modified blocks) against a dense numpy restatement of the same
    element-block factorisation, from the same BSR values."""
    print(f"case {kind} n={n} p={p} {opk} {flux}", flush=True)
    m, op, ph, b, q, eqn = _eqn(kind, n, p, opk, flux)
    linsolve.evalJacobian(m, op, eqn, fp64_blocks=False)
    vals = linsolve.getJacobianValues(eqn)
    rowptr, colind = linsolve.getSparsityPattern(eqn)
    pc = linsolve.HIPILU0PC(eqn)
    pc.calcPC()
    rhs = np.random.default_rng(7).standard_normal(q.shape)
    x = np.zeros_like(rhs)
    pc.applyPC(m, op, eqn, None, 0.0, rhs, x)
    ref = _dilu_reference(vals, rowptr, colind, m, op, rhs)
    err = np.abs(x - ref).max() / np.abs(ref).max()
    print("ilu0 max rel err", err)
    assert err < 2e-3
    eqn.free()


def test_ilu0_fewer_iterations_than_block_jacobi():
    """Two Newton steps into a diagonal-E p=3 split-form solve, the remaining
    error is smooth and element block-Jacobi needs many GMRES(30) iterations;
    natural-order ILU(0) needs several times fewer on the same system."""
    m, op, ph = problems.make_problem("vortex", 12, 3, "diage", "ir", vortex_Min=0.5)
    eqn = euler.EulerData(m, op, ph)
    eqn.set_q(problems.vortex_noise_state(m, ph, seed=3))
    linsolve.newton(m, op, eqn, linsolve.NewtonOptions(res_abstol=1e-10, itermax=2))
    euler.evalResidual(m, op, eqn)
    rhs = -eqn.res.copy()
    its, rels = {}, {}
    for name, pc in (("ilu0", linsolve.HIPILU0PC(eqn)), ("bjacobi", linsolve.HIPBJacobiPC(eqn))):
        lo = linsolve.HIPBSRLO(eqn)
        ls = linsolve.StandardLinearSolver(pc, lo, reltol=1e-2, restart=30, itermax=3000)
        ls.calcPCandLO()
        x = np.zeros_like(rhs)
        its[name] = ls.linearSolve(rhs, x)
        Ax = np.zeros_like(x)
        lo.applyLinearOperator(m, op, eqn, None, None, 0.0, x, Ax)
        rels[name] = np.linalg.norm(Ax - rhs) / np.linalg.norm(rhs)
        print(name, "iterations", its[name], "true rel residual", rels[name])
    # ILU(0) reaches the tolerance; block-Jacobi may stop at the iteration cap
    assert rels["ilu0"] < 1.1e-2
    assert its["ilu0"] * 3 <= its["bjacobi"]
    eqn.free()


@pytest.mark.parametrize("kind,n,p,opk,flux", [("vortex", 8, 3, "diage", "ir"), ("cube", 2, 2, "omega", "roe")])
def test_newton_with_ilu0(kind, n, p, opk, flux):
    m, op, ph = problems.make_problem(kind, n, p, opk, flux, Ma=0.5, aoa_deg=2.0)
    eqn = euler.EulerData(m, op, ph)
    q0 = problems.vortex_noise_state(m, ph, seed=0) if kind == "vortex" else problems.freestream_bump_state(m, ph)
    eqn.set_q(q0)
    rep = linsolve.newton(m, op, eqn, linsolve.NewtonOptions(res_abstol=1e-10, pc_kind=abi.PDES_PC_ILU0))
    print("ILU0 newton iterations", rep.iterations, "krylov", rep.total_krylov_iters, "norm", rep.res_norm)
    assert rep.converged and rep.res_norm <= 1e-10
    eqn.free()


def test_ilu0_graph_replays_match_eager(monkeypatch):
    """The level-by-level ILU(0) apply (PDES_ILU_LEVELS) replays a captured
    HIP graph of its level launches.  Every replay, with other device work and
    synchronisations in between, gives the same bits as the directly launched
    sweeps of a context created with PDES_NO_GRAPHS."""
    m, op, ph = problems.make_problem("vortex", 6, 3, "diage", "ir")
    q = problems.vortex_noise_state(m, ph, seed=5)
    ctxs = []
    monkeypatch.setenv("PDES_ILU_LEVELS", "1")
    for eager in (False, True):
        if eager:
            monkeypatch.setenv("PDES_NO_GRAPHS", "1")
        else:
            monkeypatch.delenv("PDES_NO_GRAPHS", raising=False)
        e = euler.EulerData(m, op, ph)
        e.set_q(q)
        linsolve.evalJacobian(m, op, e)
        pc = linsolve.HIPILU0PC(e)
        pc.calcPC()
        ctxs.append((e, pc))
    rng = np.random.default_rng(11)
    for k in range(4):
        b = rng.standard_normal(q.shape)
        out = []
        for e, pc in ctxs:
            euler.evalResidual(m, op, e)          # other work between replays
            x = np.zeros_like(b)
            pc.applyPC(m, op, e, None, 0.0, b, x)
            out.append(x)
        assert np.array_equal(out[0], out[1]), f"replay {k} differs from the eager sweeps"
    for e, pc in ctxs:
        e.free()


DEP_CASES = [("vortex", 6, 2, "diage", "ir", False), ("vortex", 5, 3, "diage", "ir", True),
             ("vortex", 4, 4, "diage", "ir", False), ("vortex", 5, 2, "omega", "roe", False),
             ("vortex", 4, 1, "omega", "roe", True), ("cube", 2, 2, "omega", "roe", False),
             ("cube", 2, 1, "omega", "roe", False), ("vortex", 128, 2, "diage", "roe", False)]


@pytest.mark.parametrize("kind,n,p,opk,flux,fp64", DEP_CASES)
def test_ilu0_dependency_sweeps_match_level_sweeps(monkeypatch, kind, n, p, opk, flux, fp64):
    """The dependency-driven ILU(0) sweeps (one launch per triangle, elements
    in ticket order, each loading its lower or upper neighbours' unknowns
    until none holds the pending sentinel; forced with PDES_ILU_DEP) give the
    bits of the level-by-level sweeps (forced with PDES_ILU_LEVELS), over
    repeated applications (the sweep vector returns to the sentinel after
    each, nothing is cleared) with other device work in between, for one-lane rows (diagonal-E) and several lanes
    per row (dense faces), FP32 and FP64 blocks, up to 128 x 128 x 2
    triangles.  Unforced, a cost model picks one of the two per application,
    so either choice gives the same result."""
    print(f"case {kind} n={n} p={p} {opk} {flux} fp64={fp64}", flush=True)
    m, op, ph = problems.make_problem(kind, n, p, opk, flux, Ma=0.5, aoa_deg=2.0)
    q = problems.vortex_noise_state(m, ph, seed=5) if kind == "vortex" else problems.freestream_bump_state(m, ph)
    ctxs = []
    for levels in (False, True):
        monkeypatch.setenv("PDES_ILU_LEVELS" if levels else "PDES_ILU_DEP", "1")
        monkeypatch.delenv("PDES_ILU_DEP" if levels else "PDES_ILU_LEVELS", raising=False)
        e = euler.EulerData(m, op, ph)
        e.set_q(q)
        linsolve.evalJacobian(m, op, e, fp64_blocks=fp64)
        pc = linsolve.HIPILU0PC(e)
        pc.calcPC()
        ctxs.append((e, pc))
    rng = np.random.default_rng(13)
    for k in range(3):
        b = rng.standard_normal(q.shape)
        out = []
        for e, pc in ctxs:
            euler.evalResidual(m, op, e)
            x = np.zeros_like(b)
            pc.applyPC(m, op, e, None, 0.0, b, x)
            out.append(x)
        assert np.all(np.isfinite(out[0]))
        assert np.array_equal(out[0], out[1]), f"application {k}: dependency sweeps differ from level sweeps"
    for e, pc in ctxs:
        e.free()


@pytest.mark.parametrize("kind,n,p,opk,flux", [JCASES[1], JCASES[5], JCASES[8], JCASES[10]])
def test_coloring_jacobian_matches_elementwise(kind, n, p, opk, flux):
    """calcJacobianSparse on the device (distance-2 colouring, one derivative
    evaluation per colour/node/dof) equals the element-wise Jacobian kernel."""
    print(f"case {kind} n={n} p={p} {opk} {flux}", flush=True)
    m, op, ph, b, q, eqn = _eqn(kind, n, p, opk, flux)
    linsolve.evalJacobian(m, op, eqn, fp64_blocks=True)
    a = linsolve.getJacobianValues(eqn, fp64_blocks=True).copy()
    linsolve.calcJacobianSparse(m, op, eqn, fp64_blocks=True)
    c = linsolve.getJacobianValues(eqn, fp64_blocks=True)
    err = np.abs(a - c).max()
    assert err <= 1e-12 * np.abs(a).max(), err
    eqn.free()


@pytest.mark.parametrize("kind,n,p", [("vortex", 3, 1), ("vortex", 3, 2), ("vortex", 2, 3), ("cube", 1, 1),
                                      ("cube", 1, 2)])
def test_entropy_stable_elementwise_jacobian_matches_coloring(kind, n, p):
    """The element-wise derivative of the entropy-stable Omega scheme (split
    volume, entropy-conservative face pairs, entropy penalty through dw/dq)
    equals the colouring assembly, which differentiates the whole residual."""
    print(f"case {kind} n={n} p={p} omega es", flush=True)
    m, op, ph, b = setup(kind, n, p, "omega", "es", Ma=0.5, aoa_deg=2.0)
    q = state_for(kind, m, ph)
    eqn = euler.EulerData(m, op, ph)
    eqn.set_q(q)
    linsolve.evalJacobian(m, op, eqn, fp64_blocks=True)
    a = linsolve.getJacobianValues(eqn, fp64_blocks=True).copy()
    linsolve.calcJacobianSparse(m, op, eqn, fp64_blocks=True)
    c = linsolve.getJacobianValues(eqn, fp64_blocks=True)
    err = np.abs(a - c).max()
    print("max block difference", err, "scale", np.abs(a).max(), flush=True)
    assert err <= 1e-11 * np.abs(a).max(), err
    eqn.free()


@pytest.mark.parametrize("kind,n,p", [("vortex", 3, 2), ("cube", 1, 2)])
def test_entropy_stable_omega_jacobian_and_newton(oracle, kind, n, p):
    """Entropy-stable Omega scheme (split-form volume + ESLF element face
    integral): its element-wise Jacobian gives J v equal to the oracle's
    complex-step J v, and Newton with ILU(0) converges."""
    m, op, ph, b = setup(kind, n, p, "omega", "es", Ma=0.5, aoa_deg=2.0)
    q = state_for(kind, m, ph)
    eqn = euler.EulerData(m, op, ph)
    eqn.set_q(q)
    linsolve.evalJacobian(m, op, eqn, fp64_blocks=True)
    v = np.random.default_rng(3).standard_normal(q.shape)
    lo = linsolve.HIPBSRLO(eqn, fp64_blocks=True)
    out = np.zeros_like(v)
    lo.applyLinearOperator(m, op, eqn, None, None, 0.0, v, out)
    ref = oracle.jv_complex_step(b, q, v)
    assert np.abs(out - ref).max() <= 1e-11 * np.abs(ref).max()
    eqn.set_q(problems.vortex_noise_state(m, ph, seed=1) if kind == "vortex" else problems.freestream_bump_state(m, ph))
    rep = linsolve.newton(m, op, eqn, linsolve.NewtonOptions(res_abstol=1e-10, pc_kind=abi.PDES_PC_ILU0))
    print("ES Newton iterations", rep.iterations, "krylov", rep.total_krylov_iters, "norm", rep.res_norm)
    assert rep.converged and rep.res_norm <= 1e-10
    eqn.free()


def test_ilu0_dependency_wait_limit_reports_an_error(monkeypatch):
    """A dependency-driven ILU(0) sweep whose waits give up (forced here by a
    one-poll bound, PDES_ILU_WAIT_POLLS=1) still ends every launch, and the
    solver reports the invalid preconditioner application as an error
    instead of returning its result; the same solve with the normal bound
    succeeds."""
    from pdes_amd._lib import PdesError
    m, op, ph = problems.make_problem("vortex", 40, 2, "diage", "ir")
    q = problems.vortex_noise_state(m, ph, seed=5)
    monkeypatch.setenv("PDES_ILU_DEP", "1")
    for limited in (True, False):
        if limited:
            monkeypatch.setenv("PDES_ILU_WAIT_POLLS", "1")
        else:
            monkeypatch.delenv("PDES_ILU_WAIT_POLLS", raising=False)
        e = euler.EulerData(m, op, ph)
        e.set_q(q)
        linsolve.evalJacobian(m, op, e)
        pc = linsolve.HIPILU0PC(e)
        pc.calcPC()
        b = np.random.default_rng(2).standard_normal(q.shape)
        x = np.zeros_like(b)
        if limited:
            with pytest.raises(PdesError, match="gave up waiting"):
                for _ in range(5):
                    pc.applyPC(m, op, e, None, 0.0, b, x)
        else:
            pc.applyPC(m, op, e, None, 0.0, b, x)
            assert np.all(np.isfinite(x))
        e.free()
