"""Adjoint of the steady residual for boundary functionals, through the C ABI:
the functional derivative dJ/dq (evalFunctionalDeriv_q), the transposed
Jacobian product (applyLinearOperatorTranspose) and the adjoint solve
J^T psi = -dJ/dq (calcAdjoint), and the reverse-mode product psi^T dR/dq
(evalResidual_revq)."""
import numpy as np
import pytest
import scipy.sparse as sp

from pdes_amd import abi, euler, linsolve, problems
from helpers import setup, state_for

pytestmark = pytest.mark.gpu

CASES = [("vortex", 3, 1, "omega"), ("vortex", 3, 2, "omega"), ("vortex", 3, 2, "diage"), ("cube", 1, 2, "omega"),
         ("vortex", 3, 2, "omega-es")]


def _problem(kind, n, p, opk):
    # "omega-es": split-form volume with the entropy-stable element face integral
    flux = {"diage": "ir", "omega-es": "es"}.get(opk, "roe")
    m, op, ph = problems.make_problem(kind, n, p, opk.split("-")[0], flux, Ma=0.5, aoa_deg=3.0)
    q = problems.vortex_noise_state(m, ph, seed=2) if kind == "vortex" else problems.freestream_bump_state(m, ph)
    return m, op, ph, q


def _bsr(eqn, m):
    rowptr, colind = linsolve.getSparsityPattern(eqn, fp64_blocks=True)
    vals = linsolve.getJacobianValues(eqn, fp64_blocks=True)
    nd = m.dim + 2
    n = (len(rowptr) - 1) * nd
    return sp.bsr_matrix((vals, colind, rowptr), shape=(n, n)).tocsr()


@pytest.mark.parametrize("kind,n,p,opk", CASES)
def test_functional_derivative_matches_complex_step(oracle, kind, n, p, opk):
    """dJ/dq from forward-mode integrands equals the complex-step derivative of
    the oracle's functional, dof by dof (the reference tests evalFunctionalDeriv
    the same way, test/euler/test_adjoint.jl)."""
    print(f"case {kind} n={n} p={p} {opk}", flush=True)
    m, op, ph, q = _problem(kind, n, p, opk)
    eqn = euler.EulerData(m, op, ph)
    eqn.set_q(q)
    bcs = [0, 2] if m.numBC > 2 else [0]
    h = 1e-30
    for name, key in [("lift", "lift"), ("drag", "drag"), ("entropyflux", "entropy"), ("massflow", "mass")]:
        g = linsolve.evalFunctionalDeriv_q(eqn, name, bcs)
        # only the dofs of elements with a face on the listed BC groups can be nonzero
        els = set()
        for bc in bcs:
            for bi in range(m.bndry_offsets[bc], m.bndry_offsets[bc + 1]):
                els.add(int(m.bndryfaces[bi][0]))
        ref = np.zeros_like(q)
        for e in sorted(els):
            for i in range(q.shape[1]):
                for c in range(q.shape[2]):
                    qc = q.astype(np.complex128)
                    qc[e, i, c] += 1j * h
                    ref[e, i, c] = oracle.functional(m, op, ph, qc, bcs)[key].imag / h
        scale = np.abs(ref).max()
        assert scale > 0, name
        err = np.abs(g - ref).max()
        print(name, "max |dJ/dq - complex step|", err, "scale", scale, flush=True)
        assert err <= 1e-11 * scale, (name, err, scale)
    eqn.free()


@pytest.mark.parametrize("kind,n,p,opk", CASES)
def test_transposed_product(kind, n, p, opk):
    """J^T x equals scipy's transpose of the downloaded BSR matrix, and
    y . (J x) = x . (J^T y)."""
    m, op, ph, q = _problem(kind, n, p, opk)
    eqn = euler.EulerData(m, op, ph)
    eqn.set_q(q)
    linsolve.evalJacobian(m, op, eqn, fp64_blocks=True)
    J = _bsr(eqn, m)
    rng = np.random.default_rng(5)
    x = rng.standard_normal(q.shape)
    y = rng.standard_normal(q.shape)
    jt_y = np.zeros_like(y)
    linsolve.applyLinearOperatorTranspose(eqn, y, jt_y)
    ref = (J.T @ y.ravel()).reshape(y.shape)
    assert np.abs(jt_y - ref).max() <= 1e-13 * np.abs(ref).max()
    lo = linsolve.HIPBSRLO(eqn, fp64_blocks=True)
    jx = np.zeros_like(x)
    lo.applyLinearOperator(m, op, eqn, None, None, 0.0, x, jx)
    a, b = float((y * jx).sum()), float((x * jt_y).sum())
    assert abs(a - b) <= 1e-12 * (abs(a) + abs(b)), (a, b)
    eqn.free()


@pytest.mark.parametrize("kind,n,p,opk", CASES)
@pytest.mark.parametrize("pc", [abi.PDES_PC_BJACOBI, abi.PDES_PC_ILU0])
def test_adjoint_matches_dense_solve(kind, n, p, opk, pc):
    """psi from GMRES on J^T (preconditioner built from J^T) equals a dense
    solve of J^T psi = -dJ/dq; afterwards the context holds J again."""
    print(f"case {kind} n={n} p={p} {opk} pc={pc}", flush=True)
    m, op, ph, q = _problem(kind, n, p, opk)
    eqn = euler.EulerData(m, op, ph)
    eqn.set_q(q)
    bcs = [0]
    opts = linsolve.NewtonOptions(pc_kind=pc, fp64_blocks=True, gmres_restart=60, krylov_itermax=2000)
    psi, its, res = linsolve.calcAdjoint(m, op, eqn, "drag", bcs, opts, rtol=1e-12)
    print("adjoint GMRES iterations", its, "residual", res, flush=True)
    g = linsolve.evalFunctionalDeriv_q(eqn, "drag", bcs)
    after = linsolve.getJacobianValues(eqn, fp64_blocks=True).copy()
    linsolve.evalJacobian(m, op, eqn, fp64_blocks=True)
    # the context holds J again, not J^T
    assert np.array_equal(after, linsolve.getJacobianValues(eqn, fp64_blocks=True))
    J = _bsr(eqn, m)
    # backward error of the iterative solve, then the solution itself against
    # a dense solve (these small systems have condition numbers up to ~1e6)
    r = J.T @ psi.ravel() + g.ravel()
    assert np.abs(r).max() <= 1e-10 * np.abs(g).max(), np.abs(r).max()
    ref = np.linalg.solve(J.T.toarray(), -g.ravel()).reshape(q.shape)
    assert np.abs(psi - ref).max() <= 1e-5 * np.abs(ref).max()
    eqn.free()


@pytest.mark.parametrize("kind,n,p,opk,flux", [("vortex", 3, 2, "omega", "roe"), ("vortex", 3, 2, "diage", "ir"),
                                               ("vortex", 3, 2, "omega", "es"), ("cube", 1, 1, "omega", "roe")])
def test_residual_revq_matches_complex_step(oracle, kind, n, p, opk, flux):
    """psi^T dR/dq (evalResidual_revq) against the oracle's complex-step
    directional derivatives: (psi^T dR/dq) . v = psi . (dR/dq v) for several
    v; with zero_output=False the product is added to the output."""
    print(f"case {kind} n={n} p={p} {opk} {flux}", flush=True)
    m, op, ph, b = setup(kind, n, p, opk, flux, Ma=0.5, aoa_deg=2.0)
    q = state_for(kind, m, ph)
    eqn = euler.EulerData(m, op, ph)
    eqn.set_q(q)
    rng = np.random.default_rng(11)
    psi = rng.standard_normal(q.size)
    qbar = np.zeros(q.size)
    linsolve.evalResidual_revq(m, op, eqn, None, psi, qbar)
    for k in range(3):
        v = rng.standard_normal(q.shape)
        jv = oracle.jv_complex_step(b, q, v)
        a, c = float(psi @ jv.ravel()), float(qbar @ v.ravel())
        scale = float(np.abs(psi).sum() * np.abs(jv).max())
        assert abs(a - c) <= 1e-12 * scale, (k, a, c, scale)
    acc = np.ones(q.size)
    linsolve.evalResidual_revq(m, op, eqn, None, psi, acc, zero_output=False)
    assert np.array_equal(acc, np.ones(q.size) + qbar)
    eqn.free()
