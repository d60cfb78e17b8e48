"""Homotopy globalisation through the C ABI: the first-order dissipation D(q)
(calcHomotopyDiss) and its directional derivative against the oracle, the
Jacobian of (1 - lambda) R + lambda D against the oracle's complex-step
products (test/euler/test_homotopy.jl checks the same operators: zero on
uniform flow, explicit Jacobian against products), and
predictorCorrectorHomotopy reaching the Newton solution."""
import numpy as np
import pytest

from pdes_amd import abi, euler, linsolve, problems
from helpers import setup, state_for

pytestmark = pytest.mark.gpu

CASES = [("vortex", 3, 1, "omega", "roe"), ("vortex", 3, 2, "omega", "roe"), ("vortex", 2, 3, "diage", "ir"),
         ("cube", 1, 2, "omega", "roe")]


def _eqn(kind, n, p, opk, flux, seed=0):
    m, op, ph, b = setup(kind, n, p, opk, flux)
    q = state_for(kind, m, ph, seed=seed)
    q = q * (1 + 0.02 * np.random.default_rng(seed + 7).uniform(-1, 1, q.shape))
    eqn = euler.EulerData(m, op, ph)
    eqn.set_q(q)
    return m, op, ph, b, q, eqn


@pytest.mark.parametrize("kind,n,p,opk,flux", CASES)
def test_dissipation_matches_oracle(oracle, kind, n, p, opk, flux):
    print(f"case {kind} n={n} p={p} {opk}", flush=True)
    m, op, ph, b, q, eqn = _eqn(kind, n, p, opk, flux)
    d = linsolve.evalHomotopy(m, op, eqn)
    ref = oracle.homotopy(b, q)
    assert np.abs(d - ref).max() <= 1e-12 * np.abs(ref).max()
    v = np.random.default_rng(2).standard_normal(q.shape)
    jv = linsolve.evalHomotopy_jv(eqn, v)
    jref = oracle.homotopy_jv(b, q, v)
    assert np.abs(jv - jref).max() <= 1e-11 * np.abs(jref).max()
    # homotopy_addBoundaryIntegrals
    linsolve.setHomotopyBoundaryIntegrals(eqn, True)
    d = linsolve.evalHomotopy(m, op, eqn)
    ref = oracle.homotopy(b, q, bndry=True)
    assert np.abs(d - ref).max() <= 1e-12 * np.abs(ref).max()
    jv = linsolve.evalHomotopy_jv(eqn, v)
    jref = oracle.homotopy_jv(b, q, v, bndry=True)
    assert np.abs(jv - jref).max() <= 1e-11 * np.abs(jref).max()
    # uniform flow: zero (test_homotopy.jl)
    eqn.set_q(problems.freestream((m.numEl, op.numnodes), m.dim, ph))
    assert np.abs(linsolve.evalHomotopy(m, op, eqn)).max() <= 1e-13
    linsolve.setHomotopyBoundaryIntegrals(eqn, False)
    assert np.abs(linsolve.evalHomotopy(m, op, eqn)).max() <= 1e-13
    eqn.free()


@pytest.mark.parametrize("kind,n,p,opk,flux", CASES)
@pytest.mark.parametrize("lam", [0.0, 0.35, 1.0])
def test_homotopy_jacobian_matches_products(oracle, kind, n, p, opk, flux, lam):
    print(f"case {kind} n={n} p={p} {opk} lambda={lam}", flush=True)
    m, op, ph, b, q, eqn = _eqn(kind, n, p, opk, flux, seed=1)
    linsolve.calcHomotopyJacobian(m, op, eqn, lam, fp64_blocks=True)
    lo = linsolve.HIPBSRLO(eqn, fp64_blocks=True)
    v = np.random.default_rng(3).standard_normal(q.shape)
    out = np.zeros_like(v)
    lo.applyLinearOperator(m, op, eqn, None, None, 0.0, v, out)
    ref = (1 - lam) * oracle.jv_complex_step(b, q, v) + lam * oracle.homotopy_jv(b, q, v)
    assert np.abs(out - ref).max() <= 1e-11 * np.abs(ref).max()
    eqn.free()


def test_homotopy_jacobian_with_boundary_term(oracle):
    m, op, ph, b, q, eqn = _eqn("vortex", 3, 2, "omega", "roe", seed=2)
    linsolve.setHomotopyBoundaryIntegrals(eqn, True)
    lam = 0.35
    linsolve.calcHomotopyJacobian(m, op, eqn, lam, fp64_blocks=True)
    v = np.random.default_rng(8).standard_normal(q.shape)
    out = np.zeros_like(v)
    linsolve.HIPBSRLO(eqn, fp64_blocks=True).applyLinearOperator(m, op, eqn, None, None, 0.0, v, out)
    ref = (1 - lam) * oracle.jv_complex_step(b, q, v) + lam * oracle.homotopy_jv(b, q, v, bndry=True)
    assert np.abs(out - ref).max() <= 1e-11 * np.abs(ref).max()
    eqn.free()


@pytest.mark.parametrize("pc,bndry", [(abi.PDES_PC_BJACOBI, False), (abi.PDES_PC_ILU0, False),
                                     (abi.PDES_PC_ILU0, True)])
def test_predictor_corrector_reaches_newton_solution(oracle, pc, bndry):
    """Vortex, Omega p=2, Roe, from the exact solution with 5 % noise: the
    predictor-corrector path from lambda = 1 to 0 converges to ||R|| <= 1e-10
    and to the oracle's direct-solve Newton state."""
    m, op, ph, b = setup("vortex", 4, 2, "omega", "roe")
    q0 = problems.vortex_noise_state(m, ph, seed=3, amp=5e-2)
    minv = np.repeat((m.jac / op.w[None, :])[:, :, None], 4, axis=2)
    q_ref, norms = oracle.newton_direct(b, q0, minv, res_abstol=1e-11)
    eqn = euler.EulerData(m, op, ph)
    eqn.set_q(q0)
    opts = linsolve.NewtonOptions(res_abstol=1e-10, res_reltol=1e-300, itermax=60, krylov_reltol=1e-6,
                                  fp64_blocks=True, pc_kind=pc, gmres_restart=60, krylov_itermax=3000)
    rep = linsolve.predictorCorrectorHomotopy(m, op, eqn, opts, addBoundaryIntegrals=bndry)
    print("predictor iterations", rep.iterations, "residual", rep.res_norm, "krylov", rep.total_krylov_iters,
          flush=True)
    assert rep.converged and rep.res_norm <= 1e-10
    assert rep.iterations >= 2
    q = eqn.get_q()
    assert np.abs(q - q_ref).max() <= 1e-8 * np.abs(q_ref).max()
    eqn.free()
