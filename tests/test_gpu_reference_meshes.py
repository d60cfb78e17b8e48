"""Device residual and Jacobian on the reference's own mesh files
(tests/golden/meshes, copied from src/mesh_files): the unstructured 2D vortex
annulus gvortex10, the square squarevortex_small and the 48-tet cube
tet8cube, whose interfaces use four different face orientations.  Parity with
the oracle as for the generated meshes: residual entrywise within 1e-12 of
its floating-point scale, the element-wise FP64 Jacobian's product within
1e-11 of the complex-step directional derivative."""
import os

import numpy as np
import pytest

from pdes_amd import abi, euler, linsolve, problems
from helpers import assert_parity, state_for

pytestmark = pytest.mark.gpu

MESHES = os.path.join(os.path.dirname(__file__), "golden", "meshes")
CASES = [("gvortex10.smb", 1, "omega", "roe"), ("gvortex10.smb", 3, "omega", "roe"),
         ("gvortex10.smb", 2, "diage", "ir"), ("gvortex10.smb", 2, "omega", "es"),
         ("squarevortex_small0.smb", 4, "omega", "roe"), ("squarevortex_small0.smb", 4, "diage", "ir"),
         ("tet8cube0.smb", 1, "omega", "roe"), ("tet8cube0.smb", 2, "omega", "roe"),
         ("tet8cube0.smb", 2, "omega", "es"),
         # one element (no interior face) and two elements (one interface)
         ("Test1el0.smb", 1, "omega", "roe"), ("Test1el0.smb", 3, "diage", "ir"),
         ("Test2el0.smb", 2, "omega", "roe"), ("Test2el0.smb", 2, "omega", "es")]


@pytest.mark.parametrize("name,p,opk,flux", CASES)
def test_residual_and_jacobian_on_reference_mesh(oracle, name, p, opk, flux):
    print(f"case {name} p={p} {opk} {flux}", flush=True)
    bc = "freestream" if name.startswith(("tet", "Test")) else "vortex"
    m, op, ph = problems.smb_problem(os.path.join(MESHES, name), p, opk, flux, bc, Ma=0.5, aoa_deg=2.0)
    b = abi.make_descriptors(m, op, ph)
    q = state_for("cube" if bc == "freestream" else "vortex", m, ph)
    if name.startswith("Test"):
        q = q * (1 + 0.01 * np.random.default_rng(9).uniform(-1, 1, q.shape))
    ref = oracle.residual(b, q)
    scale = oracle.residual_scale(b, q)
    eqn = euler.EulerData(m, op, ph)
    eqn.set_q(q)
    euler.evalResidual(m, op, eqn, None)
    assert_parity(eqn.res, ref, scale, 1e-12, f"{name} p={p} {opk} {flux}")
    linsolve.evalJacobian(m, op, eqn, fp64_blocks=True)
    v = np.random.default_rng(4).standard_normal(q.shape)
    out = np.zeros_like(v)
    linsolve.HIPBSRLO(eqn, fp64_blocks=True).applyLinearOperator(m, op, eqn, None, None, 0.0, v, out)
    jv = oracle.jv_complex_step(b, q, v)
    assert np.abs(out - jv).max() <= 1e-11 * np.abs(jv).max()
    eqn.free()


@pytest.mark.parametrize("name", ["Test1el0.smb", "Test2el0.smb"])
@pytest.mark.parametrize("pc", [abi.PDES_PC_BJACOBI, abi.PDES_PC_ILU0])
def test_newton_on_one_and_two_elements(oracle, name, pc):
    """Newton from a 1 %-perturbed free stream on the smallest meshes (free
    stream on every boundary face) converges to the oracle's direct-solve
    state; the homotopy dissipation matches the oracle there too."""
    print(f"case {name} pc={pc}", flush=True)
    m, op, ph = problems.smb_problem(os.path.join(MESHES, name), 2, "omega", "roe", "freestream",
                                     Ma=0.5, aoa_deg=2.0)
    b = abi.make_descriptors(m, op, ph)
    q0 = problems.freestream((m.numEl, op.numnodes), 2, ph)
    q0 = q0 * (1 + 0.01 * np.random.default_rng(10).uniform(-1, 1, q0.shape))
    minv = np.repeat((m.jac / op.w[None, :])[:, :, None], 4, axis=2)
    q_ref, norms = oracle.newton_direct(b, q0, minv, res_abstol=1e-12)
    eqn = euler.EulerData(m, op, ph)
    eqn.set_q(q0)
    d = linsolve.evalHomotopy(m, op, eqn)
    dref = oracle.homotopy(b, q0)
    assert np.abs(d - dref).max() <= 1e-12 * np.abs(dref).max()
    rep = linsolve.newton(m, op, eqn, linsolve.NewtonOptions(res_abstol=1e-12, krylov_reltol=1e-12,
                                                             fp64_blocks=True, pc_kind=pc))
    assert rep.converged
    q = eqn.get_q()
    assert np.abs(q - q_ref).max() <= 1e-10 * np.abs(q_ref).max()
    eqn.free()
