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
         ("tet8cube0.smb", 2, "omega", "es")]


@pytest.mark.parametrize("name,p,opk,flux", CASES)
def test_residual_and_jacobian_on_reference_mesh(oracle, name, p, opk, flux):
    print(f"case {name} p={p} {opk} {flux}", flush=True)
    bc = "freestream" if name.startswith("tet") else "vortex"
    m, op, ph = problems.smb_problem(os.path.join(MESHES, name), p, opk, flux, bc, Ma=0.5, aoa_deg=2.0)
    b = abi.make_descriptors(m, op, ph)
    q = state_for("cube" if bc == "freestream" else "vortex", m, ph)
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
