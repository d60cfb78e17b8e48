"""Golden residual fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py): the oracle must keep reproducing them, and the
HIP residual and Jv must match them at the residual parity tolerance."""
import os

import numpy as np
import pytest

from pdes_amd import abi, euler, problems
from helpers import assert_parity

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = {
    "vortex_4x4_omega_p1_roe": ("vortex", 4, 1, "omega", "roe"),
    "vortex_3x3_omega_p3_roe": ("vortex", 3, 3, "omega", "roe"),
    "vortex_3x3_diage_p2_ir": ("vortex", 3, 2, "diage", "ir"),
    "vortex_2x2_diage_p4_ir": ("vortex", 2, 4, "diage", "ir"),
    "cube_1_omega_p2_roe": ("cube", 1, 2, "omega", "roe"),
    "vortex_3x3_omega_p2_es": ("vortex", 3, 2, "omega", "es"),
}


def _load(name):
    kind, n, p, opk, flux = CASES[name]
    m, op, ph = problems.make_problem(kind, n, p, opk, flux, Ma=0.5, aoa_deg=2.0)
    with np.load(os.path.join(HERE, name + ".npz")) as z:
        return m, op, ph, {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_reproduces_golden(oracle, name):
    m, op, ph, g = _load(name)
    b = abi.make_descriptors(m, op, ph)
    assert_parity(oracle.residual(b, g["q"]), g["res"], g["scale"], 1e-13, name)
    jv = oracle.jv_complex_step(b, g["q"], g["v"])
    assert np.abs(jv - g["jv"]).max() <= 1e-12 * (1 + np.abs(g["jv"]).max())


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_hip_matches_golden(name):
    print(f"golden case {name}", flush=True)
    m, op, ph, g = _load(name)
    eqn = euler.EulerData(m, op, ph)
    eqn.set_q(g["q"])
    euler.evalResidual(m, op, eqn)
    assert_parity(eqn.res, g["res"], g["scale"], 1e-12, name)
    jv = euler.evaldRdqProduct(eqn, g["v"])
    assert np.abs(jv - g["jv"]).max() <= 1e-11 * (1 + np.abs(g["jv"]).max())
    eqn.free()
