"""Discretisation accuracy of the device Newton solve on the reference's own
convergence meshes (squarevortex_small/large from src/mesh_files, read by the
PUMI reader): steady isentropic vortex, Omega operators, Roe faces, vortex BC
on all boundaries, Newton from the exact solution to ||R|| <= 1e-12, error
= calcNorm(|q - q_exact|) with the mass matrix.  The reference's windows
(test/euler/convergence/p{1..4}/conservative_dg/runtests.jl:26-41): slope
within 0.1 and coarse error within x/÷1.25 of
  p=1: 2.00, 0.0120   p=2: 3.00, 8.77e-4   p=3: 3.84, 6.87e-5   p=4: 4.87, 5.86e-6.
At p=4 the degree-7 15-node symmetric cubature is a one-parameter family, and
the member the reference's operator package uses is not reproducible without
it: there the operator differs, and the test checks the design order only
(slope above 4.5)."""
import os

import numpy as np
import pytest

from pdes_amd import abi, euler, linsolve, problems, pumi

pytestmark = pytest.mark.gpu

MESHES = os.path.join(os.path.dirname(__file__), "golden", "meshes")
PUBLISHED = {1: (2.00, 0.0120), 2: (3.00, 8.77e-4), 3: (3.84, 6.87e-5), 4: (4.87, 5.86e-6)}


def _solve(name, p):
    m, op, ph = problems.smb_vortex_problem(os.path.join(MESHES, name), p, bc_tags=[[0, 1, 2, 3]])
    qe = problems.vortex_exact(m.coords, ph)
    eqn = euler.EulerData(m, op, ph)
    eqn.set_q(qe)
    rep = linsolve.newton(m, op, eqn, linsolve.NewtonOptions(res_abstol=1e-12, krylov_reltol=1e-12,
                                                             fp64_blocks=True, pc_kind=abi.PDES_PC_ILU0,
                                                             gmres_restart=60, krylov_itermax=5000, itermax=10))
    q = eqn.get_q()
    eqn.free()
    assert rep.converged and rep.res_norm <= 1e-12, (name, p, rep.res_norm)
    M = (op.w[None, :] / m.jac)[:, :, None]
    return m, op, ph, q, qe, float(np.sqrt(((q - qe) ** 2 * M).sum())), pumi.mesh_h(m, op)


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_vortex_convergence_on_reference_meshes(p):
    print(f"case squarevortex small/large p={p}", flush=True)
    errs, hs = [], []
    for name in ("squarevortex_small0.smb", "squarevortex_large0.smb"):
        *_, err, h = _solve(name, p)
        errs.append(err)
        hs.append(h)
    slope = float(np.polyfit(np.log(hs), np.log(errs), 1)[0])
    s_ref, e_ref = PUBLISHED[p]
    print(f"p={p} errors {errs[0]:.4e} {errs[1]:.4e} slope {slope:.3f} (reference {e_ref:.3g}, {s_ref})",
          flush=True)
    if p <= 3:
        assert abs(slope - s_ref) < 0.1, slope
        assert e_ref / 1.25 < errs[0] < e_ref * 1.25, errs[0]
    else:
        assert slope > 4.5, slope


def test_converged_state_matches_oracle(oracle):
    """The device's converged p=2 state on squarevortex_small equals the
    oracle's Newton (sparse direct solve) state."""
    m, op, ph, q, qe, err, h = _solve("squarevortex_small0.smb", 2)
    b = abi.make_descriptors(m, op, ph)
    minv = np.repeat((m.jac / op.w[None, :])[:, :, None], 4, axis=2)
    q_ref, norms = oracle.newton_direct(b, qe, minv, res_abstol=1e-12, itermax=10)
    assert np.abs(q - q_ref).max() <= 1e-10 * np.abs(q_ref).max()
