"""Boundary functionals on the device (calcBndryFunctional,
src/solver/euler/boundary_functional.jl:82-191; functionals.jl:19-160):
every kind of pdes_eval_functional against the oracle to 1e-10, and the
closed forms that fix their orientation on straight boundary sides."""
import ctypes as C

import numpy as np
import pytest

from pdes_amd import euler, problems
from pdes_amd._lib import check, lib

pytestmark = pytest.mark.gpu

G = 1.4
KINDS = [(1, "lift"), (2, "drag"), (3, "entropy"), (4, "mass"), (5, "lift_coef"), (6, "drag_coef")]


def _eval(eqn, kind, bcs):
    bc = np.asarray(bcs, dtype=np.int32)
    val = np.zeros(3)
    check(lib().pdes_eval_functional(eqn.ctx, kind, bc.ctypes.data, len(bc),
                                     val.ctypes.data_as(C.POINTER(C.c_double))))
    return val


@pytest.mark.parametrize("kind,n,p,opk,flux", [("vortex", 6, 2, "omega", "roe"), ("vortex", 5, 2, "diage", "ir"),
                                              ("vortex", 4, 3, "omega", "es"), ("cube", 2, 2, "omega", "roe")])
def test_every_functional_kind_matches_oracle(oracle, kind, n, p, opk, flux):
    """Kinds 0 (force vector), 1 lift, 2 drag, 3 entropy flux, 4 mass flow,
    5 lift coefficient, 6 drag coefficient, over all BC groups and over each
    group alone, within 1e-10 of the oracle (relative to 1 + |value|)."""
    print(f"case {kind} n={n} p={p} {opk} {flux}", flush=True)
    m, op, ph = problems.make_problem(kind, n, p, opk, flux, Ma=0.6, aoa_deg=3.0)
    q = problems.vortex_noise_state(m, ph) if kind == "vortex" else problems.freestream_bump_state(m, ph)
    eqn = euler.EulerData(m, op, ph)
    eqn.set_q(q)
    for bcs in [list(range(m.numBC))] + [[g] for g in range(m.numBC)]:
        ref = oracle.functional(m, op, ph, q, bcs)
        force = _eval(eqn, 0, bcs)[:m.dim]
        assert np.abs(force - ref["force"]).max() <= 1e-10 * (1 + np.abs(ref["force"]).max())
        for kid, key in KINDS:
            if key == "mass" and m.dim == 3:
                continue        # MassFlowData's integrand is two-dimensional (boundary_functional.jl:404)
            v = _eval(eqn, kid, bcs)[0]
            assert abs(v - ref[key]) <= 1e-10 * (1 + abs(ref[key])), (bcs, key, v, ref[key])
    eqn.free()


@pytest.mark.parametrize("kind,n,p", [("vortex", 4, 2), ("cube", 2, 1)])
def test_functionals_straight_side_closed_form_device(kind, n, p):
    """Uniform state q on straight sides (outward unit normal n_hat, length or
    area L): force = p(q_t) L n_hat with the normal momentum removed, lift and
    drag its rotations by the angle of attack, coefficients divided by
    0.5 rho_free Ma^2, mass flow (m . n_hat) L, entropy flux (m . n_hat) L U
    with U = -rho s/(gamma - 1)."""
    m, op, ph = problems.make_problem(kind, n, p, "omega", "roe", Ma=0.7, aoa_deg=17.0)
    dim = m.dim
    rho, pr = 1.3, 0.9
    u = np.array([0.3, -0.45, 0.2][:dim])
    qc = np.concatenate([[rho], rho * u, [pr / (G - 1) + 0.5 * rho * (u @ u)]])
    eqn = euler.EulerData(m, op, ph)
    eqn.set_q(np.broadcast_to(qc, (m.numEl, op.numnodes, dim + 2)).copy())
    a = np.radians(17.0)
    qdyn = 0.5 * 1.0 * 0.7 ** 2
    U = -rho * (np.log(pr) - G * np.log(rho)) / (G - 1)
    if kind == "vortex":
        sides = [(2, np.array([0.0, -1.0]), 2.0), (3, np.array([-1.0, 0.0]), 2.0)]
    else:
        sides = [(2 * ax + s, np.eye(3)[ax] * (1.0 if s else -1.0), 1.0) for ax in range(3) for s in range(2)]
    for bc, nh, L in sides:
        mn = rho * (u @ nh)
        mt = rho * u - nh * mn
        force = (G - 1) * (qc[-1] - 0.5 * (mt @ mt) / rho) * L * nh
        up = force[1] if dim == 2 else force[2]
        lift = -np.sin(a) * force[0] + np.cos(a) * up
        drag = np.cos(a) * force[0] + np.sin(a) * up
        assert np.abs(_eval(eqn, 0, [bc])[:dim] - force).max() < 1e-12
        assert abs(_eval(eqn, 1, [bc])[0] - lift) < 1e-12
        assert abs(_eval(eqn, 2, [bc])[0] - drag) < 1e-12
        assert abs(_eval(eqn, 5, [bc])[0] - lift / qdyn) < 1e-11
        assert abs(_eval(eqn, 6, [bc])[0] - drag / qdyn) < 1e-11
        assert abs(_eval(eqn, 3, [bc])[0] - mn * L * U) < 1e-12
        if dim == 2:
            assert abs(_eval(eqn, 4, [bc])[0] - mn * L) < 1e-12
    eqn.free()
