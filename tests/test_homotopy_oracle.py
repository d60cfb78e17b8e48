"""The oracle's restatement of calcHomotopyDiss
(src/solver/euler/homotopy.jl:51-249), pinned by what the operator must
satisfy: zero on uniform flow; conservative (the sum over all nodes vanishes,
since Q_d 1 = 0 and interior faces cancel); dissipative (q . D(q) <= 0: the
volume part is -(Q_d q)^T Lambda H^-1 (Q_d q) and each face adds
-B lambda |qL - qR|^2 / 2); on a globally linear field the face term vanishes
and the volume term equals -sum_d Q_d^T (lambda_d (dq/dxi_d)) evaluated from
the exact gradient; on an element-wise constant field the volume term
vanishes and each element's sum equals minus the face jumps weighted by
wface lambda / 2."""
import numpy as np
import pytest

from pdes_amd import problems
from helpers import setup, state_for

CASES = [("vortex", 3, 1, "omega"), ("vortex", 3, 2, "omega"), ("vortex", 2, 3, "diage"), ("cube", 1, 2, "omega")]


def _lam(ph, q, d):
    """getLambdaMax (euler_funcs.jl:1887-1910), vectorised over nodes."""
    dim = d.shape[-1]
    rho = q[..., 0]
    u = q[..., 1:dim + 1] / rho[..., None]
    p = (ph.gamma - 1) * (q[..., dim + 1] - 0.5 * rho * (u * u).sum(-1))
    a = np.sqrt(ph.gamma * p / rho)
    un = (u * d).sum(-1)
    au = np.abs(un)
    au = np.where(au > 1e-7, au, (un * un / 1e-7 + 1e-7) / 2)
    return au + np.sqrt((d * d).sum(-1)) * a


def _make(kind, n, p, opk):
    m, op, ph, b = setup(kind, n, p, opk, "ir" if opk == "diage" else "roe")
    return m, op, ph, b


@pytest.mark.parametrize("kind,n,p,opk", CASES)
def test_uniform_flow_and_conservation(oracle, kind, n, p, opk):
    m, op, ph, b = _make(kind, n, p, opk)
    q0 = problems.freestream((m.numEl, op.numnodes), m.dim, ph)
    assert np.abs(oracle.homotopy(b, q0)).max() <= 1e-13
    q = state_for(kind, m, ph)
    q = q * (1 + 0.05 * np.random.default_rng(1).uniform(-1, 1, q.shape))
    d = oracle.homotopy(b, q)
    scale = np.abs(d).sum(axis=(0, 1))
    assert np.all(np.abs(d.sum(axis=(0, 1))) <= 1e-13 * scale)
    assert float((q * d).sum()) < 0.0


@pytest.mark.parametrize("kind,n,p,opk", CASES)
def test_linear_field_volume_term(oracle, kind, n, p, opk):
    m, op, ph, b = _make(kind, n, p, opk)
    nd = m.dim + 2
    q0 = problems.freestream((1,), m.dim, ph)[0]
    G = 0.02 * np.random.default_rng(3).standard_normal((nd, m.dim))
    q = q0[None, None, :] + np.einsum("cx,enx->enc", G, m.coords)
    d = oracle.homotopy(b, q)
    # dq/dxi_d = sum_x dq/dx_x dx_x/dxi_d; dx/dxi = (|J| dxi/dx)^-1 |J|
    ref = np.zeros_like(q)
    for e in range(m.numEl):
        A = m.dxidx[e, 0]                     # [x][xi] = |J| dxi_xi/dx_x
        dxdxi = np.linalg.inv(A.T) / m.jac[e, 0]   # [x][xi]
        for dd in range(m.dim):
            dq = G @ dxdxi[:, dd]             # (nd,)
            lam = _lam(ph, q[e], m.dxidx[e, :, :, dd])
            t2 = lam[:, None] * dq[None, :]
            ref[e] -= op.Q[:, :, dd].T @ t2
    # the face jumps cancel to rounding (qL - qR ~ 1e-16 |q|)
    assert np.abs(d - ref).max() <= 1e-11 * np.abs(ref).max()


@pytest.mark.parametrize("kind,n,p,opk", CASES)
def test_piecewise_constant_face_term(oracle, kind, n, p, opk):
    m, op, ph, b = _make(kind, n, p, opk)
    rng = np.random.default_rng(4)
    q0 = problems.freestream((1,), m.dim, ph)[0]
    qe = q0[None, :] * (1 + 0.05 * rng.uniform(-1, 1, (m.numEl, m.dim + 2)))
    q = np.repeat(qe[:, None, :], op.numnodes, axis=1)
    d = oracle.homotopy(b, q)
    ref = np.zeros((m.numEl, m.dim + 2))
    wf = op.face.wface
    for i, (eL, eR, fL, fR, o) in enumerate(m.interfaces):
        lam = _lam(ph, np.broadcast_to(0.5 * (qe[eL] + qe[eR]), (len(wf), m.dim + 2)), m.nrm_face[i])
        jump = np.outer(wf * lam, qe[eL] - qe[eR]).sum(0) / 2
        ref[eL] -= jump
        ref[eR] += jump
    assert np.abs(d.sum(axis=1) - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("kind,n,p,opk", CASES)
def test_boundary_term(oracle, kind, n, p, opk):
    """homotopy_addBoundaryIntegrals: zero on the free stream itself; on an
    element-wise constant field each element's sum gains minus the jump to
    the free stream over its boundary faces, weighted by wface lambda / 2."""
    m, op, ph, b = _make(kind, n, p, opk)
    qinf = problems.freestream((1,), m.dim, ph)[0]
    q0 = np.broadcast_to(qinf, (m.numEl, op.numnodes, m.dim + 2)).copy()
    assert np.abs(oracle.homotopy(b, q0, bndry=True)).max() <= 1e-13
    rng = np.random.default_rng(6)
    qe = qinf[None, :] * (1 + 0.05 * rng.uniform(-1, 1, (m.numEl, m.dim + 2)))
    q = np.repeat(qe[:, None, :], op.numnodes, axis=1)
    d0 = oracle.homotopy(b, q)
    d1 = oracle.homotopy(b, q, bndry=True)
    ref = np.zeros((m.numEl, m.dim + 2))
    wf = op.face.wface
    for bi, (e, f) in enumerate(m.bndryfaces):
        lam = _lam(ph, np.broadcast_to(0.5 * (qe[e] + qinf), (len(wf), m.dim + 2)), m.nrm_bndry[bi])
        ref[e] -= np.outer(wf * lam, qe[e] - qinf).sum(0) / 2
    assert np.abs((d1 - d0).sum(axis=1) - ref).max() <= 1e-12 * np.abs(ref).max()


def test_homotopy_jv_matches_finite_difference(oracle):
    m, op, ph, b = _make("vortex", 3, 2, "omega")
    q = state_for("vortex", m, ph)
    v = np.random.default_rng(5).standard_normal(q.shape)
    h = 1e-6
    for bndry in (False, True):
        jv = oracle.homotopy_jv(b, q, v, bndry=bndry)
        fd = (oracle.homotopy(b, q + h * v, bndry=bndry) - oracle.homotopy(b, q - h * v, bndry=bndry)) / (2 * h)
        assert np.abs(jv - fd).max() <= 1e-6 * np.abs(jv).max()
