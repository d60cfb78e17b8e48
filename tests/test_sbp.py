"""SBP operator identities (the properties SummationByParts.jl operators
satisfy; the reference's own tests rely on them, e.g. test_flux.jl:191-200
split-form zero element sum, test_dg.jl:112-131 uniform flow)."""

import numpy as np
import pytest

from pdes_amd import sbp
from pdes_amd.sbp import eval_monomials, eval_monomial_derivs, monomial_exponents, face_points

OPS = [(2, p, "omega") for p in (1, 2, 3, 4)] + [(2, p, "diage") for p in (1, 2, 3, 4)] + \
      [(3, p, "omega") for p in (1, 2)]
NP_EXPECTED = {(2, 1, "omega"): 3, (2, 2, "omega"): 6, (2, 3, "omega"): 10, (2, 4, "omega"): 15,
               (3, 1, "omega"): 4, (3, 2, "omega"): 10}


@pytest.mark.parametrize("dim,p,kind", OPS)
def test_sbp_identities(dim, p, kind):
    op = sbp.get_operator(dim, p, kind)
    Np = op.numnodes
    if (dim, p, kind) in NP_EXPECTED:
        assert Np == NP_EXPECTED[(dim, p, kind)]
    assert np.all(op.w > 0)
    X = op.xi.T
    exps = monomial_exponents(dim, p)
    V = eval_monomials(X, exps)
    for d in range(dim):
        Q = op.Q[:, :, d]
        E = op.E[:, :, d]
        # Q + Q^T = E
        assert np.abs(Q + Q.T - E).max() < 1e-13
        # Q 1 = 0
        assert np.abs(Q.sum(axis=1)).max() < 1e-13
        # accuracy: Q V = H V_d for degree <= p
        Vd = eval_monomial_derivs(X, exps, d)
        assert np.abs(Q @ V - np.diag(op.w) @ Vd).max() < 1e-12
        # E is symmetric and 1^T E 1 = 0 (closed boundary)
        assert np.abs(E - E.T).max() < 1e-14
        assert abs(E.sum()) < 1e-13


@pytest.mark.parametrize("dim,p,kind", OPS)
def test_cubature_exactness(dim, p, kind):
    """H integrates degree 2p-1 exactly (the SBP requirement) and the face
    rule degree 2p."""
    op = sbp.get_operator(dim, p, kind)
    exps = monomial_exponents(dim, 2 * p - 1)
    exact = sbp._simplex_monomial_integral(exps, dim)
    got = eval_monomials(op.xi.T, exps).T @ op.w
    assert np.abs(got - exact).max() < 1e-13
    f = op.face
    exps_f = monomial_exponents(dim - 1, 2 * p)
    exact_f = sbp._simplex_monomial_integral(exps_f, dim - 1) if dim == 3 else \
        np.array([(1 - (-1) ** (e[0] + 1)) / (e[0] + 1) for e in exps_f])
    got_f = eval_monomials(f.xi_face.T, exps_f).T @ f.wface
    assert np.abs(got_f - exact_f).max() < 1e-13


@pytest.mark.parametrize("dim,p,kind", OPS)
def test_face_interpolation_exact(dim, p, kind):
    """R_f reproduces degree-p polynomials at the face nodes."""
    op = sbp.get_operator(dim, p, kind)
    f = op.face
    exps = monomial_exponents(dim, p)
    V = eval_monomials(op.xi.T, exps)
    for fc in range(dim + 1):
        pts = face_points(dim, fc, f.xi_face)
        Vf = eval_monomials(pts, exps)
        R = np.zeros((f.numnodes, op.numnodes))
        R[:, f.perm[:, fc]] = f.interp.T
        assert np.abs(R @ V - Vf).max() < 1e-12


@pytest.mark.parametrize("dim,p,kind", [o for o in OPS if o[0] == 3 or o[1] <= 2])
def test_nbrperm_is_permutation(dim, p, kind):
    op = sbp.get_operator(dim, p, kind)
    nb = op.face.nbrperm
    for o in range(nb.shape[1]):
        assert sorted(nb[:, o]) == list(range(op.face.numnodes))
