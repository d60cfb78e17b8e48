"""Synthetic mesh generators: counts (numEl = 2n^2, interfaces 3n^2-2n,
boundary 4n for the split quad grid; 6n^3, 12n^3-6n^2, 12n^2 for the Kuhn
cube), geometric consistency of the face data the reference reads
(docs/src/interfaces.md:345-400) and determinism."""
import numpy as np
import pytest

from pdes_amd import mesh, sbp
from pdes_amd.sbp import face_points


def _face_xyz(m, op, e, f):
    pts = face_points(m.dim, f, op.face.xi_face)
    lam = (pts.T + 1) / 2
    v = m.vert_coords[e]
    return (v[0][:, None] + (v[1:] - v[0]).T @ lam).T


@pytest.mark.parametrize("n", [1, 3, 20])
def test_annulus_counts(n):
    op = sbp.get_operator(2, 1, "omega")
    m = mesh.quarter_annulus(n, op)
    assert m.numEl == 2 * n * n
    assert m.numInterfaces == 3 * n * n - 2 * n
    assert m.numBoundaryFaces == 4 * n
    assert list(m.bndry_offsets) == [0, n, 2 * n, 3 * n, 4 * n]


@pytest.mark.parametrize("n", [1, 2, 4])
def test_cube_counts(n):
    op = sbp.get_operator(3, 1, "omega")
    m = mesh.unit_cube(n, op)
    assert m.numEl == 6 * n ** 3
    assert m.numInterfaces == 12 * n ** 3 - 6 * n ** 2
    assert m.numBoundaryFaces == 12 * n ** 2
    # total volume
    assert abs((op.w[None, :] / m.jac).sum() - 1.0) < 1e-13


@pytest.mark.parametrize("dim,p,kind,n", [(2, 1, "omega", 4), (2, 2, "diage", 4), (2, 4, "diage", 3),
                                          (3, 1, "omega", 2), (3, 2, "omega", 2)])
def test_face_geometry(dim, p, kind, n):
    op = sbp.get_operator(dim, p, kind)
    m = mesh.quarter_annulus(n, op) if dim == 2 else mesh.unit_cube(n, op)
    # face nodes coincide through nbrperm; normals point out of elementL
    for i, (eL, eR, fL, fR, o) in enumerate(m.interfaces):
        xl = _face_xyz(m, op, eL, fL)
        xr = _face_xyz(m, op, eR, fR)
        assert np.abs(xl - xr[op.face.nbrperm[:, o]]).max() < 1e-13
        cL = m.vert_coords[eL].mean(axis=0)
        cR = m.vert_coords[eR].mean(axis=0)
        assert np.dot(m.nrm_face[i, 0], cR - cL) > 0
    # each element's faces close: sum_f sum_k wface_k nrm = 0
    tot = np.zeros((m.numEl, dim))
    for i, (eL, eR, fL, fR, o) in enumerate(m.interfaces):
        s = (op.face.wface[:, None] * m.nrm_face[i]).sum(axis=0)
        tot[eL] += s
        tot[eR] -= s
    for b, (e, f) in enumerate(m.bndryfaces):
        tot[e] += (op.face.wface[:, None] * m.nrm_bndry[b]).sum(axis=0)
    assert np.abs(tot).max() < 1e-13
    # boundary coordinates are the face nodes
    for b, (e, f) in enumerate(m.bndryfaces):
        assert np.abs(_face_xyz(m, op, e, f) - m.coords_bndry[b]).max() < 1e-13


def test_mesh_deterministic():
    op = sbp.get_operator(2, 2, "diage")
    a = mesh.quarter_annulus(7, op)
    b = mesh.quarter_annulus(7, op)
    assert np.array_equal(a.interfaces, b.interfaces)
    assert np.array_equal(a.bndryfaces, b.bndryfaces)
    assert np.array_equal(a.dxidx, b.dxidx)
