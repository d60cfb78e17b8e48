"""Synthetic DG meshes with the data the reference reads from its mesh object.

The reference builds its mesh through PumiInterface.jl / PUMI
(`src/solver/common.jl:414-482`); the field contract it relies on is
documented at `docs/src/interfaces.md:191-560` (`coords`, `dxidx`, `jac`,
`interfaces`, `bndryfaces`, `bndry_offsets`, `nrm_face`, `nrm_bndry`,
`coords_bndry`, `dofs`).  PUMI is not available, so two deterministic
generators are provided:

* `quarter_annulus(n, op)`: r in [1,3], theta in [0, pi/2], n x n quads each
  split along the same diagonal into 2 counter-clockwise triangles (2n^2
  elements, 3n^2-2n interfaces, 4n boundary faces).
* `unit_cube(n, op)`: n^3 hexes each split into 6 tets along the main
  diagonal (6n^3 elements, 12n^3-6n^2 interfaces, 12n^2 boundary faces).

Array layout: every array is the reference's Julia array with its axes
reversed, i.e. the same bytes in memory (Julia is column-major):
  q/res       (numEl, Np, ndof)       <-> q[ndof, Np, numEl]
  coords      (numEl, Np, dim)        <-> coords[dim, Np, numEl]
  dxidx       (numEl, Np, dim, dim)   <-> dxidx[xi, x, Np, numEl]
  jac         (numEl, Np)             <-> jac[Np, numEl]
  nrm_face    (numIface, Nf, dim)     <-> nrm_face[dim, Nf, numIface]
  nrm_bndry / coords_bndry (numBndry, Nf, dim)
Element, face and node indices are 0-based.

Metrics follow the reference definitions: dxidx = |J| dxi/dx with
|J| = det(dx/dxi); jac = det(dxi/dx) = 1/|J|; the scaled face normal is
sum_xi N_ref[xi, f] * dxidx[xi, :] (outward from elementL / the boundary
element), so sum_k wface_k F(q_k).nrm_k integrates over the physical face.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .sbp import SBPOperator, TRI_FACE_VERTS, TET_FACE_VERTS, TET_FACE_ORIENTS

INTERFACE_DTYPE = np.int32


@dataclass
class DGMesh:
    dim: int
    numEl: int
    numNodesPerElement: int
    numNodesPerFace: int
    numDofPerNode: int
    coords: np.ndarray
    dxidx: np.ndarray
    jac: np.ndarray
    vert_coords: np.ndarray          # (numEl, dim+1, dim)
    interfaces: np.ndarray           # (numIface, 5) elementL, elementR, faceL, faceR, orient(0-based)
    bndryfaces: np.ndarray           # (numBndry, 2) element, face
    bndry_offsets: np.ndarray        # (numBC+1,)
    nrm_face: np.ndarray
    nrm_bndry: np.ndarray
    coords_bndry: np.ndarray
    bc_names: list = field(default_factory=list)
    name: str = ""

    @property
    def numInterfaces(self):
        return self.interfaces.shape[0]

    @property
    def numBoundaryFaces(self):
        return self.bndryfaces.shape[0]

    @property
    def numBC(self):
        return len(self.bndry_offsets) - 1

    @property
    def numNodes(self):
        return self.numEl * self.numNodesPerElement

    @property
    def numDof(self):
        return self.numNodes * self.numDofPerNode

    def dofs(self):
        """dofs[e, j, k]: element-major dof numbering (0-based)."""
        return np.arange(self.numDof, dtype=np.int64).reshape(
            self.numEl, self.numNodesPerElement, self.numDofPerNode)


def _affine_metrics(vert: np.ndarray, dim: int):
    """vert: (numEl, dim+1, dim) -> dxdxi (numEl, dim, dim) [x, xi]"""
    cols = [(vert[:, k + 1, :] - vert[:, 0, :]) / 2.0 for k in range(dim)]
    return np.stack(cols, axis=2)


def _finish(dim, op: SBPOperator, vert, interfaces, bfaces, boffsets, bc_names, name):
    numEl = vert.shape[0]
    Np = op.numnodes
    face = op.face
    Nf = face.numnodes
    ndof = dim + 2
    # volume node coordinates: x = v0 + sum_k (v_{k+1}-v0)(xi_k+1)/2
    lam = (op.xi + 1.0) / 2.0                        # (dim, Np)
    coords = vert[:, None, 0, :] + np.einsum("ekd,kn->end",
                                             vert[:, 1:, :] - vert[:, None, 0, :], lam)
    dxdxi = _affine_metrics(vert, dim)              # (numEl, x, xi)
    detJ = np.linalg.det(dxdxi)
    assert np.all(detJ > 0), "negatively oriented element"
    dxidx_phys = np.linalg.inv(dxdxi)               # [xi, x]
    scaled = dxidx_phys * detJ[:, None, None]       # |J| dxi/dx  [xi, x]
    dxidx = np.repeat(np.transpose(scaled, (0, 2, 1))[:, None, :, :], Np, axis=1)  # [e, j, x, xi]
    jac = np.repeat((1.0 / detJ)[:, None], Np, axis=1)

    def scaled_normal(elems, faces):
        # sum_xi N_ref[xi, f] dxidx[xi, x]
        nref = face.normal[:, faces].T               # (n, dim_xi)
        return np.einsum("nk,nkx->nx", nref, scaled[elems])

    # face-node coordinates in the reference element, per face
    from .sbp import face_points
    ref_face_pts = [face_points(dim, f, face.xi_face) for f in range(dim + 1)]

    iface = np.asarray(interfaces, dtype=INTERFACE_DTYPE).reshape(-1, 5)
    nI = iface.shape[0]
    nrm_face = np.repeat(scaled_normal(iface[:, 0], iface[:, 2])[:, None, :], Nf, axis=1) \
        if nI else np.zeros((0, Nf, dim))
    bf = np.asarray(bfaces, dtype=INTERFACE_DTYPE).reshape(-1, 2)
    nrm_bndry = np.repeat(scaled_normal(bf[:, 0], bf[:, 1])[:, None, :], Nf, axis=1)
    coords_bndry = np.zeros((bf.shape[0], Nf, dim))
    for f in range(dim + 1):
        sel = bf[:, 1] == f
        if not np.any(sel):
            continue
        lamf = (ref_face_pts[f].T + 1.0) / 2.0       # (dim, Nf)
        v = vert[bf[sel, 0]]
        coords_bndry[sel] = v[:, None, 0, :] + np.einsum("ekd,kn->end", v[:, 1:, :] - v[:, None, 0, :], lamf)
    return DGMesh(dim=dim, numEl=numEl, numNodesPerElement=Np, numNodesPerFace=Nf,
                  numDofPerNode=ndof, coords=np.ascontiguousarray(coords),
                  dxidx=np.ascontiguousarray(dxidx), jac=np.ascontiguousarray(jac),
                  vert_coords=vert, interfaces=iface, bndryfaces=bf,
                  bndry_offsets=np.asarray(boffsets, dtype=np.int32),
                  nrm_face=np.ascontiguousarray(nrm_face), nrm_bndry=np.ascontiguousarray(nrm_bndry),
                  coords_bndry=coords_bndry, bc_names=list(bc_names), name=name)


def _build_faces(elem_verts, face_verts, dim, orient_fn):
    """Generic interface/boundary detection from element vertex lists."""
    numEl = elem_verts.shape[0]
    nfaces = len(face_verts)
    table = {}
    interfaces = []
    for e in range(numEl):
        for f in range(nfaces):
            key = tuple(sorted(int(elem_verts[e, v]) for v in face_verts[f]))
            if key in table:
                eL, fL = table.pop(key)
                gl = [int(elem_verts[eL, v]) for v in face_verts[fL]]
                gr = [int(elem_verts[e, v]) for v in face_verts[f]]
                interfaces.append((eL, e, fL, f, orient_fn(gl, gr)))
            else:
                table[key] = (e, f)
    interfaces.sort(key=lambda t: (t[0], t[2]))
    boundary = sorted(table.values())
    return interfaces, boundary


def _orient_2d(gl, gr):
    assert gl[0] == gr[1] and gl[1] == gr[0], "2D neighbours must traverse the edge oppositely"
    return 0


def _orient_3d(gl, gr):
    order = tuple(gl.index(v) for v in gr)
    return TET_FACE_ORIENTS.index(order)


def quarter_annulus(n: int, op: SBPOperator, r_in=1.0, r_out=3.0, name=None) -> DGMesh:
    assert op.dim == 2
    r = r_in + (r_out - r_in) * np.arange(n + 1) / n
    th = (math.pi / 2) * np.arange(n + 1) / n
    vid = lambda i, j: j * (n + 1) + i
    pts = np.zeros(((n + 1) * (n + 1), 2))
    for j in range(n + 1):
        for i in range(n + 1):
            pts[vid(i, j)] = (r[i] * math.cos(th[j]), r[i] * math.sin(th[j]))
    ev = np.zeros((2 * n * n, 3), dtype=np.int64)
    for j in range(n):
        for i in range(n):
            q = j * n + i
            A, B, C, D = vid(i, j), vid(i + 1, j), vid(i + 1, j + 1), vid(i, j + 1)
            ev[2 * q] = (A, B, C)
            ev[2 * q + 1] = (A, C, D)
    interfaces, boundary = _build_faces(ev, TRI_FACE_VERTS, 2, _orient_2d)
    # boundary groups: r = r_in, r = r_out, theta = 0, theta = pi/2
    def group(e, f):
        a, b = TRI_FACE_VERTS[f]
        va, vb = ev[e, a], ev[e, b]
        ia, ja = va % (n + 1), va // (n + 1)
        ib, jb = vb % (n + 1), vb // (n + 1)
        if ia == 0 and ib == 0:
            return 0
        if ia == n and ib == n:
            return 1
        if ja == 0 and jb == 0:
            return 2
        if ja == n and jb == n:
            return 3
        raise AssertionError("not a boundary edge")
    groups = [[] for _ in range(4)]
    for e, f in boundary:
        groups[group(e, f)].append((e, f))
    bfaces = [bf for g in groups for bf in g]
    offsets = np.cumsum([0] + [len(g) for g in groups])
    vert = pts[ev]
    return _finish(2, op, vert, interfaces, bfaces, offsets,
                   ["r_inner", "r_outer", "theta_0", "theta_90"],
                   name or f"quarter_annulus_{n}x{n}")


def unit_cube(n: int, op: SBPOperator, name=None) -> DGMesh:
    assert op.dim == 3
    vid = lambda i, j, k: (k * (n + 1) + j) * (n + 1) + i
    pts = np.zeros(((n + 1) ** 3, 3))
    for k in range(n + 1):
        for j in range(n + 1):
            for i in range(n + 1):
                pts[vid(i, j, k)] = (i / n, j / n, k / n)
    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    ev = np.zeros((6 * n ** 3, 4), dtype=np.int64)
    e = 0
    for k in range(n):
        for j in range(n):
            for i in range(n):
                base = np.array([i, j, k])
                for pm in perms:
                    v = [base.copy()]
                    cur = base.copy()
                    for ax in pm:
                        cur = cur.copy()
                        cur[ax] += 1
                        v.append(cur)
                    ids = [vid(*vv) for vv in v]
                    X = pts[ids]
                    if np.linalg.det((X[1:] - X[0]).T) < 0:
                        ids[1], ids[2] = ids[2], ids[1]
                    ev[e] = ids
                    e += 1
    interfaces, boundary = _build_faces(ev, TET_FACE_VERTS, 3, _orient_3d)
    def group(e, f):
        X = pts[[ev[e, v] for v in TET_FACE_VERTS[f]]]
        for ax in range(3):
            if np.all(X[:, ax] == 0.0):
                return 2 * ax
            if np.all(X[:, ax] == 1.0):
                return 2 * ax + 1
        raise AssertionError("not a boundary face")
    groups = [[] for _ in range(6)]
    for e_, f in boundary:
        groups[group(e_, f)].append((e_, f))
    bfaces = [bf for g in groups for bf in g]
    offsets = np.cumsum([0] + [len(g) for g in groups])
    vert = pts[ev]
    return _finish(3, op, vert, interfaces, bfaces, offsets,
                   ["x0", "x1", "y0", "y1", "z0", "z1"], name or f"unit_cube_{n}")
