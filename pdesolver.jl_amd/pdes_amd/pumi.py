"""Reader for serial PUMI `.smb` simplex meshes, the on-disk mesh format the
reference loads through PumiInterface.jl (`src/solver/common.jl:414-482`,
options `smb_name` / `dmg_name`; its test meshes are in `src/mesh_files`).

Layout of a serial (one-part) file, every word big-endian; unsigned words are
32 bits and reals are IEEE binary64:

  magic, version, dim, number of parts          4 unsigned
  entity counts per type                         8 unsigned, types in the order
                                                 vertex, edge, triangle, quad,
                                                 hex, prism, pyramid, tet
  downward adjacency of every non-vertex type    count x (2, 3, 4, 6, 5, 5, 4)
                                                 ids of the entities one
                                                 dimension down (edge -> 2
                                                 vertices, triangle -> 3
                                                 edges, tet -> 4 triangles)
  vertex coordinates                             3 reals per vertex
  vertex parametric coordinates                  2 reals per vertex
  remote-copy counts                             2 unsigned (zero for one part)
  geometric classification                       (model dimension, model tag)
                                                 per entity, types in order
  tags and periodic matches                      not read

The boundary groups of the reference's input (`BC1 => [0, 1, 2, 3]`) list the
tags of the model edges (2D) or model faces (3D) that boundary faces are
classified on; `load_smb` groups boundary faces the same way.  Only
triangle and tetrahedron meshes of one part are read; a partitioned file, a
mesh with other element types, a truncated file or an inconsistent
classification raises ValueError.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import mesh as meshmod
from .sbp import SBPOperator, TET_FACE_VERTS, TRI_FACE_VERTS

_NTYPES = 8
_DOWN = (0, 2, 3, 4, 6, 5, 5, 4)
_TYPE_DIM = (0, 1, 2, 2, 3, 3, 3, 3)
_EDGE, _TRI, _TET = 1, 2, 7


@dataclass
class SmbMesh:
    """The raw content of a serial `.smb` file."""
    version: int
    dim: int
    counts: np.ndarray                 # (8,) entities per type
    down: dict                         # type -> (count, ndown) downward ids
    coords: np.ndarray                 # (numVert, 3)
    classification: dict               # type -> (count, 2) (model dim, model tag)


class _Words:
    def __init__(self, data: bytes):
        self.data = data
        self.pos = 0

    def take(self, n: int, dtype: str) -> np.ndarray:
        size = np.dtype(dtype).itemsize * n
        if self.pos + size > len(self.data):
            raise ValueError(f"truncated .smb file: need {size} bytes at offset {self.pos}, have "
                             f"{len(self.data) - self.pos}")
        out = np.frombuffer(self.data, dtype=dtype, count=n, offset=self.pos)
        self.pos += size
        return out


def read_smb(path: str) -> SmbMesh:
    """Parse a serial `.smb` file (layout in the module docstring)."""
    with open(path, "rb") as fh:
        w = _Words(fh.read())
    magic, version, dim, nparts = (int(v) for v in w.take(4, ">u4"))
    if magic != 0 or dim not in (2, 3):
        raise ValueError(f"{path}: not a PUMI .smb mesh (magic {magic}, dim {dim})")
    if nparts != 1:
        raise ValueError(f"{path}: partitioned mesh ({nparts} parts); only one-part files are read")
    counts = w.take(_NTYPES, ">u4").astype(np.int64)
    down = {}
    for t in range(1, _NTYPES):
        if counts[t]:
            down[t] = w.take(int(counts[t]) * _DOWN[t], ">u4").astype(np.int64).reshape(-1, _DOWN[t])
            lower = {_EDGE: 0, _TRI: _EDGE, _TET: _TRI}.get(t)
            if lower is not None and down[t].size and down[t].max() >= counts[lower]:
                raise ValueError(f"{path}: downward id out of range for type {t}")
    nv = int(counts[0])
    coords = w.take(3 * nv, ">f8").reshape(nv, 3).copy()
    w.take(2 * nv, ">f8")                 # parametric coordinates
    w.take(2, ">u4")                      # remote-copy counts of a one-part file
    classification = {}
    for t in range(_NTYPES):
        if counts[t]:
            cls = w.take(2 * int(counts[t]), ">u4").astype(np.int64).reshape(-1, 2)
            if cls[:, 0].max() > dim or cls[:, 0].min() < _TYPE_DIM[t]:
                raise ValueError(f"{path}: classification of type {t} is not a model entity of "
                                 f"dimension {_TYPE_DIM[t]}..{dim}")
            classification[t] = cls
    return SmbMesh(version=version, dim=dim, counts=counts, down=down, coords=coords,
                   classification=classification)


def _tri_vertices(edges_of_tri: np.ndarray, edge_verts: np.ndarray) -> np.ndarray:
    """Vertices of each triangle from its three edges: vertex k is shared by
    edge k and edge k-1."""
    ev = edge_verts[edges_of_tri]                     # (n, 3 edges, 2)
    out = np.empty((edges_of_tri.shape[0], 3), dtype=np.int64)
    for k in range(3):
        a, b = ev[:, k, :], ev[:, k - 1, :]
        shared = np.where((a[:, 0] == b[:, 0]) | (a[:, 0] == b[:, 1]), a[:, 0], a[:, 1])
        out[:, k] = shared
    if np.any(out[:, 0] == out[:, 1]) or np.any(out[:, 1] == out[:, 2]) or np.any(out[:, 0] == out[:, 2]):
        raise ValueError("degenerate triangle in .smb connectivity")
    return out


def _tet_vertices(tris_of_tet: np.ndarray, tri_verts: np.ndarray) -> np.ndarray:
    """Vertices of each tetrahedron: those of its first face, then the vertex
    of its second face that is not on the first."""
    f0 = tri_verts[tris_of_tet[:, 0]]
    f1 = tri_verts[tris_of_tet[:, 1]]
    apex = np.full(tris_of_tet.shape[0], -1, dtype=np.int64)
    for k in range(3):
        off = (f1[:, k, None] != f0).all(axis=1)
        apex = np.where(off & (apex < 0), f1[:, k], apex)
    if np.any(apex < 0):
        raise ValueError("degenerate tetrahedron in .smb connectivity")
    return np.concatenate([f0, apex[:, None]], axis=1)


def element_vertices(sm: SmbMesh):
    """(vertex coordinates (numVert, dim), element vertex ids (numEl, dim+1)
    positively oriented, facet vertex ids (numFacet, dim), facet
    classification (numFacet, 2))."""
    dim = sm.dim
    other = [t for t in range(3, _NTYPES) if sm.counts[t] and t != _TET]
    if other or (dim == 2 and sm.counts[_TET]):
        raise ValueError("only triangle (2D) and tetrahedron (3D) meshes are read")
    pts = sm.coords[:, :dim]
    edge_verts = sm.down[_EDGE]
    tri_verts = _tri_vertices(sm.down[_TRI], edge_verts)
    if dim == 2:
        ev = tri_verts
        facets, fcls = edge_verts, sm.classification[_EDGE]
    else:
        ev = _tet_vertices(sm.down[_TET], tri_verts)
        facets, fcls = tri_verts, sm.classification[_TRI]
    X = pts[ev]
    det = np.linalg.det(np.transpose(X[:, 1:, :] - X[:, :1, :], (0, 2, 1)))
    if np.any(det == 0.0):
        raise ValueError("zero-volume element in .smb mesh")
    neg = det < 0
    ev = ev.copy()
    ev[neg, 1], ev[neg, 2] = ev[neg, 2], ev[neg, 1].copy()
    return pts, ev, facets, fcls


def load_smb(path: str, op: SBPOperator, bc_tags=None, name: str | None = None) -> meshmod.DGMesh:
    """DGMesh of a serial simplex `.smb` file for operator op.

    bc_tags: list of boundary groups, each a list of model-entity tags (the
    reference's `BCn => [tags]`); boundary faces are ordered by group, then by
    (element, face).  Default: one group per model tag, in tag order.  Every
    boundary face must be classified on a listed tag."""
    sm = read_smb(path)
    if sm.dim != op.dim:
        raise ValueError(f"{path}: {sm.dim}D mesh for a {op.dim}D operator")
    pts, ev, facets, fcls = element_vertices(sm)
    dim = sm.dim
    face_verts = TRI_FACE_VERTS if dim == 2 else TET_FACE_VERTS
    orient = meshmod._orient_2d if dim == 2 else meshmod._orient_3d
    interfaces, boundary = meshmod._build_faces(ev, face_verts, dim, orient)
    tag_of = {tuple(sorted(int(v) for v in f)): (int(c[0]), int(c[1])) for f, c in zip(facets, fcls)}
    btag = []
    for e, f in boundary:
        mdim, tag = tag_of[tuple(sorted(int(ev[e, v]) for v in face_verts[f]))]
        if mdim != dim - 1:
            raise ValueError(f"{path}: boundary face of element {e} is classified on a model entity of "
                             f"dimension {mdim}")
        btag.append(tag)
    if bc_tags is None:
        bc_tags = [[t] for t in sorted(set(btag))]
    group_of = {}
    for g, tags in enumerate(bc_tags):
        for t in tags:
            if t in group_of:
                raise ValueError(f"model tag {t} listed in two boundary groups")
            group_of[t] = g
    groups = [[] for _ in bc_tags]
    for (e, f), t in zip(boundary, btag):
        if t not in group_of:
            raise ValueError(f"{path}: boundary faces on model tag {t} are in no boundary group")
        groups[group_of[t]].append((e, f))
    bfaces = [bf for g in groups for bf in g]
    offsets = np.cumsum([0] + [len(g) for g in groups])
    names = ["BC" + "_".join(str(t) for t in tags) for tags in bc_tags]
    return meshmod._finish(dim, op, pts[ev], interfaces, bfaces, offsets, names,
                           name or path.rsplit("/", 1)[-1])


def min_node_dist(op: SBPOperator) -> float:
    """Smallest distance between two volume nodes of the reference element."""
    x = op.xi.T
    d = np.linalg.norm(x[:, None, :] - x[None, :, :], axis=2)
    return float(d[~np.eye(len(x), dtype=bool)].min())


def mesh_h(m: meshmod.DGMesh, op: SBPOperator) -> float:
    """calcMeshH (src/Utils/Utils.jl:538-552): the mean over volume nodes of
    jac^(-1/dim), times the reference element's minimum node distance."""
    return float(np.mean(1.0 / m.jac.ravel() ** (1.0 / m.dim)) * min_node_dist(op))
