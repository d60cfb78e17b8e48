"""Multidimensional SBP operators on triangles and tetrahedra.

The reference obtains these from the external SummationByParts.jl package
(`src/solver/common.jl:276-390`: `getTriSBPOmega`, `getTriSBPDiagE`,
`getTetSBPOmega`, `TriFace`, `getTriFaceForDiagE`, `TetFace`).  That package is
not available here, so the operators are constructed from their defining
properties:

* a cubature rule (nodes, positive weights w) exact for degree 2p on the
  reference element, H = diag(w);
* boundary operators E_d = sum_f R_f^T B_f N_{f,d} R_f built from a face
  cubature exact for degree 2p (B_f = diag(wface)) and an interpolation R_f
  that is exact for degree p;
* Q_d = S_d + E_d/2 with S_d skew-symmetric and Q_d V = H V_d for the degree-p
  polynomial basis V (accuracy).

Omega operators (no nodes on the boundary) have exactly dim(P_p) nodes, so Q_d
follows uniquely from Q_d = H V_d V^{-1}; diagonal-E operators put the face
cubature nodes (Lobatto, vertices included) among the volume nodes, R_f is a
node selection (a "SparseFace") and S_d is the minimum-norm skew solution of the
accuracy equations, the same construction SummationByParts.jl uses.

Reference elements follow SummationByParts.jl: triangle (-1,-1),(1,-1),(-1,1);
tetrahedron (-1,-1,-1),(1,-1,-1),(-1,1,-1),(-1,-1,1).  Exact node positions of
SummationByParts.jl cannot be reproduced without the package; the operators
here satisfy the same identities (tested) and have the same node counts for
the Omega family.
"""
from __future__ import annotations

import itertools
import math
from dataclasses import dataclass, field

import numpy as np

TRI_VTX = np.array([[-1.0, -1.0], [1.0, -1.0], [-1.0, 1.0]])
TET_VTX = np.array([[-1.0, -1.0, -1.0], [1.0, -1.0, -1.0],
                    [-1.0, 1.0, -1.0], [-1.0, -1.0, 1.0]])
# local vertex lists of each face, face node order runs along these vertices
TRI_FACE_VERTS = [(0, 1), (1, 2), (2, 0)]
TET_FACE_VERTS = [(0, 1, 2), (0, 1, 3), (1, 2, 3), (0, 2, 3)]


# ---------------------------------------------------------------------------
# polynomial helpers
# ---------------------------------------------------------------------------
def monomial_exponents(dim: int, degree: int):
    out = []
    for tot in range(degree + 1):
        if dim == 1:
            out.append((tot,))
        elif dim == 2:
            for a in range(tot, -1, -1):
                out.append((a, tot - a))
        else:
            for a in range(tot, -1, -1):
                for b in range(tot - a, -1, -1):
                    out.append((a, b, tot - a - b))
    return out


def eval_monomials(x: np.ndarray, exps):
    """x: (n, dim) -> V (n, len(exps))"""
    V = np.ones((x.shape[0], len(exps)))
    for k, e in enumerate(exps):
        for d, ed in enumerate(e):
            if ed:
                V[:, k] *= x[:, d] ** ed
    return V


def eval_monomial_derivs(x: np.ndarray, exps, d: int):
    V = np.zeros((x.shape[0], len(exps)))
    for k, e in enumerate(exps):
        if e[d] == 0:
            continue
        col = np.full(x.shape[0], float(e[d]))
        for dd, ed in enumerate(e):
            p = ed - 1 if dd == d else ed
            if p:
                col = col * x[:, dd] ** p
        V[:, k] = col
    return V


def _simplex_monomial_integral(exps, dim):
    """Integral of prod x_d^{e_d} over the reference simplex with vertices
    -1 and unit axis points 1 (x = -1 + 2u on the unit simplex)."""
    out = np.zeros(len(exps))
    for k, e in enumerate(exps):
        # expand prod (-1 + 2u_d)^{e_d}
        total = 0.0
        ranges = [range(ed + 1) for ed in e]
        for js in itertools.product(*ranges):
            coef = 1.0
            for ed, j in zip(e, js):
                coef *= math.comb(ed, j) * (2.0 ** j) * ((-1.0) ** (ed - j))
            num = 1.0
            for j in js:
                num *= math.factorial(j)
            total += coef * num / math.factorial(sum(js) + dim)
        out[k] = total * 2.0 ** dim
    return out


def gauss_legendre(n: int):
    x, w = np.polynomial.legendre.leggauss(n)
    return x, w


def gauss_lobatto(n: int):
    """Legendre-Gauss-Lobatto nodes/weights on [-1,1], n >= 2."""
    if n == 2:
        return np.array([-1.0, 1.0]), np.array([1.0, 1.0])
    P = np.polynomial.legendre.Legendre.basis(n - 1)
    inner = np.sort(P.deriv().roots().real)
    x = np.concatenate([[-1.0], inner, [1.0]])
    w = 2.0 / (n * (n - 1) * P(x) ** 2)
    return x, w


# ---------------------------------------------------------------------------
# symmetric cubature construction
# ---------------------------------------------------------------------------
def _tri_orbit(kind, params):
    """barycentric points of an S3 orbit on the triangle"""
    if kind == "S3":
        return [(1 / 3, 1 / 3, 1 / 3)]
    if kind == "S21":
        a, = params
        b = 1 - 2 * a
        return [(b, a, a), (a, b, a), (a, a, b)]
    if kind == "S111":
        a, b = params
        c = 1 - a - b
        return [(a, b, c), (b, a, c), (c, a, b), (a, c, b), (b, c, a), (c, b, a)]
    raise ValueError(kind)


def _tet_orbit(kind, params):
    if kind == "S31":
        a, = params
        b = 1 - 3 * a
        return [(b, a, a, a), (a, b, a, a), (a, a, b, a), (a, a, a, b)]
    if kind == "S22":
        a, = params
        b = 0.5 - a
        return [(a, a, b, b), (a, b, a, b), (a, b, b, a),
                (b, a, a, b), (b, a, b, a), (b, b, a, a)]
    raise ValueError(kind)


def _bary_to_xy(bary, vtx):
    return np.asarray(bary) @ vtx


# Symmetric cubature rules, solved once from the moment equations with
# S3 / S4 orbit structure (positive weights, interior nodes, P_p-unisolvent for
# Omega).  Orbit parameters are barycentric; weights are for the reference
# elements (triangle area 2, tetrahedron volume 4/3).
#   kind, params, weight
_TRI_RULES = {
    # degree 2, 3 nodes
    "omega1": [("S21", [1.0 / 6.0], 2.0 / 3.0)],
    # degree 4, 6 nodes
    "omega2": [("S21", [0.44594849091596483], 0.4467631793560227),
               ("S21", [0.09157621350977081], 0.21990348731064396)],
    # degree 5, 10 nodes
    "omega3": [("S3", [], 0.39798292209670516),
               ("S21", [0.06971279004579463], 0.11660931639753905),
               ("S111", [0.3118989426278028, 0.07108551687495031], 0.2086981881184463)],
    # degree 7, 15 nodes
    "omega4": [("S21", [0.4129526725826635], 0.21530040871177467),
               ("S21", [0.18203687581708256], 0.14812676012851939),
               ("S21", [0.05960782497411116], 0.09025472230929828),
               ("S111", [0.6570723648743855, 0.030037466090090715], 0.10649238775853717)],
}
# diagonal-E rules: weights of the vertex orbit, of the edge-node orbits
# (Lobatto points grouped by |t|, increasing) and the interior orbits.
_DIAGE_RULES = {
    1: ([0.10000000000000006, 0.2666666666666665],
        [("S3", [], 0.9000000000000006)]),
    2: ([0.025044506019598862, 0.10729520100155707],
        [("S21", [0.21285435711180833], 0.4270317586439537)]),
    3: ([0.009130264572197825, 0.06201621057209457, 0.04537037037036954],
        [("S21", [0.14200508409677567], 0.20607267198227722),
         ("S21", [0.42438602517187946], 0.298706778799358)]),
    4: ([0.004361575619977844, 0.027867771486113656, 0.022078990549402808],
        [("S21", [0.26533138048460764], 0.15984019211004485),
         ("S21", [0.10367750814276627], 0.11358876265562022),
         ("S111", [0.08827396060212293, 0.3246404722647804], 0.14449130610499544)]),
}
_TET_RULES = {
    # degree 2, 4 nodes
    "omega1": [("S31", [(5.0 - math.sqrt(5.0)) / 20.0], 1.0 / 3.0)],
    # degree 3, 10 nodes
    "omega2": [("S31", [0.13272552099387736], 0.21527667422051405),
               ("S22", [0.07494631606636805], 0.07870443940854618)],
}


# ---------------------------------------------------------------------------
# operator containers
# ---------------------------------------------------------------------------
@dataclass
class SBPFace:
    """Face operator, mirrors the fields of SummationByParts' DenseFace /
    SparseFace used by the reference: `numnodes`, `wface`, `interp`
    (stencilsize x numnodes), `perm` (stencilsize x numfaces, volume node index
    of each stencil entry), `nbrperm` (numnodes x norient), `normal`
    (dim x numfaces scaled reference normals), plus the face node parametric
    coordinates `xi_face` ((dim-1) x numnodes)."""
    dim: int
    numnodes: int
    wface: np.ndarray
    interp: np.ndarray
    perm: np.ndarray
    nbrperm: np.ndarray
    normal: np.ndarray
    xi_face: np.ndarray
    dense: bool

    @property
    def stencilsize(self):
        return self.interp.shape[0]


@dataclass
class SBPOperator:
    """Volume operator: `numnodes`, `w` (H diagonal), `Q` (Np x Np x dim),
    `E` (Np x Np x dim), reference node coordinates `xi` (dim x Np),
    `vtx` reference vertices, and the face operator."""
    dim: int
    degree: int
    kind: str                  # "omega" | "diage"
    numnodes: int
    w: np.ndarray
    Q: np.ndarray
    E: np.ndarray
    xi: np.ndarray
    vtx: np.ndarray
    face: SBPFace
    D: np.ndarray = field(default=None)

    @property
    def S(self):
        return 0.5 * (self.Q - np.transpose(self.Q, (1, 0, 2)))

    @property
    def numfaces(self):
        return self.dim + 1


def _face_geometry(dim):
    if dim == 2:
        vtx, fv = TRI_VTX, TRI_FACE_VERTS
        normals = []
        for a, b in fv:
            t = vtx[b] - vtx[a]
            # outward normal for a counter-clockwise element, scaled by |f|/2
            normals.append(np.array([t[1], -t[0]]) / 2.0)
        return vtx, fv, np.array(normals).T
    vtx, fv = TET_VTX, TET_FACE_VERTS
    centroid = vtx.mean(axis=0)
    normals = []
    for f in fv:
        a, b, c = (vtx[i] for i in f)
        n = np.cross(b - a, c - a) / 2.0          # area-weighted normal (area of face)
        if np.dot(n, a - centroid) < 0:
            n = -n
        normals.append(n / 2.0)                   # reference face triangle has area 2
    return vtx, fv, np.array(normals).T


def face_points(dim, f, xi_face):
    """Reference-element coordinates of the face nodes of face f.
    xi_face: (dim-1, Nf) parametric coordinates on the reference face
    ([-1,1] in 2D, the reference triangle in 3D)."""
    if dim == 2:
        a, b = TRI_FACE_VERTS[f]
        s = xi_face[0]
        return np.outer((1 - s) / 2, TRI_VTX[a]) + np.outer((1 + s) / 2, TRI_VTX[b])
    a, b, c = TET_FACE_VERTS[f]
    # barycentric coordinates of the face reference triangle
    r, s = xi_face
    l2 = (r + 1) / 2
    l3 = (s + 1) / 2
    l1 = 1 - l2 - l3
    return (np.outer(l1, TET_VTX[a]) + np.outer(l2, TET_VTX[b]) + np.outer(l3, TET_VTX[c]))


def _tri_face_perm_2d(Nf):
    """For a 2D interface the neighbour traverses the edge in the opposite
    direction: node j of elementL matches node Nf-1-j of elementR."""
    return np.arange(Nf)[::-1].reshape(Nf, 1).copy()


TET_FACE_ORIENTS = list(itertools.permutations(range(3)))


def _tet_face_nbrperm(xi_face):
    """nbrperm for tet faces.  elementR lists the vertices of the shared face
    in its own order; orient o (1..6) selects the permutation
    TET_FACE_ORIENTS[o-1] that gives elementR's face vertices as positions in
    elementL's face vertex list.  Column o-1 maps elementL face node j to the
    elementR face node at the same physical point."""
    r, s = xi_face
    l = np.stack([1 - (r + 1) / 2 - (s + 1) / 2, (r + 1) / 2, (s + 1) / 2], axis=1)
    Nf = l.shape[0]
    nbr = np.zeros((Nf, len(TET_FACE_ORIENTS)), dtype=np.int64)
    for o, order in enumerate(TET_FACE_ORIENTS):
        for j in range(Nf):
            # barycentrics of L's node j with respect to R's vertex order
            target = l[j, list(order)]
            d = np.abs(l - target).sum(axis=1)
            k = int(np.argmin(d))
            assert d[k] < 1e-12
            nbr[j, o] = k
    return nbr


def _make_omega_face(dim, p, vol_xi, V_inv, exps):
    vtx, fv, normals = _face_geometry(dim)
    if dim == 2:
        s, wf = gauss_legendre(p + 1)
        xi_face = s.reshape(1, -1)
        nbrperm = _tri_face_perm_2d(p + 1)
    else:
        xi_face, wf = tri_face_cubature(2 * p)
        nbrperm = _tet_face_nbrperm(xi_face)
    Np = vol_xi.shape[1]
    # interpolation from all volume nodes to face 0's nodes; by symmetry of the
    # construction we build one interp per face and check they agree after perm
    pts0 = face_points(dim, 0, xi_face)
    interp0 = (eval_monomials(pts0, exps) @ V_inv).T     # (Np, Nf)
    perm = np.zeros((Np, dim + 1), dtype=np.int64)
    perm[:, 0] = np.arange(Np)
    # for other faces find the node permutation under the symmetry map that
    # sends face 0 to face f (affine map of the element onto itself)
    for f in range(1, dim + 1):
        pts = face_points(dim, f, xi_face)
        interp_f = (eval_monomials(pts, exps) @ V_inv).T
        # find perm so that interp_f[perm, :] == interp0  (node relabelling)
        A = _find_symmetry_perm(dim, 0, f, vol_xi)
        perm[:, f] = A
        assert np.allclose(interp_f[A, :], interp0, atol=1e-12), "face symmetry perm failed"
    return SBPFace(dim=dim, numnodes=len(wf), wface=np.asarray(wf), interp=interp0,
                   perm=perm, nbrperm=nbrperm, normal=normals, xi_face=xi_face, dense=True)


def _affine_map_face(dim, f0, f):
    """Affine map of the reference element onto itself that maps the vertex
    list of face f0 onto that of face f (in order), returned as (A, b) with
    x' = A x + b."""
    if dim == 2:
        vtx, fv = TRI_VTX, TRI_FACE_VERTS
        src = list(fv[f0]) + [3 - sum(fv[f0])]
        dst = list(fv[f]) + [3 - sum(fv[f])]
    else:
        vtx, fv = TET_VTX, TET_FACE_VERTS
        src = list(fv[f0]) + [6 - sum(fv[f0])]
        dst = list(fv[f]) + [6 - sum(fv[f])]
    X = np.hstack([vtx[src], np.ones((dim + 1, 1))])
    Y = vtx[dst]
    M = np.linalg.solve(X, Y)          # (dim+1, dim)
    A = M[:dim, :].T
    b = M[dim, :]
    return A, b


def _find_symmetry_perm(dim, f0, f, vol_xi):
    A, b = _affine_map_face(dim, f0, f)
    X = vol_xi.T
    Y = X @ A.T + b
    Np = X.shape[0]
    perm = np.zeros(Np, dtype=np.int64)
    for i in range(Np):
        d = np.abs(X - Y[i]).sum(axis=1)
        k = int(np.argmin(d))
        assert d[k] < 1e-10, "node set not symmetric"
        perm[i] = k
    return perm


def tri_face_cubature(q):
    """Symmetric cubature on the reference triangle used as tet face rule."""
    if q <= 2:
        orb = _TRI_RULES["omega1"]
    elif q <= 4:
        orb = _TRI_RULES["omega2"]
    else:
        raise NotImplementedError("tet face cubature degree > 4")
    pts = []
    ws = []
    for kind, par, w in orb:
        for bary in _tri_orbit(kind, par):
            pts.append(_bary_to_xy(bary, TRI_VTX))
            ws.append(w)
    pts = np.array(pts).T
    return pts, np.array(ws)


def _omega_tri_nodes(p):
    """Interior (Omega) cubature on the reference triangle with dim(P_p)
    nodes, exact for degree 2p-1 at least (2p for p <= 2)."""
    return _TRI_RULES["omega%d" % p]


def _omega_tet_nodes(p):
    return _TET_RULES["omega%d" % p]


def _orbits_to_nodes(orb, orbit_fn, vtx):
    pts, ws = [], []
    for kind, par, w in orb:
        for bary in orbit_fn(kind, par):
            pts.append(_bary_to_xy(bary, vtx))
            ws.append(w)
    return np.array(pts), np.array(ws)


def _build_Q(dim, xi, w, E, exps, exact_square):
    """Q_d = S_d + E_d/2 with S_d skew and Q_d V = H V_d; returns Q, D."""
    Np = xi.shape[1]
    X = xi.T
    V = eval_monomials(X, exps)
    H = np.diag(w)
    Q = np.zeros((Np, Np, dim))
    D = np.zeros((Np, Np, dim))
    if exact_square:
        Vinv = np.linalg.inv(V)
        for d in range(dim):
            Vd = eval_monomial_derivs(X, exps, d)
            HD = H @ Vd @ Vinv
            S = 0.5 * (HD - HD.T)
            Q[:, :, d] = S + 0.5 * E[:, :, d]
            D[:, :, d] = np.diag(1.0 / w) @ Q[:, :, d]
        return Q, D
    # minimum-norm skew solution of S V = H V_d - E V / 2
    il = np.tril_indices(Np, -1)
    nunk = len(il[0])
    for d in range(dim):
        Vd = eval_monomial_derivs(X, exps, d)
        W = H @ Vd - 0.5 * E[:, :, d] @ V
        nrow = Np * V.shape[1]
        A = np.zeros((nrow, nunk))
        rhs = np.zeros(nrow)
        for k in range(V.shape[1]):
            for i in range(Np):
                r = k * Np + i
                rhs[r] = W[i, k]
                # (S V)_{ik} = sum_j S_ij V_jk ; S_ij = s (i>j), -s_ji (i<j)
                for u, (ii, jj) in enumerate(zip(*il)):
                    if ii == i:
                        A[r, u] += V[jj, k]
                    elif jj == i:
                        A[r, u] -= V[ii, k]
        s, *_ = np.linalg.lstsq(A, rhs, rcond=None)
        res = A @ s - rhs
        assert np.max(np.abs(res)) < 1e-11, f"SBP accuracy equations inconsistent {np.max(np.abs(res))}"
        S = np.zeros((Np, Np))
        S[il] = s
        S = S - S.T
        Q[:, :, d] = S + 0.5 * E[:, :, d]
        D[:, :, d] = np.diag(1.0 / w) @ Q[:, :, d]
    return Q, D


def _assemble_E(dim, face: SBPFace, Np):
    E = np.zeros((Np, Np, dim))
    for f in range(dim + 1):
        R = np.zeros((face.numnodes, Np))
        R[:, face.perm[:, f]] = face.interp.T
        for d in range(dim):
            E[:, :, d] += R.T @ np.diag(face.wface * face.normal[d, f]) @ R
    return E


def getTriSBPOmega(p: int) -> SBPOperator:
    orb = _omega_tri_nodes(p)
    X, w = _orbits_to_nodes(orb, _tri_orbit, TRI_VTX)
    xi = X.T.copy()
    exps = monomial_exponents(2, p)
    V = eval_monomials(X, exps)
    Vinv = np.linalg.inv(V)
    face = _make_omega_face(2, p, xi, Vinv, exps)
    E = _assemble_E(2, face, xi.shape[1])
    Q, D = _build_Q(2, xi, w, E, exps, exact_square=True)
    return SBPOperator(dim=2, degree=p, kind="omega", numnodes=xi.shape[1], w=w, Q=Q,
                       E=E, xi=xi, vtx=TRI_VTX.copy(), face=face, D=D)


def getTetSBPOmega(p: int) -> SBPOperator:
    orb = _omega_tet_nodes(p)
    X, w = _orbits_to_nodes(orb, _tet_orbit, TET_VTX)
    xi = X.T.copy()
    exps = monomial_exponents(3, p)
    V = eval_monomials(X, exps)
    Vinv = np.linalg.inv(V)
    face = _make_omega_face(3, p, xi, Vinv, exps)
    E = _assemble_E(3, face, xi.shape[1])
    Q, D = _build_Q(3, xi, w, E, exps, exact_square=True)
    return SBPOperator(dim=3, degree=p, kind="omega", numnodes=xi.shape[1], w=w, Q=Q,
                       E=E, xi=xi, vtx=TET_VTX.copy(), face=face, D=D)


def getTriSBPDiagE(p: int) -> SBPOperator:
    """Diagonal-E triangle operator with vertices: p+2 Lobatto nodes per edge
    (vertices shared), interior orbits sized so the cubature is exact for
    degree 2p (p=1: degree 3)."""
    s, wl = gauss_lobatto(p + 2)
    inner = s[1:-1]
    # boundary node groups (fixed positions, one weight per S3 orbit)
    verts = [TRI_VTX[i] for i in range(3)]
    edge_groups = []         # list of lists of points
    half = len(inner) // 2
    edge_nodes_by_face = []  # per face, the list of points in face order
    for f, (a, b) in enumerate(TRI_FACE_VERTS):
        pts = [((1 - t) / 2) * TRI_VTX[a] + ((1 + t) / 2) * TRI_VTX[b] for t in inner]
        edge_nodes_by_face.append(pts)
    # orbit groups by |t|
    groups = {}
    for k, t in enumerate(inner):
        key = round(abs(t), 12)
        groups.setdefault(key, []).append(k)
    fixed_groups = [verts]
    group_keys = sorted(groups.keys())
    for key in group_keys:
        pts = []
        for f in range(3):
            for k in groups[key]:
                pts.append(edge_nodes_by_face[f][k])
        fixed_groups.append(pts)
    wf, orb = _DIAGE_RULES[p]
    # node list: vertices, then edge nodes face by face, then interior orbits
    X = [v for v in verts]
    W = [wf[0]] * 3
    key_w = {key: wf[1 + i] for i, key in enumerate(group_keys)}
    face_node_idx = []
    for f in range(3):
        idx = []
        for k, t in enumerate(inner):
            idx.append(len(X))
            X.append(edge_nodes_by_face[f][k])
            W.append(key_w[round(abs(t), 12)])
        face_node_idx.append(idx)
    for kind, par, w in orb:
        for bary in _tri_orbit(kind, par):
            X.append(_bary_to_xy(bary, TRI_VTX))
            W.append(w)
    X = np.array(X)
    w = np.array(W)
    xi = X.T.copy()
    Np = xi.shape[1]
    Nf = p + 2
    perm = np.zeros((Nf, 3), dtype=np.int64)
    for f, (a, b) in enumerate(TRI_FACE_VERTS):
        perm[:, f] = [a] + face_node_idx[f] + [b]
    _, _, normals = _face_geometry(2)
    face = SBPFace(dim=2, numnodes=Nf, wface=wl, interp=np.eye(Nf), perm=perm,
                   nbrperm=_tri_face_perm_2d(Nf), normal=normals,
                   xi_face=s.reshape(1, -1), dense=False)
    E = _assemble_E(2, face, Np)
    exps = monomial_exponents(2, p)
    Q, D = _build_Q(2, xi, w, E, exps, exact_square=False)
    return SBPOperator(dim=2, degree=p, kind="diage", numnodes=Np, w=w, Q=Q, E=E, xi=xi,
                       vtx=TRI_VTX.copy(), face=face, D=D)


_CACHE = {}


def get_operator(dim: int, p: int, kind: str = "omega") -> SBPOperator:
    key = (dim, p, kind)
    if key not in _CACHE:
        if dim == 2 and kind == "omega":
            op = getTriSBPOmega(p)
        elif dim == 2 and kind == "diage":
            op = getTriSBPDiagE(p)
        elif dim == 3 and kind == "omega":
            op = getTetSBPOmega(p)
        else:
            raise NotImplementedError(key)
        _CACHE[key] = op
    return _CACHE[key]
