"""Synthetic boundary meshes, boundary conditions and the BEM DOF model.

This module produces the flat arrays that the reference solver reads from its
``Network`` objects (``WATERS``): node coordinates, outward-oriented linear
triangles, a Dirichlet/Neumann tag per face, and the per-(point, face) DOF
numbering of ``setNodeFaceIndices`` with multiple nodes and CORNER points.

Reference behaviour mirrored here:
  * point classes Neumann / Dirichlet / CORNER and ``setIsMultipleNode``
    (20 degree rule): bem/BEM_setBoundaryTypes.hpp:433-440, 560-566
  * ``isNeumannID_BEM`` / ``isDirichletID_BEM`` / ``pf2ID``:
    bem/BEM_setBoundaryTypes.hpp:29-54
  * one DOF per (point, face-key); Morton (Z-order, 21 bits per axis) DOF
    numbering with a stable tie-break: bem/BEM_solveBVP_GMRES_FMM.hpp:3152-3246
    (the reference breaks ties by pointer value; here by point then face index)

DOF kinds (``dof_kind``): 0 = Dirichlet (unknown phi_n, known phi),
1 = Neumann (unknown phi, known phi_n), 2 = CORNER Neumann face DOF (unknown
phi, constraint row phi = phi_D, known phi_n).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

DIRICHLET, NEUMANN, CORNER_NEUMANN, ROBIN = 0, 1, 2, 3


@dataclasses.dataclass
class Mesh:
    X: np.ndarray            # (n_nodes, 3) float64
    tri: np.ndarray          # (n_faces, 3) int32, outward (out of the fluid) orientation
    face_dirichlet: np.ndarray  # (n_faces,) bool
    face_group: np.ndarray   # (n_faces,) int32 patch id (informational)

    @property
    def n_nodes(self) -> int:
        return int(self.X.shape[0])

    @property
    def n_faces(self) -> int:
        return int(self.tri.shape[0])

    def face_normals(self) -> np.ndarray:
        a, b, c = (self.X[self.tri[:, k]] for k in range(3))
        cr = np.cross(b - a, c - a)
        return cr / np.linalg.norm(cr, axis=1, keepdims=True)

    def face_areas(self) -> np.ndarray:
        a, b, c = (self.X[self.tri[:, k]] for k in range(3))
        return 0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1)


@dataclasses.dataclass
class DofModel:
    n_dof: int
    face_dof: np.ndarray     # (n_faces, 3) int32
    dof_kind: np.ndarray     # (n_dof,) int32
    dof_point: np.ndarray    # (n_dof,) int32
    dof_face: np.ndarray     # (n_dof,) int32, -1 for the point's null-key DOF
    point_class: np.ndarray  # (n_nodes,) int8: 0 Dirichlet, 1 Neumann, 2 CORNER
    point_multiple: np.ndarray  # (n_nodes,) bool


# ---------------------------------------------------------------------------
# mesh assembly from structured patches, welded by coordinates
# ---------------------------------------------------------------------------
class _Builder:
    def __init__(self, weld_tol: float = 1e-9):
        self.tol = weld_tol
        self.key2id: dict = {}
        self.pts: list = []
        self.tris: list = []
        self.dir: list = []
        self.grp: list = []

    def _ids(self, P: np.ndarray) -> np.ndarray:
        flat = P.reshape(-1, 3)
        keys = np.round(flat / self.tol).astype(np.int64)
        out = np.empty(flat.shape[0], dtype=np.int64)
        for r in range(flat.shape[0]):
            k = (int(keys[r, 0]), int(keys[r, 1]), int(keys[r, 2]))
            i = self.key2id.get(k)
            if i is None:
                i = len(self.pts)
                self.key2id[k] = i
                self.pts.append(flat[r])
            out[r] = i
        return out.reshape(P.shape[:2])

    def grid(self, P: np.ndarray, outward, dirichlet: bool, group: int, mask=None):
        """P: (nu+1, nv+1, 3) grid of points; quads split with alternating diagonals.

        outward: callable((k,3) points) -> (k,3) approximate outward normals, used to orient faces.
        mask: optional (nu, nv) bool array, False = skip that quad.
        """
        nu, nv = P.shape[0] - 1, P.shape[1] - 1
        if mask is None:
            mask = np.ones((nu, nv), bool)
        used = np.zeros((nu + 1, nv + 1), bool)
        used[:-1, :-1] |= mask; used[1:, :-1] |= mask; used[1:, 1:] |= mask; used[:-1, 1:] |= mask
        ids = np.full((nu + 1, nv + 1), -1, dtype=np.int64)
        ids[used] = self._ids(P[used][:, None, :])[:, 0]
        I, J = np.nonzero(mask)
        a, b, c, d = ids[I, J], ids[I + 1, J], ids[I + 1, J + 1], ids[I, J + 1]
        even = (I + J) % 2 == 0
        t1 = np.where(even[:, None], np.stack([a, b, c], 1), np.stack([a, b, d], 1))
        t2 = np.where(even[:, None], np.stack([a, c, d], 1), np.stack([b, c, d], 1))
        T = np.concatenate([t1, t2], 0)
        X = np.asarray(self.pts)
        x0, x1, x2 = X[T[:, 0]], X[T[:, 1]], X[T[:, 2]]
        nrm = np.cross(x1 - x0, x2 - x0)
        ref = outward((x0 + x1 + x2) / 3.0)
        flip = np.einsum("ij,ij->i", nrm, ref) < 0
        T[flip] = T[flip][:, [0, 2, 1]]
        self.tris.append(T)
        self.dir.append(np.full(T.shape[0], dirichlet))
        self.grp.append(np.full(T.shape[0], group))

    def mesh(self) -> Mesh:
        return Mesh(np.asarray(self.pts, dtype=np.float64), np.concatenate(self.tris).astype(np.int32),
                    np.concatenate(self.dir).astype(bool), np.concatenate(self.grp).astype(np.int32))


def _plane(o, u, v, nu, nv):
    s = np.linspace(0.0, 1.0, nu + 1)[:, None, None]
    t = np.linspace(0.0, 1.0, nv + 1)[None, :, None]
    return np.asarray(o)[None, None, :] + s * np.asarray(u)[None, None, :] + t * np.asarray(v)[None, None, :]


def box_tank(Lx: float, Ly: float, Lz: float, nx: int, ny: int, nz: int, x0=(0.0, 0.0, 0.0)) -> Mesh:
    """Closed rectangular tank [x0, x0+L]; the top face z = z0+Lz is the free surface (Dirichlet).

    nx, ny, nz are cell counts; the 4x1x1 m tank with 42x11x11 cells has N = 2,092
    nodes and F = 4,180 triangles.
    """
    b = _Builder()
    ox, oy, oz = x0
    def out(n):
        return lambda p: np.broadcast_to(np.asarray(n, dtype=float), p.shape)
    b.grid(_plane((ox, oy, oz), (Lx, 0, 0), (0, Ly, 0), nx, ny), out((0, 0, -1)), False, 0)
    b.grid(_plane((ox, oy, oz + Lz), (Lx, 0, 0), (0, Ly, 0), nx, ny), out((0, 0, 1)), True, 1)
    b.grid(_plane((ox, oy, oz), (Lx, 0, 0), (0, 0, Lz), nx, nz), out((0, -1, 0)), False, 2)
    b.grid(_plane((ox, oy + Ly, oz), (Lx, 0, 0), (0, 0, Lz), nx, nz), out((0, 1, 0)), False, 3)
    b.grid(_plane((ox, oy, oz), (0, Ly, 0), (0, 0, Lz), ny, nz), out((-1, 0, 0)), False, 4)
    b.grid(_plane((ox + Lx, oy, oz), (0, Ly, 0), (0, 0, Lz), ny, nz), out((1, 0, 0)), False, 5)
    return b.mesh()


def _square_ring(a: float, m: int) -> np.ndarray:
    """Points on the boundary of the square [-a,a]^2, m segments per side, counter-clockwise."""
    s = np.linspace(-a, a, m + 1)[:-1]
    pts = [np.stack([s, np.full_like(s, -a)], 1), np.stack([np.full_like(s, a), s], 1),
           np.stack([-s, np.full_like(s, a)], 1), np.stack([np.full_like(s, -a), -s], 1)]
    return np.concatenate(pts, 0)


def cylinder_tank(Lx=20.0, Ly=4.0, Lz=2.0, R=0.5, draft=1.0, h=0.036, a=None, cx=0.0, cy=0.0) -> Mesh:
    """Tank [-Lx/2,Lx/2]x[-Ly/2,Ly/2]x[0,Lz] with a vertical surface-piercing cylinder.

    The cylinder (radius R, axis at (cx, cy)) occupies z in [Lz-draft, Lz] and is
    excluded from the fluid. The free surface z = Lz (Dirichlet) is an O-grid
    inside a square of half-side ~a around the cylinder whose edges lie on the
    lines of the structured free-surface grid outside it; all other faces are Neumann.
    """
    if a is None:
        a = 2.0 * R
    nx = int(round(Lx / h)); ny = int(round(Ly / h)); nz = int(round(Lz / h))
    hx, hy = Lx / nx, Ly / ny
    x0, y0 = -Lx / 2, -Ly / 2
    gx = lambda i: x0 + (i / nx) * Lx  # same arithmetic as _plane
    gy = lambda j: y0 + (j / ny) * Ly
    i_lo, i_hi = int(round((cx - a - x0) / hx)), int(round((cx + a - x0) / hx))
    j_lo, j_hi = int(round((cy - a - y0) / hy)), int(round((cy + a - y0) / hy))
    b = _Builder()
    const = lambda n: (lambda p: np.broadcast_to(np.asarray(n, dtype=float), p.shape))
    b.grid(_plane((x0, y0, 0.0), (Lx, 0, 0), (0, Ly, 0), nx, ny), const((0, 0, -1)), False, 0)
    b.grid(_plane((x0, y0, 0.0), (Lx, 0, 0), (0, 0, Lz), nx, nz), const((0, -1, 0)), False, 2)
    b.grid(_plane((x0, y0 + Ly, 0.0), (Lx, 0, 0), (0, 0, Lz), nx, nz), const((0, 1, 0)), False, 3)
    b.grid(_plane((x0, y0, 0.0), (0, Ly, 0), (0, 0, Lz), ny, nz), const((-1, 0, 0)), False, 4)
    b.grid(_plane((x0 + Lx, y0, 0.0), (0, Ly, 0), (0, 0, Lz), ny, nz), const((1, 0, 0)), False, 5)
    # free surface minus the O-grid square (cells [i_lo, i_hi) x [j_lo, j_hi))
    P = _plane((x0, y0, Lz), (Lx, 0, 0), (0, Ly, 0), nx, ny)
    mask = np.ones((nx, ny), bool)
    mask[i_lo:i_hi, j_lo:j_hi] = False
    b.grid(P, const((0, 0, 1)), True, 1, mask=mask)
    # square boundary through the grid nodes, counter-clockwise from (i_lo, j_lo)
    ring = [(gx(i), gy(j_lo)) for i in range(i_lo, i_hi)] + [(gx(i_hi), gy(j)) for j in range(j_lo, j_hi)] + \
           [(gx(i), gy(j_hi)) for i in range(i_hi, i_lo, -1)] + [(gx(i_lo), gy(j)) for j in range(j_hi, j_lo, -1)]
    sq = np.asarray(ring + ring[:1], dtype=float)
    rel = sq - np.array([cx, cy])[None, :]
    ang = np.arctan2(rel[:, 1], rel[:, 0])
    circ = np.stack([R * np.cos(ang), R * np.sin(ang)], 1)
    circ[-1] = circ[0]
    half = min(gx(i_hi) - cx, cx - gx(i_lo), gy(j_hi) - cy, cy - gy(j_lo))
    nr = max(2, int(round((half - R) / h)))
    t = np.linspace(0.0, 1.0, nr + 1)
    G = np.zeros((sq.shape[0], nr + 1, 3))
    G[:, :, 0] = (cx + circ[:, None, 0]) * (1 - t[None, :]) + sq[:, None, 0] * t[None, :]
    G[:, :, 1] = (cy + circ[:, None, 1]) * (1 - t[None, :]) + sq[:, None, 1] * t[None, :]
    G[:, :, 2] = Lz
    b.grid(G, const((0, 0, 1)), True, 1)
    # cylinder side: out of the fluid = towards the axis
    nzc = max(2, int(round(draft / h)))
    zc = np.linspace(Lz - draft, Lz, nzc + 1)
    S = np.zeros((sq.shape[0], nzc + 1, 3))
    S[:, :, 0] = cx + circ[:, None, 0]
    S[:, :, 1] = cy + circ[:, None, 1]
    S[:, :, 2] = zc[None, :]
    b.grid(S, lambda p: -np.stack([p[:, 0] - cx, p[:, 1] - cy, np.zeros(len(p))], 1), False, 6)
    # cylinder bottom disk (z = Lz - draft): out of the fluid = +z; scaled inner square + annulus
    sc = 0.5 * R / half
    sqb = np.array([cx, cy])[None, :] + sc * rel
    nrb = max(2, int(round(0.5 * R / h)))
    t = np.linspace(0.0, 1.0, nrb + 1)
    D = np.zeros((sq.shape[0], nrb + 1, 3))
    D[:, :, 0] = (cx + circ[:, None, 0]) * (1 - t[None, :]) + sqb[:, None, 0] * t[None, :]
    D[:, :, 1] = (cy + circ[:, None, 1]) * (1 - t[None, :]) + sqb[:, None, 1] * t[None, :]
    D[:, :, 2] = Lz - draft
    b.grid(D, const((0, 0, 1)), False, 7)
    Q = _plane((sqb[0, 0], sqb[0, 1], Lz - draft), (sqb[i_hi - i_lo, 0] - sqb[0, 0], 0, 0),
               (0, sqb[(i_hi - i_lo) + (j_hi - j_lo), 1] - sqb[0, 1], 0), i_hi - i_lo, j_hi - j_lo)
    b.grid(Q, const((0, 0, 1)), False, 7)
    return b.mesh()


# ---------------------------------------------------------------------------
# DOF model
# ---------------------------------------------------------------------------
def _part1by2(v: np.ndarray) -> np.ndarray:
    v = v.astype(np.uint64) & np.uint64(0x1FFFFF)
    v = (v | (v << np.uint64(32))) & np.uint64(0x1F00000000FFFF)
    v = (v | (v << np.uint64(16))) & np.uint64(0x1F0000FF0000FF)
    v = (v | (v << np.uint64(8))) & np.uint64(0x100F00F00F00F00F)
    v = (v | (v << np.uint64(4))) & np.uint64(0x10C30C30C30C30C3)
    v = (v | (v << np.uint64(2))) & np.uint64(0x1249249249249249)
    return v


def morton_keys(X: np.ndarray) -> np.ndarray:
    """21-bit/axis Morton keys in the point bounding box (GMRES_FMM.hpp:3164-3222)."""
    lo, hi = X.min(0), X.max(0)
    maxv = float((1 << 21) - 1)
    q = np.zeros(X.shape, dtype=np.uint64)
    for a in range(3):
        if hi[a] > lo[a]:
            t = np.clip((X[:, a] - lo[a]) / (hi[a] - lo[a]), 0.0, 1.0)
            q[:, a] = (t * maxv + 0.5).astype(np.uint64)
    return _part1by2(q[:, 0]) | (_part1by2(q[:, 1]) << np.uint64(1)) | (_part1by2(q[:, 2]) << np.uint64(2))


def build_dofs(mesh: Mesh, multiple_angle_deg: float = 20.0) -> DofModel:
    n, F = mesh.n_nodes, mesh.n_faces
    tri = mesh.tri
    fn = mesh.face_normals()
    fa = mesh.face_areas()
    # incidence
    order = np.argsort(tri.reshape(-1), kind="stable")
    ptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(ptr, tri.reshape(-1) + 1, 1)
    ptr = np.cumsum(ptr)
    inc_face = (order // 3).astype(np.int64)
    nd = mesh.face_dirichlet
    all_neu = np.ones(n, bool); all_dir = np.ones(n, bool)
    np.logical_and.at(all_neu, tri.reshape(-1), np.repeat(~nd, 3))
    np.logical_and.at(all_dir, tri.reshape(-1), np.repeat(nd, 3))
    pclass = np.where(all_dir, 0, np.where(all_neu, 1, 2)).astype(np.int8)
    # point normal: area-weighted mean of face normals; multiple node if any face deviates > angle
    pn = np.zeros((n, 3))
    np.add.at(pn, tri.reshape(-1), np.repeat(fn * fa[:, None], 3, axis=0))
    pn /= np.linalg.norm(pn, axis=1, keepdims=True)
    cosang = np.einsum("ij,ij->i", fn[inc_face], pn[np.repeat(np.arange(n), np.diff(ptr))])
    dev = cosang < math.cos(math.radians(multiple_angle_deg))
    multiple = np.zeros(n, bool)
    np.logical_or.at(multiple, np.repeat(np.arange(n), np.diff(ptr)), dev)
    multiple |= pclass == 2
    # DOF keys: (point, face) for Neumann faces of multiple Neumann/CORNER points; else (point, -1)
    entries = []  # (point, face, kind)
    for p in range(n):
        pc = pclass[p]
        neu_or_corner = pc in (1, 2)
        has_null = (neu_or_corner and not multiple[p]) or pc in (0, 2)
        if has_null:
            entries.append((p, -1, 0 if pc in (0, 2) else 1))
        if neu_or_corner and multiple[p]:
            for f in inc_face[ptr[p]:ptr[p + 1]]:
                if not nd[f]:
                    entries.append((p, int(f), 2 if pc == 2 else 1))
    ent = np.asarray(entries, dtype=np.int64)
    keys = morton_keys(mesh.X)[ent[:, 0]]
    srt = np.lexsort((ent[:, 1], ent[:, 0], keys))
    ent = ent[srt]
    n_dof = ent.shape[0]
    lookup = {(int(p), int(f)): i for i, (p, f, _) in enumerate(ent)}
    face_dof = np.empty((F, 3), dtype=np.int32)
    for f in range(F):
        for k in range(3):
            p = int(tri[f, k])
            key = (p, f) if ((pclass[p] in (1, 2)) and multiple[p] and not nd[f]) else (p, -1)
            face_dof[f, k] = lookup[key]
    return DofModel(n_dof, face_dof, ent[:, 2].astype(np.int32), ent[:, 0].astype(np.int32),
                    ent[:, 1].astype(np.int32), pclass, multiple)


# ---------------------------------------------------------------------------
# true quadratic elements: edge-midpoint nodes and their DOFs
# ---------------------------------------------------------------------------
def to_quadratic(mesh: Mesh, project=None) -> Mesh:
    """The same surface with 6-node faces: (p0, p1, p2, m01, m12, m20), the midpoint nodes appended after the
    vertices (one per edge). ``project`` maps the (k, 3) straight-edge midpoints and the (k, 2) end vertex ids to
    their positions on the curved surface (default: the straight-edge midpoints). The midpoint X is the
    reference's networkLine::X_mid (Network.hpp:280), the order that of getQuadDOFIndices
    (bem/BEM_setBoundaryTypes.hpp:120-133)."""
    tri = np.asarray(mesh.tri)[:, :3]
    e = np.stack([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]], 1).reshape(-1, 2)
    key = np.sort(e, axis=1)
    uniq, inv = np.unique(key, axis=0, return_inverse=True)
    Xm = 0.5 * (mesh.X[uniq[:, 0]] + mesh.X[uniq[:, 1]])
    if project is not None:
        Xm = np.asarray(project(Xm, uniq), dtype=np.float64)
    mid = mesh.n_nodes + inv.reshape(-1, 3)
    tri6 = np.concatenate([tri, mid], 1).astype(np.int32)
    return Mesh(np.concatenate([mesh.X, Xm], 0), tri6, mesh.face_dirichlet.copy(), mesh.face_group.copy())


def vertex_mesh(mesh6: Mesh) -> Mesh:
    """The linear view of a quadratic mesh (vertices only, midpoint nodes kept unreferenced)."""
    return Mesh(mesh6.X, np.ascontiguousarray(mesh6.tri[:, :3]), mesh6.face_dirichlet, mesh6.face_group)


def build_dofs_quadratic(mesh6: Mesh, multiple_angle_deg: float = 20.0) -> DofModel:
    """DOF model of true quadratic faces (setNodeFaceIndices with use_true_quadratic_element,
    bem/BEM_setBoundaryTypes.hpp:190-233; lf2Index :113-117): the vertex DOFs of build_dofs, then per edge
    midpoint a primary DOF (Dirichlet if both faces of the edge are Dirichlet, else Neumann; Dirichlet also for
    a CORNER edge between a Dirichlet and a Neumann face) and, on a CORNER edge, a CORNER Neumann DOF for its
    Neumann face. A face's midpoint entry is the primary DOF, or the CORNER DOF for the Neumann face of a CORNER
    edge (edges are never multiple nodes otherwise, however sharp)."""
    nv = int(mesh6.tri[:, :3].max()) + 1  # vertices first, then the midpoint nodes (to_quadratic)
    if mesh6.tri[:, 3:].min() < nv:
        raise ValueError("midpoint nodes must follow the vertices")
    lin = Mesh(mesh6.X[:nv], np.ascontiguousarray(mesh6.tri[:, :3]), mesh6.face_dirichlet, mesh6.face_group)
    d = build_dofs(lin, multiple_angle_deg)
    F = mesh6.n_faces
    mids = mesh6.tri[:, 3:6]
    nd = mesh6.face_dirichlet
    nm = mesh6.n_nodes
    n_dir = np.zeros(nm, np.int64)
    n_neu = np.zeros(nm, np.int64)
    np.add.at(n_dir, mids[nd].reshape(-1), 1)
    np.add.at(n_neu, mids[~nd].reshape(-1), 1)
    used = np.flatnonzero((n_dir + n_neu) > 0)
    kinds, points, faces = [], [], []
    primary = np.full(nm, -1, np.int64)
    corner_dof = np.full(nm, -1, np.int64)
    k = d.n_dof
    for m in used:
        corner = n_dir[m] > 0 and n_neu[m] > 0
        primary[m] = k
        kinds.append(DIRICHLET if n_dir[m] > 0 else NEUMANN)
        points.append(m)
        faces.append(-1)
        k += 1
        if corner:
            corner_dof[m] = k
            kinds.append(CORNER_NEUMANN)
            points.append(m)
            faces.append(-2)  # set to the Neumann face below
            k += 1
    face_dof = np.empty((F, 6), np.int32)
    face_dof[:, :3] = d.face_dof
    dof_face = np.concatenate([d.dof_face, np.asarray(faces, np.int32)])
    for f in range(F):
        for e in range(3):
            m = int(mids[f, e])
            if corner_dof[m] >= 0 and not nd[f]:
                face_dof[f, 3 + e] = corner_dof[m]
                dof_face[corner_dof[m]] = f
            else:
                face_dof[f, 3 + e] = primary[m]
    pclass = np.concatenate([d.point_class, np.zeros(nm - lin.n_nodes, np.int8)])
    pclass[used] = np.where((n_dir[used] > 0) & (n_neu[used] > 0), 2, np.where(n_dir[used] > 0, 0, 1))
    multiple = np.concatenate([d.point_multiple, np.zeros(nm - lin.n_nodes, bool)])
    return DofModel(int(k), face_dof, np.concatenate([d.dof_kind, np.asarray(kinds, np.int32)]).astype(np.int32),
                    np.concatenate([d.dof_point, np.asarray(points, np.int32)]).astype(np.int32),
                    dof_face.astype(np.int32), pclass, multiple)


def standing_wave(mesh: Mesh, dofs: DofModel, Lx: float, n: int = 1):
    """phi = cosh(k z) cos(k x), k = n pi / Lx: harmonic, with zero normal derivative on the walls x = 0, Lx,
    y = const and the bottom z = 0 of a box tank; returns (phi, phin) boundary data (phi on Dirichlet DOFs, 0
    elsewhere as the Neumann data) and the exact (phi, phin) per DOF (phin on the free surface z = Lz only)."""
    kk = n * math.pi / Lx
    X = mesh.X[dofs.dof_point]
    phi_ex = np.cosh(kk * X[:, 2]) * np.cos(kk * X[:, 0])
    phin_ex = np.zeros(dofs.n_dof)
    top = dofs.dof_kind == DIRICHLET
    phin_ex[top] = kk * np.sinh(kk * X[top, 2]) * np.cos(kk * X[top, 0])
    return phi_ex.copy(), np.zeros(dofs.n_dof), phi_ex, phin_ex


def point_normals(mesh: Mesh) -> np.ndarray:
    fn, fa = mesh.face_normals(), mesh.face_areas()
    pn = np.zeros((mesh.n_nodes, 3))
    k = mesh.tri.shape[1]  # 3, or 6 with the edge midpoints of quadratic faces
    np.add.at(pn, mesh.tri.reshape(-1), np.repeat(fn * fa[:, None], k, axis=0))
    nrm = np.linalg.norm(pn, axis=1, keepdims=True)
    return pn / np.where(nrm > 0, nrm, 1.0)


def dof_normals(mesh: Mesh, dofs: DofModel) -> np.ndarray:
    """Normal a DOF's phi_n refers to: its face normal for per-face DOFs; for the null-key DOF
    of a CORNER point, the mean normal of its Dirichlet faces; else the point normal."""
    pn = point_normals(mesh)
    fn, fa = mesh.face_normals(), mesh.face_areas()
    dn = np.zeros((mesh.n_nodes, 3))
    w = (fn * fa[:, None]) * mesh.face_dirichlet[:, None]
    np.add.at(dn, mesh.tri.reshape(-1), np.repeat(w, mesh.tri.shape[1], axis=0))
    nd = np.linalg.norm(dn, axis=1)
    corner = dofs.point_class == 2
    pn[corner] = dn[corner] / nd[corner, None]
    out = pn[dofs.dof_point].copy()
    per_face = dofs.dof_face >= 0
    out[per_face] = fn[dofs.dof_face[per_face]]
    return out


# ---------------------------------------------------------------------------
# boundary conditions of the benchmark workloads
# ---------------------------------------------------------------------------
def free_surface_modes(rng: np.random.Generator, n_modes: int = 8, amp: float = 0.01):
    a = rng.normal(0.0, amp, n_modes)
    k = rng.integers(1, 5, n_modes)
    l = rng.integers(1, 5, n_modes)
    th = rng.uniform(0.0, 2 * math.pi, n_modes)
    return a, k, l, th


def tank_bc(mesh: Mesh, dofs: DofModel, seed: int, Lx: float, Ly: float, body_velocity_sigma: float = 0.0,
            body_group: tuple = (6, 7)):
    """phi on Dirichlet (free surface): seeded sum of 8 sine modes; phi_n = 0 on walls and bottom;
    phi_n = n.U on the body faces (groups 6, 7) with U ~ N(0, sigma) (seeded)."""
    rng = np.random.default_rng(seed)
    a, k, l, th = free_surface_modes(rng)
    X = mesh.X
    phi_pt = np.zeros(mesh.n_nodes)
    for i in range(len(a)):
        phi_pt += a[i] * np.sin(math.pi * (k[i] * X[:, 0] / Lx + l[i] * X[:, 1] / Ly) + th[i])
    phi = phi_pt[dofs.dof_point].copy()
    phin = np.zeros(dofs.n_dof)
    if body_velocity_sigma > 0.0:
        U = rng.normal(0.0, body_velocity_sigma, 3)
        on_body = np.zeros(mesh.n_nodes, bool)
        body_faces = np.isin(mesh.face_group, body_group)
        on_body[mesh.tri[body_faces].reshape(-1)] = True
        nrm = dof_normals(mesh, dofs)
        sel = (dofs.dof_kind != DIRICHLET) & on_body[dofs.dof_point]
        # per-face DOF: only if its face is a body face
        pf = dofs.dof_face >= 0
        sel &= ~pf | np.isin(mesh.face_group[np.maximum(dofs.dof_face, 0)], body_group)
        phin[sel] = nrm[sel] @ U
    return phi, phin


def perturb_free_surface(mesh: Mesh, seed: int, sigma: float) -> Mesh:
    """z-offsets ~ N(0, sigma) on free-surface nodes that are interior to the top face."""
    rng = np.random.default_rng(seed)
    X = mesh.X.copy()
    top = np.zeros(mesh.n_nodes, bool)
    top[mesh.tri[mesh.face_dirichlet].reshape(-1)] = True
    other = np.zeros(mesh.n_nodes, bool)
    other[mesh.tri[~mesh.face_dirichlet].reshape(-1)] = True
    interior = top & ~other
    X[interior, 2] += rng.normal(0.0, sigma, int(interior.sum()))
    return Mesh(X, mesh.tri.copy(), mesh.face_dirichlet.copy(), mesh.face_group.copy())


def analytic_bc(mesh: Mesh, dofs: DofModel, kind: str = "linear"):
    """Known harmonic field: returns (phi, phin) BC arrays and the exact (phi, phin) per DOF."""
    X = mesh.X[dofs.dof_point]
    nrm = dof_normals(mesh, dofs)
    if kind == "linear":
        g = np.array([1.0, -0.5, 0.25])
        phi_ex = X @ g
        phin_ex = nrm @ g
    elif kind == "exp":
        kk = math.pi / 2.0
        phi_ex = np.exp(kk * X[:, 2]) * np.cos(kk * X[:, 0])
        grad = np.stack([-kk * np.exp(kk * X[:, 2]) * np.sin(kk * X[:, 0]), np.zeros(len(X)),
                         kk * np.exp(kk * X[:, 2]) * np.cos(kk * X[:, 0])], 1)
        phin_ex = np.einsum("ij,ij->i", grad, nrm)
    else:
        raise ValueError(kind)
    return phi_ex.copy(), phin_ex.copy(), phi_ex, phin_ex


def robin_dof_kind(mesh: Mesh, dofs: DofModel, face_robin: np.ndarray) -> np.ndarray:
    """DOF kinds with the Robin DOFs of a frequency-domain problem marked ROBIN, by the rule of
    main_freq_domain.cpp:413-426: a DOF keyed by a face is Robin if that face is a Robin (free-surface)
    face; a point's null-key DOF is Robin if the point lies on a Robin face. Robin faces must be
    non-Dirichlet in ``mesh`` (the reference makes every point Neumann there, :403-408), so the
    DOFs they carry are Neumann DOFs of ``build_dofs``; CORNER DOFs stay CORNER."""
    fr = np.asarray(face_robin, dtype=bool)
    if fr.shape != (mesh.n_faces,):
        raise ValueError("face_robin needs one entry per face")
    if (fr & mesh.face_dirichlet).any():
        raise ValueError("a Robin face must not be Dirichlet")
    on_robin = np.zeros(mesh.n_nodes, bool)
    on_robin[mesh.tri[fr].reshape(-1)] = True
    kind = dofs.dof_kind.copy()
    f = dofs.dof_face
    robin = np.where(f >= 0, fr[np.maximum(f, 0)], on_robin[dofs.dof_point])
    kind[(kind == NEUMANN) & robin] = ROBIN
    return kind
