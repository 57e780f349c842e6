"""The reference's input format: Wavefront OBJ surfaces and the JSON case files that name them.

A case directory holds ``settings.json`` (``"input_files": [...]``) and one JSON file per body
(``{"name", "type": "Fluid" | "RigidBody" | ..., "objfile": [path]}``; bem/examples/goring1979,
read by bem/BEM_inputfile_reader.hpp). This module turns such a case into the flat arrays of
``mesh.Mesh``: the fluid surface, oriented out of the fluid, with a Dirichlet/Neumann tag per face.

Boundary types follow ``setBoundaryTypes`` (bem/BEM_setBoundaryTypes.hpp:447-470): a fluid face is
Neumann if every one of its points is in contact with a rigid-body face, or if its centroid lies
inside a (closed) rigid body; every other face is Dirichlet (free surface). The reference's contact test
(``setContactRange`` / ``addContactFaces``, lib/include/Network.hpp:2955-2990, with normal-angle
checks in lib/src/networkPoint.cpp:178-210) is reduced here to distance: a point is in contact when
the nearest point of some rigid-body triangle is closer than ``contact_factor`` times the mean length
of the point's own edges and the triangle is parallel to the fluid face within 60 degrees (the
role of the reference's normal-angle checks). The Goring 1979 example (water against a tank and a wave maker) gets the
expected split with this rule (free surface Dirichlet, walls / bottom / paddle Neumann); cases that
rely on the angle checks (e.g. a body piercing the surface at a shallow angle) may differ.
"""
from __future__ import annotations

import dataclasses
import json
import os

import numpy as np

from .mesh import Mesh, morton_keys


def read_obj(path: str):
    """Vertices (n, 3), polygon faces (lists of 0-based vertex indices) and a group id per face.

    Accepts ``v x y z [w]``, ``f a b c ...`` with ``a``, ``a/t``, ``a/t/n`` or ``a//n`` corners and
    negative (relative) indices, ``g``/``o`` group names; other records (vt, vn, usemtl, mtllib, s,
    comments) are skipped."""
    verts, faces, groups, names = [], [], [], {}
    g = 0
    with open(path, "r", encoding="utf-8", errors="replace") as fh:
        for line in fh:
            parts = line.split()
            if not parts:
                continue
            tag = parts[0]
            if tag == "v":
                verts.append((float(parts[1]), float(parts[2]), float(parts[3])))
            elif tag == "f":
                idx = []
                for c in parts[1:]:
                    k = int(c.split("/")[0])
                    idx.append(k - 1 if k > 0 else len(verts) + k)
                if len(idx) < 3:
                    raise ValueError(f"{path}: face with fewer than 3 corners")
                faces.append(idx)
                groups.append(g)
            elif tag in ("g", "o"):
                name = " ".join(parts[1:])
                g = names.setdefault(name, len(names))
    X = np.asarray(verts, dtype=np.float64).reshape(-1, 3)
    for f in faces:
        if min(f) < 0 or max(f) >= len(X):
            raise ValueError(f"{path}: face index out of range")
    return X, faces, np.asarray(groups, dtype=np.int32)


def triangulate(X: np.ndarray, faces, weld_tol: float = 0.0):
    """Fan-triangulated faces (n_tri, 3), unused vertices dropped; with weld_tol > 0 vertices closer
    than weld_tol are merged (OBJ exporters often duplicate seam vertices). Returns (X, tri, src_face)."""
    tri, src = [], []
    for k, f in enumerate(faces):
        for j in range(1, len(f) - 1):
            tri.append((f[0], f[j], f[j + 1]))
            src.append(k)
    tri = np.asarray(tri, dtype=np.int64).reshape(-1, 3)
    src = np.asarray(src, dtype=np.int64)
    if weld_tol > 0:
        key = np.round(X / weld_tol).astype(np.int64)
        _, first, inv = np.unique(key, axis=0, return_index=True, return_inverse=True)
        tri = inv.reshape(-1)[tri]
        X = X[first]
    keep = np.zeros(len(X), bool)
    keep[tri.reshape(-1)] = True
    new = np.cumsum(keep) - 1
    tri = new[tri]
    X = X[keep]
    # degenerate triangles (welded corners) are dropped
    ok = (tri[:, 0] != tri[:, 1]) & (tri[:, 1] != tri[:, 2]) & (tri[:, 0] != tri[:, 2])
    return X, tri[ok].astype(np.int32), src[ok]


def signed_volume(X: np.ndarray, tri: np.ndarray) -> float:
    """Volume enclosed by a closed triangulated surface, positive for outward orientation."""
    a, b, c = X[tri[:, 0]], X[tri[:, 1]], X[tri[:, 2]]
    return float(np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0)


def _point_tri_dist(P: np.ndarray, A: np.ndarray, B: np.ndarray, C: np.ndarray) -> np.ndarray:
    """(n, m) distances from the points P (n, 3) to the triangles (A, B, C) (m, 3 each), by the
    region classification of Ericson, Real-Time Collision Detection, 5.1.5."""
    ab, ac = B - A, C - A
    p = P[:, None, :]
    ap, bp, cp = p - A, p - B, p - C
    d1, d2 = np.einsum("nmk,mk->nm", ap, ab), np.einsum("nmk,mk->nm", ap, ac)
    d3, d4 = np.einsum("nmk,mk->nm", bp, ab), np.einsum("nmk,mk->nm", bp, ac)
    d5, d6 = np.einsum("nmk,mk->nm", cp, ab), np.einsum("nmk,mk->nm", cp, ac)
    va = d3 * d6 - d5 * d4
    vb = d5 * d2 - d1 * d6
    vc = d1 * d4 - d3 * d2
    denom = va + vb + vc
    with np.errstate(divide="ignore", invalid="ignore"):
        v = np.where(denom != 0, vb / denom, 0.0)
        w = np.where(denom != 0, vc / denom, 0.0)
        t_ab = np.clip(d1 / np.where(d1 - d3 != 0, d1 - d3, 1.0), 0, 1)
        t_ac = np.clip(d2 / np.where(d2 - d6 != 0, d2 - d6, 1.0), 0, 1)
        t_bc = np.clip((d4 - d3) / np.where((d4 - d3) + (d5 - d6) != 0, (d4 - d3) + (d5 - d6), 1.0), 0, 1)
    interior = (va >= 0) & (vb >= 0) & (vc >= 0)
    dist = np.minimum(np.minimum(np.linalg.norm(p - (A + t_ab[..., None] * ab), axis=-1),
                                 np.linalg.norm(p - (A + t_ac[..., None] * ac), axis=-1)),
                      np.linalg.norm(p - (B + t_bc[..., None] * (C - B)), axis=-1))
    return np.where(interior, np.linalg.norm(p - (A + v[..., None] * ab + w[..., None] * ac), axis=-1), dist)


def _closest_point_distance(P: np.ndarray, A: np.ndarray, B: np.ndarray, C: np.ndarray) -> np.ndarray:
    """Distance from each point of P to the nearest of the triangles (A, B, C); chunked over P."""
    out = np.full(len(P), np.inf)
    for s in range(0, len(P), 512):
        out[s:s + 512] = _point_tri_dist(P[s:s + 512], A, B, C).min(axis=1)
    return out


def _closed(tri: np.ndarray) -> bool:
    """Every edge shared by exactly two triangles (the inside test needs a closed surface)."""
    e = np.sort(np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]]), axis=1)
    _, cnt = np.unique(e, axis=0, return_counts=True)
    return bool((cnt == 2).all())


def _inside(P: np.ndarray, X: np.ndarray, tri: np.ndarray) -> np.ndarray:
    """Points strictly inside a closed surface: parity of the crossings of a ray in a fixed generic
    direction (Moller-Trumbore)."""
    d = np.array([0.5772156649, 0.3183098862, 0.7548776662])
    d /= np.linalg.norm(d)
    A, B, C = X[tri[:, 0]], X[tri[:, 1]], X[tri[:, 2]]
    e1, e2 = B - A, C - A
    h = np.cross(d, e2)
    a = np.einsum("mk,mk->m", e1, h)
    ok = np.abs(a) > 1e-14
    f = np.where(ok, 1.0 / np.where(ok, a, 1.0), 0.0)
    out = np.zeros(len(P), bool)
    for s in range(0, len(P), 512):
        p = P[s:s + 512, None, :]
        sv = p - A
        u = f * np.einsum("nmk,mk->nm", sv, h)
        q = np.cross(sv, e1)
        v = f * np.einsum("k,nmk->nm", d, q)
        t = f * np.einsum("nmk,mk->nm", q, e2)
        hit = ok & (u >= 0) & (v >= 0) & (u + v <= 1) & (t > 1e-12)
        out[s:s + 512] = (hit.sum(axis=1) % 2) == 1
    return out


def load_case(settings_path: str, contact_factor: float = 0.5, weld_tol: float = 0.0) -> tuple:
    """A reference case (settings.json + body JSON files + OBJ surfaces) as (Mesh, info).

    The Fluid body's surface is oriented out of the fluid; its faces are tagged Dirichlet/Neumann by
    contact with the RigidBody surfaces (module docstring). ``face_group`` is 0 for Dirichlet faces
    and 1 + the index of the nearest rigid body for Neumann faces. ``info`` lists the bodies read."""
    base = os.path.dirname(os.path.abspath(settings_path))
    with open(settings_path, "r", encoding="utf-8") as fh:
        settings = json.load(fh)
    fluid, solids, info = None, [], {"bodies": []}
    for name in settings.get("input_files", []):
        with open(os.path.join(base, name), "r", encoding="utf-8") as fh:
            body = json.load(fh)
        kind = body.get("type", "")
        objs = body.get("objfile", [])
        objs = [objs] if isinstance(objs, str) else objs
        if kind not in ("Fluid", "RigidBody", "SoftBody") or not objs:
            continue  # probes, gauges and other inputs carry no surface for the BVP
        X, faces, _ = read_obj(os.path.join(base, objs[0]))
        X, tri, _ = triangulate(X, faces, weld_tol)
        info["bodies"].append({"name": body.get("name", name), "type": kind, "obj": objs[0],
                               "n_nodes": int(len(X)), "n_faces": int(len(tri))})
        if kind == "Fluid":
            if fluid is not None:
                raise ValueError("more than one Fluid body")
            fluid = (X, tri)
        else:
            solids.append((X, tri))
    if fluid is None:
        raise ValueError("no Fluid body in the case")
    return fluid_mesh(fluid[0], fluid[1], solids, contact_factor), info


def fluid_mesh(X: np.ndarray, tri: np.ndarray, solids, contact_factor: float = 0.5) -> Mesh:
    """Mesh of a fluid surface with boundary types from contact with the rigid surfaces ``solids``
    (list of (X, tri)): faces whose points all touch a solid, or whose centroid is inside one, are
    Neumann."""
    tri = np.asarray(tri, dtype=np.int32)
    if signed_volume(X, tri) < 0:
        tri = tri[:, [0, 2, 1]]
    # contact range per point: contact_factor x mean length of its edges
    e = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]])
    ln = np.linalg.norm(X[e[:, 0]] - X[e[:, 1]], axis=1)
    tot = np.bincount(e.reshape(-1), np.repeat(ln, 2), minlength=len(X))
    cnt = np.bincount(e.reshape(-1), minlength=len(X))
    rng = contact_factor * tot / np.maximum(cnt, 1)
    n_f = len(tri)
    neumann = np.zeros(n_f, bool)
    group = np.zeros(n_f, np.int32)
    best = np.full(n_f, np.inf)
    cent = X[tri].mean(axis=1)
    nf = np.cross(X[tri[:, 1]] - X[tri[:, 0]], X[tri[:, 2]] - X[tri[:, 0]])
    nf /= np.linalg.norm(nf, axis=1, keepdims=True)
    cos_par = np.cos(np.deg2rad(60.0))
    # fluid faces in Morton order of their centroids, so that a chunk of them is compact in space and
    # only the rigid triangles near its bounding box need distances
    order = np.argsort(morton_keys(cent), kind="stable")
    for k, (Xs, Ts) in enumerate(solids):
        A, B, Cc = Xs[Ts[:, 0]], Xs[Ts[:, 1]], Xs[Ts[:, 2]]
        ns = np.cross(B - A, Cc - A)
        ns /= np.linalg.norm(ns, axis=1, keepdims=True)
        tlo, thi = np.minimum(np.minimum(A, B), Cc), np.maximum(np.maximum(A, B), Cc)
        # a rigid triangle counts for a fluid face only if the two are parallel within 60 degrees (either
        # orientation: rigid OBJs come both ways): walls perpendicular to a face do not make it Neumann
        # (the rim of the free surface)
        dmax = np.full(n_f, np.inf)
        for s in range(0, n_f, 64):
            sel = order[s:s + 64]
            f = tri[sel]
            P = X[f.reshape(-1)]
            r = rng[f].max()
            lo, hi = P.min(axis=0) - r, P.max(axis=0) + r
            cand = np.nonzero(((tlo <= hi) & (thi >= lo)).all(axis=1))[0]
            if len(cand) == 0:
                continue
            D = _point_tri_dist(P, A[cand], B[cand], Cc[cand]).reshape(len(f), 3, -1)
            facing = np.abs(nf[sel] @ ns[cand].T) > cos_par
            D = np.where(facing[:, None, :], D, np.inf)
            touch = D <= rng[f][:, :, None]
            ok = touch.any(axis=2).all(axis=1)
            dmax[sel] = np.where(ok, D.min(axis=2).max(axis=1), np.inf)
        inside = np.zeros(n_f, bool)
        if _closed(Ts):  # only centroids in the body's bounding box need the ray test
            box = ((cent >= Xs.min(axis=0)) & (cent <= Xs.max(axis=0))).all(axis=1)
            inside[box] = _inside(cent[box], Xs, Ts)
        hit = np.isfinite(dmax) | inside
        dm = np.where(inside & ~np.isfinite(dmax), 0.0, dmax)
        take = hit & (dm < best)
        group[take] = k + 1
        best[take] = dm[take]
        neumann |= hit
    return Mesh(np.ascontiguousarray(X, dtype=np.float64), np.ascontiguousarray(tri, dtype=np.int32),
                ~neumann, group)


def write_obj(path: str, X: np.ndarray, tri: np.ndarray, group: str = "surface"):
    """Write a triangulated surface as OBJ (1-based indices)."""
    with open(path, "w", encoding="utf-8") as fh:
        fh.write(f"g {group}\n")
        for x in X:
            fh.write(f"v {x[0]:.17g} {x[1]:.17g} {x[2]:.17g}\n")
        for t in tri:
            fh.write(f"f {t[0] + 1} {t[1] + 1} {t[2] + 1}\n")


# ---------------------------------------------------------------------------
# checkpoint files (format version 5, bem/BEM_checkpoint.hpp:18-55; writer :127-322)
# ---------------------------------------------------------------------------
CHECKPOINT_MAGIC = b"BEMCKPT\0"
CHECKPOINT_VERSION = 5
# per face value of a multiple node: checkpoint face id (-1 = the point's null-face key), phi, phin, phit, phint
FACE_VALUE = np.dtype([("face", "<i4"), ("phi", "<f8"), ("phin", "<f8"), ("phit", "<f8"), ("phint", "<f8")])
# per bad-quality event of a face
FACE_EVENT = np.dtype([("time_step", "<i4"), ("rk_step", "<i4"), ("type", "u1"), ("value", "<f8"), ("threshold", "<f8")])
_POINT_HEAD = np.dtype([("X", "<f8", 3), ("phiphin", "<f8", 2), ("phiphin_t", "<f8", 2), ("phi_dirichlet", "<f8")])
_LINE_HEAD = np.dtype([("ends", "<u4", 2), ("X_mid", "<f8", 3), ("phiphin", "<f8", 2), ("phiphin_t", "<f8", 2),
                       ("corner_offset", "<f8", 3)])


@dataclasses.dataclass
class CheckpointFluid:
    """One fluid surface of a checkpoint. Face values of multiple nodes are kept in CSR form:
    the records of point p are ``face_values[fv_ptr[p]:fv_ptr[p + 1]]`` (dtype FACE_VALUE); likewise
    the bad-quality events of face f are ``events[ev_ptr[f]:ev_ptr[f + 1]]`` (dtype FACE_EVENT).
    ``lines`` holds the true-quadratic line data when the file has it (endpoint ids, midpoint data,
    and their face values in the same CSR form), else None."""
    name: str
    X: np.ndarray              # (n, 3)
    phiphin: np.ndarray        # (n, 2) phi, phi_n
    phiphin_t: np.ndarray      # (n, 2) phi_t, phi_nt
    phi_dirichlet: np.ndarray  # (n,)
    fv_ptr: np.ndarray         # (n + 1,) int64
    face_values: np.ndarray    # (m,) FACE_VALUE
    tri: np.ndarray            # (F, 3) int64 checkpoint point ids
    ev_ptr: np.ndarray         # (F + 1,) int64
    events: np.ndarray         # (k,) FACE_EVENT
    lines: dict | None = None


@dataclasses.dataclass
class Checkpoint:
    """Contents of a checkpoint file: header, fluid surfaces, rigid and soft bodies (name, COM,
    quaternion Q, velocity, acceleration for rigid bodies, point positions) and vortex particles
    (x, alpha, sigma, volume)."""
    time_step: int
    simulation_time: float
    dt: float
    fluids: list
    rigid: list
    soft: list
    particles: dict


class _Reader:
    def __init__(self, buf: bytes):
        self.b, self.o = buf, 0

    def take(self, dtype, count=1):
        dt = np.dtype(dtype)
        n = dt.itemsize * count
        if self.o + n > len(self.b):
            raise ValueError("checkpoint: unexpected end of file")
        a = np.frombuffer(self.b, dtype=dt, count=count, offset=self.o)
        self.o += n
        return a

    def val(self, dtype):
        return self.take(dtype)[0]

    def string(self):
        n = int(self.val("<u4"))
        s = self.take("u1", n).tobytes()
        return s.decode("utf-8", errors="replace")


def read_checkpoint(path: str) -> Checkpoint:
    """Read a checkpoint file written by the reference (``writeCheckpointToPath``) or by
    ``write_checkpoint``. Raises ValueError on a wrong magic, an unknown version or a short file."""
    with open(path, "rb") as fh:
        r = _Reader(fh.read())
    if r.take("u1", 8).tobytes() != CHECKPOINT_MAGIC:
        raise ValueError("checkpoint: bad magic")
    version = int(r.val("<u4"))
    if version != CHECKPOINT_VERSION:
        raise ValueError(f"checkpoint: version {version}, expected {CHECKPOINT_VERSION}")
    time_step = int(r.val("<i4"))
    sim_t = float(r.val("<f8"))
    dt = float(r.val("<f8"))

    def point_records(count):
        heads = np.empty(count, dtype=_POINT_HEAD)
        ptr = np.zeros(count + 1, dtype=np.int64)
        fvs = []
        for i in range(count):
            heads[i] = r.take(_POINT_HEAD)[0]
            k = int(r.val("<u4"))
            fvs.append(r.take(FACE_VALUE, k).copy())
            ptr[i + 1] = ptr[i] + k
        return heads, ptr, (np.concatenate(fvs) if fvs else np.empty(0, FACE_VALUE))

    fluids = []
    for _ in range(int(r.val("<u4"))):
        name = r.string()
        n = int(r.val("<u4"))
        heads, fv_ptr, fv = point_records(n)
        F = int(r.val("<u4"))
        tri = np.empty((F, 3), dtype=np.int64)
        ev_ptr = np.zeros(F + 1, dtype=np.int64)
        evs = []
        for f in range(F):
            tri[f] = r.take("<u4", 3)
            k = int(r.val("<u4"))
            evs.append(r.take(FACE_EVENT, k).copy())
            ev_ptr[f + 1] = ev_ptr[f] + k
        lines = None
        if int(r.val("u1")):
            nl = int(r.val("<u4"))
            lh = np.empty(nl, dtype=_LINE_HEAD)
            lptr = np.zeros(nl + 1, dtype=np.int64)
            lfv = []
            for i in range(nl):
                lh[i] = r.take(_LINE_HEAD)[0]
                k = int(r.val("<u4"))
                lfv.append(r.take(FACE_VALUE, k).copy())
                lptr[i + 1] = lptr[i] + k
            lines = {"ends": lh["ends"].astype(np.int64), "X_mid": lh["X_mid"].copy(), "phiphin": lh["phiphin"].copy(),
                     "phiphin_t": lh["phiphin_t"].copy(), "corner_offset": lh["corner_offset"].copy(), "fv_ptr": lptr,
                     "face_values": np.concatenate(lfv) if lfv else np.empty(0, FACE_VALUE)}
        if F and (tri.max() >= n):
            raise ValueError("checkpoint: face refers to a point id past the point count")
        fluids.append(CheckpointFluid(name, heads["X"].copy(), heads["phiphin"].copy(), heads["phiphin_t"].copy(),
                                      heads["phi_dirichlet"].copy(), fv_ptr, fv, tri, ev_ptr,
                                      np.concatenate(evs) if evs else np.empty(0, FACE_EVENT), lines))
    rigid = []
    for _ in range(int(r.val("<u4"))):
        name = r.string()
        com, q = r.take("<f8", 3).copy(), r.take("<f8", 4).copy()
        vel, acc = r.take("<f8", 6).copy(), r.take("<f8", 6).copy()
        n = int(r.val("<u4"))
        rigid.append({"name": name, "COM": com, "Q": q, "velocity": vel, "acceleration": acc,
                      "X": r.take("<f8", 3 * n).reshape(n, 3).copy()})
    soft = []
    for _ in range(int(r.val("<u4"))):
        name = r.string()
        com, q, vel = r.take("<f8", 3).copy(), r.take("<f8", 4).copy(), r.take("<f8", 6).copy()
        n = int(r.val("<u4"))
        soft.append({"name": name, "COM": com, "Q": q, "velocity": vel, "X": r.take("<f8", 3 * n).reshape(n, 3).copy()})
    npart = int(r.val("<u4"))
    P = r.take("<f8", 8 * npart).reshape(npart, 8)
    particles = {"x": P[:, 0:3].copy(), "alpha": P[:, 3:6].copy(), "sigma": P[:, 6].copy(), "volume": P[:, 7].copy()}
    if r.o != len(r.b):
        raise ValueError(f"checkpoint: {len(r.b) - r.o} trailing bytes")
    return Checkpoint(time_step, sim_t, dt, fluids, rigid, soft, particles)


def write_checkpoint(path: str, ck: Checkpoint):
    """Write ``ck`` in the same format (field order of ``writeCheckpointToPath``)."""
    out = [CHECKPOINT_MAGIC, np.array([CHECKPOINT_VERSION], "<u4").tobytes(), np.array([ck.time_step], "<i4").tobytes(),
           np.array([ck.simulation_time, ck.dt], "<f8").tobytes()]

    def u32(v):
        out.append(np.array([v], "<u4").tobytes())

    def string(s):
        b = s.encode("utf-8")
        u32(len(b))
        out.append(b)

    def f8(a, n):
        a = np.asarray(a, dtype="<f8").reshape(-1)
        if a.size != n:
            raise ValueError("checkpoint: wrong record size")
        out.append(a.tobytes())

    u32(len(ck.fluids))
    for fl in ck.fluids:
        string(fl.name)
        n = len(fl.X)
        u32(n)
        head = np.empty(1, dtype=_POINT_HEAD)
        for i in range(n):
            head["X"], head["phiphin"], head["phiphin_t"], head["phi_dirichlet"] = (fl.X[i], fl.phiphin[i], fl.phiphin_t[i],
                                                                                   fl.phi_dirichlet[i])
            out.append(head.tobytes())
            rec = fl.face_values[fl.fv_ptr[i]:fl.fv_ptr[i + 1]].astype(FACE_VALUE)
            u32(len(rec))
            out.append(rec.tobytes())
        u32(len(fl.tri))
        for f in range(len(fl.tri)):
            out.append(np.asarray(fl.tri[f], dtype="<u4").tobytes())
            ev = fl.events[fl.ev_ptr[f]:fl.ev_ptr[f + 1]].astype(FACE_EVENT)
            u32(len(ev))
            out.append(ev.tobytes())
        out.append(np.array([1 if fl.lines is not None else 0], "u1").tobytes())
        if fl.lines is not None:
            L = fl.lines
            nl = len(L["ends"])
            u32(nl)
            lh = np.empty(1, dtype=_LINE_HEAD)
            for i in range(nl):
                lh["ends"], lh["X_mid"], lh["phiphin"] = L["ends"][i], L["X_mid"][i], L["phiphin"][i]
                lh["phiphin_t"], lh["corner_offset"] = L["phiphin_t"][i], L["corner_offset"][i]
                out.append(lh.tobytes())
                rec = L["face_values"][L["fv_ptr"][i]:L["fv_ptr"][i + 1]].astype(FACE_VALUE)
                u32(len(rec))
                out.append(rec.tobytes())
    u32(len(ck.rigid))
    for b in ck.rigid:
        string(b["name"])
        f8(b["COM"], 3), f8(b["Q"], 4), f8(b["velocity"], 6), f8(b["acceleration"], 6)
        X = np.asarray(b["X"], dtype="<f8").reshape(-1, 3)
        u32(len(X))
        out.append(X.tobytes())
    u32(len(ck.soft))
    for b in ck.soft:
        string(b["name"])
        f8(b["COM"], 3), f8(b["Q"], 4), f8(b["velocity"], 6)
        X = np.asarray(b["X"], dtype="<f8").reshape(-1, 3)
        u32(len(X))
        out.append(X.tobytes())
    P = ck.particles
    npart = len(P["x"]) if P else 0
    u32(npart)
    if npart:
        rec = np.concatenate([np.asarray(P["x"]).reshape(-1, 3), np.asarray(P["alpha"]).reshape(-1, 3),
                              np.asarray(P["sigma"]).reshape(-1, 1), np.asarray(P["volume"]).reshape(-1, 1)], axis=1)
        out.append(np.ascontiguousarray(rec, dtype="<f8").tobytes())
    with open(path, "wb") as fh:
        fh.write(b"".join(out))


def checkpoint_mesh(fl: CheckpointFluid, solids=(), face_dirichlet=None, contact_factor: float = 0.5) -> Mesh:
    """The fluid surface of a checkpoint as a Mesh (point and face ids are the checkpoint's). Boundary
    types come from ``face_dirichlet`` when given, else from contact with ``solids`` as in
    ``fluid_mesh`` (the file stores positions and values, not the boundary types, which the reference
    derives again from contact at every step). The orientation is made outward (face ids unchanged)."""
    tri = np.asarray(fl.tri, dtype=np.int32)
    X = np.ascontiguousarray(fl.X, dtype=np.float64)
    if face_dirichlet is None:
        return fluid_mesh(X, tri, list(solids), contact_factor)
    if signed_volume(X, tri) < 0:
        tri = tri[:, [0, 2, 1]]
    fd = np.asarray(face_dirichlet, dtype=bool)
    if fd.shape != (len(tri),):
        raise ValueError("face_dirichlet needs one entry per face")
    return Mesh(X, np.ascontiguousarray(tri), fd, np.where(fd, 0, 1).astype(np.int32))


def _face_record(fl: CheckpointFluid, p: int, f: int):
    rec = fl.face_values[fl.fv_ptr[p]:fl.fv_ptr[p + 1]]
    hit = np.nonzero(rec["face"] == f)[0]
    return rec[hit[0]] if len(hit) else None


def checkpoint_knowns(fl: CheckpointFluid, dofs) -> tuple:
    """Per-DOF values of a DofModel built on ``checkpoint_mesh(fl, ...)``: (phi, phin, x0).

    phi / phin are the known boundary values in the ABI's convention (phi on Dirichlet DOFs, phi_n on
    Neumann DOFs, phi_D = phi_Dirichlet and phi_n on CORNER DOFs); x0 holds the stored unknowns (phi_n
    of Dirichlet DOFs, phi of the others) as the initial guess of solveSystemGMRES. A DOF keyed by a
    face takes that face's record of its point when there is one (phiOnFace / phinOnFace), else the
    point's phiphin."""
    n = dofs.n_dof
    phi, phin = np.zeros(n), np.zeros(n)
    for d in range(n):
        p, f = int(dofs.dof_point[d]), int(dofs.dof_face[d])
        rec = _face_record(fl, p, f) if (f >= 0 or fl.fv_ptr[p + 1] > fl.fv_ptr[p]) else None
        if rec is not None:
            phi[d], phin[d] = rec["phi"], rec["phin"]
        else:
            phi[d], phin[d] = fl.phiphin[p]
    kind = np.asarray(dofs.dof_kind)
    x0 = np.where(kind == 0, phin, phi)
    known_phi = np.where(kind == 2, fl.phi_dirichlet[np.asarray(dofs.dof_point)], phi)
    return known_phi, phin, x0


def checkpoint_from_bvp(mesh: Mesh, dofs, phi: np.ndarray, phin: np.ndarray, name: str = "water", time_step: int = 0,
                        simulation_time: float = 0.0, dt: float = 0.0, phi_dirichlet=None) -> Checkpoint:
    """A checkpoint of one fluid surface from per-DOF values (the inverse of ``checkpoint_knowns``):
    a point's phiphin is its null-key DOF's values (or its first DOF's), and every face-keyed DOF of a
    multiple node becomes a face record of that point."""
    n = mesh.n_nodes
    pp = np.zeros((n, 2))
    seen = np.zeros(n, bool)
    recs = [[] for _ in range(n)]
    for d in range(dofs.n_dof):
        p, f = int(dofs.dof_point[d]), int(dofs.dof_face[d])
        if f < 0 or not seen[p]:
            pp[p] = (phi[d], phin[d])
            seen[p] |= f < 0
        if bool(dofs.point_multiple[p]):
            recs[p].append((f, phi[d], phin[d], 0.0, 0.0))
    fv_ptr = np.zeros(n + 1, dtype=np.int64)
    fv_ptr[1:] = np.cumsum([len(r) for r in recs])
    fv = np.array([r for rr in recs for r in rr], dtype=FACE_VALUE) if fv_ptr[-1] else np.empty(0, FACE_VALUE)
    if phi_dirichlet is None:  # phi_D of CORNER points: the value their CORNER DOFs carry in phi
        pd = np.zeros(n)
        corner = np.asarray(dofs.dof_kind) == 2
        pd[np.asarray(dofs.dof_point)[corner]] = np.asarray(phi)[corner]
    else:
        pd = np.asarray(phi_dirichlet, dtype=np.float64)
    F = mesh.n_faces
    fl = CheckpointFluid(name, mesh.X.copy(), pp, np.zeros((n, 2)), pd, fv_ptr, fv, mesh.tri.astype(np.int64),
                         np.zeros(F + 1, dtype=np.int64), np.empty(0, FACE_EVENT))
    return Checkpoint(time_step, simulation_time, dt, [fl], [], [],
                      {"x": np.empty((0, 3)), "alpha": np.empty((0, 3)), "sigma": np.empty(0), "volume": np.empty(0)})
