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
