"""The product's host plan (libbemfmm's plan.cpp/operators.cpp) executed by a host emulator of the
device pipeline, against the oracle. Pins the tree/list construction, the float128 translation
operators in real-field packed and level-scaled form, the P2M/L2P matrices and the near quadrature
without a GPU (the GPU tests then pin the kernels)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from bem_for_nonlinear_waves_amd import mesh as M

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "emu", "_build", "libplanemu.so")
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


@pytest.fixture(scope="module")
def emu():
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "emu")], check=True)
    L = ctypes.CDLL(LIB)
    L.emu_create.restype = ctypes.c_void_p
    L.emu_create.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _ip, _ip, ctypes.c_int, _ip, _ip, _dp, _dp, ctypes.c_int,
                             ctypes.c_double, ctypes.c_int, ctypes.c_int]
    L.emu_destroy.argtypes = [ctypes.c_void_p]
    for f in ("emu_matvec",):
        getattr(L, f).argtypes = [ctypes.c_void_p, _dp, _dp]
    L.emu_rhs.argtypes = [ctypes.c_void_p, _dp]
    L.emu_alpha.argtypes = [ctypes.c_void_p, _dp]
    L.emu_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]
    L.emu_operator.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
    L.emu_rotation_size.argtypes = [ctypes.c_int]
    L.emu_rotation_size.restype = ctypes.c_int
    return L


def _p(a, t):
    return a.ctypes.data_as(t)


def _create(L, m, d, phi, phin, p, leaf, ml=7):
    keep = [np.ascontiguousarray(m.X), np.ascontiguousarray(m.tri, np.int32), np.ascontiguousarray(d.face_dof, np.int32),
            np.ascontiguousarray(d.dof_kind, np.int32), np.ascontiguousarray(d.dof_point, np.int32),
            np.ascontiguousarray(phi), np.ascontiguousarray(phin)]
    X, tri, fd, kd, pt, ph, phn = keep
    h = L.emu_create(m.n_nodes, _p(X, _dp), m.n_faces, _p(tri, _ip), _p(fd, _ip), d.n_dof, _p(kd, _ip), _p(pt, _ip),
                     _p(ph, _dp), _p(phn, _dp), p, 0.25, leaf, ml)
    return h, keep


@pytest.mark.parametrize("case,p,leaf", [("box", 8, 32), ("box", 10, 32), ("cyl", 10, 32)])
def test_emulated_plan_matches_oracle(emu, oracle_mod, case, p, leaf):
    if case == "box":
        m = M.perturb_free_surface(M.box_tank(4.0, 1.0, 1.0, 42, 11, 11), 1, 0.01)
        d = M.build_dofs(m)
        phi, phin = M.tank_bc(m, d, 1, 4.0, 1.0)
    else:
        m = M.perturb_free_surface(M.cylinder_tank(h=0.2), 3, 0.002)
        d = M.build_dofs(m)
        phi, phin = M.tank_bc(m, d, 3, 20.0, 4.0, body_velocity_sigma=0.1)
    h, keep = _create(emu, m, d, phi, phin, p, leaf)
    o = oracle_mod.Oracle(m, d, phi, phin, p=p, leaf_max=leaf)
    st = np.zeros(6, np.int64)
    emu.emu_stats(h, st.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    so = o.stats()
    assert (st[1], st[2], st[3], st[4]) == (so["n_buckets"], so["n_leaves"], so["n_m2l_pairs"], so["near_nnz"])
    n = d.n_dof
    a = np.empty(n)
    emu.emu_alpha(h, _p(a, _dp))
    assert np.abs(a - o.alpha()).max() <= 1e-11 * 2 * np.pi
    b = np.empty(n)
    emu.emu_rhs(h, _p(b, _dp))
    bo = o.rhs()
    assert np.linalg.norm(b - bo) <= 1e-11 * np.linalg.norm(bo)
    x = np.random.default_rng(2).standard_normal(n)
    y = np.empty(n)
    emu.emu_matvec(h, _p(x, _dp), _p(y, _dp))
    yo = o.matvec(x)
    assert np.linalg.norm(y - yo) <= 1e-11 * np.linalg.norm(yo)
    emu.emu_destroy(h)


def test_m2l_operator_reflection_symmetry(emu):
    """Mirror x -> -x maps the source/target offset (dx,dy,dz) -> (-dx,dy,dz); in real-field packing this is
    a diagonal sign change (-1)^m on Re and -(-1)^m on Im of both input and output (phi -> pi - phi)."""
    P = 6
    C = (P + 1) ** 2
    A = np.empty(C * C)
    B = np.empty(C * C)
    emu.emu_operator(P, 2, 3, 1, -2, _p(A, _dp))
    emu.emu_operator(P, 2, -3, 1, -2, _p(B, _dp))
    A = A.reshape(C, C).T  # [out, in]
    B = B.reshape(C, C).T
    s = np.empty(C)
    for n in range(P + 1):
        s[n * n] = 1.0
        for m in range(1, n + 1):
            s[n * n + 2 * m - 1] = (-1) ** m
            s[n * n + 2 * m] = -((-1) ** m)
    assert np.abs(np.diag(s) @ A @ np.diag(s) - B).max() <= 1e-13 * np.abs(A).max()


def _rotation_blocks(emu, P, o):
    """(D, T, H^T) as dense (P+1)^2 matrices [out, in] from the compact factors of M2L offset o:
    D = Y_D Z, H^T = Z^T Y_H^T with Z the per-order z-rotation, Y the cos/sin-type blocks of each degree."""
    C = (P + 1) ** 2
    f = np.empty(emu.emu_rotation_size(P))
    emu.emu_operator(P, 3, o[0], o[1], o[2], _p(f, _dp))

    def col(n, m, part):  # packed index of (n, m): real part, or imaginary part for m > 0
        return n * n if m == 0 else n * n + 2 * m - 1 + part
    k = 0
    Z = np.eye(C)
    for m in range(1, P + 1):
        c, s = f[k], f[k + 1]
        k += 2
        for n in range(m, P + 1):
            re, im = col(n, m, 0), col(n, m, 1)
            Z[re, re], Z[re, im], Z[im, re], Z[im, im] = c, -s, s, c

    def degree_blocks(k):
        Y = np.zeros((C, C))
        for n in range(P + 1):
            cos_idx = [col(n, m, 0) for m in range(n + 1)]
            sin_idx = [col(n, m, 1) for m in range(1, n + 1)]
            for idx in (cos_idx, sin_idx):
                w = len(idx)
                Y[np.ix_(idx, idx)] = f[k:k + w * w].reshape(w, w)
                k += w * w
        return Y, k
    YD, k = degree_blocks(k)
    T = np.zeros((C, C))
    for m in range(P + 1):
        w = P + 1 - m
        blk = f[k:k + w * w].reshape(w, w)
        k += w * w
        for part in ((0,) if m == 0 else (0, 1)):
            for j in range(m, P + 1):
                for n in range(m, P + 1):
                    T[col(j, m, part), col(n, m, part)] = blk[j - m, n - m]
    YHt, k = degree_blocks(k)
    assert k == f.size
    return YD @ Z, T, Z.T @ YHt


@pytest.mark.parametrize("P", [5, 6, 8, 10])
def test_m2l_rotation_factors(emu, P):
    """The rotation M2L factors (used by the FP32 operator's M2L kernel) reproduce the dense float128-built
    M2L operator: T(o) = H^T T(|o| z) D for offsets in every octant, on the axes and near the x axis (the
    other branch of the rotation frame). On the +z axis the rotations are the identity."""
    C = (P + 1) ** 2
    offsets = [(2, 3, -1), (4, 0, 0), (-4, 0, 0), (0, 0, 5), (0, 0, -4), (0, 3, 0), (-3, 2, 2), (1, -6, 3),
               (2, 2, 2), (-7, -7, -7), (6, 1, -1), (-2, -5, 4), (3, -3, 0)]
    for o in offsets:
        A = np.empty(C * C)
        emu.emu_operator(P, 2, o[0], o[1], o[2], _p(A, _dp))
        A = A.reshape(C, C).T  # [out, in]
        D, T, Ht = _rotation_blocks(emu, P, o)
        err = np.abs(Ht @ T @ D - A).max() / np.abs(A).max()
        assert err <= 1e-12, (o, err)
    D, T, Ht = _rotation_blocks(emu, P, (0, 0, 3))
    assert np.abs(D - np.eye(C)).max() <= 1e-12 and np.abs(Ht - np.eye(C)).max() <= 1e-12
