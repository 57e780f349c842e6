"""ctypes binding of the CPU oracle (test infrastructure only; never imported by the product).

The oracle is an FP64 CPU restatement of the reference's FMM-GMRES BVP path
(see oracle.cpp for the per-stage citations). ``build()`` compiles it with the
Makefile next to this file.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class OracleStats(ctypes.Structure):
    _fields_ = [("n_levels", ctypes.c_int), ("n_buckets", ctypes.c_int), ("n_leaves", ctypes.c_int),
                ("n_m2l_pairs", ctypes.c_longlong), ("near_nnz", ctypes.c_longlong), ("L", ctypes.c_double),
                ("scale", ctypes.c_double)]


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB_PATH) or os.path.getmtime(_LIB_PATH) < os.path.getmtime(
            os.path.join(_HERE, "oracle.cpp")):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def use_native() -> str:
    """Build (on this machine) and select the -O3 -march=native variant of the library, as the timed CPU
    baseline is built; call before the first lib()."""
    global _LIB_PATH
    if _lib is not None:
        raise RuntimeError("oracle library already loaded")
    import hashlib
    import platform
    cpu = platform.node()
    try:
        with open("/proc/cpuinfo") as f:
            cpu += "".join(l for l in f if l.startswith(("model name", "flags")))
    except OSError:
        pass
    path = os.path.join(_HERE, "_build", f"liboracle_native_{hashlib.sha256(cpu.encode()).hexdigest()[:12]}.so")
    subprocess.run(["make", "-s", "-C", _HERE, "native", f"NATIVE={path}"], check=True)
    _LIB_PATH = path
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_create.restype = ctypes.c_void_p
        L.oracle_create.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _ip, _ip, ctypes.c_int, _ip, _ip, _dp, _dp,
                                    ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.oracle_tri6_shape.argtypes = [ctypes.c_double, ctypes.c_double, _dp, _dp, _dp]
        L.oracle_create_npf.restype = ctypes.c_void_p
        L.oracle_create_npf.argtypes = L.oracle_create.argtypes + [ctypes.c_int]
        L.oracle_destroy.argtypes = [ctypes.c_void_p]
        L.oracle_set_robin.argtypes = [ctypes.c_void_p, _dp]
        L.oracle_matvec.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.oracle_rhs.argtypes = [ctypes.c_void_p, _dp]
        L.oracle_alpha.argtypes = [ctypes.c_void_p, _dp]
        L.oracle_dense.argtypes = [ctypes.c_void_p, _dp]
        L.oracle_solve.restype = ctypes.c_int
        L.oracle_solve.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_int, _ip, _dp,
                                   ctypes.c_int]
        L.oracle_get_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(OracleStats)]
        L.oracle_near_row.restype = ctypes.c_int
        L.oracle_near_row.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, _dp, _dp, ctypes.c_int]
        L.oracle_set_m2l_cache.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.oracle_set_schwarz.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int64), _ip, _ip]
        L.oracle_schwarz_apply.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.oracle_diag_inv.argtypes = [ctypes.c_void_p, _dp]
        L.oracle_ilu_build.restype = ctypes.c_longlong
        L.oracle_ilu_build.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.oracle_ilu_rows.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong), _ip, _dp]
        L.oracle_ilu_factored.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong), _ip, _dp]
        L.oracle_ilu_apply.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.oracle_ilu_face_rule.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.oracle_ilu_kind.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                      ctypes.c_double]
        L.oracle_ilu_factored_nnz.restype = ctypes.c_longlong
        L.oracle_ilu_factored_nnz.argtypes = [ctypes.c_void_p]
        L.oracle_schwarz_ref_build.restype = ctypes.c_int
        L.oracle_schwarz_ref_build.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.oracle_schwarz_ref_blocks.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong), _ip, _ip]
        L.oracle_schwarz_ref_apply.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.oracle_sph_harmonics.argtypes = [_dp, ctypes.c_int, _dp, _dp]
        L.oracle_dR.argtypes = [_dp, _dp, ctypes.c_int, _dp, _dp]
        L.oracle_translate_chain.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int, _dp]
        _lib = L
    return _lib


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _p(a, t):
    return a.ctypes.data_as(t)


def tri6_shape(t0: float, t1: float):
    """TriShape<6> and its derivatives in t0 and t1 at (t0, t1), as the oracle evaluates them."""
    N, D0, D1 = np.zeros(6), np.zeros(6), np.zeros(6)
    lib().oracle_tri6_shape(float(t0), float(t1), _p(N, _dp), _p(D0, _dp), _p(D1, _dp))
    return N, D0, D1


class Oracle:
    """CPU FP64 FMM-GMRES problem: same inputs as the product's C ABI."""

    def __init__(self, mesh, dofs, phi, phin, p=10, theta=0.25, leaf_max=50, max_level=7, q=6):
        L = lib()
        self._keep = [_d(mesh.X), _i(mesh.tri), _i(dofs.face_dof), _i(dofs.dof_kind), _i(dofs.dof_point), _d(phi),
                      _d(phin)]
        X, tri, fd, kind, pt, ph, phn = self._keep
        self.n_dof = int(dofs.n_dof)
        npf = int(np.shape(mesh.tri)[1])  # 3 linear, 6 true quadratic faces
        if np.shape(dofs.face_dof)[1] != npf:
            raise ValueError("face_dof needs one entry per face node")
        self.h = L.oracle_create_npf(mesh.n_nodes, _p(X, _dp), mesh.n_faces, _p(tri, _ip), _p(fd, _ip), self.n_dof,
                                     _p(kind, _ip), _p(pt, _ip), _p(ph, _dp), _p(phn, _dp), p, theta, leaf_max,
                                     max_level, q, npf)
        if not self.h:
            raise RuntimeError("oracle_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_destroy(self.h)
            self.h = None

    def matvec(self, x):
        x = _d(x)
        y = np.empty(self.n_dof)
        lib().oracle_matvec(self.h, _p(x, _dp), _p(y, _dp))
        return y

    def rhs(self):
        b = np.empty(self.n_dof)
        lib().oracle_rhs(self.h, _p(b, _dp))
        return b

    def alpha(self):
        a = np.empty(self.n_dof)
        lib().oracle_alpha(self.h, _p(a, _dp))
        return a

    def dense(self):
        A = np.empty((self.n_dof, self.n_dof))
        lib().oracle_dense(self.h, _p(A, _dp))
        return A

    def solve(self, x0=None, tol=1e-10, max_iter=1000, restart=1000, precond=0):
        """GMRES as the reference's run_gmres; precond=1: left scaling by the operator diagonal;
        precond=2: restricted additive Schwarz over the blocks given to set_schwarz."""
        x0 = np.zeros(self.n_dof) if x0 is None else _d(x0)
        x = np.empty(self.n_dof)
        it = ctypes.c_int(0)
        res = ctypes.c_double(0)
        rc = lib().oracle_solve(self.h, _p(x0, _dp), _p(x, _dp), tol, max_iter, restart, ctypes.byref(it),
                                ctypes.byref(res), precond)
        return x, {"rc": rc, "iters": it.value, "residual": res.value}

    def stats(self):
        s = OracleStats()
        lib().oracle_get_stats(self.h, ctypes.byref(s))
        return {k: getattr(s, k) for k, _ in OracleStats._fields_}

    def set_robin(self, kappa):
        """Robin coefficient per DOF (physical units) for ROBIN DOFs (kind 3): phi_n = kappa phi + g."""
        self._kappa = _d(kappa)
        lib().oracle_set_robin(self.h, _p(self._kappa, _dp))

    def set_m2l_cache(self, on: bool):
        """Per-(level, lattice offset) M2L coefficients computed once (CPU-baseline mode for large N)."""
        lib().oracle_set_m2l_cache(self.h, int(bool(on)))

    def set_schwarz(self, ptr, idx, ncore):
        self._sw = [np.ascontiguousarray(ptr, np.int64), _i(idx), _i(ncore)]
        lib().oracle_set_schwarz(self.h, len(self._sw[2]), self._sw[0].ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                 _p(self._sw[1], _ip), _p(self._sw[2], _ip))

    def schwarz_apply(self, r):
        r = _d(r)
        z = np.empty(self.n_dof)
        lib().oracle_schwarz_apply(self.h, _p(r, _dp), _p(z, _dp))
        return z

    def diag_inv(self):
        d = np.empty(self.n_dof)
        lib().oracle_diag_inv(self.h, _p(d, _dp))
        return d

    def ilu_build(self, neighborhood=0, kring=1, kind=0, omega=1.0, drop_tol=1e-3, max_entries=50, pivot_min=1e-12):
        """The reference's sparse preconditioner matrix (BUCKETS = 0 / K-RING = 1) and its factorisation (kind 0
        ILU(0), 1 MILU(0), 2 ILUT); returns the matrix's nnz."""
        lib().oracle_ilu_kind(self.h, kind, omega, drop_tol, max_entries, pivot_min)
        nnz = lib().oracle_ilu_build(self.h, neighborhood, kring)
        if nnz < 0:
            raise RuntimeError("oracle ILU(0) failed")
        self._ilu_nnz = int(nnz)
        return self._ilu_nnz

    def _ilu_get(self, fn, nnz):
        ptr = np.empty(self.n_dof + 1, np.int64)
        col = np.empty(nnz, np.int32)
        val = np.empty(nnz)
        fn(self.h, ptr.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), _p(col, _ip), _p(val, _dp))
        return ptr, col, val

    def ilu_rows(self):
        """(ptr, col, val) of the unfactored sparse matrix (columns ascending per row)."""
        return self._ilu_get(lib().oracle_ilu_rows, self._ilu_nnz)

    def ilu_factored(self):
        """(ptr, col, val) of the factors: L strictly below the diagonal, U from it."""
        return self._ilu_get(lib().oracle_ilu_factored, int(lib().oracle_ilu_factored_nnz(self.h)))

    def ilu_apply(self, r):
        r = _d(r)
        z = np.empty(self.n_dof)
        lib().oracle_ilu_apply(self.h, _p(r, _dp), _p(z, _dp))
        return z

    def schwarz_ref_build(self, core_k=1, overlap_k=1, max_core=64, max_block=128):
        """The reference's SCHWARZ blocks on the ILU(0) matrix (ilu_build first); returns (ptr, idx, ncore)."""
        nb = lib().oracle_schwarz_ref_build(self.h, core_k, overlap_k, max_core, max_block)
        ptr = np.empty(nb + 1, np.int64)
        idx = np.empty(self.n_dof * max_block, np.int32)
        ncore = np.empty(nb, np.int32)
        lib().oracle_schwarz_ref_blocks(self.h, ptr.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), _p(idx, _ip),
                                        _p(ncore, _ip))
        return ptr, idx[:ptr[-1]].copy(), ncore

    def schwarz_ref_apply(self, r):
        r = _d(r)
        z = np.empty(self.n_dof)
        lib().oracle_schwarz_ref_apply(self.h, _p(r, _dp), _p(z, _dp))
        return z

    def near_row(self, t, cap=100000):
        idx = np.empty(cap, np.int32)
        wg = np.empty(cap)
        wgn = np.empty(cap)
        n = lib().oracle_near_row(self.h, t, _p(idx, _ip), _p(wg, _dp), _p(wgn, _dp), cap)
        return idx[:n].copy(), wg[:n].copy(), wgn[:n].copy()


def ilu_face_rule(P, x, singular, near):
    """Single- and double-layer weights (G[3], H[3]) of one face (corners P, 3x3, collapsed at P[0]) seen from x,
    as the reference's sparse preconditioner matrix integrates it."""
    P = _d(np.asarray(P).reshape(9))
    x = _d(x)
    G, H = np.empty(3), np.empty(3)
    lib().oracle_ilu_face_rule(_p(P, _dp), _p(x, _dp), int(singular), int(near), _p(G, _dp), _p(H, _dp))
    return G, H


def sph_harmonics(x, P):
    x = _d(x)
    re = np.empty((P + 1) ** 2)
    im = np.empty((P + 1) ** 2)
    lib().oracle_sph_harmonics(_p(x, _dp), P, _p(re, _dp), _p(im, _dp))
    return re + 1j * im


def grad_R_dot_n(x, nh, P):
    x, nh = _d(x), _d(nh)
    re = np.empty((P + 1) ** 2)
    im = np.empty((P + 1) ** 2)
    lib().oracle_dR(_p(x, _dp), _p(nh, _dp), P, _p(re, _dp), _p(im, _dp))
    return re + 1j * im


def translate_chain(ys, qs, nhs, c1, c2, c3, c4, x, P):
    ys, qs, nhs = _d(ys), _d(qs), _d(nhs)
    c1, c2, c3, c4, x = (_d(v) for v in (c1, c2, c3, c4, x))
    out = np.empty(4)
    lib().oracle_translate_chain(len(qs), _p(ys, _dp), _p(qs, _dp), _p(nhs, _dp), _p(c1, _dp), _p(c2, _dp),
                                 _p(c3, _dp), _p(c4, _dp), _p(x, _dp), P, _p(out, _dp))
    return out
