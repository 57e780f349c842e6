"""Restricted additive Schwarz preconditioner: oracle pins, blocking invariants, and the
GPU kernel's algorithm (entry lookup in the block-dense near field + in-place Gauss-Jordan
with partial pivoting) mirrored on the CPU against the oracle's block LU.

Pins of the oracle's Schwarz apply (it restates SchwarzAdditive_CRS,
basic_linear_systems.hpp:2870-3003, on the near-field operator of
BEM_solveBVP_GMRES_FMM.hpp:440-640):
  * blocks of one DOF reduce it to the Jacobi scaling by the operator diagonal;
  * on a mesh whose tree is a single box there is no far field, so one block holding every
    DOF is the whole operator and M^-1 A = I (to rounding), and GMRES converges at once.
"""
import ctypes

import numpy as np
import pytest

from test_plan_emulator import _create, _dp, _ip, _p, emu  # noqa: F401 (emu is a fixture)


def _blocks(emu, d, seed, max_core, max_block, overlap):
    emu.emu_schwarz_create.restype = ctypes.c_void_p
    emu.emu_schwarz_create.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _ip, _ip, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int]
    emu.emu_schwarz_sizes.restype = ctypes.c_int
    emu.emu_schwarz_sizes.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]
    emu.emu_schwarz_get.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), _ip, _ip]
    emu.emu_schwarz_destroy.argtypes = [ctypes.c_void_p]
    fd = np.ascontiguousarray(d.face_dof, np.int32)
    pt = np.ascontiguousarray(d.dof_point, np.int32)
    sd = np.ascontiguousarray(seed, np.int32)
    h = emu.emu_schwarz_create(d.n_dof, len(fd), _p(fd, _ip), _p(pt, _ip), _p(sd, _ip), max_core, max_block, overlap)
    nidx = ctypes.c_int64(0)
    nb = emu.emu_schwarz_sizes(h, ctypes.byref(nidx))
    ptr = np.zeros(nb + 1, np.int64)
    idx = np.zeros(nidx.value, np.int32)
    ncore = np.zeros(nb, np.int32)
    emu.emu_schwarz_get(h, ptr.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), _p(idx, _ip), _p(ncore, _ip))
    return h, ptr, idx, ncore


def _dof_graph(d):
    adj = [set() for _ in range(d.n_dof)]
    for f in d.face_dof.reshape(-1, 3):
        for a in f:
            adj[a].update(int(b) for b in f if b != a)
    by_pt = {}
    for i, p in enumerate(d.dof_point):
        by_pt.setdefault(int(p), []).append(i)
    for ds in by_pt.values():
        for a in ds:
            adj[a].update(b for b in ds if b != a)
    return adj


def test_blocking_invariants(emu, small_tank):
    m, d, phi, phin = small_tank
    seed = np.arange(d.n_dof)
    h, ptr, idx, ncore = _blocks(emu, d, seed, 64, 128, 1)
    emu.emu_schwarz_destroy(h)
    nb = len(ncore)
    assert nb >= d.n_dof // 64
    cores = np.concatenate([idx[ptr[b]:ptr[b] + ncore[b]] for b in range(nb)])
    assert np.array_equal(np.sort(cores), np.arange(d.n_dof))  # every DOF in exactly one core
    sizes = np.diff(ptr)
    assert sizes.max() <= 128 and ncore.max() <= 64 and np.all(ncore >= 1) and np.all(sizes >= ncore)
    adj = _dof_graph(d)
    for b in range(0, nb, max(1, nb // 25)):
        blk = idx[ptr[b]:ptr[b + 1]]
        assert len(set(blk.tolist())) == len(blk)
        core = set(blk[:ncore[b]].tolist())
        ring = set().union(*(adj[u] for u in core)) - core
        assert set(blk[ncore[b]:].tolist()) <= ring  # overlap = one ring of neighbours of the core
        if sizes[b] < 128:
            assert set(blk[ncore[b]:].tolist()) == ring
    # determinism
    h2, ptr2, idx2, ncore2 = _blocks(emu, d, seed, 64, 128, 1)
    emu.emu_schwarz_destroy(h2)
    assert np.array_equal(ptr, ptr2) and np.array_equal(idx, idx2) and np.array_equal(ncore, ncore2)


def test_oracle_schwarz_singletons_is_jacobi(oracle_mod, tiny_tank):
    m, d, phi, phin = tiny_tank
    o = oracle_mod.Oracle(m, d, phi, phin, p=8, leaf_max=16)
    n = d.n_dof
    o.set_schwarz(np.arange(n + 1), np.arange(n), np.ones(n))
    r = np.random.default_rng(4).standard_normal(n)
    z = o.schwarz_apply(r)
    assert np.allclose(z, o.diag_inv() * r, rtol=1e-13, atol=0)


def test_oracle_schwarz_single_block_inverts_near_only_operator(oracle_mod, tiny_tank):
    m, d, phi, phin = tiny_tank
    o = oracle_mod.Oracle(m, d, phi, phin, p=8, leaf_max=100000)
    assert o.stats()["n_m2l_pairs"] == 0  # one box: the whole operator is near field
    n = d.n_dof
    o.set_schwarz(np.array([0, n]), np.arange(n), np.array([n]))
    x = np.random.default_rng(5).standard_normal(n)
    z = o.schwarz_apply(o.matvec(x))
    assert np.linalg.norm(z - x) <= 1e-10 * np.linalg.norm(x)
    xs, info = o.solve(tol=1e-10, max_iter=20, precond=2)
    assert info["rc"] == 0 and info["iters"] <= 3


def test_kernel_algorithm_mirror_matches_oracle(emu, oracle_mod, small_tank):
    """Serial mirror of k_schwarz_build/k_schwarz_apply (emulator) vs the oracle's block LU."""
    m, d, phi, phin = small_tank
    emu.emu_schwarz_apply.argtypes = [ctypes.c_void_p, ctypes.c_void_p, _dp, _dp]
    he, keep = _create(emu, m, d, phi, phin, 8, 32)
    hs, ptr, idx, ncore = _blocks(emu, d, np.arange(d.n_dof), 64, 128, 1)
    o = oracle_mod.Oracle(m, d, phi, phin, p=8, leaf_max=32)
    o.set_schwarz(ptr, idx, ncore)
    r = np.random.default_rng(6).standard_normal(d.n_dof)
    z_e = np.empty(d.n_dof)
    emu.emu_schwarz_apply(he, hs, _p(r, _dp), _p(z_e, _dp))
    z_o = o.schwarz_apply(r)
    assert np.linalg.norm(z_e - z_o) <= 1e-10 * np.linalg.norm(z_o)
    emu.emu_schwarz_destroy(hs)
    emu.emu_destroy(he)


def test_oracle_schwarz_fewer_iterations_than_jacobi(emu, oracle_mod, tiny_tank):
    m, d, phi, phin = tiny_tank
    o = oracle_mod.Oracle(m, d, phi, phin, p=8, leaf_max=16)
    hs, ptr, idx, ncore = _blocks(emu, d, np.arange(d.n_dof), 64, 128, 1)
    emu.emu_schwarz_destroy(hs)
    o.set_schwarz(ptr, idx, ncore)
    xj, ij = o.solve(tol=1e-10, max_iter=600, precond=1)
    xs, isw = o.solve(tol=1e-10, max_iter=600, precond=2)
    assert ij["rc"] == 0 and isw["rc"] == 0
    assert isw["iters"] < ij["iters"]
    A, b = o.dense(), o.rhs()
    L = o.stats()["L"]
    x_lu = np.linalg.solve(A, b)
    x_lu = np.where(d.dof_kind == 0, x_lu / L, x_lu)
    assert np.linalg.norm(xs - x_lu) / np.linalg.norm(x_lu) <= 1e-7
