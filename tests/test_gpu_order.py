"""Internal Morton DOF order (solveGMRES's reindex, BEM_solveBVP_GMRES_FMM.hpp:3164-3246) on the GPU.

* The order the library computes (GPU keys + radix sort) equals the reference rule restated in numpy
  (`mesh.morton_keys`: 21-bit quantisation in the point bounding box, bit interleave), stably sorted so
  that equal keys keep the caller's order.
* Results do not depend on the caller's DOF numbering: a randomly renumbered problem gives the same
  matvec, RHS, alpha and solution (mapped back) to 1e-12 relative, with and without the reorder, and
  with FP32+refinement + Schwarz; the ABI's arrays are always in the caller's numbering.
* Out-of-range indices are rejected before any kernel runs.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _renumber(d, phi, phin, seed):
    """The same DOF model with DOF i renamed sigma[i]."""
    from dataclasses import replace
    n = d.n_dof
    sigma = np.random.default_rng(seed).permutation(n).astype(np.int32)
    kind = np.empty(n, np.int32)
    point = np.empty(n, np.int32)
    face = np.empty(n, np.int32)
    kind[sigma], point[sigma], face[sigma] = d.dof_kind, d.dof_point, d.dof_face
    ph, phn = np.empty(n), np.empty(n)
    ph[sigma], phn[sigma] = phi, phin
    d2 = replace(d, face_dof=sigma[d.face_dof].astype(np.int32), dof_kind=kind, dof_point=point, dof_face=face)
    return d2, ph, phn, sigma


def _bvp(m, d, phi, phin, **kw):
    from bem_for_nonlinear_waves_amd.bvp import BEM_BVP
    print(f"[gpu] BEM_BVP n_dof={d.n_dof} opts={kw}", flush=True)
    b = BEM_BVP(m, d, **kw)
    b.set_bc(phi, phin)
    b.prepare()
    return b


def test_order_is_the_morton_rule(small_tank):
    from bem_for_nonlinear_waves_amd import mesh as M
    m, d, phi, phin = small_tank
    d2, ph2, phn2, _ = _renumber(d, phi, phin, 3)
    g = _bvp(m, d2, ph2, phn2, p=8, leaf_max=32)
    order = g.dof_order()
    keys = M.morton_keys(m.X)[d2.dof_point]
    expect = np.argsort(keys, kind="stable")
    assert np.array_equal(order, expect)
    g.close()
    g = _bvp(m, d2, ph2, phn2, p=8, leaf_max=32, dof_reorder=0)
    assert np.array_equal(g.dof_order(), np.arange(d.n_dof))
    g.close()


@pytest.mark.parametrize("reorder", [1, 0])
def test_results_independent_of_numbering(small_tank, reorder):
    m, d, phi, phin = small_tank
    d2, ph2, phn2, sigma = _renumber(d, phi, phin, 4)
    a = _bvp(m, d, phi, phin, p=8, leaf_max=32)
    b = _bvp(m, d2, ph2, phn2, p=8, leaf_max=32, dof_reorder=reorder)
    x = np.random.default_rng(5).standard_normal(d.n_dof)
    x2 = np.empty_like(x)
    x2[sigma] = x
    ya, yb = a.compute_Ax(x), b.compute_Ax(x2)
    assert np.linalg.norm(yb[sigma] - ya) <= 1e-12 * np.linalg.norm(ya)
    ra, rb = a.rhs(), b.rhs()
    assert np.linalg.norm(rb[sigma] - ra) <= 1e-12 * np.linalg.norm(ra)
    aa, ab = a.diag_coeffs(), b.diag_coeffs()
    assert np.abs(ab[sigma] - aa).max() <= 1e-12 * np.abs(aa).max()
    sa, _ = a.solveSystemGMRES(tol=1e-11, max_iter=600)
    sb, _ = b.solveSystemGMRES(tol=1e-11, max_iter=600)
    assert np.linalg.norm(sb[sigma] - sa) <= 1e-9 * np.linalg.norm(sa)
    a.close()
    b.close()


def test_numbering_with_refinement_and_schwarz(small_tank):
    from bem_for_nonlinear_waves_amd import _lib
    m, d, phi, phin = small_tank
    d2, ph2, phn2, sigma = _renumber(d, phi, phin, 6)
    kw = dict(p=10, leaf_max=32, precision=_lib.PRECISION_FP32_REFINE, preconditioner=_lib.PRECOND_SCHWARZ)
    a = _bvp(m, d, phi, phin, **kw)
    b = _bvp(m, d2, ph2, phn2, **kw)
    sa, ia = a.solveSystemGMRES(tol=1e-11, max_iter=600)
    sb, ib = b.solveSystemGMRES(tol=1e-11, max_iter=600)
    assert ia["converged"] == 1 and ib["converged"] == 1
    assert np.linalg.norm(sb[sigma] - sa) <= 1e-8 * np.linalg.norm(sa)
    # Schwarz blocks come back in the caller's numbering and cover every DOF once
    ptr, idx, ncore = b.schwarz_blocks()
    cores = np.concatenate([idx[ptr[k]:ptr[k] + ncore[k]] for k in range(len(ncore))])
    assert np.array_equal(np.sort(cores), np.arange(d.n_dof))
    # device matvec entry point with a non-identity order (buffers from the HIP runtime the library runs on)
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so.7")
    nb = d.n_dof * 8
    px, py = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(px), ctypes.c_size_t(nb)) == 0 and hip.hipMalloc(ctypes.byref(py), ctypes.c_size_t(nb)) == 0
    x = np.random.default_rng(8).standard_normal(d.n_dof)
    y = np.empty(d.n_dof)
    assert hip.hipMemcpy(px, x.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(nb), 1) == 0  # host to device
    b.matvec_device(px.value, py.value)
    b.sync()
    assert hip.hipMemcpy(y.ctypes.data_as(ctypes.c_void_p), py, ctypes.c_size_t(nb), 2) == 0  # device to host
    hip.hipFree(px)
    hip.hipFree(py)
    assert np.linalg.norm(y - b.compute_Ax(x)) <= 1e-13 * np.linalg.norm(y)
    a.close()
    b.close()


def test_bad_indices_rejected(tiny_tank):
    from dataclasses import replace

    from bem_for_nonlinear_waves_amd import _lib
    from bem_for_nonlinear_waves_amd.bvp import BEM_BVP, BEMError
    m, d, phi, phin = tiny_tank
    fd = d.face_dof.copy()
    fd[3, 1] = d.n_dof + 5
    for bad in (replace(d, face_dof=fd), replace(d, dof_point=np.where(np.arange(d.n_dof) == 7, m.n_nodes, d.dof_point))):
        b = BEM_BVP(m, bad, p=8, leaf_max=16)
        b.set_bc(phi, phin)
        with pytest.raises(BEMError) as e:
            b.prepare()
        assert e.value.code == _lib.E_INVALID
        b.close()
