"""bemfmm_matvec_device_noalloc (the allocation-free matvec) against bemfmm_matvec_device on the GPU.

Two calls in a row on different contexts, the second smaller than the first: each result equals the
regular entry point's on the same device vectors to 1e-12 relative (both are the FP64 operator; the M2L sums
of the allocation-free call are accumulated in another order, so the last bits differ). A problem whose work
memory exceeds the per-GPU array is refused with BEMFMM_E_INVALID and a message that says so.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _tank(Lx, nx, ny, seed):
    from bem_for_nonlinear_waves_amd import mesh as M
    m = M.box_tank(Lx, 1.0, 1.0, nx, ny, ny)
    m = M.perturb_free_surface(m, seed, 0.01)
    d = M.build_dofs(m)
    phi, phin = M.tank_bc(m, d, seed, Lx, 1.0)
    return m, d, phi, phin


class _Dev:
    """Device vectors of the caller, through the HIP runtime the library is linked with."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a, np.float64)
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(a.nbytes)) == 0
        assert self.hip.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), 1) == 0
        self.ptrs.append(p)
        return p

    def get(self, p, n):
        a = np.empty(n)
        assert self.hip.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), p, ctypes.c_size_t(a.nbytes), 2) == 0
        return a

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _compare(bvp, n, seed):
    lib = bvp.lib
    dev = _Dev()
    x = dev.put(np.random.default_rng(seed).standard_normal(n))
    y_ref = dev.put(np.zeros(n))
    y = dev.put(np.full(n, np.nan))
    assert lib.bemfmm_matvec_device(bvp.ctx, x, y_ref) == 0
    rc = lib.bemfmm_matvec_device_noalloc(bvp.ctx, x, y)
    assert rc == 0, lib.bemfmm_last_error(bvp.ctx).decode()
    assert lib.bemfmm_sync(bvp.ctx) == 0
    a, b = dev.get(y, n), dev.get(y_ref, n)
    dev.free()
    assert np.all(np.isfinite(a))
    rel = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    print(f"[gpu] noalloc n_dof={n}: rel diff vs bemfmm_matvec_device {rel:.2e}", flush=True)
    assert rel <= 1e-12


def test_noalloc_matvec_matches_regular_two_calls_second_smaller():
    from bem_for_nonlinear_waves_amd.bvp import BEM_BVP
    # first call: the 4 x 1 x 1 m tank (2,092 nodes, p = 8); second: a smaller 2 x 1 x 1 m tank (p = 10)
    for (Lx, nx, ny, seed, p) in [(4.0, 42, 11, 1, 8), (2.0, 12, 6, 2, 10)]:
        m, d, phi, phin = _tank(Lx, nx, ny, seed)
        print(f"[gpu] noalloc case: tank {Lx} m, {m.n_nodes} nodes, p={p}", flush=True)
        bvp = BEM_BVP(m, d, p=p, leaf_max=32)
        bvp.set_bc(phi, phin)
        bvp.prepare()
        _compare(bvp, d.n_dof, seed)
        _compare(bvp, d.n_dof, seed + 10)  # a second call on the same context reuses the same array
        bvp.close()


def test_noalloc_matvec_refuses_a_problem_beyond_the_array():
    from bem_for_nonlinear_waves_amd import _lib
    from bem_for_nonlinear_waves_amd.bvp import BEM_BVP
    # 16k nodes in leaves of <= 4 faces down to level 8: 18.8k boxes, about 70 MiB of work memory at p = 10
    m, d, phi, phin = _tank(4.0, 120, 30, 3)
    print(f"[gpu] noalloc refusal case: {m.n_nodes} nodes, leaf 4, depth 8", flush=True)
    bvp = BEM_BVP(m, d, p=10, leaf_max=4, max_level=8)
    bvp.set_bc(phi, phin)
    bvp.prepare()
    dev = _Dev()
    x, y = dev.put(np.zeros(d.n_dof)), dev.put(np.zeros(d.n_dof))
    rc = bvp.lib.bemfmm_matvec_device_noalloc(bvp.ctx, x, y)
    msg = bvp.lib.bemfmm_last_error(bvp.ctx).decode()
    dev.free()
    bvp.close()
    assert rc == _lib.E_INVALID
    assert "MiB" in msg and "allocation-free" in msg
