"""GPU parity of the FP32 + FP64-iterative-refinement mode (BEMFMM_PRECISION_FP32_REFINE).

The FP32 operator (near field, P2M and M2M/M2L/L2L on FP32 MFMA, FP64 vectors) is checked
against the FP64 oracle matvec to 2e-5 relative (FP32 rounding of the matrices and of the
translation accumulation; the FMM truncation error is common to both sides). The solutions
of the refinement loop must meet the same FP64 residual criterion as the FP64 path and agree
with dense LU of the oracle's operator, and with the FP64 GPU solve, to 1e-8 relative on
phi (Neumann DOFs) and phi_n (Dirichlet DOFs).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _gpu_bvp(m, d, phi, phin, **kw):
    from bem_for_nonlinear_waves_amd.bvp import BEM_BVP
    print(f"[gpu] BEM_BVP n_nodes={m.n_nodes} n_dof={d.n_dof} opts={kw}", flush=True)
    b = BEM_BVP(m, d, **kw)
    b.set_bc(phi, phin)
    b.prepare()
    return b


@pytest.mark.parametrize("p,kernel", [(10, 1), (10, 2), (8, 2)])
def test_fp32_operator_vs_oracle(small_tank, oracle_mod, p, kernel):
    """M2L of the FP32 operator on dense matrix-core tiles (kernel 1) or by rotation (kernel 2)."""
    from bem_for_nonlinear_waves_amd import _lib
    m, d, phi, phin = small_tank
    g = _gpu_bvp(m, d, phi, phin, p=p, leaf_max=32, precision=_lib.PRECISION_FP32_REFINE, m2l_kernel=kernel)
    assert g.stats()["m2l_rotation"] == (1 if kernel == 2 else 0)
    o = oracle_mod.Oracle(m, d, phi, phin, p=p, leaf_max=32)
    rng = np.random.default_rng(11)
    for _ in range(2):
        x = rng.standard_normal(d.n_dof)
        y32, y_o = g.compute_Ax_fp32(x), o.matvec(x)
        rel = np.linalg.norm(y32 - y_o) / np.linalg.norm(y_o)
        print(f"[gpu] fp32 operator rel err vs oracle = {rel:.3e}")
        assert rel <= 2e-5
        assert rel >= 1e-12  # the FP32 path really runs in FP32
        # the FP64 operator of the same context is unchanged by the FP32 copies
        y64 = g.compute_Ax(x)
        assert np.linalg.norm(y64 - y_o) <= 1e-10 * np.linalg.norm(y_o)
    # linear up to FP32 rounding
    x1, x2 = rng.standard_normal(d.n_dof), rng.standard_normal(d.n_dof)
    y1, y2, y12 = g.compute_Ax_fp32(x1), g.compute_Ax_fp32(x2), g.compute_Ax_fp32(x1 + x2)
    assert np.linalg.norm(y12 - y1 - y2) <= 1e-5 * np.linalg.norm(y12)
    g.close()


@pytest.mark.parametrize("precond", [0, 1])
def test_fp32_refine_solve_vs_dense_lu(tiny_tank, oracle_mod, precond):
    from bem_for_nonlinear_waves_amd import _lib
    m, d, phi, phin = tiny_tank
    g = _gpu_bvp(m, d, phi, phin, p=10, leaf_max=16, precision=_lib.PRECISION_FP32_REFINE, preconditioner=precond)
    x_g, info = g.solveSystemGMRES(tol=1e-12, max_iter=1200)
    hist = g.refinement_history()
    print(f"[gpu] refine precond={precond}: iters={info['iters']} steps={info['refinement_steps']} "
          f"residual={info['residual']:.3e} history={np.array2string(hist, precision=2)}")
    assert info["converged"] == 1 and info["residual"] <= 1e-12
    assert info["refinement_steps"] >= 1 and hist[-1] == pytest.approx(info["residual"])
    o = oracle_mod.Oracle(m, d, phi, phin, p=10, leaf_max=16)
    A, b = o.dense(), o.rhs()
    L = o.stats()["L"]
    x_lu = np.linalg.solve(A, b)
    x_lu = np.where(d.dof_kind == 0, x_lu / L, x_lu)
    rel = np.linalg.norm(x_g - x_lu) / np.linalg.norm(x_lu)
    print(f"[gpu] refine precond={precond}: rel_vs_dense_LU={rel:.3e}")
    assert rel <= 1e-8
    g.close()


def test_fp32_refine_matches_fp64_path(small_tank):
    """Same tank, Jacobi: FP32+refinement and the FP64 path give the same answer (to 1e-8)."""
    from bem_for_nonlinear_waves_amd import _lib
    m, d, phi, phin = small_tank
    out = {}
    for prec in (_lib.PRECISION_FP64, _lib.PRECISION_FP32_REFINE):
        g = _gpu_bvp(m, d, phi, phin, p=10, leaf_max=32, precision=prec, preconditioner=_lib.PRECOND_JACOBI)
        x, info = g.solveSystemGMRES(tol=1e-11, max_iter=2000)
        print(f"[gpu] precision={prec}: iters={info['iters']} matvecs={info['n_matvec']} "
              f"steps={info['refinement_steps']} residual={info['residual']:.3e}")
        assert info["converged"] == 1
        out[prec] = x
        g.close()
    x64, x32 = out[_lib.PRECISION_FP64], out[_lib.PRECISION_FP32_REFINE]
    rel = np.linalg.norm(x32 - x64) / np.linalg.norm(x64)
    print(f"[gpu] fp32_refine vs fp64 solution rel = {rel:.3e}")
    assert rel <= 1e-8


def test_fp32_matvec_requires_refine_mode(tiny_tank):
    from bem_for_nonlinear_waves_amd import _lib
    from bem_for_nonlinear_waves_amd.bvp import BEMError
    m, d, phi, phin = tiny_tank
    g = _gpu_bvp(m, d, phi, phin, p=8, leaf_max=16)
    with pytest.raises(BEMError) as e:
        g.compute_Ax_fp32(np.zeros(d.n_dof))
    assert e.value.code == _lib.E_STATE
    g.close()


def test_rotation_m2l_matches_dense_tiles(small_tank):
    """The rotation M2L (H^T T(|o| z) D per pair, FP32 factors) and the dense matrix-core M2L give the same
    FP32 operator to FP32 rounding; the default picks the rotation for p = 10; orders without a compiled
    rotation kernel are refused when it is requested and fall back to the dense tiles by default."""
    from bem_for_nonlinear_waves_amd import _lib
    from bem_for_nonlinear_waves_amd.bvp import BEM_BVP, BEMError
    m, d, phi, phin = small_tank
    kw = dict(p=10, leaf_max=32, precision=_lib.PRECISION_FP32_REFINE)
    a = _gpu_bvp(m, d, phi, phin, m2l_kernel=1, **kw)
    b = _gpu_bvp(m, d, phi, phin, **kw)
    assert b.stats()["m2l_rotation"] == 1
    rng = np.random.default_rng(21)
    for _ in range(3):
        x = rng.standard_normal(d.n_dof)
        ya, yb = a.compute_Ax_fp32(x), b.compute_Ax_fp32(x)
        rel = np.linalg.norm(ya - yb) / np.linalg.norm(ya)
        print(f"[gpu] rotation vs dense FP32 M2L rel = {rel:.3e}")
        assert rel <= 5e-6
    # a re-prepare with moved nodes reuses the resident factors and stays consistent
    from bem_for_nonlinear_waves_amd import mesh as M
    m2 = M.perturb_free_surface(m, 9, 0.001)
    for g in (a, b):
        g.set_geometry(m2, d)
        g.set_bc(phi, phin)
        g.prepare()
    x = rng.standard_normal(d.n_dof)
    assert np.linalg.norm(a.compute_Ax_fp32(x) - b.compute_Ax_fp32(x)) <= 5e-6 * np.linalg.norm(a.compute_Ax_fp32(x))
    a.close()
    b.close()
    c = BEM_BVP(m, d, p=9, leaf_max=32, precision=_lib.PRECISION_FP32_REFINE, m2l_kernel=2)
    c.set_bc(phi, phin)
    with pytest.raises(BEMError) as e:
        c.prepare()
    assert e.value.code == _lib.E_INVALID
    c.close()
    c = _gpu_bvp(m, d, phi, phin, p=9, leaf_max=32, precision=_lib.PRECISION_FP32_REFINE)
    assert c.stats()["m2l_rotation"] == 0
    c.close()
