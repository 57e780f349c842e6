"""FP32 inner operator at a lower expansion order (inner_p < p, FP32 + FP64 refinement with the rotation M2L).

The packed expansions are degree-major and the translation operators are triangular in the degree (M2M: a parent
degree takes child degrees up to it; L2L: a child degree takes parent degrees from it up), so running the M2L at
inner_p on the p-order multipoles and zeroing the locals above inner_p gives the far field of the inner_p FMM
exactly; the rigid-mode diagonal stays the p-order one. Checked here: the FP32 inner operator equals the FP64
oracle at inner_p with the oracle's p-order diagonal on the unknown-phi rows (to 2e-5, the FP32 operator bound;
measured ~2e-9), the FP64 operator is untouched (1e-10 against the oracle at p), solves reach dense LU of the
p-order operator to 1e-8, and the option is refused outside the rotation M2L's orders and the FP32 mode.
"""
import numpy as np
import pytest

from bem_for_nonlinear_waves_amd import _lib

pytestmark = pytest.mark.gpu


def _bvp(m, d, phi, phin, **kw):
    from bem_for_nonlinear_waves_amd.bvp import BEM_BVP
    print(f"[gpu] inner order: n_dof={d.n_dof} opts={kw}", flush=True)
    b = BEM_BVP(m, d, **kw)
    b.set_bc(phi, phin)
    b.prepare()
    return b


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("p,inner", [(10, 8), (10, 6), (10, 5), (8, 6)])
def test_inner_operator_is_the_lower_order_fmm(small_tank, oracle_mod, p, inner):
    m, d, phi, phin = small_tank
    g = _bvp(m, d, phi, phin, p=p, leaf_max=32, precision=_lib.PRECISION_FP32_REFINE, inner_p=inner,
             preconditioner=_lib.PRECOND_JACOBI)
    assert g.stats()["inner_p"] == inner and g.stats()["m2l_rotation"] == 1
    o_in = oracle_mod.Oracle(m, d, phi, phin, p=inner, leaf_max=32)
    o_p = oracle_mod.Oracle(m, d, phi, phin, p=p, leaf_max=32)
    uphi = np.isin(d.dof_kind, (1, 3))  # rows whose unknown phi meets the rigid-mode diagonal
    x = np.random.default_rng(p + inner).standard_normal(d.n_dof)
    ref = o_in.matvec(x) - np.where(uphi, (o_p.alpha() - o_in.alpha()) * x, 0.0)
    e32 = _rel(g.compute_Ax_fp32(x), ref)
    e64 = _rel(g.compute_Ax(x), o_p.matvec(x))
    print(f"[gpu] p={p} inner_p={inner}: FP32 inner operator vs oracle {e32:.2e}, FP64 operator vs oracle {e64:.2e}",
          flush=True)
    assert e32 <= 2e-5 and e64 <= 1e-10
    g.close()


@pytest.mark.parametrize("inner", [8, 6, 5])
def test_inner_order_solve_reaches_dense_lu(tiny_tank, oracle_mod, inner):
    m, d, phi, phin = tiny_tank
    g = _bvp(m, d, phi, phin, p=10, leaf_max=16, precision=_lib.PRECISION_FP32_REFINE, inner_p=inner,
             preconditioner=_lib.PRECOND_SCHWARZ, schwarz_max_core=16, schwarz_max_block=64)
    x, info = g.solveSystemGMRES(tol=1e-12, max_iter=1200)
    assert info["converged"] == 1, info
    o = oracle_mod.Oracle(m, d, phi, phin, p=10, leaf_max=16)
    x_lu = np.linalg.solve(o.dense(), o.rhs())
    x_lu = np.where(d.dof_kind == 0, x_lu / o.stats()["L"], x_lu)
    rel = _rel(x, x_lu)
    print(f"[gpu] inner_p={inner}: iters={info['iters']} steps={info['refinement_steps']} "
          f"residual={info['residual']:.2e} rel_vs_dense_LU={rel:.2e}", flush=True)
    assert rel <= 1e-8
    g.close()


def test_inner_order_outside_its_range_is_refused(tiny_tank):
    from bem_for_nonlinear_waves_amd.bvp import BEM_BVP, BEMError
    m, d, phi, phin = tiny_tank
    bad = [dict(p=10, inner_p=10, precision=_lib.PRECISION_FP32_REFINE),  # not below p
           dict(p=10, inner_p=7, precision=_lib.PRECISION_FP32_REFINE),   # no rotation kernel at 7
           dict(p=10, inner_p=4, precision=_lib.PRECISION_FP32_REFINE),   # nor at 4
           dict(p=10, inner_p=8),                                         # FP64 mode
           dict(p=10, inner_p=8, precision=_lib.PRECISION_FP32_REFINE, m2l_kernel=1)]  # dense FP32 M2L
    for kw in bad:
        b = BEM_BVP(m, d, leaf_max=16, **kw)
        b.set_bc(phi, phin)
        with pytest.raises(BEMError):
            b.prepare()
        b.close()
