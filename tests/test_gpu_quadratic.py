"""True quadratic faces on the GPU (element_order=2; isTrueQuadraticElement, bem/BEM_solveBVP_GMRES_FMM.hpp:
1311-1657, midpoint rows :111-144, DOFs bem/BEM_setBoundaryTypes.hpp:190-233) against the FP64 CPU oracle:
matvec, alpha and right-hand side to 1e-10 (the tolerance of the linear faces' parity), with the tree built on the
host and on the GPU, on a rough box tank (one edge midpoint without a DOF on one face) and on the cylinder tank
with its edge midpoints on the cylinder; GMRES solutions (FP64, and the FP32 operator with FP64 refinement) equal
to dense LU of the oracle's operator to 1e-8; the standing wave cosh(k z) cos(k x) reproduced on the walls to the
accuracy the oracle's dense solve gives.
"""
import math

import numpy as np
import pytest

from bem_for_nonlinear_waves_amd import mesh as M

pytestmark = pytest.mark.gpu


def _rough_box():
    m = M.to_quadratic(M.perturb_free_surface(M.box_tank(2.0, 1.0, 1.0, 8, 4, 4), 2, 0.02))
    d = M.build_dofs_quadratic(m)
    d.face_dof[5, 4] = -1  # this face's midpoint weight goes to the edge's two vertices
    phi, phin = M.tank_bc(m, d, 3, 2.0, 1.0)
    return m, d, phi, phin


def _cylinder():
    R = 0.25
    m = M.cylinder_tank(Lx=3.0, Ly=2.0, Lz=1.0, R=R, draft=0.5, h=0.15)

    def on_cylinder(Xm, ends):
        r = np.hypot(m.X[ends, 0], m.X[ends, 1])
        hit = np.all(np.abs(r - R) < 1e-9, axis=1)
        out = Xm.copy()
        s = R / np.hypot(Xm[hit, 0], Xm[hit, 1])
        out[hit, 0] *= s
        out[hit, 1] *= s
        return out

    q = M.to_quadratic(m, on_cylinder)
    d = M.build_dofs_quadratic(q)
    phi, phin = M.tank_bc(q, d, 5, 3.0, 1.0)
    return q, d, phi, phin


def _gpu_bvp(m, d, phi, phin, **kw):
    from bem_for_nonlinear_waves_amd.bvp import BEM_BVP
    print(f"[gpu] quadratic BEM_BVP n_nodes={m.n_nodes} n_faces={m.n_faces} n_dof={d.n_dof} opts={kw}", flush=True)
    b = BEM_BVP(m, d, element_order=2, **kw)
    b.set_bc(phi, phin)
    b.prepare()
    return b


def _rel(u, v):
    return float(np.linalg.norm(u - v) / np.linalg.norm(v))


@pytest.mark.parametrize("case,p,tree", [("rough box", 8, 0), ("rough box", 10, 1), ("cylinder", 10, 1),
                                         ("cylinder", 8, 0)])
def test_quadratic_operator_parity(oracle_mod, case, p, tree):
    m, d, phi, phin = _rough_box() if case == "rough box" else _cylinder()
    leaf = 16 if case == "rough box" else 32
    g = _gpu_bvp(m, d, phi, phin, p=p, leaf_max=leaf, tree_on_gpu=tree)
    o = oracle_mod.Oracle(m, d, phi, phin, p=p, leaf_max=leaf)
    st, so = g.stats(), o.stats()
    assert st["tree_on_gpu"] == tree
    assert (st["n_boxes"], st["n_m2l_pairs"], st["near_nnz"]) == (so["n_buckets"], so["n_m2l_pairs"], so["near_nnz"])
    x = np.random.default_rng(p).standard_normal(d.n_dof)
    e = (_rel(g.compute_Ax(x), o.matvec(x)), _rel(g.diag_coeffs(), o.alpha()), _rel(g.rhs(), o.rhs()))
    print(f"[gpu] quadratic {case} p={p} tree_on_gpu={tree}: matvec {e[0]:.2e} alpha {e[1]:.2e} rhs {e[2]:.2e}",
          flush=True)
    assert max(e) <= 1e-10
    g.close()


@pytest.mark.parametrize("precision,precond", [(0, 1), (0, 2), (1, 1), (1, 2)])
def test_quadratic_solve_vs_dense_lu(oracle_mod, precision, precond):
    m, d, phi, phin = _rough_box()
    g = _gpu_bvp(m, d, phi, phin, p=10, leaf_max=16, precision=precision, preconditioner=precond,
                 schwarz_max_core=16, schwarz_max_block=64)
    x_g, info = g.solveSystemGMRES(tol=1e-12, max_iter=1200)
    assert info["converged"] == 1, info
    o = oracle_mod.Oracle(m, d, phi, phin, p=10, leaf_max=16)
    A, b = o.dense(), o.rhs()
    x_lu = np.linalg.solve(A, b)
    x_lu = np.where(d.dof_kind == 0, x_lu / o.stats()["L"], x_lu)
    rel = _rel(x_g, x_lu)
    print(f"[gpu] quadratic solve precision={precision} precond={precond}: iters={info['iters']} "
          f"residual={info['residual']:.3e} rel_vs_dense_LU={rel:.3e}", flush=True)
    assert rel <= 1e-8
    g.close()


def test_quadratic_standing_wave(oracle_mod):
    m = M.to_quadratic(M.box_tank(2.0, 1.0, 1.0, 8, 4, 4))
    d = M.build_dofs_quadratic(m)
    phi, phin, phi_ex, phin_ex = M.standing_wave(m, d, 2.0)
    g = _gpu_bvp(m, d, phi, phin, p=10, leaf_max=16, preconditioner=2)
    x_g, info = g.solveSystemGMRES(tol=1e-12, max_iter=1200)
    assert info["converged"] == 1, info
    o = oracle_mod.Oracle(m, d, phi, phin, p=10, leaf_max=16)
    x_lu = np.linalg.solve(o.dense(), o.rhs())
    walls = d.dof_kind != 0
    err_g = math.sqrt(np.mean((x_g[walls] - phi_ex[walls]) ** 2)) / np.abs(phi_ex[walls]).max()
    err_o = math.sqrt(np.mean((x_lu[walls] - phi_ex[walls]) ** 2)) / np.abs(phi_ex[walls]).max()
    print(f"[gpu] quadratic standing wave, phi on the walls (rms): GPU {err_g:.2e}, oracle dense {err_o:.2e}",
          flush=True)
    assert abs(err_g - err_o) <= 1e-3 * err_o and err_g < 2e-4
    g.close()
