"""Parity at the workloads the benchmark configurations name (4 x 1 x 1 m tank, linear triangles, Dirichlet top
with phi = seeded sum of 8 sine modes, Neumann walls and bottom with phi_n = 0):

* config 1 exactly: 42 x 11 x 11 cells, N = 2,092 nodes, flat top, p = 8, leaf <= 32 sources, theta = 0.25,
  GMRES tol 1e-10;
* config 2's 5k-node subsample: 70 x 18 x 18 cells with the top's interior nodes offset by N(0, 0.02 m) in z;
* config 2 at full size (210 x 53 x 53 cells, ~50k nodes) at its stated p = 8, through properties that hold at
  any size (the oracle takes minutes per matvec there).

Each GPU solve (FP64, and FP32 operator + FP64 refinement) is compared with the FP64 oracle's own GMRES
(the reference's run_gmres restated) and with the dense LU solution of the oracle's FMM-consistent operator:
relative error <= 1e-8 on phi (Neumann DOFs) and phi_n (Dirichlet DOFs), the stated target. Both sides solve
to the configuration's tolerance with Jacobi scaling (the oracle's GMRES without a preconditioner needs far
more iterations than these tests can afford).
"""
import numpy as np
import pytest

from bem_for_nonlinear_waves_amd import mesh as M

pytestmark = pytest.mark.gpu


def _tank(nx, ny, nz, seed, sigma):
    m = M.box_tank(4.0, 1.0, 1.0, nx, ny, nz)
    if sigma > 0.0:
        m = M.perturb_free_surface(m, seed, sigma)
    d = M.build_dofs(m)
    phi, phin = M.tank_bc(m, d, seed, 4.0, 1.0)
    return m, d, phi, phin


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _gpu_solves(m, d, phi, phin, p, tol):
    from bem_for_nonlinear_waves_amd import _lib
    from bem_for_nonlinear_waves_amd.bvp import BEM_BVP
    out = {}
    for prec in (_lib.PRECISION_FP64, _lib.PRECISION_FP32_REFINE):
        print(f"[gpu] config solve n_nodes={m.n_nodes} n_dof={d.n_dof} p={p} precision={prec}", flush=True)
        b = BEM_BVP(m, d, p=p, leaf_max=32, precision=prec, preconditioner=_lib.PRECOND_JACOBI)
        x, info = b.solveGMRES(phi, phin, tol=tol, max_iter=2000)
        print(f"[gpu]   iters={info['iters']} steps={info['refinement_steps']} residual={info['residual']:.3e}",
              flush=True)
        assert info["converged"] == 1, info
        out[prec] = x
        b.close()
    return out


def _check_against_oracle(oracle_mod, m, d, phi, phin, p, tol, xs):
    o = oracle_mod.Oracle(m, d, phi, phin, p=p, leaf_max=32)
    L = o.stats()["L"]
    A, b = o.dense(), o.rhs()
    x_lu = np.linalg.solve(A, b)
    x_lu = np.where(d.dof_kind == 0, x_lu / L, x_lu)  # Dirichlet unknowns are L phi_n in the scaled system
    del A
    x_o, oi = o.solve(tol=tol, max_iter=3000, precond=1)
    assert oi["rc"] == 0, oi
    print(f"[oracle] iters={oi['iters']} residual={oi['residual']:.3e}; oracle GMRES vs dense LU "
          f"{_rel(x_o, x_lu):.3e}", flush=True)
    for prec, x in xs.items():
        e_lu, e_o = _rel(x, x_lu), _rel(x, x_o)
        # per unknown kind as well: phi on Neumann DOFs, phi_n on Dirichlet DOFs
        dn, nn = d.dof_kind == 0, d.dof_kind != 0
        e_phin, e_phi = _rel(x[dn], x_lu[dn]), _rel(x[nn], x_lu[nn])
        print(f"[gpu] precision={prec}: vs dense LU {e_lu:.3e} (phi_n {e_phin:.3e}, phi {e_phi:.3e}); "
              f"vs oracle GMRES {e_o:.3e}", flush=True)
        assert e_lu <= 1e-8 and e_o <= 1e-8 and e_phin <= 1e-8 and e_phi <= 1e-8


def test_config1_tank_2092_p8_tol1e10(oracle_mod):
    m, d, phi, phin = _tank(42, 11, 11, 1, 0.0)
    assert m.n_nodes == 2092
    xs = _gpu_solves(m, d, phi, phin, 8, 1e-10)
    _check_against_oracle(oracle_mod, m, d, phi, phin, 8, 1e-10, xs)


def test_config2_subsample_5k_rough_top_p8(oracle_mod):
    m, d, phi, phin = _tank(70, 18, 18, 2, 0.02)
    assert 5_000 <= m.n_nodes <= 6_000
    xs = _gpu_solves(m, d, phi, phin, 8, 1e-10)
    _check_against_oracle(oracle_mod, m, d, phi, phin, 8, 1e-10, xs)


def test_config2_50k_p8_linearity_and_refinement_equals_fp64():
    from bem_for_nonlinear_waves_amd import _lib
    from bem_for_nonlinear_waves_amd.bvp import BEM_BVP
    m, d, phi, phin = _tank(210, 53, 53, 2, 0.02)
    assert 48_000 <= m.n_nodes <= 52_000
    xs = {}
    for prec in (_lib.PRECISION_FP64, _lib.PRECISION_FP32_REFINE):
        print(f"[gpu] config 2 full size n_nodes={m.n_nodes} p=8 precision={prec}", flush=True)
        b = BEM_BVP(m, d, p=8, leaf_max=32, precision=prec, preconditioner=_lib.PRECOND_SCHWARZ)
        b.set_bc(phi, phin)
        b.prepare()
        if prec == _lib.PRECISION_FP64:
            rng = np.random.default_rng(23)
            x1, x2 = rng.standard_normal(d.n_dof), rng.standard_normal(d.n_dof)
            y1, y2, y12 = b.compute_Ax(x1), b.compute_Ax(x2), b.compute_Ax(2.0 * x1 - 0.5 * x2)
            assert np.linalg.norm(y12 - (2.0 * y1 - 0.5 * y2)) <= 1e-12 * np.linalg.norm(y12)
        x, info = b.solveSystemGMRES(tol=1e-10, max_iter=1000)
        print(f"[gpu]   iters={info['iters']} steps={info['refinement_steps']} residual={info['residual']:.3e}",
              flush=True)
        assert info["converged"] == 1
        xs[prec] = x
        b.close()
    rel = _rel(xs[_lib.PRECISION_FP32_REFINE], xs[_lib.PRECISION_FP64])
    print(f"[gpu] config 2 full size: FP32+refinement vs FP64 {rel:.3e}", flush=True)
    assert rel <= 1e-9
