"""GPU parity of the preconditioners (through bemfmm_precond_apply and bemfmm_solve).

Jacobi: z = r / diag(A) against the oracle's operator diagonal, to 1e-12 relative.
Schwarz (restricted additive, blocks read back from the library): the GPU's in-LDS Gauss-Jordan
inverses applied to r against the oracle's block LU solves on the same blocks, to 1e-10 relative
(both invert the same near-field blocks; pivoting order and summation order differ). Solutions:
1e-8 relative against dense LU of the oracle's operator, and the Schwarz GMRES iteration count
within 2 of the oracle's on the same blocks.
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


def _dense_lu(o, d):
    A, b = o.dense(), o.rhs()
    L = o.stats()["L"]
    x = np.linalg.solve(A, b)
    return np.where(d.dof_kind == 0, x / L, x)


def test_jacobi_apply_parity(small_tank, oracle_mod):
    from bem_for_nonlinear_waves_amd import _lib
    m, d, phi, phin = small_tank
    g = _gpu_bvp(m, d, phi, phin, p=8, leaf_max=32, preconditioner=_lib.PRECOND_JACOBI)
    o = oracle_mod.Oracle(m, d, phi, phin, p=8, leaf_max=32)
    r = np.random.default_rng(21).standard_normal(d.n_dof)
    z_g, z_o = g.precond_apply(r), o.diag_inv() * r
    assert np.linalg.norm(z_g - z_o) <= 1e-12 * np.linalg.norm(z_o)
    g.close()


def test_schwarz_apply_parity(small_tank, oracle_mod):
    from bem_for_nonlinear_waves_amd import _lib
    m, d, phi, phin = small_tank
    g = _gpu_bvp(m, d, phi, phin, p=8, leaf_max=32, preconditioner=_lib.PRECOND_SCHWARZ)
    ptr, idx, ncore = g.schwarz_blocks()
    st = g.stats()
    print(f"[gpu] schwarz blocks={st['n_precond_blocks']} avg={st['precond_avg_block']:.1f} "
          f"fallback={st['precond_fallback_blocks']}")
    assert st["n_precond_blocks"] == len(ncore) and st["precond_fallback_blocks"] == 0
    cores = np.concatenate([idx[ptr[b]:ptr[b] + ncore[b]] for b in range(len(ncore))])
    assert np.array_equal(np.sort(cores), np.arange(d.n_dof))
    assert np.diff(ptr).max() <= 128 and ncore.max() <= 64
    o = oracle_mod.Oracle(m, d, phi, phin, p=8, leaf_max=32)
    o.set_schwarz(ptr, idx, ncore)
    rng = np.random.default_rng(22)
    for _ in range(2):
        r = rng.standard_normal(d.n_dof)
        z_g, z_o = g.precond_apply(r), o.schwarz_apply(r)
        rel = np.linalg.norm(z_g - z_o) / np.linalg.norm(z_o)
        print(f"[gpu] schwarz apply rel err vs oracle = {rel:.3e}")
        assert rel <= 1e-10
    g.close()


def test_schwarz_solve_parity(tiny_tank, oracle_mod):
    from bem_for_nonlinear_waves_amd import _lib
    m, d, phi, phin = tiny_tank
    g = _gpu_bvp(m, d, phi, phin, p=10, leaf_max=16, preconditioner=_lib.PRECOND_SCHWARZ)
    x_g, info = g.solveSystemGMRES(tol=1e-12, max_iter=600)
    o = oracle_mod.Oracle(m, d, phi, phin, p=10, leaf_max=16)
    o.set_schwarz(*g.schwarz_blocks())
    x_o, oi = o.solve(tol=1e-12, max_iter=600, precond=2)
    _, oj = o.solve(tol=1e-12, max_iter=600, precond=1)
    x_lu = _dense_lu(o, d)
    rel = np.linalg.norm(x_g - x_lu) / np.linalg.norm(x_lu)
    print(f"[gpu] schwarz: gpu iters={info['iters']} oracle iters={oi['iters']} (jacobi {oj['iters']}) "
          f"rel_vs_dense_LU={rel:.3e}")
    assert info["converged"] == 1 and abs(info["iters"] - oi["iters"]) <= 2 and info["iters"] < oj["iters"]
    assert rel <= 1e-8
    assert np.linalg.norm(x_g - x_o) / np.linalg.norm(x_o) <= 1e-8
    g.close()


def test_schwarz_with_fp32_refinement(small_tank):
    """FP32 + refinement with the Schwarz preconditioner reaches the FP64 Jacobi answer (to 1e-8)."""
    from bem_for_nonlinear_waves_amd import _lib
    m, d, phi, phin = small_tank
    ref = _gpu_bvp(m, d, phi, phin, p=10, leaf_max=32, preconditioner=_lib.PRECOND_JACOBI)
    x64, i64 = ref.solveSystemGMRES(tol=1e-11, max_iter=2000)
    ref.close()
    g = _gpu_bvp(m, d, phi, phin, p=10, leaf_max=32, preconditioner=_lib.PRECOND_SCHWARZ,
                 precision=_lib.PRECISION_FP32_REFINE)
    x, info = g.solveSystemGMRES(tol=1e-11, max_iter=2000)
    rel = np.linalg.norm(x - x64) / np.linalg.norm(x64)
    print(f"[gpu] schwarz+fp32_refine: iters={info['iters']} steps={info['refinement_steps']} "
          f"(jacobi fp64 iters={i64['iters']}) rel={rel:.3e}")
    assert info["converged"] == 1 and rel <= 1e-8
    g.close()


def test_schwarz_option_validation(tiny_tank):
    from bem_for_nonlinear_waves_amd import _lib
    from bem_for_nonlinear_waves_amd.bvp import BEM_BVP, BEMError
    m, d, phi, phin = tiny_tank
    b = BEM_BVP(m, d, p=8, leaf_max=16, preconditioner=_lib.PRECOND_SCHWARZ, schwarz_max_block=256)
    b.set_bc(phi, phin)
    with pytest.raises(BEMError) as e:
        b.prepare()
    assert e.value.code == _lib.E_INVALID
    b.close()
