"""GPU parity: the HIP path (through the C ABI) against the FP64 CPU oracle.

Tolerances (FP64 path): matvec / RHS / alpha agree to 1e-10 relative (both
evaluate the same FMM operator; the GPU uses lattice-exact translation operators
computed in float128 while the oracle evaluates GR97 per box pair, and the sums
run in a different order). Solutions: GMRES to tol 1e-12 on both sides agree to
1e-8 relative on phi (Neumann DOFs) and phi_n (Dirichlet DOFs), the target of
the FP64 oracle comparison.
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


@pytest.mark.parametrize("p,tile", [(8, 0), (10, 0), (10, 16), (10, 64)])
def test_operator_parity_small_tank(small_tank, oracle_mod, p, tile):
    m, d, phi, phin = small_tank
    g = _gpu_bvp(m, d, phi, phin, p=p, leaf_max=32, m2l_tile_pairs=tile)
    o = oracle_mod.Oracle(m, d, phi, phin, p=p, leaf_max=32)
    st, so = g.stats(), o.stats()
    assert st["n_boxes"] == so["n_buckets"] and st["n_leaves"] == so["n_leaves"]
    assert st["n_m2l_pairs"] == so["n_m2l_pairs"] and st["near_nnz"] == so["near_nnz"]
    a_g, a_o = g.diag_coeffs(), o.alpha()
    assert np.abs(a_g - a_o).max() <= 1e-10 * np.abs(a_o).max()
    b_g, b_o = g.rhs(), o.rhs()
    assert np.linalg.norm(b_g - b_o) <= 1e-10 * np.linalg.norm(b_o)
    rng = np.random.default_rng(7)
    for _ in range(2):
        x = rng.standard_normal(d.n_dof)
        y_g, y_o = g.compute_Ax(x), o.matvec(x)
        assert np.linalg.norm(y_g - y_o) <= 1e-10 * np.linalg.norm(y_o)
    g.close()


def test_matvec_linearity_and_determinism(small_tank):
    m, d, phi, phin = small_tank
    g = _gpu_bvp(m, d, phi, phin, p=8, leaf_max=32)
    rng = np.random.default_rng(3)
    x1, x2 = rng.standard_normal(d.n_dof), rng.standard_normal(d.n_dof)
    y1, y2, y12 = g.compute_Ax(x1), g.compute_Ax(x2), g.compute_Ax(2.0 * x1 - 0.5 * x2)
    assert np.linalg.norm(y12 - (2.0 * y1 - 0.5 * y2)) <= 1e-12 * np.linalg.norm(y12)
    # M2L/M2M/L2L accumulate with atomics: repeated evaluations agree to rounding
    assert np.linalg.norm(g.compute_Ax(x1) - y1) <= 1e-13 * np.linalg.norm(y1)
    g.close()


def test_solve_parity_tiny_tank(tiny_tank, oracle_mod):
    m, d, phi, phin = tiny_tank
    g = _gpu_bvp(m, d, phi, phin, p=10, leaf_max=16)
    x_g, info = g.solveSystemGMRES(tol=1e-12, max_iter=600)
    assert info["converged"] == 1, info
    o = oracle_mod.Oracle(m, d, phi, phin, p=10, leaf_max=16)
    # FMM-consistent dense operator of the same discretisation, solved by LU
    A, b = o.dense(), o.rhs()
    L = o.stats()["L"]
    x_lu = np.linalg.solve(A, b)
    x_lu = np.where(d.dof_kind == 0, x_lu / L, x_lu)
    rel = np.linalg.norm(x_g - x_lu) / np.linalg.norm(x_lu)
    print(f"[gpu] tiny tank: iters={info['iters']} residual={info['residual']:.3e} rel_vs_dense_LU={rel:.3e}")
    assert rel <= 1e-8
    x_o, oi = o.solve(tol=1e-12, max_iter=600)
    rel_o = np.linalg.norm(x_g - x_o) / np.linalg.norm(x_o)
    print(f"[gpu] tiny tank: oracle iters={oi['iters']} rel_vs_oracle_gmres={rel_o:.3e}")
    assert rel_o <= 1e-8
    g.close()


def test_analytic_harmonic_field():
    """Linear harmonic field phi = x - y/2 + z/4: the BIE reproduces it up to discretisation error."""
    from bem_for_nonlinear_waves_amd import mesh as M
    m = M.box_tank(2.0, 1.0, 1.0, 12, 6, 6)
    d = M.build_dofs(m)
    phi, phin, phi_ex, phin_ex = M.analytic_bc(m, d, "linear")
    g = _gpu_bvp(m, d, phi, phin, p=10, leaf_max=16)
    x, info = g.solveSystemGMRES(tol=1e-10, max_iter=600)
    ex = np.where(d.dof_kind == 0, phin_ex, phi_ex)
    assert info["converged"] == 1
    assert np.linalg.norm(x - ex) / np.linalg.norm(ex) < 1e-3
    g.close()


def test_error_paths(tiny_tank):
    from bem_for_nonlinear_waves_amd import _lib
    from bem_for_nonlinear_waves_amd.bvp import BEM_BVP, BEMError
    m, d, phi, phin = tiny_tank
    b = BEM_BVP(m, d, p=8, leaf_max=16)
    with pytest.raises(BEMError) as e:
        b.compute_Ax(np.zeros(d.n_dof))  # before prepare
    assert e.value.code == _lib.E_STATE
    bad = phi.copy()
    bad[np.argmax(d.dof_kind == 0)] = np.nan
    with pytest.raises(BEMError) as e:
        b.set_bc(bad, phin)
    assert e.value.code == _lib.E_NONFINITE
    b.set_bc(phi, phin)
    b.prepare()
    x, info = b.solveSystemGMRES(tol=1e-14, max_iter=3)
    assert info["converged"] == 0 and info["iters"] <= 4
    with pytest.raises(BEMError) as e:
        b.solveSystemGMRES(tol=1e-14, max_iter=3, raise_on_nonconvergence=True)
    assert e.value.code == _lib.E_NOT_CONVERGED
    b.close()


def test_jacobi_solve_parity(tiny_tank, oracle_mod):
    """Left Jacobi scaling by the operator diagonal (G_ii / -Gn_ii - alpha_i): same solution, fewer iterations."""
    from bem_for_nonlinear_waves_amd import _lib
    m, d, phi, phin = tiny_tank
    g = _gpu_bvp(m, d, phi, phin, p=10, leaf_max=16, preconditioner=_lib.PRECOND_JACOBI)
    x_g, info = g.solveSystemGMRES(tol=1e-12, max_iter=600)
    o = oracle_mod.Oracle(m, d, phi, phin, p=10, leaf_max=16)
    x_o, oi = o.solve(tol=1e-12, max_iter=600, precond=1)
    A, b = o.dense(), o.rhs()
    L = o.stats()["L"]
    x_lu = np.linalg.solve(A, b)
    x_lu = np.where(d.dof_kind == 0, x_lu / L, x_lu)
    print(f"[gpu] jacobi: gpu iters={info['iters']} oracle iters={oi['iters']}")
    assert info["converged"] == 1 and abs(info["iters"] - oi["iters"]) <= 2
    assert np.linalg.norm(x_g - x_lu) / np.linalg.norm(x_lu) <= 1e-8
    assert np.linalg.norm(x_g - x_o) / np.linalg.norm(x_o) <= 1e-8
    g.close()


@pytest.mark.parametrize("precision,precond", [(0, 1), (1, 1), (0, 2), (1, 2)])
def test_graph_replays_match_eager(small_tank, precision, precond):
    """The matvec replayed from its captured HIP graph equals the eager launches, on every replay, with
    other (eager) work and synchronisations between the replays. In the solves the graph's input is
    written by the basis normalisation and its output read in place by the preconditioner (Jacobi,
    Schwarz), which the eager path does with ordinary buffers."""
    m, d, phi, phin = small_tank
    # preconditioned: the unpreconditioned solve takes ~400 iterations here, where last-bit differences
    # (M2M/L2L atomics) already move the stopping iteration
    print(f"[gpu] graph vs eager precision={precision} precond={precond}", flush=True)
    g = _gpu_bvp(m, d, phi, phin, p=10, leaf_max=32, precision=precision, use_graphs=1, preconditioner=precond)
    e = _gpu_bvp(m, d, phi, phin, p=10, leaf_max=32, precision=precision, use_graphs=0, preconditioner=precond)
    mv = (lambda b, x: b.compute_Ax_fp32(x)) if precision == 1 else (lambda b, x: b.compute_Ax(x))
    rng = np.random.default_rng(12)
    for it in range(4):
        x = rng.standard_normal(d.n_dof)
        yg, ye = mv(g, x), mv(e, x)
        assert np.linalg.norm(yg - ye) <= 1e-12 * np.linalg.norm(ye), it
        b = g.rhs()  # eager evaluation between replays
        assert np.linalg.norm(b - e.rhs()) <= 1e-12 * np.linalg.norm(b)
    xg, ig = g.solveSystemGMRES(tol=1e-11, max_iter=600)
    xe, ie = e.solveSystemGMRES(tol=1e-11, max_iter=600)
    assert abs(ig["iters"] - ie["iters"]) <= 2 and np.linalg.norm(xg - xe) <= 1e-9 * np.linalg.norm(xe)
    g.close()
    e.close()


def test_reprepare_reuses_operator_table(small_tank, oracle_mod):
    """A second prepare of the same context after the free surface moved (an RK stage) uploads no
    translation operators it already holds (lattice offsets mostly unchanged) and matches a fresh
    context and the oracle."""
    from bem_for_nonlinear_waves_amd import mesh as M
    from bem_for_nonlinear_waves_amd.bvp import BEM_BVP
    m, d, phi, phin = small_tank
    g = _gpu_bvp(m, d, phi, phin, p=10, leaf_max=32)
    first = g.stats()["n_ops_uploaded"]
    m2 = M.perturb_free_surface(m, 9, 0.001)
    g2 = BEM_BVP(m2, d, p=10, leaf_max=32)
    g2.set_bc(phi, phin)
    g2.prepare()
    # re-prepare the first context with the moved geometry
    g.set_geometry(m2, d)
    g.set_bc(phi, phin)
    g.prepare()
    st = g.stats()
    print(f"[gpu] operators uploaded: first prepare {first}, after moving nodes {st['n_ops_uploaded']}")
    assert first > 16 and st["n_ops_uploaded"] <= first // 10
    x = np.random.default_rng(13).standard_normal(d.n_dof)
    y1, y2 = g.compute_Ax(x), g2.compute_Ax(x)
    assert np.linalg.norm(y1 - y2) <= 1e-13 * np.linalg.norm(y2)
    o = oracle_mod.Oracle(m2, d, phi, phin, p=10, leaf_max=32)
    assert np.linalg.norm(y1 - o.matvec(x)) <= 1e-10 * np.linalg.norm(y2)
    g.close()
    g2.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_phi_t_second_bvp_on_resident_plan(small_tank, precision):
    """The phi_t BVP of an RK stage (solveSystemGMRES(is_time_derivative=true, x0)) on the prepared
    context: no re-prepare (no operator upload, no setup time), and the same solution as a context
    prepared from scratch with those knowns; a warm start from a nearby solution converges too."""
    import time
    from bem_for_nonlinear_waves_amd import mesh as M
    m, d, phi, phin = small_tank
    kw = dict(p=10, leaf_max=32, precision=precision, preconditioner=1)
    g = _gpu_bvp(m, d, phi, phin, **kw)
    x_phi, info = g.solveSystemGMRES(tol=1e-11, max_iter=600)
    assert info["converged"] == 1
    uploaded = g.stats()["n_ops_uploaded"]
    phit, phint = M.tank_bc(m, d, 17, 4.0, 1.0)
    t = time.perf_counter()
    xt, it = g.solveSystemGMRES_phi_t(phit, phint, tol=1e-11, max_iter=600)
    t_phit = time.perf_counter() - t
    assert it["converged"] == 1 and g.stats()["n_ops_uploaded"] == uploaded
    fresh = _gpu_bvp(m, d, phit, phint, **kw)
    xf, _ = fresh.solveSystemGMRES(tol=1e-11, max_iter=600)
    print(f"[gpu] phi_t solve on the resident plan {t_phit*1e3:.1f} ms (setup of a fresh context {fresh.time_setup*1e3:.1f} ms)")
    assert np.linalg.norm(xt - xf) <= 1e-8 * np.linalg.norm(xf)
    xw, iw = g.solveSystemGMRES_phi_t(phit, phint, x0=xf + 1e-6 * np.abs(xf).max(), tol=1e-11, max_iter=600)
    assert iw["converged"] == 1 and np.linalg.norm(xw - xf) <= 1e-8 * np.linalg.norm(xf)
    g.close()
    fresh.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_near_field_beside_fmm_matches_sequential(small_tank, precision):
    """The near-field columns on the second stream (overlap_near=1, default) give the same matvec as
    the single GEMV after L2L (overlap_near=0), eager and replayed from the captured graph."""
    m, d, phi, phin = small_tank
    x = np.random.default_rng(31).standard_normal(d.n_dof)
    ys = []
    for ov, gr in ((0, 0), (1, 0), (1, 1)):
        g = _gpu_bvp(m, d, phi, phin, p=10, leaf_max=32, precision=precision, overlap_near=ov, use_graphs=gr)
        mv = g.compute_Ax_fp32 if precision == 1 else g.compute_Ax
        ys.append([mv(x) for _ in range(3)])
        g.close()
    ref = ys[0][0]
    for run in ys:
        for y in run:
            # the two passes add the same terms in another grouping: last-bit differences only
            assert np.linalg.norm(y - ref) <= 1e-12 * np.linalg.norm(ref)
