"""ROBIN DOFs: the frequency-domain free surface phi_n = kappa phi (+ g), main_freq_domain.cpp:403-426, 543-555.

Pins (oracle, CPU):
* the DOF rule of `mesh.robin_dof_kind` (face-keyed DOFs by their face, null-key DOFs by their point);
* a linear harmonic field with an inhomogeneous Robin top (g = d phi/dn - kappa phi) and Neumann walls
  is reproduced as closely as the all-Dirichlet/Neumann linear field of test_oracle;
* the standing wave phi = e^{kz} cos(kx) satisfies phi_z = k phi on a flat top: the reference's
  homogeneous form (g = 0, kappa = k) with exact Neumann walls converges to it under refinement;
* the FMM-consistent dense operator has the Robin columns kappa_j G - Gn (its product with x equals the
  FMM matvec) and its diagonal is the Jacobi diagonal.
On the GPU the matvec, right-hand side and solves (Jacobi, Schwarz, FP32 + refinement) match the oracle.
"""
import dataclasses

import numpy as np
import pytest

from bem_for_nonlinear_waves_amd import mesh as M

KAPPA = 1.7


def _robin_problem(nx, ny, nz, field):
    m0 = M.box_tank(2.0, 1.0, 1.0, nx, ny, nz)
    top = np.isclose(m0.X[m0.tri].mean(axis=1)[:, 2], m0.X[:, 2].max()) & (m0.face_normals()[:, 2] > 0.999)
    m = M.Mesh(m0.X, m0.tri, np.zeros(m0.n_faces, bool), m0.face_group)  # every face Neumann-type
    d = M.build_dofs(m)
    d = dataclasses.replace(d, dof_kind=M.robin_dof_kind(m, d, top))
    _, _, phi_ex, phin_ex = M.analytic_bc(m, d, field)
    robin = d.dof_kind == M.ROBIN
    kappa = np.where(robin, KAPPA if field == "linear" else np.pi / 2.0, 0.0)
    g = np.where(robin, phin_ex - kappa * phi_ex, phin_ex)  # Robin: g; Neumann: phi_n
    return m, d, top, kappa, phi_ex.copy(), g, phi_ex, phin_ex


def test_robin_dof_rule():
    m, d, top, kappa, *_ = _robin_problem(8, 4, 4, "linear")
    robin = d.dof_kind == M.ROBIN
    assert robin.any() and (d.dof_kind == M.NEUMANN).any() and not (d.dof_kind == M.DIRICHLET).any()
    keyed = d.dof_face >= 0
    assert np.array_equal(robin[keyed], top[d.dof_face[keyed]])
    on_top = np.zeros(m.n_nodes, bool)
    on_top[m.tri[top].reshape(-1)] = True
    assert np.array_equal(robin[~keyed], on_top[d.dof_point[~keyed]])
    with pytest.raises(ValueError):
        M.robin_dof_kind(M.box_tank(2.0, 1.0, 1.0, 8, 4, 4), d, top)  # the box tank's top is Dirichlet


def test_robin_linear_field_reproduced(oracle_mod):
    m, d, top, kappa, phi, g, phi_ex, _ = _robin_problem(12, 6, 6, "linear")
    o = oracle_mod.Oracle(m, d, phi, g, p=10, leaf_max=16)
    o.set_robin(kappa)
    x, info = o.solve(tol=1e-10, max_iter=800, precond=1)
    assert info["rc"] == 0 and info["residual"] <= 1e-10
    assert np.linalg.norm(x - phi_ex) / np.linalg.norm(phi_ex) < 1e-3  # every unknown is phi here


def test_robin_standing_wave_converges(oracle_mod):
    errs = []
    for n in (6, 12):
        m, d, top, kappa, phi, g, phi_ex, _ = _robin_problem(2 * n, n, n, "exp")
        robin = d.dof_kind == M.ROBIN
        assert np.abs(g[robin]).max() < 1e-12  # phi_z = k phi on the flat top: the homogeneous form
        o = oracle_mod.Oracle(m, d, phi, np.where(robin, 0.0, g), p=10, leaf_max=16)
        o.set_robin(kappa)
        x, info = o.solve(tol=1e-10, max_iter=800, precond=1)
        assert info["rc"] == 0 and info["residual"] <= 1e-10
        errs.append(np.linalg.norm(x - phi_ex) / np.linalg.norm(phi_ex))
    print("standing wave relative errors", errs)
    assert errs[1] < 0.5 * errs[0] and errs[1] < 2e-2


def test_robin_dense_operator(oracle_mod):
    m, d, top, kappa, phi, g, *_ = _robin_problem(8, 4, 4, "linear")
    o = oracle_mod.Oracle(m, d, phi, g, p=10, leaf_max=16)
    o.set_robin(kappa)
    A = o.dense()
    x = np.random.default_rng(3).standard_normal(d.n_dof)
    y = o.matvec(x)
    assert np.linalg.norm(A @ x - y) <= 1e-5 * np.linalg.norm(y)
    o2 = oracle_mod.Oracle(m, d, phi, g, p=10, leaf_max=16)  # kappa = 0: the Robin columns lose kappa G
    A0 = o2.dense()
    robin = d.dof_kind == M.ROBIN
    dA = A - A0
    assert np.abs(dA[:, ~robin]).max() == 0.0 and np.abs(dA[:, robin]).max() > 0.0
    assert np.allclose(1.0 / o.diag_inv(), np.diag(A), rtol=1e-12, atol=0)


@pytest.mark.gpu
@pytest.mark.parametrize("precond,precision", [(1, 0), (2, 0), (2, 1)])
def test_robin_gpu_parity(oracle_mod, precond, precision):
    from bem_for_nonlinear_waves_amd.bvp import BEM_BVP
    m, d, top, kappa, phi, g, phi_ex, _ = _robin_problem(12, 6, 6, "linear")
    o = oracle_mod.Oracle(m, d, phi, g, p=10, leaf_max=16)
    o.set_robin(kappa)
    print(f"[gpu] robin n_dof={d.n_dof} precond={precond} precision={precision}", flush=True)
    b = BEM_BVP(m, d, p=10, leaf_max=16, preconditioner=precond, precision=precision)
    b.set_bc(phi, g)
    b.set_robin(kappa)
    b.prepare()
    x = np.random.default_rng(8).standard_normal(d.n_dof)
    y, yo = b.compute_Ax(x), o.matvec(x)
    assert np.linalg.norm(y - yo) <= 1e-10 * np.linalg.norm(yo)
    rb, ro = b.rhs(), o.rhs()
    assert np.linalg.norm(rb - ro) <= 1e-10 * np.linalg.norm(ro)
    xs, info = b.solveSystemGMRES(tol=1e-10, max_iter=800)
    assert info["converged"] == 1
    if precond == 2:
        o.set_schwarz(*b.schwarz_blocks())
    xo, _ = o.solve(tol=1e-10, max_iter=800, precond=precond)
    assert np.linalg.norm(xs - xo) <= 1e-8 * np.linalg.norm(xo)
    assert np.linalg.norm(xs - phi_ex) / np.linalg.norm(phi_ex) < 1e-3
    b.close()


@pytest.mark.gpu
def test_complex_data_on_robin_problem(oracle_mod):
    """Complex boundary data (the diffraction problem's form): phi = e^{kz} e^{ikx} is harmonic with
    phi_z = k phi on the flat top, so a homogeneous Robin top (kappa = k) with the exact complex phi_n on
    the walls reproduces it; the complex solve equals the oracle's real and imaginary solves."""
    from bem_for_nonlinear_waves_amd.bvp import BEM_BVP
    m, d, top, kappa, *_ = _robin_problem(24, 12, 12, "exp")
    k = np.pi / 2.0
    X = m.X[d.dof_point]
    nrm = M.dof_normals(m, d)
    ph = np.exp(k * X[:, 2]) * np.exp(1j * k * X[:, 0])
    grad = np.stack([1j * k * ph, np.zeros(len(X)), k * ph], axis=1)
    phn = np.einsum("ij,ij->i", grad, nrm)
    robin = d.dof_kind == M.ROBIN
    g = np.where(robin, 0.0, phn)
    print(f"[gpu] complex robin n_dof={d.n_dof}", flush=True)
    b = BEM_BVP(m, d, p=10, leaf_max=16, preconditioner=2)
    b.set_bc(ph.real, g.real)
    b.set_robin(kappa)
    x, infos = b.solveGMRES_complex(ph, g, tol=1e-10, max_iter=800)
    assert all(i["converged"] == 1 for i in infos)
    err = np.linalg.norm(x - ph) / np.linalg.norm(ph)
    print(f"[gpu] complex standing wave error {err:.3e}", flush=True)
    assert err < 3e-2
    parts = []
    for part in (np.real, np.imag):
        o = oracle_mod.Oracle(m, d, part(ph), part(g), p=10, leaf_max=16)
        o.set_robin(kappa)
        parts.append(o.solve(tol=1e-10, max_iter=800, precond=1)[0])
    xo = parts[0] + 1j * parts[1]
    assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo)
    b.close()


@pytest.mark.gpu
def test_robin_error_paths():
    """A non-finite kappa on a ROBIN DOF is rejected; a new kappa on a prepared context requires prepare
    again (the operator and its preconditioner depend on it); DOF kinds past 3 are rejected."""
    from bem_for_nonlinear_waves_amd import _lib
    from bem_for_nonlinear_waves_amd.bvp import BEM_BVP, BEMError
    m, d, top, kappa, phi, g, *_ = _robin_problem(8, 4, 4, "linear")
    print(f"[gpu] robin error paths n_dof={d.n_dof}", flush=True)
    b = BEM_BVP(m, d, p=8, leaf_max=16)
    bad = kappa.copy()
    bad[np.argmax(d.dof_kind == M.ROBIN)] = np.nan
    with pytest.raises(BEMError) as e:
        b.set_robin(bad)
    assert e.value.code == _lib.E_NONFINITE
    b.set_bc(phi, g)
    b.set_robin(kappa)
    b.prepare()
    y1 = b.compute_Ax(np.ones(d.n_dof))
    b.set_robin(2.0 * kappa)
    with pytest.raises(BEMError) as e:
        b.compute_Ax(np.ones(d.n_dof))
    assert e.value.code == _lib.E_STATE
    b.prepare()
    y2 = b.compute_Ax(np.ones(d.n_dof))
    assert np.linalg.norm(y2 - y1) > 1e-6 * np.linalg.norm(y1)  # the new kappa is in the operator
    b.close()
    d4 = dataclasses.replace(d, dof_kind=np.where(d.dof_kind == M.ROBIN, 4, d.dof_kind).astype(np.int32))
    b4 = BEM_BVP(m, d4, p=8, leaf_max=16)
    with pytest.raises(BEMError) as e:
        b4.set_bc(phi, g)
        b4.prepare()
    assert e.value.code == _lib.E_INVALID
    b4.close()
