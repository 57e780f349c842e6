"""Pins the CPU oracle (oracle/oracle.cpp) before it is trusted as the checker.

Pinning status per oracle function:
  quadrature tables (GW5x5, Dunavant 1/3/6/7) — Gauss-Legendre closed form and the
      reference's own known-answer test (monomial exactness, lib/include/dunavant_rules.hpp:172-257)
  spherical harmonics Y_n^m — scipy.special.sph_harm_y (independent library), all |m| <= n <= 12
  grad R_n^m . n (P2M double layer) — central finite differences of R_n^m
  P2M / M2M / M2L / L2L / L2P (GR97 coefficients) — a point-charge and point-dipole chain converges
      exponentially to the direct sums 1/r and d/dn(1/r) as p grows
  octree + MAC + lists + near field + far field — FMM matvec equals the FMM-consistent dense operator
      (near weights + direct sums over far P2M sources) to truncation error, which falls with p
  rigid-mode alpha — geometric solid angle: 2 pi on smooth surface, pi on edges, pi/2 at box corners
  GMRES (MGS + Givens, reference stopping rule) — dense LU solution of the same operator
  the whole BIE (signs, DOF model, scaling) — a linear harmonic field is reproduced to discretisation
      error; coordinate-scaling invariance; DOF-permutation invariance
Not pinned by the reference itself (it cannot be built here: missing AAA tables, float128_t, LAPACK):
  agreement with the reference's own numerical phi/phi_n — "parity unpinned" against the reference binary.
"""
import math

import numpy as np
import pytest

from bem_for_nonlinear_waves_amd import mesh as M


def _gw5_table():
    # values as printed in lib/include/basic_constants.hpp:1326-1350 (same as oracle.cpp kGW5)
    t = [0.046910077030668, 0.230765344947158, 0.5, 0.769234655052841, 0.953089922969332]
    tab = np.array([
        [0.046910077030668, 0.046910077030668, 0.0140335872156063], [0.046910077030668, 0.230765344947158, 0.0283499999999991],
        [0.046910077030668, 0.5, 0.0336962680968792], [0.046910077030668, 0.769234655052841, 0.0283499999999991],
        [0.046910077030668, 0.953089922969332, 0.0140335872156063], [0.230765344947158, 0.046910077030668, 0.0283499999999991],
        [0.230765344947158, 0.230765344947158, 0.0572713510559978], [0.230765344947158, 0.5, 0.0680716331376877],
        [0.230765344947158, 0.769234655052841, 0.0572713510559978], [0.230765344947158, 0.953089922969332, 0.0283499999999991],
        [0.5, 0.046910077030668, 0.0336962680968792], [0.5, 0.230765344947158, 0.0680716331376877],
        [0.5, 0.5, 0.0809086419753086], [0.5, 0.769234655052841, 0.0680716331376877],
        [0.5, 0.953089922969332, 0.0336962680968792], [0.769234655052841, 0.046910077030668, 0.0283499999999991],
        [0.769234655052841, 0.230765344947158, 0.0572713510559978], [0.769234655052841, 0.5, 0.0680716331376877],
        [0.769234655052841, 0.769234655052841, 0.0572713510559978], [0.769234655052841, 0.953089922969332, 0.0283499999999991],
        [0.953089922969332, 0.046910077030668, 0.0140335872156063], [0.953089922969332, 0.230765344947158, 0.0283499999999991],
        [0.953089922969332, 0.5, 0.0336962680968792], [0.953089922969332, 0.769234655052841, 0.0283499999999991],
        [0.953089922969332, 0.953089922969332, 0.0140335872156063]])
    return t, tab


def test_gw5_is_gauss_legendre_product():
    _, tab = _gw5_table()
    x, w = np.polynomial.legendre.leggauss(5)
    x, w = 0.5 * (x + 1.0), 0.5 * w
    exp = np.array([[x[i], x[j], w[i] * w[j]] for i in range(5) for j in range(5)])
    assert np.abs(tab - exp).max() < 2e-14


def test_collapsed_gw5_integrates_singular_kernel():
    """sum ww (1-t0) J f(x(t0,t1)) integrates 1/r over a triangle with r measured from the collapsed vertex."""
    _, tab = _gw5_table()
    X = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    J = 1.0
    s = 0.0
    for t0, t1, ww in tab:
        N = np.array([t0, t1 * (1 - t0), (t0 - 1) * (t1 - 1)])
        xq = N @ X
        s += ww * (1 - t0) * J / np.linalg.norm(xq)
    # exact: integral over the unit right triangle of 1/r from the right-angle vertex = sqrt(2) asinh(1)
    assert abs(s - math.sqrt(2.0) * math.asinh(1.0)) < 1e-3


@pytest.mark.parametrize("q,deg", [(1, 1), (3, 2), (6, 4), (7, 5)])
def test_dunavant_monomial_exactness(q, deg):
    rules = {
        1: [[1 / 3, 1 / 3, 1 / 3, 0.5]],
        3: [[2 / 3, 1 / 6, 1 / 6, 1 / 6], [1 / 6, 2 / 3, 1 / 6, 1 / 6], [1 / 6, 1 / 6, 2 / 3, 1 / 6]],
        6: [[0.816847572980459, 0.091576213509771, 0.091576213509771, 0.054975871827661],
            [0.091576213509771, 0.816847572980459, 0.091576213509771, 0.054975871827661],
            [0.091576213509771, 0.091576213509771, 0.816847572980459, 0.054975871827661],
            [0.108103018168070, 0.445948490915965, 0.445948490915965, 0.111690794839006],
            [0.445948490915965, 0.108103018168070, 0.445948490915965, 0.111690794839006],
            [0.445948490915965, 0.445948490915965, 0.108103018168070, 0.111690794839006]],
        7: [[1 / 3, 1 / 3, 1 / 3, 0.1125],
            [0.059715871789770, 0.470142064105115, 0.470142064105115, 0.066197076394253],
            [0.470142064105115, 0.059715871789770, 0.470142064105115, 0.066197076394253],
            [0.470142064105115, 0.470142064105115, 0.059715871789770, 0.066197076394253],
            [0.797426985353087, 0.101286507323456, 0.101286507323456, 0.062969590272414],
            [0.101286507323456, 0.797426985353087, 0.101286507323456, 0.062969590272414],
            [0.101286507323456, 0.101286507323456, 0.797426985353087, 0.062969590272414]],
    }
    R = np.array(rules[q])
    for a in range(deg + 1):
        for b in range(deg + 1 - a):
            for c in range(deg + 1 - a - b):
                got = np.sum(R[:, 3] * R[:, 0] ** a * R[:, 1] ** b * R[:, 2] ** c)
                exact = math.factorial(a) * math.factorial(b) * math.factorial(c) / math.factorial(a + b + c + 2)
                assert abs(got - exact) < 1e-12, (q, a, b, c)


def test_sph_harmonics_vs_scipy(oracle_mod):
    from scipy.special import sph_harm_y
    P = 12
    rng = np.random.default_rng(11)
    for _ in range(5):
        x = rng.standard_normal(3)
        Y = oracle_mod.sph_harmonics(x, P)
        th = math.atan2(math.hypot(x[0], x[1]), x[2])
        ph = math.atan2(x[1], x[0])
        for n in range(P + 1):
            for m in range(-n, n + 1):
                ref = math.sqrt(4 * math.pi / (2 * n + 1)) * sph_harm_y(n, abs(m), th, ph)
                if m < 0:
                    ref = np.conj(ref)
                assert abs(Y[n * (n + 1) + m] - ref) < 1e-13


def test_grad_solid_harmonic_finite_difference(oracle_mod):
    P = 10
    rng = np.random.default_rng(5)
    for x in [rng.standard_normal(3) * 0.4, np.array([0.0, 0.0, 0.35])]:
        nh = rng.standard_normal(3)
        nh /= np.linalg.norm(nh)

        def R(y):
            Y = oracle_mod.sph_harmonics(y, P)
            r = np.linalg.norm(y)
            return np.array([Y[n * (n + 1) + m] * r ** n for n in range(P + 1) for m in range(-n, n + 1)])

        h = 1e-5
        fd = (R(x + h * nh) - R(x - h * nh)) / (2 * h)
        g = oracle_mod.grad_R_dot_n(x, nh, P)
        assert np.all(np.isfinite(g))
        assert np.abs(fd - g).max() < 1e-7 * max(1.0, np.abs(g).max())


def test_translation_chain_converges_to_direct_sum(oracle_mod):
    rng = np.random.default_rng(0)
    c1 = np.array([0.1, 0.05, -0.02])
    c2 = np.zeros(3)
    c3 = np.array([1.2, 0.7, -0.4])
    c4 = c3 + np.array([0.05, -0.05, 0.05])
    ys = c1 + 0.08 * rng.standard_normal((5, 3))
    qs = rng.standard_normal(5)
    nhs = rng.standard_normal((5, 3))
    nhs /= np.linalg.norm(nhs, axis=1)[:, None]
    x = c4 + np.array([0.03, 0.02, -0.04])
    d = x - ys
    r = np.linalg.norm(d, axis=1)
    single = np.sum(qs / r)
    double = np.sum(qs * np.einsum("ij,ij->i", d, nhs) / r ** 3)
    errs = []
    for P in (4, 8, 12, 16):
        out = oracle_mod.translate_chain(ys, qs, nhs, c1, c2, c3, c4, x, P)
        errs.append((abs(out[0] - single) / abs(single), abs(out[1] - double) / abs(double)))
        assert abs(out[2]) < 1e-14 and abs(out[3]) < 1e-14  # real fields stay real
    for a, b in zip(errs, errs[1:]):
        assert b[0] < a[0] * 0.1 and b[1] < a[1] * 0.1
    assert errs[-1][0] < 1e-12 and errs[-1][1] < 1e-11


def test_alpha_is_solid_angle(oracle_mod, tiny_tank):
    m, d, phi, phin = tiny_tank
    o = oracle_mod.Oracle(m, d, phi, phin, p=10, leaf_max=16)
    a = o.alpha()
    # unperturbed box: count incident planes per point
    box = M.box_tank(2.0, 1.0, 1.0, 12, 6, 6)
    db = M.build_dofs(box)
    ob = oracle_mod.Oracle(box, db, *M.tank_bc(box, db, 2, 2.0, 1.0), p=10, leaf_max=16)
    ab = ob.alpha()
    X = box.X[db.dof_point]
    nfaces_on = sum(np.isclose(X[:, k], v) for k, v in [(0, 0), (0, 2), (1, 0), (1, 1), (2, 0), (2, 1)])
    exp = np.select([nfaces_on == 1, nfaces_on == 2, nfaces_on == 3], [2 * math.pi, math.pi, math.pi / 2])
    assert np.abs(ab - exp).max() < 2e-4 * 2 * math.pi
    assert np.all(np.isfinite(a))


def test_fmm_matches_dense_operator(oracle_mod, small_tank):
    m, d, phi, phin = small_tank
    rng = np.random.default_rng(1)
    x = rng.standard_normal(d.n_dof)
    errs = []
    for p in (4, 8):
        o = oracle_mod.Oracle(m, d, phi, phin, p=p, leaf_max=32)
        A = o.dense()
        y = o.matvec(x)
        errs.append(np.linalg.norm(y - A @ x) / np.linalg.norm(A @ x))
    assert errs[1] < errs[0] * 0.05
    assert errs[1] < 1e-8


def test_gmres_matches_dense_lu(oracle_mod, tiny_tank):
    m, d, phi, phin = tiny_tank
    o = oracle_mod.Oracle(m, d, phi, phin, p=10, leaf_max=16)
    x, info = o.solve(tol=1e-12, max_iter=600)
    assert info["rc"] == 0 and info["residual"] <= 1e-12
    A, b = o.dense(), o.rhs()
    L = o.stats()["L"]
    xl = np.linalg.solve(A, b)
    xl = np.where(d.dof_kind == 0, xl / L, xl)
    assert np.linalg.norm(x - xl) / np.linalg.norm(xl) < 1e-9


def test_linear_field_reproduced(oracle_mod):
    m = M.box_tank(2.0, 1.0, 1.0, 12, 6, 6)
    d = M.build_dofs(m)
    phi, phin, phi_ex, phin_ex = M.analytic_bc(m, d, "linear")
    o = oracle_mod.Oracle(m, d, phi, phin, p=10, leaf_max=16)
    x, info = o.solve(tol=1e-10, max_iter=600)
    ex = np.where(d.dof_kind == 0, phin_ex, phi_ex)
    assert np.linalg.norm(x - ex) / np.linalg.norm(ex) < 1e-3


def test_scaling_invariance(oracle_mod, tiny_tank):
    """The solver works in coordinates scaled by the bbox diagonal: scaling the geometry by s leaves phi
    unchanged and scales the computed phi_n by 1/s (GMRES_FMM.hpp:3076-3091, 3462-3486)."""
    m, d, phi, phin = tiny_tank
    s = 3.0
    m2 = M.Mesh(m.X * s, m.tri, m.face_dirichlet, m.face_group)
    o1 = oracle_mod.Oracle(m, d, phi, phin, p=8, leaf_max=16)
    o2 = oracle_mod.Oracle(m2, d, phi, phin / s, p=8, leaf_max=16)
    b1, b2 = o1.rhs(), o2.rhs()
    assert np.linalg.norm(b1 - b2) <= 1e-12 * np.linalg.norm(b1)


def test_dof_permutation_invariance(oracle_mod, tiny_tank):
    m, d, phi, phin = tiny_tank
    rng = np.random.default_rng(4)
    perm = rng.permutation(d.n_dof)  # new index of old DOF i is inv[i]
    inv = np.empty_like(perm)
    inv[perm] = np.arange(d.n_dof)
    d2 = M.DofModel(d.n_dof, inv[d.face_dof].astype(np.int32), d.dof_kind[perm], d.dof_point[perm], d.dof_face[perm],
                    d.point_class, d.point_multiple)
    o1 = oracle_mod.Oracle(m, d, phi, phin, p=8, leaf_max=16)
    o2 = oracle_mod.Oracle(m, d2, phi[perm], phin[perm], p=8, leaf_max=16)
    x = rng.standard_normal(d.n_dof)
    y1 = o1.matvec(x)
    y2 = o2.matvec(x[perm])
    assert np.linalg.norm(y2 - y1[perm]) <= 1e-12 * np.linalg.norm(y1)


def test_planted_sign_error_is_caught(oracle_mod):
    """A flipped double-layer sign (a plausible mistake) must break the linear-field reproduction."""
    m = M.box_tank(2.0, 1.0, 1.0, 12, 6, 6)
    d = M.build_dofs(m)
    phi, phin, phi_ex, phin_ex = M.analytic_bc(m, d, "linear")
    o = oracle_mod.Oracle(m, d, phi, -phin, p=8, leaf_max=16)  # wrong normal convention
    x, _ = o.solve(tol=1e-8, max_iter=400)
    ex = np.where(d.dof_kind == 0, phin_ex, phi_ex)
    assert np.linalg.norm(x - ex) / np.linalg.norm(ex) > 1e-2


def test_cached_m2l_matches_per_pair(oracle_mod, small_tank):
    """The CPU-baseline mode (M2L coefficients cached per level and lattice offset) evaluates the
    same operator as the per-pair SimpleM2L evaluation."""
    m, d, phi, phin = small_tank
    o = oracle_mod.Oracle(m, d, phi, phin, p=8, leaf_max=32)
    x = np.random.default_rng(9).standard_normal(d.n_dof)
    y0 = o.matvec(x)
    o.set_m2l_cache(True)
    y1 = o.matvec(x)
    assert np.linalg.norm(y1 - y0) <= 1e-13 * np.linalg.norm(y0)
