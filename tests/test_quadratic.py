"""True quadratic faces (isTrueQuadraticElement, bem/BEM_solveBVP_GMRES_FMM.hpp:1311-1657, midpoint rows :111-144,
DOFs bem/BEM_setBoundaryTypes.hpp:190-233): the oracle restatement pinned by what the mathematics fixes, then the
library's host code and CPU kernel versions (tests/hipstub) against the oracle.

Pins of the oracle:
* TriShape<6> (basic_arithmetic_array_operations.hpp:1467-1566): Kronecker at its six nodes, partition of unity,
  every quadratic polynomial reproduced, the derivatives equal to central differences;
* the near-field weights of one curved face (a patch of a cylinder, midpoints on the cylinder) against adaptive
  integration (scipy) of the same integrals, sum_k N_k |dX/dt0 x dX/dt1| / r and N_k (-R . (dX/dt0 x dX/dt1)) / r^3:
  from a vertex of the face (the GW5x5 rule collapsed at it), from an edge midpoint (the two halves of the face),
  and from a far point (the 6-point Dunavant rule); the target node's own double-layer weight is dropped;
* the rigid mode: on a box tank of quadratic faces, alpha at every vertex and edge-midpoint node is the solid
  angle of the box there (2 pi on a face, pi on an edge, pi/2 at a corner);
* accuracy: on the box tank, the standing wave cosh(k z) cos(k x) (zero Neumann data) is reproduced on the walls
  more than ten times better with quadratic faces than with linear ones at the same mesh.
The library (its host plan and the CPU versions of its kernels, built against the stub HIP runtime) then equals the
oracle: matvec, right-hand side and alpha to 1e-12, with the tree built on the host and by the device driver.
"""
import math
import os
import subprocess
import sys
import textwrap

import numpy as np
import pytest

from bem_for_nonlinear_waves_amd import mesh as M

HERE = os.path.dirname(os.path.abspath(__file__))
NODES = [(1.0, 0.0), (0.0, 1.0), (0.0, 0.0), (0.5, 0.5), (0.0, 0.5), (0.5, 0.0)]  # (t0, t1) of the six nodes


def test_tri6_shape_functions(oracle_mod):
    for a, (t0, t1) in enumerate(NODES):
        N, _, _ = oracle_mod.tri6_shape(t0, t1)
        assert np.allclose(N, np.eye(6)[a], atol=1e-15)
    rng = np.random.default_rng(3)
    for _ in range(20):
        t0, t1 = rng.uniform(0, 0.5, 2)
        N, D0, D1 = oracle_mod.tri6_shape(t0, t1)
        assert abs(N.sum() - 1.0) < 1e-14
        c = rng.standard_normal(6)  # a quadratic polynomial in (t0, t1), sampled at the nodes and interpolated
        poly = lambda u, v: c[0] + c[1] * u + c[2] * v + c[3] * u * u + c[4] * u * v + c[5] * v * v
        vals = np.array([poly(u, v) for u, v in NODES])
        assert abs(N @ vals - poly(t0, t1)) < 1e-13
        h = 1e-6
        fd0 = (oracle_mod.tri6_shape(t0 + h, t1)[0] - oracle_mod.tri6_shape(t0 - h, t1)[0]) / (2 * h)
        fd1 = (oracle_mod.tri6_shape(t0, t1 + h)[0] - oracle_mod.tri6_shape(t0, t1 - h)[0]) / (2 * h)
        assert np.abs(fd0 - D0).max() < 1e-8 and np.abs(fd1 - D1).max() < 1e-8


def _cyl(theta, z, R=1.0):
    return np.array([R * math.cos(theta), R * math.sin(theta), z])


def _one_face(extra):
    """One curved 6-node face on the unit cylinder plus isolated target nodes; a Neumann DOF per node."""
    v = [(0.0, 0.0), (0.45, 0.0), (0.2, 0.35)]
    P = [_cyl(a, z) for a, z in v]
    for i, j in ((0, 1), (1, 2), (2, 0)):
        P.append(_cyl(0.5 * (v[i][0] + v[j][0]), 0.5 * (v[i][1] + v[j][1])))
    X = np.array(P + [np.asarray(e, float) for e in extra])
    n = X.shape[0]
    m = M.Mesh(X, np.arange(6, dtype=np.int32)[None, :], np.zeros(1, bool), np.zeros(1, np.int32))
    d = M.DofModel(n, np.arange(6, dtype=np.int32)[None, :], np.ones(n, np.int32), np.arange(n, dtype=np.int32),
                   np.full(n, -1, np.int32), np.ones(n, np.int8), np.zeros(n, bool))
    return m, d


def _exact_weights(O, Xs, x, split=None):
    """Adaptive integrals of N_k |cr| / r and N_k (-R . cr) / r^3 over the 6-node face Xs (scaled), in Duffy
    coordinates collapsed at vertex `split[1]` (split = ("vertex", v)) or on the two halves at edge midpoint
    `split[1]` (("mid", e)), else over the plain triangle."""
    from scipy import integrate

    def geom(b0, b1):
        N, D0, D1 = O.tri6_shape(b0, b1)
        xq = N @ Xs
        cr = np.cross(D0 @ Xs, D1 @ Xs)
        return N, xq, cr

    def integrand(b0, b1, jac, k, layer):
        N, xq, cr = geom(b0, b1)
        R = xq - x
        r = np.linalg.norm(R)
        return jac * N[k] * (np.linalg.norm(cr) / r if layer == 0 else -(R @ cr) / r ** 3)

    out = np.zeros((2, 6))
    for k in range(6):
        for layer in range(2):
            if split is None:
                f = lambda t1, t0: integrand(t0, t1, 1.0, k, layer)
                val, _ = integrate.dblquad(f, 0, 1, 0, lambda t0: 1 - t0, epsabs=1e-12, epsrel=1e-10)
            elif split[0] == "vertex":
                s = split[1]

                def f(u, t0):
                    b = [0.0, 0.0, 0.0]
                    b[s], b[(s + 1) % 3], b[(s + 2) % 3] = t0, u * (1 - t0), (1 - t0) * (1 - u)
                    return integrand(b[0], b[1], 1 - t0, k, layer)
                val, _ = integrate.dblquad(f, 0, 1, 0, 1, epsabs=1e-12, epsrel=1e-10)
            else:
                e = split[1]
                va, vb, vc = e, (e + 1) % 3, (e + 2) % 3
                val = 0.0
                for half in range(2):
                    def f(u, t0, half=half):
                        o = 1 - t0
                        b = [0.0, 0.0, 0.0]
                        if half == 0:
                            b[va], b[vb], b[vc] = 0.5 * t0, 0.5 * t0 + u * o, o * (1 - u)
                        else:
                            b[va], b[vb], b[vc] = 0.5 * t0 + o * (1 - u), 0.5 * t0, u * o
                        return integrand(b[0], b[1], 0.5 * o, k, layer)
                    v, _ = integrate.dblquad(f, 0, 1, 0, 1, epsabs=1e-12, epsrel=1e-10)
                    val += v
            out[layer, k] = val
    return out


def _dunavant6_weights(O, Xs, x):
    """The 6-point degree-4 Dunavant rule (Dunavant 1985, table: a = 0.445948490915965 with weight
    0.223381589678011, a = 0.091576213509771 with weight 0.109951743655322) applied to the same integrands."""
    out = np.zeros((2, 6))
    for a, w in ((0.445948490915965, 0.223381589678011), (0.091576213509771, 0.109951743655322)):
        for t0, t1 in ((a, a), (1 - 2 * a, a), (a, 1 - 2 * a)):
            N, D0, D1 = O.tri6_shape(t0, t1)
            cr = np.cross(D0 @ Xs, D1 @ Xs)
            R = N @ Xs - x
            r = np.linalg.norm(R)
            out[0] += 0.5 * w * N * np.linalg.norm(cr) / r
            out[1] += 0.5 * w * N * -(R @ cr) / r ** 3
    return out


# tolerances (single layer, double layer): the accuracy of the reference's rules on this face, GW5x5 on the
# Duffy-collapsed integrands (the double layer's R . cr / r^3 is the rougher one on a curved face; the two halves
# at an edge midpoint are wider angles from the target, measured 1e-4 / 9e-4 against 2e-8 / 4e-7 for GW10x10 on the
# same halves), the 6-point Dunavant rule at a point about three face sizes away (measured 6.7e-5 / 9.8e-4; there
# the oracle also equals an independent evaluation of that rule to rounding); a wrong shape function, sign,
# Jacobian or weight fails them by orders of magnitude
@pytest.mark.parametrize("target,split,tol", [(0, ("vertex", 0), (5e-6, 1e-4)), (1, ("vertex", 1), (5e-6, 1e-4)),
                                              (3, ("mid", 0), (3e-4, 5e-3)), (5, ("mid", 2), (3e-4, 5e-3)),
                                              (6, None, (2e-4, 3e-3))])
def test_quadratic_face_weights_vs_adaptive_integration(oracle_mod, target, split, tol):
    far = (0.6, 1.3, 0.9)
    m, d = _one_face([far])
    phi = np.zeros(d.n_dof)
    o = oracle_mod.Oracle(m, d, phi, phi, p=4, leaf_max=16)
    L = o.stats()["L"]
    Xs = m.X[:6] / L
    x = m.X[target] / L
    idx, wg, wgn = o.near_row(target)
    got = np.zeros((2, 6))
    for i, g, h in zip(idx, wg, wgn):
        if i < 6:
            got[0, i], got[1, i] = g, h
    ex = _exact_weights(oracle_mod, Xs, x, split)
    if target < 6:
        assert got[1, target] == 0.0  # the target node's own double-layer weight is dropped
        ex[1, target] = 0.0
    scale = np.abs(ex).max(axis=1, keepdims=True)
    err = np.abs(got - ex) / scale
    print(f"target {target}: max relative error single {err[0].max():.2e} double {err[1].max():.2e}")
    assert err[0].max() <= tol[0] and err[1].max() <= tol[1]
    if split is None:
        d6 = _dunavant6_weights(oracle_mod, Xs, x)
        assert np.abs(got - d6).max() <= 1e-13 * np.abs(d6).max()


def test_quadratic_alpha_is_solid_angle(oracle_mod):
    box = M.to_quadratic(M.box_tank(2.0, 1.0, 1.0, 8, 4, 4))
    d = M.build_dofs_quadratic(box)
    o = oracle_mod.Oracle(box, d, *M.tank_bc(box, d, 2, 2.0, 1.0), p=8, leaf_max=16)
    a = o.alpha()
    X = box.X[d.dof_point]
    on = sum(np.isclose(X[:, k], v) for k, v in [(0, 0), (0, 2), (1, 0), (1, 1), (2, 0), (2, 1)])
    exp = np.select([on == 1, on == 2, on == 3], [2 * math.pi, math.pi, math.pi / 2])
    err = np.abs(a - exp).max() / (2 * math.pi)
    print(f"alpha vs solid angle: {err:.2e} (vertices and edge midpoints)")
    # corners are exact to 5e-10; the largest error (2.0e-3 of 2 pi, against 8.5e-5 with linear faces) is at nodes
    # of a wall half a cell from the next wall, whose faces the rule of non-adjacent faces (Dunavant) sees from a
    # distance of half their size
    assert err < 3e-3


def test_quadratic_standing_wave_beats_linear(oracle_mod):
    errs = {}
    for order in (1, 2):
        m = M.box_tank(2.0, 1.0, 1.0, 8, 4, 4)
        if order == 2:
            m = M.to_quadratic(m)
            d = M.build_dofs_quadratic(m)
        else:
            d = M.build_dofs(m)
        phi, phin, phi_ex, phin_ex = M.standing_wave(m, d, 2.0)
        o = oracle_mod.Oracle(m, d, phi, phin, p=8, leaf_max=16)
        A, b = o.dense(), o.rhs()
        x = np.linalg.solve(A, b)
        walls = d.dof_kind != 0
        errs[order] = np.sqrt(np.mean((x[walls] - phi_ex[walls]) ** 2)) / np.abs(phi_ex[walls]).max()
    print(f"standing wave, phi on the walls (rms): linear {errs[1]:.2e}, quadratic {errs[2]:.2e}")
    assert errs[2] < errs[1] / 10


SCRIPT = textwrap.dedent("""
    import sys
    sys.path.insert(0, {root!r}); sys.path.insert(0, {oracle!r})
    import numpy as np
    from bem_for_nonlinear_waves_amd import _lib
    _lib.LIB_PATH = {lib!r}
    import oracle as O
    from bem_for_nonlinear_waves_amd import mesh as M
    from bem_for_nonlinear_waves_amd.bvp import BEM_BVP, BEMError
    m = M.to_quadratic(M.perturb_free_surface(M.box_tank(2.0, 1.0, 1.0, 8, 4, 4), 2, 0.02))
    d = M.build_dofs_quadratic(m)
    d.face_dof[5, 4] = -1  # one edge midpoint without a DOF on one face: its weight goes to the edge's vertices
    phi, phin = M.tank_bc(m, d, 3, 2.0, 1.0)
    o = O.Oracle(m, d, phi, phin, p=8, leaf_max=16)
    x = np.random.default_rng(1).standard_normal(d.n_dof)
    yo, ao, bo = o.matvec(x), o.alpha(), o.rhs()
    for g in (0, 1):
        b = BEM_BVP(m, d, p=8, leaf_max=16, element_order=2, tree_on_gpu=g, preconditioner=2, schwarz_max_core=16,
                    schwarz_max_block=64)
        b.set_bc(phi, phin)
        b.prepare()
        rel = lambda u, v: float(np.linalg.norm(u - v) / np.linalg.norm(v))
        e = (rel(b.compute_Ax(x), yo), rel(b.diag_coeffs(), ao), rel(b.rhs(), bo))
        print("tree_on_gpu", g, "matvec %.2e alpha %.2e rhs %.2e" % e, flush=True)
        assert max(e) <= 1e-12, e
        b.close()
    b = BEM_BVP(m, d, p=8, leaf_max=16, element_order=2, preconditioner=3)
    b.set_bc(phi, phin)
    try:
        b.prepare()
        raise SystemExit("ILU(0) accepted for quadratic faces")
    except BEMError:
        pass
    print("ok")
""")


def test_library_host_code_equals_oracle_on_quadratic_faces():
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "hipstub"), "lib"], check=True)
    root = os.path.dirname(HERE)
    r = subprocess.run([sys.executable, "-c", SCRIPT.format(root=root, oracle=os.path.join(root, "oracle"),
                                                            lib=os.path.join(HERE, "hipstub", "_build",
                                                                             "libbemfmm_stub.so"))],
                       capture_output=True, text=True, timeout=900, env=dict(os.environ, OMP_NUM_THREADS="4"))
    print(r.stdout)
    assert r.returncode == 0, r.stderr[-4000:]
    assert r.stdout.strip().endswith("ok")
