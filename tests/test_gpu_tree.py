"""The plan's tree stage on the GPU (tree.hip: level-synchronous radix sorts on path keys) against the host
builder that follows the reference's generateTree / buildNearAndM2LLists (plan.cpp plan_tree_lists): equal tree
digests (level offsets, parents, leaf flags, box centres, depth-first face order, face ranges, target boxes,
near and M2L lists), equal box / pair / near-entry counts, and matvecs equal to rounding, from small tanks to the 209k
cylinder tank of the benchmark; and a re-prepare after moved nodes gives the host builder's plan of the moved
mesh. Equality is exact: the box arithmetic is shared and compiled without FMA contraction on both sides.
"""
import time

import numpy as np
import pytest

from bem_for_nonlinear_waves_amd import mesh as M

pytestmark = pytest.mark.gpu


def _prepared(m, d, phi, phin, g, **kw):
    from bem_for_nonlinear_waves_amd.bvp import BEM_BVP
    b = BEM_BVP(m, d, tree_on_gpu=g, **kw)
    b.set_bc(phi, phin)
    t = time.perf_counter()
    b.prepare()
    return b, time.perf_counter() - t


CASES = [
    ("rough tank 12x6x6 leaf 16", lambda: M.perturb_free_surface(M.box_tank(2.0, 1.0, 1.0, 12, 6, 6), 2, 0.02), 16, 7),
    ("rough tank 12x6x6 leaf 2 depth 3", lambda: M.perturb_free_surface(M.box_tank(2.0, 1.0, 1.0, 12, 6, 6), 2, 0.02),
     2, 3),
    ("flat tank 42x11x11 leaf 32", lambda: M.box_tank(4.0, 1.0, 1.0, 42, 11, 11), 32, 7),
    ("cylinder tank h 0.08 leaf 8", lambda: M.cylinder_tank(Lx=3.0, Ly=2.0, Lz=1.0, R=0.25, draft=0.5, h=0.08), 8, 7),
]


@pytest.mark.parametrize("name,make,leaf,lvl", CASES, ids=[c[0] for c in CASES])
def test_gpu_tree_equals_host_tree(name, make, leaf, lvl):
    m = make()
    d = M.build_dofs(m)
    phi, phin = M.tank_bc(m, d, 3, 2.0, 1.0)
    print(f"[gpu] tree {name}: {m.n_nodes} nodes {m.n_faces} faces", flush=True)
    x = np.random.default_rng(7).standard_normal(d.n_dof)
    out = []
    for g in (0, 1):
        b, t = _prepared(m, d, phi, phin, g, p=6, leaf_max=leaf, max_level=lvl, preconditioner=1, m2l_tile_pairs=16)
        st = b.stats()
        out.append((b.tree_digest(), st["n_boxes"], st["n_leaves"], st["n_m2l_pairs"], st["near_nnz"], b.compute_Ax(x),
                    st["tree_on_gpu"]))
        print(f"[gpu]   {'GPU' if g else 'host'} tree: digest {out[-1][0]:016x} boxes {st['n_boxes']} "
              f"pairs {st['n_m2l_pairs']} nnz {st['near_nnz']} tree {st['t_tree']:.4f}s lists {st['t_lists']:.4f}s",
              flush=True)
        b.close()
    assert out[0][6] == 0 and out[1][6] == 1
    assert out[0][:5] == out[1][:5]
    # same plan, same tiles: the matvecs agree to rounding (the M2L reduction's launch order is timing-free, but
    # two contexts may pick different kernel instances for their first launches)
    rel = np.linalg.norm(out[0][5] - out[1][5]) / np.linalg.norm(out[0][5])
    print(f"[gpu]   matvec host plan vs GPU plan: {rel:.2e} (bit-equal: {np.array_equal(out[0][5], out[1][5])})",
          flush=True)
    assert rel <= 1e-13


def test_gpu_tree_at_209k_and_after_moved_nodes():
    """The benchmark's cylinder tank (209,016 nodes, leaf 32, depth 7): the GPU plan equals the host plan, and
    after a re-prepare with the free surface moved by N(0, 1e-3 h) it equals the host plan of the moved mesh."""
    import bench
    m, d, phi, phin = bench.make_workload(0.036)
    print(f"[gpu] tree 209k: {m.n_nodes} nodes {m.n_faces} faces", flush=True)
    kw = dict(p=8, leaf_max=32, max_level=7)
    bh, th = _prepared(m, d, phi, phin, 0, **kw)
    sh = bh.stats()
    hh = bh.tree_digest()
    bh.close()
    bg, tg = _prepared(m, d, phi, phin, 1, **kw)
    sg = bg.stats()
    print(f"[gpu]   host: boxes {sh['n_boxes']} pairs {sh['n_m2l_pairs']} nnz {sh['near_nnz']} tree+lists "
          f"{sh['t_tree'] + sh['t_lists']:.3f}s; GPU: tree+lists {sg['t_tree'] + sg['t_lists']:.3f}s", flush=True)
    assert (sh["n_boxes"], sh["n_leaves"], sh["n_m2l_pairs"], sh["near_nnz"]) == (
        sg["n_boxes"], sg["n_leaves"], sg["n_m2l_pairs"], sg["near_nnz"])
    assert bg.tree_digest() == hh
    mk = M.perturb_free_surface(m, 17, 1e-3 * 0.036)
    bg.set_geometry(mk, d)
    bg.set_bc(phi, phin)
    t = time.perf_counter()
    bg.prepare()
    tr = time.perf_counter() - t
    hg = bg.tree_digest()
    sg = bg.stats()
    bg.close()
    bh, _ = _prepared(mk, d, phi, phin, 0, **kw)
    print(f"[gpu]   re-prepare after moved nodes {tr:.3f}s (tree {sg['t_tree']:.3f} lists {sg['t_lists']:.3f}); "
          f"digest equal to the host plan of the moved mesh: {bh.tree_digest() == hg}", flush=True)
    assert bh.tree_digest() == hg
    bh.close()
