"""Benchmark: FMM-accelerated GMRES BVP solves on one MI355X.

Workload (one GPU): 20 x 4 x 2 m tank with a surface-piercing vertical cylinder
(R = 0.5 m, draft 1 m), linear triangles at h = 0.036 m -> ~209k nodes, ~223k DOFs;
free surface Dirichlet (seeded 8-mode phi), walls/bottom Neumann phi_n = 0, body
Neumann phi_n = n.U with U ~ N(0, 0.1) (seed 3); p = 10, theta = 0.25,
leaf <= 32 faces, depth 7 (reference default), GMRES tol 1e-10 (absolute, on the
left-preconditioned residual as in the reference), Jacobi scaling by the operator
diagonal (without it GMRES does not reach 1e-10 within 2000 iterations at this size).

A "step" is one BVP solve with the plan (tree, operators, near field, alpha)
resident in HBM: RHS + GMRES to tolerance, every matvec on the GPU. `value` is
solves/s over the timed steps; matvec Mnode/s (N / t_matvec) is timed separately
on device-resident vectors. Setup time (plan + near quadrature + alpha) is
reported beside it. The CPU baseline is the FP64 oracle restatement of the
reference algorithm on a bounded sample (see cpu_baseline.sample).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.append(ROOT)

METRIC = "BVP solves/s & FMM matvec Mnode/s at N=200k; p50/p99 solve latency"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def make_workload(h: float, seed: int = 3):
    from bem_for_nonlinear_waves_amd import mesh as M
    m = M.cylinder_tank(Lx=20.0, Ly=4.0, Lz=2.0, R=0.5, draft=1.0, h=h)
    m = M.perturb_free_surface(m, seed, 0.002)
    d = M.build_dofs(m)
    phi, phin = M.tank_bc(m, d, seed, 20.0, 4.0, body_velocity_sigma=0.1)
    return m, d, phi, phin


def cpu_baseline(p, leaf_max, max_level, gpu_iters, n_nodes_full, budget_s=25.0):
    """Oracle (FP64 OpenMP restatement) matvecs on a smaller tank of the same family; Mnode/s and the
    solves/s it implies at the full size with the GPU's iteration count (FMM cost is O(N))."""
    from oracle import oracle as O
    O.build()
    m, d, phi, phin = make_workload(0.12)
    t = time.perf_counter()
    o = O.Oracle(m, d, phi, phin, p=p, leaf_max=leaf_max, max_level=max_level)
    t_setup = time.perf_counter() - t
    x = np.random.default_rng(0).standard_normal(d.n_dof)
    n = 0
    t0 = time.perf_counter()
    while True:
        o.matvec(x)
        n += 1
        if time.perf_counter() - t0 > budget_s / 2 or n >= 20:
            break
    tm = (time.perf_counter() - t0) / n
    mnode = m.n_nodes / tm / 1e6
    t_solve_full = (gpu_iters + 3) * tm * (n_nodes_full / m.n_nodes)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    try:
        model = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        model = [l.split(":", 1)[1].strip() for l in model.splitlines() if l.startswith("Model name")][0]
    except Exception:
        model = "unknown"
    return {"value": 1.0 / t_solve_full, "unit": "solves/s", "cores": threads, "kind": "port",
            "matvec_mnode_per_s": mnode, "cpu": model,
            "sample": (f"oracle FP64 FMM matvec x{n} on the same tank family at h=0.12 ({m.n_nodes} nodes, "
                       f"setup {t_setup:.1f}s, {tm*1e3:.1f} ms/matvec); solves/s extrapolated linearly in N to "
                       f"{n_nodes_full} nodes at the GPU's {gpu_iters}+3 matvecs per solve")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--h", type=float, default=0.036)
    ap.add_argument("--p", type=int, default=10)
    ap.add_argument("--leaf", type=int, default=32)
    ap.add_argument("--max-level", type=int, default=7)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--max-iter", type=int, default=2000)
    ap.add_argument("--matvecs", type=int, default=20)
    ap.add_argument("--precond", choices=["none", "jacobi"], default="jacobi")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-only", action="store_true",
                    help="setup + a 256 MiB device copy (counter calibration) + the timed matvecs; no solves")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(a single BVP is solved on one MI355X; see the sharding notes in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import torch

    from bem_for_nonlinear_waves_amd import _lib
    from bem_for_nonlinear_waves_amd.bvp import BEM_BVP, device_name

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    log(f"device {device_name()}")
    t = time.perf_counter()
    m, d, phi, phin = make_workload(args.h)
    log(f"mesh: {m.n_nodes} nodes, {m.n_faces} faces, {d.n_dof} DOFs ({time.perf_counter() - t:.1f}s)")
    pc = _lib.PRECOND_JACOBI if args.precond == "jacobi" else _lib.PRECOND_NONE
    bvp = BEM_BVP(m, d, p=args.p, leaf_max=args.leaf, max_level=args.max_level, preconditioner=pc)
    bvp.set_bc(phi, phin)
    t = time.perf_counter()
    bvp.prepare()
    t_setup = time.perf_counter() - t
    st = bvp.stats()
    log(f"setup {t_setup:.2f}s: boxes={st['n_boxes']} leaves={st['n_leaves']} m2l_pairs={st['n_m2l_pairs']} "
        f"near_nnz={st['near_nnz']} device={st['device_bytes']/1e9:.1f}GB "
        f"(tree {st['t_tree']:.2f} lists {st['t_lists']:.2f} near_struct {st['t_near_struct']:.2f} "
        f"p2m {st['t_p2m']:.2f} ops {st['t_ops']:.2f} upload {st['t_upload']:.2f} near_quad {st['t_near_quad']:.2f} "
        f"alpha {st['t_alpha']:.2f})")

    # ---- matvec throughput on device-resident vectors, per-stage HIP event timing ----
    n = d.n_dof
    xs = torch.randn(n, dtype=torch.float64, device="cuda")
    ys = torch.empty(n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    bvp.profile()  # enable event recording on the library's stream
    for _ in range(2):
        bvp.matvec_device(xs.data_ptr(), ys.data_ptr())
    bvp.sync()
    t0 = time.perf_counter()
    for _ in range(args.matvecs):
        bvp.matvec_device(xs.data_ptr(), ys.data_ptr())
    bvp.sync()
    t_mv = (time.perf_counter() - t0) / args.matvecs
    stage = np.zeros(7)
    for _ in range(5):
        bvp.matvec_device(xs.data_ptr(), ys.data_ptr())
        stage += bvp.profile()  # synchronises on the last event of this matvec
    stage /= 5
    names = ["scatter", "p2m", "m2m", "m2l", "l2l", "near_l2p", "assemble"]
    log("matvec %.3f ms (%.2f Mnode/s); stages ms: %s" % (t_mv * 1e3, m.n_nodes / t_mv / 1e6,
        " ".join(f"{k}={v:.3f}" for k, v in zip(names, stage))))

    if args.profile_only:
        cal = torch.empty(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
        cal2 = torch.empty_like(cal)
        cal.fill_(1.0)
        torch.cuda.synchronize()
        cal2.copy_(cal)  # 256 MiB read + 256 MiB written: FETCH_SIZE/WRITE_SIZE calibration
        torch.cuda.synchronize()
        log("profile-only: done")
        return

    # ---- solves ----
    lat = []
    iters = []
    x0 = np.zeros(n)
    for i in range(args.warmup):
        x, info = bvp.solveSystemGMRES(x0, tol=args.tol, max_iter=args.max_iter, restart=args.max_iter)
        log(f"warmup solve {i}: iters={info['iters']} residual={info['residual']:.2e} {info['time_solve']:.2f}s")
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        ts = time.perf_counter()
        x, info = bvp.solveSystemGMRES(x0, tol=args.tol, max_iter=args.max_iter, restart=args.max_iter)
        lat.append(time.perf_counter() - ts)
        iters.append(info["iters"])
        log(f"solve {i}: iters={info['iters']} matvecs={info['n_matvec']} residual={info['residual']:.2e} "
            f"converged={info['converged']} {lat[-1]:.3f}s")
    torch.cuda.synchronize()
    t_total = time.perf_counter() - t_start
    solves_per_s = args.steps / t_total

    # ---- roofline of the dominant kernel (from the live HIP-event stage timings) ----
    C = (args.p + 1) ** 2
    m2l_flops = st["n_m2l_pairs"] * 2 * C * C * 2.0
    near_rows = m.n_nodes
    near_bytes = (st["near_nnz"] + near_rows * C) * 2 * 8.0  # dense G/Gn weights incl. L2P columns (FP64)
    dom = int(np.argmax(stage))
    if names[dom] == "m2l":
        ach = m2l_flops / (stage[dom] * 1e-3) / 1e12
        roof = {"kernel": "k_translate (M2L)", "bound": "mfma", "achieved": ach, "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": ach / FP64_PEAK_TFLOPS, "traffic": None}
    else:
        ach = near_bytes / (stage[dom] * 1e-3) / 1e9
        roof = {"kernel": f"k_block_gemv ({names[dom]})", "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": None}
    roof["stage_ms"] = dict(zip(names, [round(float(v), 4) for v in stage]))
    roof["m2l_tflops"] = m2l_flops / (stage[3] * 1e-3) / 1e12
    roof["near_gbs"] = near_bytes / (stage[5] * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(args.p, args.leaf, args.max_level, int(np.median(iters)), m.n_nodes)
            log(f"cpu baseline: {cpu['sample']}: {cpu['matvec_mnode_per_s']:.3f} Mnode/s")
        except Exception as e:  # the baseline never decides the GPU result
            log(f"cpu baseline failed: {e}")

    lat_ms = np.array(lat) * 1e3
    out = {
        "metric": METRIC,
        "value": solves_per_s,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_total / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"cylinder tank 20x4x2 m, R=0.5 draft=1, h={args.h} (configs[2] geometry), "
                               f"FP64 FMM-GMRES", "n_nodes": m.n_nodes, "n_faces": m.n_faces, "n_dof": n,
                   "p": args.p, "theta": 0.25, "leaf_max": args.leaf, "max_level": args.max_level, "tol": args.tol,
                   "precision": "fp64", "preconditioner": args.precond},
        "matvec_mnode_per_s": m.n_nodes / t_mv / 1e6,
        "matvec_ms": t_mv * 1e3,
        "p50_solve_ms": float(np.percentile(lat_ms, 50)),
        "p99_solve_ms": float(np.percentile(lat_ms, 99)),
        "gmres_iters": iters,
        "setup_s": t_setup,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
