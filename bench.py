"""Benchmark: FMM-accelerated GMRES BVP solves on one MI355X.

Workload (one GPU): 20 x 4 x 2 m tank with a surface-piercing vertical cylinder
(R = 0.5 m, draft 1 m), linear triangles at h = 0.036 m -> ~209k nodes, ~223k DOFs;
free surface Dirichlet (seeded 8-mode phi), walls/bottom Neumann phi_n = 0, body
Neumann phi_n = n.U with U ~ N(0, 0.1) (seed 3); p = 10, theta = 0.25,
leaf <= 32 faces, depth 7 (reference default), GMRES tol 1e-10 (absolute, on the
left-preconditioned residual as in the reference).

Default solver: FP32 near field / P2M / translations (FP32 MFMA) inside FP64
iterative refinement, restricted additive Schwarz preconditioner (blocks of <= 128
DOFs around cores of <= 64, the reference's SCHWARZ defaults). The FP64 path with
the same preconditioner is timed beside it (`fp64_path`); `--precond jacobi` gives
the diagonal scaling (about 4x the iterations at this size).

A "step" is one BVP solve with the plan (tree, operators, near field, alpha,
preconditioner) resident in HBM: RHS + GMRES to tolerance, every matvec on the GPU.
`value` is solves/s over the timed steps; matvec Mnode/s (N / t_matvec) is timed
separately on device-resident vectors. Setup time (plan + near quadrature + alpha +
preconditioner) is reported beside it. The CPU baseline is the FP64 oracle
restatement of the reference algorithm on a bounded sample (see cpu_baseline.sample).
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
PROFILES = os.path.join(ROOT, "profiles")


def device_code_hash() -> str:
    """sha256 (first 16 hex digits) of the device code (kernels.hip, kernels.hpp): counter figures under
    profiles/ are used only for the build they were measured on."""
    import hashlib
    h = hashlib.sha256()
    for f in ("kernels.hip", "kernels.hpp"):
        with open(os.path.join(ROOT, "bem_for_nonlinear_waves_amd", "csrc", f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def workload_key(args, fp32) -> str:
    return (f"cylinder_h{args.h}_p{args.p}_leaf{args.leaf}_lvl{args.max_level}_{'fp32' if fp32 else 'fp64'}"
            if not args.case else "case")


def measured_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` from profiles/traffic_<workload>.txt: lines
    `build <hash>`, `calibration fetch <f> write <w>` (factors from a plain device copy of known size) and
    `<kernel> fetch_kib <F> write_kib <W> launches <n>` (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, KiB summed
    over the launches). None unless the file exists and was measured on this device code."""
    path = os.path.join(PROFILES, f"traffic_{workload}.txt")
    if not os.path.exists(path):
        return None
    build, fscale, wscale, rows = None, 1.0, 1.0, {}
    for line in open(path):
        t = line.split()
        if not t or t[0].startswith("#"):
            continue
        if t[0] == "build":
            build = t[1]
        elif t[0] == "calibration":
            fscale, wscale = float(t[2]), float(t[4])
        elif len(t) >= 7 and t[1] == "fetch_kib":
            rows[t[0]] = (float(t[2]), float(t[4]), float(t[6]))
    if build != device_code_hash() or kernel not in rows:
        return None
    f, w, n = rows[kernel]
    return {"bytes": (f * fscale + w * wscale) * 1024.0 / n, "source": os.path.relpath(path, ROOT)}


def case_workload(path: str, seed: int = 3):
    """A case in the reference's input format (settings.json + body JSON + OBJ): the fluid surface with
    boundary types by contact; synthetic knowns: phi from seeded free-surface modes over the bounding
    box, phi_n = n.U on the faces of the second and later rigid bodies (wave makers, floating bodies)."""
    from bem_for_nonlinear_waves_amd import inputs
    from bem_for_nonlinear_waves_amd import mesh as M
    m, info = inputs.load_case(path)
    d = M.build_dofs(m)
    ext = m.X.max(axis=0) - m.X.min(axis=0)
    groups = tuple(range(2, 2 + max(0, len(info["bodies"]) - 2)))
    phi, phin = M.tank_bc(m, d, seed, float(ext[0]), float(ext[1]), body_velocity_sigma=0.1, body_group=groups)
    return m, d, phi, phin


def make_workload(h: float, seed: int = 3):
    from bem_for_nonlinear_waves_amd import mesh as M
    m = M.cylinder_tank(Lx=20.0, Ly=4.0, Lz=2.0, R=0.5, draft=1.0, h=h)
    m = M.perturb_free_surface(m, seed, 0.002)
    d = M.build_dofs(m)
    phi, phin = M.tank_bc(m, d, seed, 20.0, 4.0, body_velocity_sigma=0.1)
    return m, d, phi, phin


def cpu_baseline(m, d, phi, phin, p, leaf_max, max_level, matvecs_per_solve, n_timed=3):
    """The FP64 oracle (OpenMP restatement of the reference's FMM matvec) on the bench's own workload at its
    full size, built on this machine with -O3 -march=native: setup, one matvec that also builds the M2L
    coefficients per (level, offset), then n_timed timed matvecs; solves/s at the FP64 path's matvec count."""
    from oracle import oracle as O
    O.use_native()
    t = time.perf_counter()
    o = O.Oracle(m, d, phi, phin, p=p, leaf_max=leaf_max, max_level=max_level)
    o.set_m2l_cache(True)
    x = np.random.default_rng(0).standard_normal(d.n_dof)
    o.matvec(x)
    t_setup = time.perf_counter() - t
    t0 = time.perf_counter()
    for _ in range(n_timed):
        o.matvec(x)
    tm = (time.perf_counter() - t0) / n_timed
    del o
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        model = [l.split(":", 1)[1].strip() for l in out.splitlines() if l.startswith("Model name")][0]
    except Exception:
        model = "unknown"
    return {"value": 1.0 / (matvecs_per_solve * tm), "unit": "solves/s", "cores": threads, "kind": "port",
            "matvec_mnode_per_s": m.n_nodes / tm / 1e6, "matvec_s": tm, "setup_s": t_setup, "cpu": model,
            "omp_proc_bind": os.environ.get("OMP_PROC_BIND", ""),
            "sample": (f"measured at {m.n_nodes:,} nodes (the bench workload): oracle built -O3 -march=native, "
                       f"{threads} OpenMP threads; setup + first matvec (M2L coefficients cached per level and "
                       f"offset) {t_setup:.1f} s, then {n_timed} timed matvecs of {tm:.2f} s; solves/s at "
                       f"{matvecs_per_solve} matvecs per solve (the FP64 GPU path's count)")}


def kernel_key(name: str) -> str:
    """`k_m2l_rot<10>` from a demangled dispatch name such as
    `void (anonymous namespace)::k_m2l_rot<10>(int, float const*, ...)`."""
    n = name.replace("void ", "").replace("(anonymous namespace)::", "")
    if "rocprim::" in n:  # library sorts and scans of the setup: one line per algorithm stage
        for stage in ("radix_sort_onesweep_iteration", "radix_sort_onesweep_global_offsets", "radix_sort_block_sort",
                      "merge_sort_block_merge", "merge_mergepath_partition", "scan_impl", "transform", "reduce"):
            if stage in n:
                return f"rocprim::{stage}"
        return "rocprim::other"
    depth, out = 0, []
    for ch in n:
        if ch == "(" and depth == 0:
            break
        depth += ch == "<"
        depth -= ch == ">"
        out.append(ch)
    return "".join(out).replace(" ", "")


def write_traffic(fetch_dir: str, write_dir: str, workload: str) -> str:
    """profiles/traffic_<workload>.txt from two rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE; csv output) over
    `bench.py --profile-only`: KiB summed per kernel and the dispatch count, plus the calibration factors of the
    256 MiB device copy that the same run makes (its bytes read and written over what the counters report)."""
    import csv
    import glob

    def load(d, counter):
        rows = {}
        files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        if not files:
            raise SystemExit(f"no counter_collection.csv under {d}")
        copy = 0.0
        for f in files:
            for r in csv.DictReader(open(f)):
                if r.get("Counter_Name") != counter:
                    continue
                v = float(r["Counter_Value"])
                if "copyBuffer" in r["Kernel_Name"]:
                    copy = max(copy, v)
                    continue
                k = kernel_key(r["Kernel_Name"])
                kib, n = rows.get(k, (0.0, 0))
                rows[k] = (kib + v, n + 1)
        return rows, copy

    fetch, fcopy = load(fetch_dir, "FETCH_SIZE")
    write, wcopy = load(write_dir, "WRITE_SIZE")
    cal_kib = 256 * 1024.0
    lines = [f"# HBM traffic per kernel, workload {workload}: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE over",
             "# `python3 bench.py --profile-only` (KiB summed over the dispatches of each kernel); calibration = bytes",
             "# of the run's 256 MiB device copy over the counter's reading (fetch, write)",
             f"build {device_code_hash()}",
             f"calibration fetch {cal_kib / fcopy if fcopy else 1.0:.4f} write {cal_kib / wcopy if wcopy else 1.0:.4f}"]
    for k in sorted(fetch, key=lambda k: -fetch[k][0]):
        if k in write and fetch[k][1] == write[k][1]:
            lines.append(f"{k} fetch_kib {fetch[k][0]:.1f} write_kib {write[k][0]:.1f} launches {fetch[k][1]}")
    path = os.path.join(PROFILES, f"traffic_{workload}.txt")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


def cpu_baseline_child(args, matvecs_per_solve):
    """The CPU baseline in a child process of its own, with its OpenMP threads bound close to each other
    (OMP_PROC_BIND is read by libgomp when it is first loaded, and binding the request threads of the GPU
    process would crowd them onto the same cores)."""
    env = dict(os.environ, OMP_PROC_BIND="close")
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-baseline-only", str(matvecs_per_solve),
           "--h", str(args.h), "--p", str(args.p), "--leaf", str(args.leaf), "--max-level", str(args.max_level)]
    if args.case:
        cmd += ["--case", args.case]
    r = subprocess.run(cmd, capture_output=True, text=True, env=env)
    if r.returncode != 0:
        raise RuntimeError(f"cpu baseline process failed ({r.returncode}): {r.stderr[-500:]}")
    return json.loads(r.stdout.strip().splitlines()[-1])


STAGES = ["scatter", "p2m", "m2m", "m2l", "m2l_reduce", "l2l", "near_l2p", "assemble"]


def _log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def setup_bvp(m, d, phi, phin, args, precision):
    from bem_for_nonlinear_waves_amd import _lib
    from bem_for_nonlinear_waves_amd.bvp import BEM_BVP
    pc = {"none": _lib.PRECOND_NONE, "jacobi": _lib.PRECOND_JACOBI, "schwarz": _lib.PRECOND_SCHWARZ}[args.precond]
    prec = _lib.PRECISION_FP32_REFINE if precision == "fp32_refine" else _lib.PRECISION_FP64
    bvp = BEM_BVP(m, d, p=args.p, leaf_max=args.leaf, max_level=args.max_level, preconditioner=pc, precision=prec,
                  refine_eta=args.refine_eta, m2l_tile_pairs=args.m2l_tile_pairs,
                  m2l_kernel={"auto": 0, "dense": 1, "rotation": 2, "plane-wave": 3, "plane-wave-fp64": 4}[args.m2l_kernel],
                  plane_wave_digits=args.plane_wave_digits,
                  schwarz_max_core=args.schwarz_core, schwarz_max_block=args.schwarz_block,
                  schwarz_overlap=args.schwarz_overlap, use_graphs=args.graphs, refine_recycle=args.refine_recycle,
                  inner_p=_inner_p(args) if precision == "fp32_refine" else 0)
    bvp.set_bc(phi, phin)
    t = time.perf_counter()
    bvp.prepare()
    t_setup = time.perf_counter() - t
    st = bvp.stats()
    _log(f"[{precision}] M2L tile: {st['m2l_tile_pairs']} pairs; operators uploaded {st['n_ops_uploaded']}; "
         f"FP32 M2L by rotation: {bool(st['m2l_rotation'])}; near GEMV rows/wave {st['near_rows_per_wave']}, "
         f"P2M loads in flight {st['p2m_loads']} (FP64, FP32 operator)")
    if st["n_precond_blocks"]:
        _log(f"[{precision}] schwarz blocks={st['n_precond_blocks']} avg={st['precond_avg_block']:.1f} "
             f"fallback={st['precond_fallback_blocks']}")
    _log(f"[{precision}] setup {t_setup:.2f}s: boxes={st['n_boxes']} leaves={st['n_leaves']} "
         f"m2l_pairs={st['n_m2l_pairs']} near_nnz={st['near_nnz']} device={st['device_bytes']/1e9:.1f}GB "
         f"(tree {st['t_tree']:.2f} lists {st['t_lists']:.2f} near_struct {st['t_near_struct']:.2f} "
         f"p2m {st['t_p2m']:.2f} ops {st['t_ops']:.2f} upload {st['t_upload']:.2f} near_quad {st['t_near_quad']:.2f} "
         f"alpha {st['t_alpha']:.2f} tune {st['t_tune']:.2f} precond {st['t_precond']:.2f} fp32 {st['t_fp32']:.2f})")
    return bvp, st, t_setup


SETUP_STAGES = ("t_tree", "t_lists", "t_near_struct", "t_p2m", "t_ops", "t_upload", "t_near_quad", "t_alpha", "t_tune",
                "t_precond", "t_fp32")


def time_reprepare(bvp, m, d, args, stages=3, seed=11):
    """Re-prepare after the nodes moved, as every RK stage of a time-stepping session does: free-surface nodes
    displaced by N(0, 1e-3 h), then bemfmm_set_geometry + bemfmm_prepare on the same context (tree, lists, near
    structure and quadrature, P2M, alpha and preconditioner rebuilt; operators and buffers reused)."""
    from bem_for_nonlinear_waves_amd import mesh as M
    times, mk = [], m
    for k in range(stages):
        mk = M.perturb_free_surface(mk, seed + k, 1e-3 * args.h)
        phi, phin = M.tank_bc(mk, d, seed + k, 20.0, 4.0, body_velocity_sigma=0.1)
        bvp.set_geometry(mk, d)
        bvp.set_bc(phi, phin)
        t = time.perf_counter()
        bvp.prepare()
        times.append(time.perf_counter() - t)
    st = bvp.stats()
    out = {"s": times, "tree_on_gpu": int(st["tree_on_gpu"]),
           "stages_s": {k[2:]: round(float(st[k]), 4) for k in SETUP_STAGES}}
    _log(f"re-prepare after moved nodes: {' '.join(f'{t:.3f}' for t in times)} s; last: " +
         " ".join(f"{k} {v:.3f}" for k, v in out["stages_s"].items()))
    return out


def time_matvec(bvp, n, n_nodes, reps, fp32, whole_only=False):
    """Matvec on device-resident vectors: wall time per matvec and per-stage HIP-event times (ms). whole_only: only
    the stage-timed configuration (every kernel alone on the stream, the near-field GEMV as one launch), for the
    profiler's per-launch figures."""
    from bem_for_nonlinear_waves_amd import devbuf
    mv = bvp.matvec_fp32_device if fp32 else bvp.matvec_device
    xd = devbuf.DeviceArray.from_host(np.random.default_rng(0).standard_normal(n))
    yd = devbuf.DeviceArray(n)
    xs, ys = xd, yd
    bvp.profile()  # enable event recording on the library's stream (eager launches)
    for _ in range(2):
        mv(xs.address, ys.address)
    bvp.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        mv(xs.address, ys.address)
    bvp.sync()
    t_mv = (time.perf_counter() - t0) / reps
    stage = np.zeros(8)
    for _ in range(5):
        mv(xs.address, ys.address)
        stage += bvp.profile()  # synchronises on the last event of this matvec
    stage /= 5
    bvp.profile_off()  # the configuration the solves run in (no event recording)
    if whole_only:
        xd.free()
        yd.free()
        return t_mv, stage
    for _ in range(2):
        mv(xs.address, ys.address)
    bvp.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        mv(xs.address, ys.address)
    bvp.sync()
    t_run = (time.perf_counter() - t0) / reps
    _log("%s matvec %.3f ms (%.3f ms with stage events; %.2f Mnode/s); stages ms: %s" % (
        "fp32" if fp32 else "fp64", t_run * 1e3, t_mv * 1e3, n_nodes / t_run / 1e6,
        " ".join(f"{k}={v:.3f}" for k, v in zip(STAGES, stage))))
    xd.free()
    yd.free()
    return t_run, stage


def run_solves(bvp, n, args, label):
    x0 = np.zeros(n)
    for i in range(args.warmup):
        x, info = bvp.solveSystemGMRES(x0, tol=args.tol, max_iter=args.max_iter, restart=args.max_iter)
        _log(f"[{label}] warmup solve {i}: iters={info['iters']} steps={info['refinement_steps']} "
             f"residual={info['residual']:.2e} {info['time_solve']:.2f}s")
        if info['refinement_steps']:
            _log(f"[{label}] FP64 residual at each refinement step: "
                 + " ".join(f"{r:.2e}" for r in bvp.refinement_history()))
            if os.environ.get("BENCH_GMRES_HISTORY"):
                _log(f"[{label}] GMRES estimates: " + " ".join(f"{r:.1e}" for r in bvp.residual_history()))
    lat, infos = [], []
    t_start = time.perf_counter()
    for i in range(args.steps):
        ts = time.perf_counter()
        x, info = bvp.solveSystemGMRES(x0, tol=args.tol, max_iter=args.max_iter, restart=args.max_iter)
        lat.append(time.perf_counter() - ts)
        infos.append(info)
        _log(f"[{label}] solve {i}: iters={info['iters']} matvecs={info['n_matvec']} "
             f"steps={info['refinement_steps']} residual={info['residual']:.2e} converged={info['converged']} "
             f"{lat[-1]:.3f}s")
    t_total = time.perf_counter() - t_start
    return x, lat, infos, t_total


def make_batch_problem(rng, n_target: int):
    """Tank of about n_target nodes: aspect ratio Lx/Ly ~ U[1, 8], a surface-piercing cylinder with
    probability 1/4; free surface Dirichlet (seeded 8-mode phi), body phi_n = n.U."""
    from bem_for_nonlinear_waves_amd import mesh as M
    a = float(rng.uniform(1.0, 8.0))
    seed = int(rng.integers(1 << 30))
    if rng.uniform() < 0.25:
        Lx, Ly, Lz = 2.0 * a, 2.0, 1.0
        h = math.sqrt(2.0 * (Lx * Ly + Ly * Lz + Lx * Lz) / n_target)
        m = M.cylinder_tank(Lx=Lx, Ly=Ly, Lz=Lz, R=0.25, draft=0.5, h=h)
        sigma = 0.1
    else:
        Lx, Ly, Lz = a, 1.0, 1.0
        h = math.sqrt(2.0 * (Lx * Ly + Ly * Lz + Lx * Lz) / n_target)
        m = M.box_tank(Lx, Ly, Lz, max(2, round(Lx / h)), max(2, round(Ly / h)), max(2, round(Lz / h)))
        sigma = 0.0
    m = M.perturb_free_surface(m, seed, 0.05 * h)
    d = M.build_dofs(m)
    phi, phin = M.tank_bc(m, d, seed, Lx, Ly, body_velocity_sigma=sigma)
    return m, d, phi, phin, h, seed, (Lx, Ly)


def _inner_p(args) -> int:
    """Expansion order of the FP32 operator's M2L: --inner-p when it applies (FP32 + refinement, the rotation
    M2L, below --p), else 0 (the plan's order)."""
    ok = args.precision == "fp32_refine" and args.m2l_kernel in ("auto", "rotation") and 0 < args.inner_p < args.p
    return args.inner_p if ok else 0


def _solver_kw(args):
    from bem_for_nonlinear_waves_amd import _lib
    return dict(p=args.p, leaf_max=args.leaf, max_level=args.max_level, tol=args.tol, max_iter=args.max_iter,
                restart=args.max_iter, refine_eta=args.refine_eta,
                precision=_lib.PRECISION_FP32_REFINE if args.precision == "fp32_refine" else _lib.PRECISION_FP64,
                preconditioner={"none": _lib.PRECOND_NONE, "jacobi": _lib.PRECOND_JACOBI,
                                "schwarz": _lib.PRECOND_SCHWARZ}[args.precond], use_graphs=args.graphs,
                inner_p=_inner_p(args))


def run_batch(args):
    """64 independent tank BVPs (N log-uniform in [2k, 40k]) submitted at once to the request loop;
    every request includes its setup (new geometry). Also timed one at a time (one slot)."""
    from bem_for_nonlinear_waves_amd.serve import BVPServer
    rng = np.random.default_rng(4)
    t = time.perf_counter()
    sizes = np.exp(rng.uniform(np.log(2000), np.log(40000), args.batch)).astype(int)
    probs = [make_batch_problem(rng, int(n)) for n in sizes]
    n_nodes = [p[0].n_nodes for p in probs]
    _log(f"batch: {len(probs)} tanks, nodes {min(n_nodes)}..{max(n_nodes)} (sum {sum(n_nodes)}), "
         f"{sum(p[6][1] == 2.0 for p in probs)} with a cylinder ({time.perf_counter() - t:.1f}s)")
    kw = _solver_kw(args)

    def run(slots):
        srv = BVPServer(slots=slots)
        w = [srv.submit(m, d, phi, phin, **kw) for m, d, phi, phin, *_ in probs[:2]]  # warm-up (operator cache)
        for tk in w:
            srv.wait(tk)
        srv.close()
        srv = BVPServer(slots=slots)
        t0 = time.perf_counter()
        tickets = [srv.submit(m, d, phi, phin, **kw) for m, d, phi, phin, *_ in probs]
        res = [srv.wait(tk) for tk in tickets]
        wall = time.perf_counter() - t0
        st = srv.stats()
        srv.close()
        conv = sum(int(i["converged"]) for _, i in res)
        svc = np.array([i["service_s"] for _, i in res]) * 1e3
        _log(f"batch slots={slots}: {wall:.2f}s ({len(probs) / wall:.2f} solves/s) converged {conv}/{len(probs)} "
             f"latency p50 {st['p50_latency_s'] * 1e3:.0f} ms p99 {st['p99_latency_s'] * 1e3:.0f} ms, "
             f"service p50 {np.percentile(svc, 50):.0f} ms")
        return wall, st, svc, conv, [x for x, _ in res]

    wall, st, svc, conv, xs = run(args.slots)
    seq = None
    if not args.no_sequential:
        wall1, st1, svc1, conv1, xs1 = run(1)
        diff = max(float(np.linalg.norm(a - b) / np.linalg.norm(b)) for a, b in zip(xs, xs1))
        seq = {"solves_per_s": len(probs) / wall1, "wall_s": wall1, "speedup": wall1 / wall,
               "max_rel_diff_vs_mix": diff}
        _log(f"batch: {wall1 / wall:.2f}x over one slot; results alone vs in the mix differ by <= {diff:.1e}")
    out = {
        "metric": METRIC, "value": len(probs) / wall, "unit": "solves/s", "n_gpus": 1, "steps": len(probs),
        "warmup": 2, "ms_per_step": wall / len(probs) * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64" if args.precision == "fp32_refine" else "f64", "data": "synthetic",
        "config": {"workload": f"batch of {len(probs)} tank BVPs, N log-uniform in [2k, 40k], aspect 1-8, "
                               f"1/4 with a cylinder, setup included in every solve",
                   "slots": args.slots, "p": args.p, "leaf_max": args.leaf, "tol": args.tol,
                   "precision": args.precision, "preconditioner": args.precond, "total_nodes": int(sum(n_nodes)),
                   "inner_p": _inner_p(args)},
        "p50_solve_ms": st["p50_latency_s"] * 1e3, "p99_solve_ms": st["p99_latency_s"] * 1e3,
        "p50_service_ms": float(np.percentile(svc, 50)), "p99_service_ms": float(np.percentile(svc, 99)),
        "converged": conv, "peak_device_gb": st["peak_bytes"] / 1e9, "one_slot": seq,
    }
    print(json.dumps(out))


def run_serve(args):
    """Time-stepping sessions arriving as a Poisson process: each session (a tank of N nodes) runs
    `stages` BVPs in order, moving its free-surface nodes between stages (seeded N(0, 1e-3 h)
    displacements); the request loop keeps each session's context resident. Latency = submit -> result."""
    import threading

    from bem_for_nonlinear_waves_amd import mesh as M
    from bem_for_nonlinear_waves_amd.serve import BVPServer
    rng = np.random.default_rng(5)
    t = time.perf_counter()
    sizes = rng.uniform(args.serve_min_n, args.serve_max_n, args.sessions).astype(int)
    arrivals = np.cumsum(rng.exponential(1.0 / args.rate, args.sessions))
    sessions = []
    for si, n in enumerate(sizes):
        m, d, phi, phin, h, seed, (Lx, Ly) = make_batch_problem(rng, int(n))
        stages = []
        mk = m
        for k in range(args.stages):
            mk = M.perturb_free_surface(mk, seed + 1 + k, 1e-3 * h)
            ph, phn = M.tank_bc(mk, d, seed + 1 + k, Lx, Ly, body_velocity_sigma=0.1)
            stages.append((mk, ph, phn))
        sessions.append((d, stages))
    _log(f"serve: {args.sessions} sessions x {args.stages} stages, nodes {sizes.min()}..{sizes.max()}, "
         f"arrivals over {arrivals[-1]:.2f}s ({time.perf_counter() - t:.1f}s to build)")
    kw = _solver_kw(args)
    srv = BVPServer(slots=args.slots)
    w = srv.submit(sessions[0][1][0][0], sessions[0][0], sessions[0][1][0][1], sessions[0][1][0][2], **kw)
    srv.wait(w)  # warm-up (operator cache)
    srv.close()
    srv = BVPServer(slots=args.slots)
    lat, svc, conv = [], [], []
    lock = threading.Lock()
    t0 = time.perf_counter()

    def session(si):
        d, stages = sessions[si]
        time.sleep(max(0.0, arrivals[si] - (time.perf_counter() - t0)))
        for mk, ph, phn in stages:
            x, info = srv.wait(srv.submit(mk, d, ph, phn, session=si, **kw))
            with lock:
                lat.append(info["latency_s"])
                svc.append(info["service_s"])
                conv.append(int(info["converged"]))
        srv.close_session(si)

    th = [threading.Thread(target=session, args=(i,)) for i in range(args.sessions)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    lat_ms, svc_ms = np.array(lat) * 1e3, np.array(svc) * 1e3
    _log(f"serve: {len(lat)} solves in {wall:.2f}s ({len(lat) / wall:.2f} solves/s), latency p50 "
         f"{np.percentile(lat_ms, 50):.0f} ms p99 {np.percentile(lat_ms, 99):.0f} ms, converged {sum(conv)}/{len(conv)}")
    out = {
        "metric": METRIC, "value": len(lat) / wall, "unit": "solves/s", "n_gpus": 1, "steps": len(lat), "warmup": 1,
        "ms_per_step": wall / len(lat) * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if args.precision == "fp32_refine" else "f64", "data": "synthetic",
        "config": {"workload": f"serving: {args.sessions} time-stepping sessions (Poisson arrivals at {args.rate}/s), "
                               f"N uniform in [{args.serve_min_n}, {args.serve_max_n}], {args.stages} stages each "
                               f"(one BVP per stage, free-surface nodes moved between stages, setup included)",
                   "slots": args.slots, "p": args.p, "leaf_max": args.leaf, "tol": args.tol,
                   "precision": args.precision, "preconditioner": args.precond, "inner_p": _inner_p(args)},
        "p50_solve_ms": float(np.percentile(lat_ms, 50)), "p99_solve_ms": float(np.percentile(lat_ms, 99)),
        "p50_service_ms": float(np.percentile(svc_ms, 50)), "p99_service_ms": float(np.percentile(svc_ms, 99)),
        "converged": int(sum(conv)), "peak_device_gb": st["peak_bytes"] / 1e9,
    }
    print(json.dumps(out))


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
    ap.add_argument("--precond", choices=["none", "jacobi", "schwarz"], default="schwarz")
    ap.add_argument("--precision", choices=["fp32_refine", "fp64"], default="fp32_refine")
    ap.add_argument("--refine-eta", type=float, default=3e-7)
    ap.add_argument("--inner-p", type=int, default=6, choices=[0, 5, 6, 8],
                    help="fp32_refine: expansion order of the FP32 operator's M2L (0: --p); the FP64 residuals keep --p")
    ap.add_argument("--refine-recycle", type=int, default=1, choices=[0, 1],
                    help="refinement steps after the first keep the earlier steps' Krylov directions (library option)")
    ap.add_argument("--m2l-tile-pairs", type=int, default=0, choices=[0, 16, 32, 64])
    ap.add_argument("--case", default=None,
                    help="solve a case in the reference's input format (path to its settings.json) instead of the cylinder tank")
    # Schwarz blocks: small cores with many overlap rings (blocks still <= 128 DOFs) measured best on the
    # 209k tank: core 8 / 6 rings 39 iterations (7.90 solves/s), 16/5 40, 32/3 41, 48/2 43; the reference
    # default 64/128/1 takes more
    ap.add_argument("--schwarz-core", type=int, default=8, help="Schwarz core size (reference default 64)")
    ap.add_argument("--schwarz-block", type=int, default=128, help="Schwarz block size (reference default 128)")
    ap.add_argument("--schwarz-overlap", type=int, default=6, help="Schwarz overlap rings (reference default 1)")
    ap.add_argument("--m2l-kernel", default="auto", choices=["auto", "dense", "rotation", "plane-wave", "plane-wave-fp64"],
                    help="M2L of the FP32 operator: rotation where compiled (p = 8, 10), dense matrix-core tiles, or "
                         "the plane-wave composed operators on the matrix-core tiles (plane-wave-fp64: for the FP64 "
                         "operator too)")
    ap.add_argument("--plane-wave-digits", type=int, default=10, choices=[6, 10])
    ap.add_argument("--no-fp64-compare", action="store_true",
                    help="skip the plain FP64 solves timed beside the FP32+refinement ones")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--schedule", default="auto", choices=["auto", "spin", "yield"],
                    help="how host threads wait for the GPU (hipSetDeviceFlags before the first context)")
    ap.add_argument("--graphs", type=int, default=0, choices=[0, 1],
                    help="replay each context's matvec from a HIP graph (library option use_graphs)")
    ap.add_argument("--mode", choices=["solve", "batch", "serve"], default="solve",
                    help="solve: the 209k-node BVP (default); batch: 64 tanks at once; serve: time-stepping sessions")
    ap.add_argument("--slots", type=int, default=8, help="batch/serve: concurrent requests")
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--no-sequential", action="store_true", help="batch: skip the one-slot comparison")
    ap.add_argument("--sessions", type=int, default=8)
    ap.add_argument("--stages", type=int, default=40, help="serve: RK4 stages (BVPs) per session")
    ap.add_argument("--rate", type=float, default=50.0, help="serve: session arrivals per second")
    ap.add_argument("--serve-min-n", type=int, default=5000)
    ap.add_argument("--serve-max-n", type=int, default=200000)
    ap.add_argument("--profile-only", action="store_true",
                    help="setup + the stage-timed matvecs (every kernel alone, the near field as one launch) + a "
                         "256 MiB device copy (counter calibration); no solves")
    ap.add_argument("--write-traffic", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"), default=None,
                    help="summarise two rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE) of --profile-only into "
                         "profiles/traffic_<workload>.txt and exit")
    ap.add_argument("--cpu-baseline-only", type=int, default=0, metavar="MATVECS_PER_SOLVE",
                    help="(child process) time the CPU baseline on the workload and print it as JSON")
    args = ap.parse_args()
    if args.cpu_baseline_only:
        m, d, phi, phin = case_workload(args.case) if args.case else make_workload(args.h)
        print(json.dumps(cpu_baseline(m, d, phi, phin, args.p, args.leaf, args.max_level, args.cpu_baseline_only)))
        return
    if args.write_traffic:
        fp32 = args.precision == "fp32_refine"
        print(write_traffic(args.write_traffic[0], args.write_traffic[1], workload_key(args, fp32)))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(a single BVP is solved on one MI355X; see the sharding notes in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    # no torch here: it bundles a HIP runtime of its own, which would be the one libbemfmm binds to if torch
    # were imported first; device vectors come from the library's runtime (devbuf)
    if args.schedule != "auto":  # how host threads wait for the GPU (before any context exists)
        from bem_for_nonlinear_waves_amd import devbuf
        flag = {"spin": 1, "yield": 2}[args.schedule]
        rc = devbuf.hip().hipSetDeviceFlags(flag)
        _log(f"hipSetDeviceFlags({args.schedule}) -> {rc}")
    from bem_for_nonlinear_waves_amd.bvp import device_name

    _log(f"device {device_name()}")
    if args.mode == "batch":
        return run_batch(args)
    if args.mode == "serve":
        return run_serve(args)
    t = time.perf_counter()
    _log(f"building the workload ({'case ' + args.case if args.case else f'cylinder tank, h={args.h}'})")
    m, d, phi, phin = case_workload(args.case) if args.case else make_workload(args.h)
    n = d.n_dof
    _log(f"mesh: {m.n_nodes} nodes, {m.n_faces} faces, {n} DOFs ({time.perf_counter() - t:.1f}s)")
    fp32 = args.precision == "fp32_refine"
    bvp, st, t_setup = setup_bvp(m, d, phi, phin, args, args.precision)
    t_mv64, stage64 = time_matvec(bvp, n, m.n_nodes, args.matvecs, False, args.profile_only)
    t_mv, stage = (time_matvec(bvp, n, m.n_nodes, args.matvecs, True, args.profile_only) if fp32 else (t_mv64, stage64))

    if args.profile_only:
        from bem_for_nonlinear_waves_amd import devbuf
        cal = devbuf.DeviceArray.from_host(np.ones(32 * 1024 * 1024))
        cal2 = devbuf.DeviceArray(cal.n)
        devbuf.synchronize()
        cal2.copy_from(cal)  # 256 MiB read + 256 MiB written: FETCH_SIZE/WRITE_SIZE calibration
        devbuf.synchronize()
        _log("profile-only: done")
        return

    x, lat, infos, t_total = run_solves(bvp, n, args, args.precision)
    solves_per_s = args.steps / t_total
    iters = [i["iters"] for i in infos]
    reprep = time_reprepare(bvp, m, d, args) if not args.case else None

    fp64_cmp = None
    if fp32 and not args.no_fp64_compare:
        bvp.close()
        b64, _, t_setup64 = setup_bvp(m, d, phi, phin, args, "fp64")
        x64, lat64, infos64, t64 = run_solves(b64, n, args, "fp64")
        b64.close()
        rel = float(np.linalg.norm(x - x64) / np.linalg.norm(x64))
        fp64_cmp = {"solves_per_s": args.steps / t64, "p50_solve_ms": float(np.percentile(np.array(lat64) * 1e3, 50)),
                    "gmres_iters": [i["iters"] for i in infos64], "matvecs": [i["n_matvec"] for i in infos64],
                    "speedup": solves_per_s / (args.steps / t64),
                    "rel_diff_solution_vs_fp64": rel, "matvec_ms": t_mv64 * 1e3}
        _log(f"fp32_refine vs fp64: {fp64_cmp['speedup']:.2f}x solves/s, solution rel diff {rel:.2e}")

    # ---- roofline of the dominant kernel of the matvec the solves run (live HIP-event stage timings) ----
    C = (args.p + 1) ** 2
    esz = 4.0 if fp32 else 8.0
    mfma_peak = 157.3 if fp32 else FP64_PEAK_TFLOPS
    rot = bool(fp32 and st.get("m2l_rotation", 0))
    # algorithmic multiply-adds per pair and layer: dense C x C operator, or the rotation factors
    # (D = Y Z and H^T = Z^T Y^T: 4 per order m >= 1 of each degree for Z, (n+1)^2 + n^2 per degree for Y;
    # axial: (p+1)^2 for m = 0 and 2 (p+1-m)^2 for m > 0)
    # the FP32 operator's M2L runs at the inner order when one is set (its source gathers are C_m coefficients)
    pp = (st.get("inner_p", 0) or args.p) if rot else args.p
    C_m = (pp + 1) ** 2
    rot_macs = (2 * sum(4 * k + (k + 1) ** 2 + k ** 2 for k in range(pp + 1)) + (pp + 1) ** 2
                + 2 * sum((pp + 1 - k) ** 2 for k in range(1, pp + 1)))
    m2l_flops = st["n_m2l_pairs"] * 2 * (rot_macs if rot else C * C) * 2.0
    # near field + L2P: the FP64 operator reads G and Gn per entry (and the L2P columns of both layers' rows); the
    # FP32 operator's is folded for A x, one weight per entry and 2 C L2P columns per row (single and double layer)
    # (the folded L2P columns stop at the inner order when one is set: its locals are zero above it)
    C_l = (st.get("inner_p", 0) + 1) ** 2 if (fp32 and st.get("inner_p", 0)) else C
    near_bytes = ((st["near_nnz"] + m.n_nodes * 2 * C_l) * 4.0 if fp32 else (st["near_nnz"] + m.n_nodes * C) * 2 * 8.0)
    dom = int(np.argmax(stage))
    # boxes that receive M2L: each target's local expansion is written once (2 C doubles)
    m2l_targets = st.get("n_m2l_targets", st["n_boxes"])
    if STAGES[dom] == "m2l":
        ach = m2l_flops / (stage[dom] * 1e-3) / 1e12
        if rot:
            # 2.84e10 FP32 flop against 2.2 GB of source gathers per launch is ~13 flop/B, below the ridge point
            # (157.3 TFLOP/s / 8 TB/s = 19.7 flop/B): the rotation M2L sits on the bandwidth side of the roofline.
            # Its algorithmic bytes are the gathers of the source multipoles, each pair's two FP32 columns
            # (2 C x 4 B); the per-pair staging it writes for the reduction is not algorithmic work.
            rot_bytes = st["n_m2l_pairs"] * 2 * C_m * 4.0
            achb = rot_bytes / (stage[dom] * 1e-3) / 1e9
            roof = {"kernel": "k_m2l_rot (M2L by rotation, FP32 FMA)", "bound": "hbm", "achieved": achb,
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achb / HBM_PEAK_GBS, "traffic": None,
                    "algorithmic_bytes": rot_bytes, "valu_tflops": ach, "valu_frac": ach / mfma_peak}
        else:
            roof = {"kernel": "k_translate_mfma (M2L)", "bound": "mfma", "achieved": ach, "peak": mfma_peak,
                    "unit": "TFLOP/s", "frac": ach / mfma_peak, "traffic": None}
    else:
        ach = near_bytes / (stage[dom] * 1e-3) / 1e9
        roof = {"kernel": ("k_block_gemv_u (near field + L2P folded for A x, FP32)" if fp32
                           else "k_block_gemv (near field + L2P, FP64)"), "bound": "hbm", "achieved": ach,
                "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": None, "algorithmic_bytes": near_bytes}
    # the M2L stage as a whole (translation + per-box reduction): algorithmic bytes = source gathers + each
    # target's locals written once
    m2l_alg = st["n_m2l_pairs"] * 2 * C_m * esz + m2l_targets * 2 * C * 8.0
    roof["m2l_stage"] = {"ms": float(stage[3] + stage[4]), "algorithmic_bytes": m2l_alg,
                         "achieved_gbs": m2l_alg / ((stage[3] + stage[4]) * 1e-3) / 1e9}
    # HBM bytes per launch from the rocprofv3 counter passes of this device code (profiles/), if they exist
    if rot and STAGES[dom] == "m2l":
        tkey = f"k_m2l_rot<{pp}>"
    elif STAGES[dom] == "m2l":
        tkey = f"k_translate_mfma<double,{(C + 63) // 64},true,{st['m2l_tile_pairs']}>"
    else:
        tkey = (f"k_block_gemv_u<{st['near_rows_per_wave'][1]}>" if fp32
                else f"k_block_gemv<double,{st['near_rows_per_wave'][0]}>")
    traffic = measured_traffic(tkey, workload_key(args, fp32))
    if traffic is not None:
        roof["traffic"] = traffic["bytes"]
        roof["traffic_source"] = traffic["source"]
    roof["stage_ms"] = dict(zip(STAGES, [round(float(v), 4) for v in stage]))
    roof["m2l_tflops"] = m2l_flops / (stage[3] * 1e-3) / 1e12
    if rot:  # the same M2L as dense operator multiply-adds per second
        roof["m2l_dense_equivalent_tflops"] = st["n_m2l_pairs"] * 2 * C_m * C_m * 2.0 / (stage[3] * 1e-3) / 1e12
    roof["near_gbs"] = near_bytes / (stage[6] * 1e-3) / 1e9
    roof["fp64_stage_ms"] = dict(zip(STAGES, [round(float(v), 4) for v in stage64]))

    cpu = None
    if not args.no_cpu_baseline:
        try:
            # matvecs of one FP64 solve on the GPU: the iterations + the FP64 residual checks
            mv_fp64 = (int(np.median(fp64_cmp["matvecs"])) if fp64_cmp else int(np.median([i["n_matvec"] for i in infos])))
            cpu = cpu_baseline_child(args, mv_fp64)
            _log(f"cpu baseline: {cpu['sample']}: {cpu['matvec_mnode_per_s']:.4f} Mnode/s")
        except Exception as e:  # the baseline never decides the GPU result
            _log(f"cpu baseline failed: {e}")

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
        "dtype": "f32+f64" if fp32 else "f64",
        "data": "synthetic",
        "config": {"workload": (f"case {args.case}, " if args.case else f"cylinder tank 20x4x2 m, R=0.5 draft=1, h={args.h}, ")
                               + ("FP32 FMM-GMRES with FP64 iterative refinement" if fp32 else "FP64 FMM-GMRES"),
                   "n_nodes": m.n_nodes, "n_faces": m.n_faces, "n_dof": n,
                   "p": args.p, "theta": 0.25, "leaf_max": args.leaf, "max_level": args.max_level, "tol": args.tol,
                   "precision": args.precision, "preconditioner": args.precond,
                   "schwarz": [args.schwarz_core, args.schwarz_block, args.schwarz_overlap],
                   "inner_p": _inner_p(args)},
        "matvec_mnode_per_s": m.n_nodes / t_mv / 1e6,
        "matvec_ms": t_mv * 1e3,
        "matvec_fp64_ms": t_mv64 * 1e3,
        "p50_solve_ms": float(np.percentile(lat_ms, 50)),
        "p99_solve_ms": float(np.percentile(lat_ms, 99)),
        "gmres_iters": iters,
        "refinement_steps": [i["refinement_steps"] for i in infos],
        "residual": [i["residual"] for i in infos],
        "setup_s": t_setup,
        "launch_shapes": {"m2l_tile_pairs": st["m2l_tile_pairs"], "near_rows_per_wave": st["near_rows_per_wave"],
                          "p2m_loads": st["p2m_loads"]},
        "reprepare": reprep,
        "fp64_path": fp64_cmp,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
