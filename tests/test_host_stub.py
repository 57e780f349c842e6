"""The library's host code under sanitizers, with several contexts at once (no GPU).

tests/hipstub builds bemfmm.cpp, server.cpp and the plan against a stub HIP runtime (streams as host
threads, device memory as host memory, graphs as recorded operation lists) and CPU versions of the
kernels, into a driver linked with AddressSanitizer (+UBSan) or ThreadSanitizer. The stub records every
use of the runtime that its documentation or the stream-capture rules do not allow; in particular a
device-wide synchronisation (hipFree and hipHostFree synchronise the device) while another thread is
capturing a stream into a graph. Here FP32+refinement sessions (re-prepared between stages) run first
alone, then four at once on their own threads and through the request loop: the sanitizers must stay
silent, the stub must record no violation, and every result in the mix must equal its result alone.
"""
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
STUB = os.path.join(HERE, "hipstub")
BUILD = os.path.join(STUB, "_build")


@pytest.fixture(scope="module")
def harness():
    subprocess.run(["make", "-s", "-C", STUB, "-j2"], check=True)
    return BUILD


def _write(path, specs, n_stages=2):
    from bem_for_nonlinear_waves_amd import mesh as M
    with open(path, "wb") as f:
        np.array([len(specs)], np.int32).tofile(f)
        for k, (kind, a, b, c) in enumerate(specs):
            if kind == "box":
                m = M.box_tank(a, 1.0, 1.0, b, c, c)
                Lx, Ly, sigma = a, 1.0, 0.0
            else:
                m = M.cylinder_tank(Lx=a, Ly=2.0, Lz=1.0, R=0.25, draft=0.5, h=b)
                Lx, Ly, sigma = a, 2.0, 0.1
            m = M.perturb_free_surface(m, 10 + k, 0.01)
            d = M.build_dofs(m)
            stages, mk = [], m
            for s in range(n_stages):
                if s:
                    mk = M.perturb_free_surface(mk, 100 + 10 * k + s, 1e-3)
                phi, phin = M.tank_bc(mk, d, 10 + k + s, Lx, Ly, body_velocity_sigma=sigma)
                stages.append((mk.X, phi, phin))
            np.array([m.n_nodes, m.n_faces, d.n_dof, n_stages], np.int32).tofile(f)
            for X, _, _ in stages:
                np.ascontiguousarray(X, np.float64).tofile(f)
            for a_ in (m.tri, d.face_dof, d.dof_kind, d.dof_point):
                np.ascontiguousarray(a_, np.int32).tofile(f)
            for _, phi, _ in stages:
                np.ascontiguousarray(phi, np.float64).tofile(f)
            for _, _, phin in stages:
                np.ascontiguousarray(phin, np.float64).tofile(f)


def _run(exe, *args, timeout=900):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", OMP_NUM_THREADS="2")
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=timeout, env=env)
    out = r.stdout + r.stderr
    print(out[-3000:])
    return r.returncode, out


def test_stub_records_a_device_sync_during_another_threads_capture(harness):
    rc, out = _run(os.path.join(harness, "driver_asan"), "selftest")
    assert rc == 0 and "selftest: violations 1" in out


@pytest.mark.parametrize("mode", ["threads", "server"])
def test_concurrent_fp32_refine_contexts_asan(harness, tmp_path, mode):
    path = str(tmp_path / "problems.bin")
    _write(path, [("box", 2.0, 16, 8), ("box", 3.0, 21, 7), ("cyl", 3.0, 0.16, 0), ("box", 1.5, 12, 9)])
    rc, out = _run(os.path.join(harness, "driver_asan"), path, mode, "fp32", "4", "8")
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out
    assert "violations 0" in out
    assert rc == 0  # also: every result in the mix equals its result alone to 1e-10


def test_concurrent_fp32_refine_server_tsan(harness, tmp_path):
    path = str(tmp_path / "problems.bin")
    _write(path, [("box", 2.0, 12, 6), ("box", 3.0, 18, 6), ("box", 1.5, 10, 7)])
    rc, out = _run(os.path.join(harness, "driver_tsan"), path, "server", "fp32", "3", "8", timeout=1500)
    assert "WARNING: ThreadSanitizer" not in out
    assert "violations 0" in out
    assert rc == 0
