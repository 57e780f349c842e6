"""Python mirror of the request loop (include/bemfmm_server.h): many BVPs on one GPU.

``BVPServer.submit`` queues one solve (the arrays of ``BEM_BVP`` plus the solver arguments)
and returns a ticket; ``wait`` returns its solution and diagnostics. Requests of one session
run in order on a context kept resident between them (the time-stepping driver's BVP per RK
stage, bem/main_time_domain.cpp:2095, and its phi_t BVP, bem/BEM_solveBVP.hpp:999).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .bvp import BEMError
from .mesh import DofModel, Mesh


class ServerOpts(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("slots", ctypes.c_int), ("hbm_budget_bytes", ctypes.c_double),
                ("omp_threads_per_slot", ctypes.c_int)]


class Problem(ctypes.Structure):
    _fields_ = [("n_nodes", ctypes.c_int), ("X", _lib._dp), ("n_faces", ctypes.c_int), ("tri", _lib._ip),
                ("face_dof", _lib._ip), ("n_dof", ctypes.c_int), ("dof_kind", _lib._ip), ("dof_point", _lib._ip),
                ("phi", _lib._dp), ("phin", _lib._dp), ("x0", _lib._dp), ("opts", _lib.Opts), ("tol", ctypes.c_double),
                ("max_iter", ctypes.c_int), ("restart", ctypes.c_int)]


class ServerStats(ctypes.Structure):
    _fields_ = [("completed", ctypes.c_int64), ("failed", ctypes.c_int64), ("queued", ctypes.c_int64),
                ("evicted", ctypes.c_int64), ("running", ctypes.c_int), ("resident_sessions", ctypes.c_int),
                ("p50_latency_s", ctypes.c_double), ("p99_latency_s", ctypes.c_double),
                ("used_bytes", ctypes.c_double), ("peak_bytes", ctypes.c_double), ("pooled_contexts", ctypes.c_int)]


def default_opts() -> ServerOpts:
    o = ServerOpts()
    _lib.load().bemfmm_server_default_opts(ctypes.byref(o))
    return o


class BVPServer:
    def __init__(self, slots: int = 4, hbm_budget_bytes: float = 200e9, device: int = 0, omp_threads_per_slot: int = 0):
        self.lib = _lib.load()
        o = default_opts()
        o.device, o.slots, o.hbm_budget_bytes, o.omp_threads_per_slot = device, slots, hbm_budget_bytes, \
            omp_threads_per_slot
        self.srv = self.lib.bemfmm_server_create(ctypes.byref(o))
        if not self.srv:
            raise BEMError(_lib.E_INVALID, "bemfmm_server_create failed (invalid options)")
        self._n_dof = {}

    def submit(self, mesh: Mesh, dofs: DofModel, phi, phin, session: int = -1, x0=None, tol: float = 1e-10,
               max_iter: int = 1000, restart: int = 1000, p: int = 10, theta: float = 0.25, leaf_max: int = 50,
               max_level: int = 7, q: int = 6, precision: int = _lib.PRECISION_FP64,
               preconditioner: int = _lib.PRECOND_NONE, refine_eta: float = 3e-7, use_graphs: int = 0,
               inner_p: int = 0) -> int:
        o = _lib.Opts()
        self.lib.bemfmm_default_opts(ctypes.byref(o))
        o.p, o.theta, o.leaf_max, o.max_level, o.q = p, theta, leaf_max, max_level, q
        o.precision, o.preconditioner, o.refine_eta = precision, preconditioner, refine_eta
        o.use_graphs, o.inner_p = use_graphs, inner_p
        keep = [np.ascontiguousarray(mesh.X, np.float64), np.ascontiguousarray(mesh.tri, np.int32),
                np.ascontiguousarray(dofs.face_dof, np.int32), np.ascontiguousarray(dofs.dof_kind, np.int32),
                np.ascontiguousarray(dofs.dof_point, np.int32), np.ascontiguousarray(phi, np.float64),
                np.ascontiguousarray(phin, np.float64)]
        x0a = None if x0 is None else np.ascontiguousarray(x0, np.float64)
        X, tri, fd, kd, pt, ph, phn = keep
        pr = Problem(mesh.n_nodes, X.ctypes.data_as(_lib._dp), mesh.n_faces, tri.ctypes.data_as(_lib._ip),
                     fd.ctypes.data_as(_lib._ip), int(dofs.n_dof), kd.ctypes.data_as(_lib._ip),
                     pt.ctypes.data_as(_lib._ip), ph.ctypes.data_as(_lib._dp), phn.ctypes.data_as(_lib._dp),
                     None if x0a is None else x0a.ctypes.data_as(_lib._dp), o, tol, max_iter, restart)
        t = self.lib.bemfmm_server_submit(self.srv, ctypes.byref(pr), session)
        if t <= 0:
            raise BEMError(int(t), "bemfmm_server_submit failed")
        self._n_dof[t] = int(dofs.n_dof)
        return t

    def wait(self, ticket: int):
        n = self._n_dof.pop(ticket)
        x = np.empty(n)
        r = _lib.Result()
        times = np.zeros(3)
        rc = self.lib.bemfmm_server_wait(self.srv, ticket, x.ctypes.data_as(_lib._dp), ctypes.byref(r),
                                         times.ctypes.data_as(_lib._dp))
        if rc != _lib.OK:
            raise BEMError(rc, self.lib.bemfmm_server_last_error(self.srv).decode())
        info = {k: getattr(r, k) for k, _ in _lib.Result._fields_}
        info.update(queue_s=times[0], service_s=times[1], latency_s=times[2])
        return x, info

    def close_session(self, session: int):
        rc = self.lib.bemfmm_server_close_session(self.srv, session)
        if rc != _lib.OK:
            raise BEMError(rc, "bemfmm_server_close_session failed")

    def stats(self) -> dict:
        s = ServerStats()
        self.lib.bemfmm_server_get_stats(self.srv, ctypes.byref(s))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def close(self):
        if getattr(self, "srv", None):
            self.lib.bemfmm_server_destroy(self.srv)
            self.srv = None

    def __del__(self):
        self.close()
