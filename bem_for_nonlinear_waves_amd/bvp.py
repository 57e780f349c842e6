"""Host-side mirror of the reference's BEM_BVP GMRES/FMM entry point.

``BEM_BVP`` wraps one ``bemfmm_ctx`` (include/bemfmm.h). Method names follow the
reference (bem/BEM_solveBVP_GMRES_FMM.hpp): ``solveGMRES`` (:3056) performs the
setup then ``solveSystemGMRES`` (:2007); ``compute_Ax`` corresponds to
``return_A_dot_v`` without preconditioner (:2124); ``diag_coeffs`` is the
rigid-mode alpha (:2674-2716). Errors raise ``BEMError`` carrying the C return
code, the analogue of the reference's ``error_message`` / non-converged flag.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

from . import _lib
from .mesh import DofModel, Mesh


class BEMError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"bemfmm error {code}: {msg}")
        self.code = code


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _ptr(a, t):
    return a.ctypes.data_as(t)


def is_available() -> bool:
    return bool(_lib.load().bemfmm_is_available())


def device_name() -> str:
    return _lib.load().bemfmm_device_name().decode()


class BEM_BVP:
    """One BVP on one GPU: geometry + DOF model + boundary values -> FMM-GMRES solution."""

    def __init__(self, mesh: Mesh, dofs: DofModel, p: int = 10, theta: float = 0.25, leaf_max: int = 50,
                 max_level: int = 7, q: int = 6, precision: int = _lib.PRECISION_FP64,
                 preconditioner: int = _lib.PRECOND_NONE, refine_eta: float = 3e-7, schwarz_max_core: int = 64,
                 schwarz_max_block: int = 128, schwarz_overlap: int = 1, m2l_tile_pairs: int = 0, dof_reorder: int = 1,
                 use_graphs: int = 0, m2l_kernel: int = 0, overlap_near: int = 1, refine_recycle: int = 1,
                 ilu_neighborhood: int = 0, ilu_kring: int = 1, schwarz_variant: int = 0, schwarz_core_k: int = 1,
                 milu_omega: float = 1.0, ilut_drop_tol: float = 1e-3, ilut_max_entries: int = 50,
                 ilut_pivot_min: float = 1e-12, plane_wave_digits: int = 10, tree_on_gpu: int = 1,
                 element_order: int = 1, near_rows_per_wave: int = 0, p2m_loads: int = 0, inner_p: int = 0,
                 device: int = 0):
        self.lib = _lib.load()
        o = _lib.Opts()
        self.lib.bemfmm_default_opts(ctypes.byref(o))
        o.p, o.theta, o.leaf_max, o.max_level, o.q = p, theta, leaf_max, max_level, q
        o.precision, o.preconditioner, o.refine_eta = precision, preconditioner, refine_eta
        o.schwarz_max_core, o.schwarz_max_block, o.schwarz_overlap = schwarz_max_core, schwarz_max_block, schwarz_overlap
        o.m2l_tile_pairs, o.dof_reorder, o.use_graphs = m2l_tile_pairs, dof_reorder, use_graphs
        o.m2l_kernel, o.overlap_near, o.refine_recycle = m2l_kernel, overlap_near, refine_recycle
        o.ilu_neighborhood, o.ilu_kring = ilu_neighborhood, ilu_kring
        o.schwarz_variant, o.schwarz_core_k = schwarz_variant, schwarz_core_k
        o.milu_omega, o.ilut_drop_tol, o.ilut_max_entries, o.ilut_pivot_min = (milu_omega, ilut_drop_tol,
                                                                              ilut_max_entries, ilut_pivot_min)
        o.plane_wave_digits = plane_wave_digits
        o.tree_on_gpu = tree_on_gpu
        o.element_order = element_order
        o.near_rows_per_wave, o.p2m_loads, o.inner_p = near_rows_per_wave, p2m_loads, inner_p
        self.element_order = element_order
        self.ctx = self.lib.bemfmm_ctx_create(device, ctypes.byref(o))
        if not self.ctx:
            raise BEMError(_lib.E_DEVICE, "bemfmm_ctx_create failed (no usable GPU?)")
        self.set_geometry(mesh, dofs)
        self.time_setup = 0.0

    def set_geometry(self, mesh: Mesh, dofs: DofModel):
        """New nodes / DOF model (an RK stage moves the free surface); call prepare() before solving.
        The context keeps its device operator table across geometries."""
        self.n_dof = int(dofs.n_dof)
        npf = 6 if self.element_order == 2 else 3
        if np.shape(mesh.tri)[1] != npf or np.shape(dofs.face_dof)[1] != npf:
            raise ValueError(f"element_order {self.element_order} needs {npf} nodes and DOF entries per face")
        self._geom = [_d(mesh.X), _i(mesh.tri), _i(dofs.face_dof), _i(dofs.dof_kind), _i(dofs.dof_point)]
        X, tri, fd, kd, pt = self._geom
        self._check(self.lib.bemfmm_set_geometry(self.ctx, mesh.n_nodes, _ptr(X, _lib._dp), mesh.n_faces,
                                                 _ptr(tri, _lib._ip), _ptr(fd, _lib._ip), self.n_dof,
                                                 _ptr(kd, _lib._ip), _ptr(pt, _lib._ip)))
        self.prepared = False

    def _check(self, rc: int):
        if rc != _lib.OK:
            raise BEMError(rc, self.lib.bemfmm_last_error(self.ctx).decode())

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.bemfmm_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        self.close()

    # -- setup ------------------------------------------------------------------
    def set_bc(self, phi, phin):
        self._bc = [_d(phi), _d(phin)]
        self._check(self.lib.bemfmm_set_bc(self.ctx, _ptr(self._bc[0], _lib._dp), _ptr(self._bc[1], _lib._dp)))

    def set_robin(self, kappa):
        """Robin coefficient per DOF (physical units) for ROBIN DOFs (kind 3: phi_n = kappa phi + g, g passed
        as their phin in set_bc): the frequency-domain free surface phin = kappa phi, kappa = omega^2 / g
        (main_freq_domain.cpp:543-555). The operator depends on it: call prepare afterwards."""
        self._robin = _d(kappa)
        self._check(self.lib.bemfmm_set_robin(self.ctx, _ptr(self._robin, _lib._dp)))
        self.prepared = False

    def prepare(self):
        t = time.perf_counter()
        self._check(self.lib.bemfmm_prepare(self.ctx))
        self.time_setup = time.perf_counter() - t
        self.prepared = True

    # -- operators ----------------------------------------------------------------
    def compute_Ax(self, x) -> np.ndarray:
        x = _d(x)
        y = np.empty(self.n_dof)
        self._check(self.lib.bemfmm_matvec(self.ctx, _ptr(x, _lib._dp), _ptr(y, _lib._dp)))
        return y

    def compute_Ax_fp32(self, x) -> np.ndarray:
        """The FP32 operator of the refinement loop (context created with PRECISION_FP32_REFINE)."""
        x = _d(x)
        y = np.empty(self.n_dof)
        self._check(self.lib.bemfmm_matvec_fp32(self.ctx, _ptr(x, _lib._dp), _ptr(y, _lib._dp)))
        return y

    def matvec_fp32_device(self, x_ptr: int, y_ptr: int):
        self._check(self.lib.bemfmm_matvec_fp32_device(self.ctx, ctypes.c_void_p(x_ptr), ctypes.c_void_p(y_ptr)))

    def matvec_device(self, x_ptr: int, y_ptr: int):
        self._check(self.lib.bemfmm_matvec_device(self.ctx, ctypes.c_void_p(x_ptr), ctypes.c_void_p(y_ptr)))

    def sync(self):
        self._check(self.lib.bemfmm_sync(self.ctx))

    def rhs(self) -> np.ndarray:
        b = np.empty(self.n_dof)
        self._check(self.lib.bemfmm_rhs(self.ctx, _ptr(b, _lib._dp)))
        return b

    def diag_coeffs(self) -> np.ndarray:
        a = np.empty(self.n_dof)
        self._check(self.lib.bemfmm_alpha(self.ctx, _ptr(a, _lib._dp)))
        return a

    def stats(self) -> dict:
        s = _lib.Stats()
        self._check(self.lib.bemfmm_get_stats(self.ctx, ctypes.byref(s)))
        return {k: (list(getattr(s, k)) if isinstance(getattr(s, k), ctypes.Array) else getattr(s, k))
                for k, _ in _lib.Stats._fields_}

    def profile_off(self):
        self._check(self.lib.bemfmm_matvec_profile(self.ctx, None))

    def profile(self):
        out = np.zeros(8)
        self._check(self.lib.bemfmm_matvec_profile(self.ctx, _ptr(out, _lib._dp)))
        return out

    def residual_history(self) -> np.ndarray:
        n = self.lib.bemfmm_residual_history(self.ctx, None, 0)
        out = np.empty(max(n, 0))
        if n > 0:
            self.lib.bemfmm_residual_history(self.ctx, _ptr(out, _lib._dp), n)
        return out

    def precond_apply(self, r) -> np.ndarray:
        """z = M^-1 r with the context's preconditioner."""
        r = _d(r)
        z = np.empty(self.n_dof)
        self._check(self.lib.bemfmm_precond_apply(self.ctx, _ptr(r, _lib._dp), _ptr(z, _lib._dp)))
        return z

    def dof_order(self) -> np.ndarray:
        """order[i] = caller DOF at internal position i (Morton order unless dof_reorder=0)."""
        out = np.empty(self.n_dof, np.int32)
        self._check(self.lib.bemfmm_dof_order(self.ctx, _ptr(out, _lib._ip)))
        return out

    def schwarz_blocks(self):
        """(ptr, idx, ncore) of the Schwarz blocks (empty without the Schwarz preconditioner)."""
        nb = self.lib.bemfmm_schwarz_blocks(self.ctx, None, None, None)
        if nb < 0:
            self._check(nb)
        ptr = np.zeros(nb + 1, np.int64)
        ncore = np.zeros(max(nb, 1), np.int32)
        self.lib.bemfmm_schwarz_blocks(self.ctx, ptr.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), None, None)
        idx = np.zeros(max(int(ptr[-1]), 1), np.int32)
        self.lib.bemfmm_schwarz_blocks(self.ctx, ptr.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                       _ptr(idx, _lib._ip), _ptr(ncore, _lib._ip))
        return ptr, idx[:int(ptr[-1])], ncore[:nb]

    def tree_digest(self) -> int:
        """64-bit digest of the prepared tree, face order, target boxes and interaction lists."""
        out = ctypes.c_uint64(0)
        self._check(self.lib.bemfmm_tree_digest(self.ctx, ctypes.byref(out)))
        return int(out.value)

    def refinement_history(self) -> np.ndarray:
        n = self.lib.bemfmm_refinement_history(self.ctx, None, 0)
        out = np.empty(max(n, 0))
        if n > 0:
            self.lib.bemfmm_refinement_history(self.ctx, _ptr(out, _lib._dp), n)
        return out

    # -- solve --------------------------------------------------------------------
    def solveSystemGMRES(self, x0=None, tol: float = 1e-10, max_iter: int = 1000, restart: int = 1000,
                         raise_on_nonconvergence: bool = False):
        x0 = np.zeros(self.n_dof) if x0 is None else _d(x0)
        x = np.empty(self.n_dof)
        r = _lib.Result()
        rc = self.lib.bemfmm_solve(self.ctx, _ptr(x0, _lib._dp), _ptr(x, _lib._dp), tol, max_iter, restart,
                                   ctypes.byref(r))
        info = {k: getattr(r, k) for k, _ in _lib.Result._fields_}
        info["time_setup"] = self.time_setup
        if rc == _lib.E_NOT_CONVERGED and not raise_on_nonconvergence:
            return x, info
        self._check(rc)
        return x, info

    def solveSystemGMRES_phi_t(self, phit, phint, x0=None, tol: float = 1e-10, max_iter: int = 1000,
                               restart: int = 1000, raise_on_nonconvergence: bool = False):
        """Second BVP of an RK stage, solveSystemGMRES(is_time_derivative=true, x0)
        (BEM_solveBVP.hpp:922-1010, BEM_solveBVP_GMRES_FMM.hpp:2007): the same operator with the time
        derivatives as knowns (phi_t on Dirichlet DOFs, phin_t on Neumann DOFs). Tree, operators, near
        field, alpha and preconditioner stay resident: only the right-hand side is rebuilt. The context
        keeps these knowns until the next set_bc."""
        self.set_bc(phit, phint)
        return self.solveSystemGMRES(x0, tol, max_iter, restart, raise_on_nonconvergence)

    def solveGMRES_complex(self, phi, phin, tol: float = 1e-10, max_iter: int = 1000, restart: int = 1000):
        """Complex boundary data on a real operator (real kappa on ROBIN DOFs): the frequency-domain
        diffraction BVP of main_freq_domain.cpp:650-687 (incident-wave phi_n on the body, Robin free
        surface), which the reference assembles as a dense complex system (solve_linear_bvp,
        BEM_freqency_domain.hpp:277), solved as two real BVPs on one prepared plan: x = x(Re) + i x(Im).
        Prepares the context if needed; returns (x, [info_re, info_im])."""
        phi, phin = np.asarray(phi, dtype=complex), np.asarray(phin, dtype=complex)
        self.set_bc(phi.real, phin.real)
        if not getattr(self, "prepared", False):
            self.prepare()
        xr, ir = self.solveSystemGMRES(None, tol, max_iter, restart)
        self.set_bc(phi.imag, phin.imag)
        xi, ii = self.solveSystemGMRES(None, tol, max_iter, restart)
        return xr + 1j * xi, [ir, ii]

    def solveGMRES(self, phi, phin, x0=None, tol: float = 1e-10, max_iter: int = 1000, restart: int = 1000):
        """Setup (scaling, tree, operators, near field, alpha) + GMRES, as BEM_BVP::solveGMRES."""
        self.set_bc(phi, phin)
        self.prepare()
        return self.solveSystemGMRES(x0, tol, max_iter, restart)
