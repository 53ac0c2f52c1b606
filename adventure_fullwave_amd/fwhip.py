"""ctypes front end of libfwhip.so (C ABI: include/fwhip.h).

Python-side mirror of the reference's solver entry point for this path,
hddm_solver_CS(PartMesh*, OPTIONS, set_hddm_func) (lib/hddm/hddm_solver.h:25-183):
`Solver(model).solve(eps, max_loop)` runs LOOP_1ST / main loop / LOOP_LAST on
one GPU and returns the per-subdomain solutions, the owned interface values
and the residual history the reference logs to calc_log/log_g_HDDM_Full_Wave.

The library is loaded from this package directory; there is no CPU fallback:
if libfwhip.so is missing or no GPU is visible, the calls raise.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfwhip.so")
_LIB = None

FWH_E = {-1: "FWH_E_ARG", -2: "FWH_E_HIP", -3: "FWH_E_MESH", -4: "FWH_E_STATE", -5: "FWH_E_LIMIT"}

EXPORTS = ("fwh_create", "fwh_destroy", "fwh_last_error", "fwh_load", "fwh_get_info", "fwh_solve",
           "fwh_first", "fwh_iterate", "fwh_last", "fwh_get_stats", "fwh_get_x", "fwh_get_u",
           "fwh_get_resid", "fwh_schur_apply", "fwh_ldl_solve_all", "fwh_time_kernels", "fwh_get_sub",
           "fwh_element_batch", "fwh_cocg_icc_all", "fwh_schur_form", "fwh_post", "fwh_set_load",
           "fwh_grad_eps", "fwh_load_batch", "fwh_select", "fwh_element_batch_static", "fwh_sky_factor",
           "fwh_sky_solve", "fwh_release_factors", "fwh_set_factor_budget", "fwh_aij_cocg",
           "fwh_aij_cocg_pc")
# the reference-shaped call surface (include/fwh_ref.h)
REF_EXPORTS = ("hddm_solver_CS_hip", "hddm_solver_CS_hip_status", "fwh_ref_desc_from_partmesh",
               "fwh_ref_desc_free", "advlas_skyline_ldl_decomposite", "advlas_skyline_ldl_solve",
               "mymt_CS_AIJ_Solver_COCG", "mymt_CS_AIJ_Solver_COCG_nmpcm", "set_hddm_func_fullwave_hip")


class FwhOpts(ctypes.Structure):
    _fields_ = [("hddm_eps", ctypes.c_double), ("hddm_max_loop", ctypes.c_int),
                ("plain", ctypes.c_int), ("graph_iters", ctypes.c_int), ("schur", ctypes.c_int)]


class FwhStats(ctypes.Structure):
    _fields_ = [("loops", ctypes.c_int), ("iters", ctypes.c_int), ("converged", ctypes.c_int),
                ("r0", ctypes.c_double), ("final_resid", ctypes.c_double), ("ms_first", ctypes.c_double),
                ("ms_loop", ctypes.c_double), ("ms_last", ctypes.c_double), ("sub_solves", ctypes.c_long)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class FwhInfo(ctypes.Structure):
    _fields_ = [("nsub", ctypes.c_int), ("npart", ctypes.c_int), ("n_iface", ctypes.c_int),
                ("n_dof", ctypes.c_long), ("n_slots", ctypes.c_long), ("nnz", ctypes.c_long),
                ("nlen", ctypes.c_long), ("n_elem", ctypes.c_long), ("n_cpl", ctypes.c_long),
                ("n_grow", ctypes.c_long), ("max_n", ctypes.c_int), ("ms_assemble", ctypes.c_double),
                ("ms_factor", ctypes.c_double), ("bytes_solve", ctypes.c_double),
                ("bytes_iter", ctypes.c_double), ("flops_factor", ctypes.c_double),
                ("ms_schur_form", ctypes.c_double), ("smat_bytes", ctypes.c_double),
                ("bytes_smv", ctypes.c_double), ("solve_buckets", ctypes.c_int),
                ("solve_bucketed", ctypes.c_int), ("ms_solve_single", ctypes.c_double),
                ("ms_solve_bucketed", ctypes.c_double), ("smv_waves", ctypes.c_int),
                ("ms_smv_tune", ctypes.c_double * 3), ("solve_depth", ctypes.c_int),
                ("unit_depth", ctypes.c_int), ("ms_solve_depth", ctypes.c_double * 2),
                ("nsys", ctypes.c_int), ("smv_order", ctypes.c_int), ("ms_smv_order", ctypes.c_double * 3),
                ("schur_gemm", ctypes.c_int),
                ("factor_blocked", ctypes.c_int),
                ("ms_host_index", ctypes.c_double), ("ms_load", ctypes.c_double),
                ("ms_factor_kernel", ctypes.c_double), ("ms_schur_fwd", ctypes.c_double),
                ("ms_schur_gemm", ctypes.c_double), ("flops_schur_fwd", ctypes.c_double),
                ("flops_schur_gemm", ctypes.c_double), ("smv_split", ctypes.c_int),
                ("smv_nwork", ctypes.c_int), ("ms_smv_split", ctypes.c_double * 4),
                ("solve_waves", ctypes.c_int), ("ms_solve_waves", ctypes.c_double * 3),
                ("smv_persist", ctypes.c_int), ("smv_pgrid", ctypes.c_int),
                ("ms_smv_persist", ctypes.c_double * 2),
                ("solve_persist", ctypes.c_int), ("solve_pgrid", ctypes.c_int),
                ("solve_nlarge", ctypes.c_int), ("ms_solve_persist", ctypes.c_double * 2),
                ("factors_resident", ctypes.c_int), ("factor_chunks", ctypes.c_int),
                ("factor_bytes_released", ctypes.c_double), ("factor_chunk_bytes", ctypes.c_double),
                ("factor_need_max", ctypes.c_double), ("factor_need_total", ctypes.c_double),
                ("ms_solve_wide", ctypes.c_double)]

    def as_dict(self):
        out = {}
        for k, _ in self._fields_:
            v = getattr(self, k)
            out[k] = list(v) if isinstance(v, ctypes.Array) else v
        return out


class FwhError(RuntimeError):
    pass


def lib():
    """Load libfwhip.so (raises if it was not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise FwhError(f"{LIB_PATH} not built: run adventure_fullwave_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        I = ctypes.c_int
        D = ctypes.c_double
        L.fwh_last_error.restype = ctypes.c_char_p
        L.fwh_create.argtypes = [ctypes.POINTER(P), I]
        L.fwh_destroy.argtypes = [P]
        L.fwh_destroy.restype = None
        L.fwh_load.argtypes = [P, P]
        L.fwh_load_batch.argtypes = [P, P, I]
        L.fwh_select.argtypes = [P, I]
        L.fwh_get_info.argtypes = [P, P]
        L.fwh_solve.argtypes = [P, P, P]
        L.fwh_first.argtypes = [P, P]
        L.fwh_iterate.argtypes = [P, P, I, ctypes.POINTER(I)]
        L.fwh_last.argtypes = [P]
        L.fwh_get_stats.argtypes = [P, P]
        L.fwh_get_x.argtypes = [P, P]
        L.fwh_get_u.argtypes = [P, P]
        L.fwh_get_resid.argtypes = [P, P, I]
        L.fwh_schur_apply.argtypes = [P, P, P, I]
        L.fwh_ldl_solve_all.argtypes = [P, P, I]
        L.fwh_time_kernels.argtypes = [P, I, ctypes.POINTER(D), ctypes.POINTER(D), ctypes.POINTER(D)]
        L.fwh_schur_form.argtypes = [P]
        L.fwh_release_factors.argtypes = [P, ctypes.c_longlong]
        L.fwh_set_factor_budget.argtypes = [P, ctypes.c_longlong]
        L.fwh_post.argtypes = [P, P, P, P, P]
        L.fwh_set_load.argtypes = [P, P]
        L.fwh_grad_eps.argtypes = [P, P, P, P]
        L.fwh_get_sub.argtypes = [P, I, I, P]
        L.fwh_element_batch.argtypes = [P, I, P, P, P, D, P, P, P, P, P]
        L.fwh_element_batch_static.argtypes = [P, I, P, P, P, D, P, P, P, P, P]
        L.fwh_sky_factor.argtypes = [P, I, P, P, P, P]
        L.fwh_sky_solve.argtypes = [P, I, P, P, P, P]
        L.fwh_cocg_icc_all.argtypes = [P, P, P, P, D, D]
        for name in EXPORTS:
            if name not in ("fwh_destroy", "fwh_last_error"):
                getattr(L, name).restype = I
        _LIB = L
    return _LIB


def _check(rc: int, what: str):
    if rc != 0:
        msg = lib().fwh_last_error().decode(errors="replace")
        raise FwhError(f"{what}: {FWH_E.get(rc, rc)}: {msg}")


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data


class Context:
    """One device context (fwh_create / fwh_destroy)."""

    def __init__(self, device: int = 0):
        self.h = ctypes.c_void_p()
        _check(lib().fwh_create(ctypes.byref(self.h), device), "fwh_create")

    def close(self):
        if self.h:
            lib().fwh_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def element_batch(self, xyz, eps, sigma, omega, pml=None, src=None, joe=None):
        nel = len(eps)
        xyz = np.ascontiguousarray(xyz, np.float64).reshape(nel, 12)
        eps = np.ascontiguousarray(eps, np.float64)
        sigma = np.ascontiguousarray(sigma, np.float64)
        pml = None if pml is None else np.ascontiguousarray(pml, np.float64).reshape(nel, 6)
        src = None if src is None else np.ascontiguousarray(src, np.int32)
        joe = None if joe is None else np.ascontiguousarray(joe, np.float64).reshape(nel, 24)
        ae = np.zeros((nel, 72))
        be = np.zeros((nel, 12))
        _check(lib().fwh_element_batch(self.h, nel, _ptr(xyz), _ptr(eps), _ptr(sigma), float(omega),
                                       _ptr(pml), _ptr(src), _ptr(joe), _ptr(ae), _ptr(be)),
               "fwh_element_batch")
        return ae.view(np.complex128).reshape(nel, 6, 6), be.view(np.complex128).reshape(nel, 6)


class Solver:
    """A decomposed model resident on the GPU (fwh_load) and its HDDM solve.

    model may also be a list of models on one mesh and decomposition (a
    frequency sweep, say): they are loaded as one batch (fwh_load_batch) and
    solved together; `select(k)` picks the system that the per-system
    methods (stats, x, u, resid, post, set_load, grad_eps, schur_apply,
    ldl_solve_all, sub) address."""

    def __init__(self, model, device: int = 0, ctx: Optional[Context] = None, factor_budget: int = 0):
        """factor_budget (bytes, 0: none): a model whose factors exceed it is
        loaded streamed (fwh_set_factor_budget)"""
        self.models = list(model) if isinstance(model, (list, tuple)) else [model]
        self.model = self.models[0]
        self.ctx = ctx or Context(device)
        _check(lib().fwh_set_factor_budget(self.ctx.h, int(factor_budget)), "fwh_set_factor_budget")
        if len(self.models) == 1:
            self._desc = self.model.desc()
            _check(lib().fwh_load(self.ctx.h, ctypes.byref(self._desc)), "fwh_load")
        else:
            from .mesh import FwhMeshDesc
            descs = [m.desc() for m in self.models]
            self._desc = (FwhMeshDesc * len(descs))(*descs)
            self._desc_keep = descs
            _check(lib().fwh_load_batch(self.ctx.h, self._desc, len(descs)), "fwh_load_batch")
        self.sel = 0
        inf = FwhInfo()
        _check(lib().fwh_get_info(self.ctx.h, ctypes.byref(inf)), "fwh_get_info")
        self.info = inf.as_dict()

    @property
    def nsys(self) -> int:
        return len(self.models)

    def select(self, k: int):
        """the system that the per-system methods address"""
        _check(lib().fwh_select(self.ctx.h, int(k)), "fwh_select")
        self.sel = int(k)
        return self

    @staticmethod
    def opts(eps=1e-8, max_loop=100000, plain=False, graph_iters=16, schur=False):
        return FwhOpts(float(eps), int(max_loop), int(bool(plain)), int(graph_iters), int(bool(schur)))

    def solve(self, eps=1e-8, max_loop=100000, plain=False, graph_iters=16, schur=False):
        """schur=True: explicit subdomain Schur complements (formed once)."""
        o = self.opts(eps, max_loop, plain, graph_iters, schur)
        st = FwhStats()
        _check(lib().fwh_solve(self.ctx.h, ctypes.byref(o), ctypes.byref(st)), "fwh_solve")
        self.last_opts = o
        self._refresh_info()
        return st.as_dict()

    def first(self, eps=1e-8, max_loop=100000, plain=False, graph_iters=16, schur=False):
        self.last_opts = self.opts(eps, max_loop, plain, graph_iters, schur)
        _check(lib().fwh_first(self.ctx.h, ctypes.byref(self.last_opts)), "fwh_first")
        self._refresh_info()

    def _refresh_info(self):
        inf = FwhInfo()
        _check(lib().fwh_get_info(self.ctx.h, ctypes.byref(inf)), "fwh_get_info")
        self.info = inf.as_dict()

    def schur_form(self):
        _check(lib().fwh_schur_form(self.ctx.h), "fwh_schur_form")
        self._refresh_info()

    def release_factors(self, budget_bytes):
        """free the skyline factors after the Schur formation; the solves that
        need them recompute them in chunks of at most budget_bytes"""
        _check(lib().fwh_release_factors(self.ctx.h, int(budget_bytes)), "fwh_release_factors")
        self._refresh_info()

    def iterate(self, n):
        done = ctypes.c_int()
        _check(lib().fwh_iterate(self.ctx.h, ctypes.byref(self.last_opts), int(n), ctypes.byref(done)),
               "fwh_iterate")
        return bool(done.value)

    def last(self):
        _check(lib().fwh_last(self.ctx.h), "fwh_last")

    def stats(self):
        st = FwhStats()
        _check(lib().fwh_get_stats(self.ctx.h, ctypes.byref(st)), "fwh_get_stats")
        return st.as_dict()

    def stats_all(self):
        """statistics of every system of the batch"""
        keep = self.sel
        out = [self.select(k).stats() for k in range(self.nsys)]
        self.select(keep)
        return out

    def x(self):
        out = np.zeros(int(self.info["n_dof"]), np.complex128)
        _check(lib().fwh_get_x(self.ctx.h, _ptr(out)), "fwh_get_x")
        return out

    def x_list(self):
        x = self.x()
        off = np.concatenate([[0], np.cumsum(np.asarray(self.model.nmp, np.int64))])
        return [x[off[s]:off[s + 1]] for s in range(self.model.nsub)]

    def u(self):
        out = np.zeros(int(self.info["n_iface"]), np.complex128)
        _check(lib().fwh_get_u(self.ctx.h, _ptr(out)), "fwh_get_u")
        return out

    def resid(self, n):
        out = np.zeros(int(n))
        _check(lib().fwh_get_resid(self.ctx.h, _ptr(out), int(n)), "fwh_get_resid")
        return out

    def post(self, nu=None):
        """Per-element (rotE, E, Je) of the last solution, each (nel, 3)
        complex (HDDM_post_em_E, main/fullwave/fullwave.c:640-763); nu: per
        element 1/mu or None for 7.957747e5 everywhere."""
        nel = int(self.info["n_elem"])
        out = [np.zeros((nel, 3), np.complex128) for _ in range(3)]
        nu_p = None
        if nu is not None:
            nu = np.ascontiguousarray(nu, np.float64)
            if nu.shape != (nel,):
                raise FwhError(f"nu must have shape ({nel},)")
            nu_p = _ptr(nu)
        _check(lib().fwh_post(self.ctx.h, nu_p, *(_ptr(o) for o in out)), "fwh_post")
        return tuple(out)

    def set_load(self, f=None):
        """Per-DOF load of the next solve (fwh_get_x layout), None: the
        assembled load."""
        if f is None:
            _check(lib().fwh_set_load(self.ctx.h, None), "fwh_set_load")
            return
        f = np.ascontiguousarray(f, np.complex128)
        if f.shape != (int(self.info["n_dof"]),):
            raise FwhError("load must have n_dof entries")
        _check(lib().fwh_set_load(self.ctx.h, _ptr(f)), "fwh_set_load")

    def grad_eps(self, xE, xL):
        """dJ/deps per element from forward and adjoint solutions."""
        xE = np.ascontiguousarray(xE, np.complex128)
        xL = np.ascontiguousarray(xL, np.complex128)
        n = int(self.info["n_dof"])
        if xE.shape != (n,) or xL.shape != (n,):
            raise FwhError("solutions must have n_dof entries")
        g = np.zeros(int(self.info["n_elem"]))
        _check(lib().fwh_grad_eps(self.ctx.h, _ptr(xE), _ptr(xL), _ptr(g)), "fwh_grad_eps")
        return g

    def schur_apply(self, p, plain=False, explicit=False):
        p = np.ascontiguousarray(p, np.complex128)
        if p.shape != (int(self.info["n_iface"]),):
            raise FwhError(f"p must have shape ({int(self.info['n_iface'])},)")
        q = np.zeros_like(p)
        mode = int(bool(plain)) + 2 * int(bool(explicit))
        _check(lib().fwh_schur_apply(self.ctx.h, _ptr(p), _ptr(q), mode), "fwh_schur_apply")
        if explicit:
            self._refresh_info()
        return q

    def ldl_solve_all(self, rhs, plain=False):
        v = np.ascontiguousarray(rhs, np.complex128).copy()
        if v.shape != (int(self.info["n_dof"]),):
            raise FwhError(f"rhs must have shape ({int(self.info['n_dof'])},)")
        _check(lib().fwh_ldl_solve_all(self.ctx.h, _ptr(v), int(bool(plain))), "fwh_ldl_solve_all")
        return v

    def time_kernels(self, reps=10):
        """(ms per batched LDL^T solve, per implicit Schur step, per explicit
        Schur matvec or -1 when S_s are not formed)."""
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _check(lib().fwh_time_kernels(self.ctx.h, int(reps), ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)),
               "fwh_time_kernels")
        return a.value, b.value, c.value

    def sub(self, s, factor=True):
        """subdomain s: assembled matrix, load, profile and (factor=True) its
        skyline factor"""
        L = lib()
        sz = np.zeros(4, np.int32)
        _check(L.fwh_get_sub(self.ctx.h, s, 7, _ptr(sz)), "fwh_get_sub")
        n, nnz, nlen, ng = (int(x) for x in sz)
        val = np.zeros(max(nnz, 1), np.complex128)
        rowptr = np.zeros(n + 1, np.int32)
        col = np.zeros(max(nnz, 1), np.int32)
        f = np.zeros(n, np.complex128)
        idx1 = np.zeros(n, np.int32)
        idx2 = np.zeros(n, np.int32)
        fac = np.zeros(max(nlen, 1), np.complex128)
        parts = [(0, val), (1, rowptr), (2, col), (3, f), (4, idx1), (5, idx2)] + ([(6, fac)] if factor else [])
        for what, arr in parts:
            _check(L.fwh_get_sub(self.ctx.h, s, what, _ptr(arr)), "fwh_get_sub")
        return dict(n=n, nnz=nnz, nlen=nlen, ng=ng, val=val[:nnz], rowptr=rowptr, col=col[:nnz], f=f,
                    idx1=idx1, idx2=idx2, factor=fac[:nlen] if factor else None)

    def close(self):
        self.ctx.close()
