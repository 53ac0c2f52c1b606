"""ctypes front end of the C oracle (oracle/fw_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never the product.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class FwoOpts(ctypes.Structure):
    _fields_ = [("hddm_eps", ctypes.c_double), ("hddm_max_loop", ctypes.c_int),
                ("dom_solver", ctypes.c_int), ("sol_eps", ctypes.c_double),
                ("ic_shift", ctypes.c_double), ("max_sub", ctypes.c_int),
                ("perturb", ctypes.c_int), ("nthreads", ctypes.c_int)]


class FwoStats(ctypes.Structure):
    _fields_ = [("loops", ctypes.c_int), ("iters", ctypes.c_int), ("converged", ctypes.c_int),
                ("r0", ctypes.c_double), ("seconds_loop", ctypes.c_double),
                ("sub_solves", ctypes.c_long)]


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libfworacle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        L.fwo_element.restype = ctypes.c_int
        L.fwo_element.argtypes = [P, ctypes.c_double, ctypes.c_double, ctypes.c_double, P,
                                  ctypes.c_int, P, P, P]
        L.fwo_post_element.argtypes = [P, P, ctypes.c_double, ctypes.c_double, ctypes.c_double, P, P, P]
        L.fwo_dg_load.restype = ctypes.c_int
        L.fwo_dg_load.argtypes = [P, ctypes.c_double, ctypes.c_double, P, P]
        L.fwo_sky_factor.argtypes = [ctypes.c_int, P, P, P]
        L.fwo_sky_solve.argtypes = [ctypes.c_int, P, P, P, P]
        L.fwo_model_build.restype = P
        L.fwo_model_build.argtypes = [P]
        L.fwo_model_free.argtypes = [P]
        L.fwo_model_sizes.argtypes = [P, ctypes.c_int, P, P, P, P]
        L.fwo_model_get.argtypes = [P, ctypes.c_int, P, P, P, P, P, P, P]
        L.fwo_model_schur_sub.argtypes = [P, ctypes.c_int, P, P, P]
        L.fwo_hddm_solve.restype = ctypes.c_int
        L.fwo_hddm_solve.argtypes = [P, P, P, P, P, P]
        L.fwo_hddm_solve_load.restype = ctypes.c_int
        L.fwo_hddm_solve_load.argtypes = [P, P, P, P, P, P, P]
        L.fwo_hddm_time_iterations.restype = ctypes.c_double
        L.fwo_hddm_time_iterations.argtypes = [P, ctypes.c_int, P, ctypes.c_int]
        L.fwo_cocg_icc.restype = ctypes.c_int
        L.fwo_cocg_icc.argtypes = [ctypes.c_int, P, P, P, P, P, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_int, P]
        L.fwo_icc_make.argtypes = [ctypes.c_int, P, P, P, ctypes.c_double, P]
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data if a is not None else None


def element(xyz, eps, sigma, omega, pml=None, src=0, joe=None):
    """Element matrix (6x6 complex) and load (6 complex) of one tetrahedron."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(12)
    ae = np.zeros(72)
    be = np.zeros(12)
    J = np.zeros(24) if joe is None else np.ascontiguousarray(joe, dtype=np.float64).reshape(24)
    pm = None if pml is None else np.ascontiguousarray(pml, dtype=np.float64).reshape(6)
    rc = lib().fwo_element(_p(xyz), eps, sigma, omega, _p(pm), int(src), _p(J), _p(ae), _p(be))
    if rc != 0:
        raise ValueError(f"fwo_element failed ({rc})")
    return ae.view(np.complex128).reshape(6, 6), be.view(np.complex128)


def dg_load(xyz, sigma, omega, dv):
    """Delta-gap load (6 complex) of one tetrahedron with vertex gap voltages dv."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(12)
    dv = np.ascontiguousarray(dv, dtype=np.float64).reshape(4)
    be = np.zeros(12)
    rc = lib().fwo_dg_load(_p(xyz), float(sigma), float(omega), _p(dv), _p(be))
    if rc != 0:
        raise ValueError(f"fwo_dg_load failed ({rc})")
    return be.view(np.complex128)


def post_element(xyz, xe, nu, sigma, omega):
    """(rotE, E, Je) of one element, 3 complex each (HDDM_post_em_E)."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(12)
    x = np.ascontiguousarray(xe, dtype=np.complex128).reshape(6)
    out = [np.zeros(3, np.complex128) for _ in range(3)]
    rc = lib().fwo_post_element(_p(xyz), _p(x), float(nu), float(sigma), float(omega), *(_p(o) for o in out))
    if rc != 0:
        raise ValueError(f"fwo_post_element failed ({rc})")
    return tuple(out)


def post_model(model, x_flat, nu=None):
    """post_element over every element of a decomposed model; x_flat is the
    per-subdomain solution concatenated (2*nmp doubles each)."""
    nop = np.asarray(model.nop).reshape(-1, 10)
    nel = nop.shape[0]
    dof0 = np.concatenate([[0], np.cumsum(np.asarray(model.nmp, np.int64))])
    out = [np.zeros((nel, 3), np.complex128) for _ in range(3)]
    crd = np.asarray(model.crd).reshape(-1, 3)
    for s in range(model.nsub):
        for el in range(model.el_off[s], model.el_off[s + 1]):
            v = crd[model.ap_off[s] + nop[el, :4]]
            xe = x_flat[dof0[s] + nop[el, 4:]]
            r = post_element(v, xe, 7.957747e+05 if nu is None else nu[el], model.sigma[el], model.omega)
            for o, val in zip(out, r):
                o[el] = val
    return tuple(out)


def sky_factor(idx1, idx2, mat):
    idx1 = np.ascontiguousarray(idx1, np.int32)
    idx2 = np.ascontiguousarray(idx2, np.int32)
    m = np.ascontiguousarray(mat, np.complex128).copy()
    lib().fwo_sky_factor(len(idx1), _p(idx1), _p(idx2), _p(m))
    return m


def sky_solve(idx1, idx2, fac, rhs):
    idx1 = np.ascontiguousarray(idx1, np.int32)
    idx2 = np.ascontiguousarray(idx2, np.int32)
    fac = np.ascontiguousarray(fac, np.complex128)
    v = np.ascontiguousarray(rhs, np.complex128).copy()
    lib().fwo_sky_solve(len(idx1), _p(idx1), _p(idx2), _p(fac), _p(v))
    return v


class OracleModel:
    """Assembled + factored subdomains of a mesh description."""

    def __init__(self, model):
        self.model = model
        self._desc = model.desc()
        self.h = lib().fwo_model_build(ctypes.byref(self._desc))
        if not self.h:
            raise RuntimeError("fwo_model_build failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().fwo_model_free(self.h)
            self.h = None

    def sizes(self, s):
        v = [ctypes.c_int() for _ in range(4)]
        lib().fwo_model_sizes(self.h, s, *[ctypes.byref(x) for x in v])
        return tuple(x.value for x in v)  # n, nlen, nnz, nbc

    def get(self, s):
        n, nlen, nnz, nbc = self.sizes(s)
        rowptr = np.zeros(n + 1, np.int32)
        col = np.zeros(max(nnz, 1), np.int32)
        val = np.zeros(max(nnz, 1), np.complex128)
        f = np.zeros(n, np.complex128)
        idx1 = np.zeros(n, np.int32)
        idx2 = np.zeros(n, np.int32)
        fac = np.zeros(max(nlen, 1), np.complex128)
        lib().fwo_model_get(self.h, s, _p(rowptr), _p(col), _p(val), _p(f), _p(idx1), _p(idx2), _p(fac))
        return dict(rowptr=rowptr, col=col[:nnz], val=val[:nnz], f=f, idx1=idx1, idx2=idx2,
                    factor=fac[:nlen], n=n, nlen=nlen, nnz=nnz, nbc=nbc)

    def schur_sub(self, s, vbc):
        n = self.sizes(s)[0]
        vbc = np.ascontiguousarray(vbc, np.complex128)
        x = np.zeros(n, np.complex128)
        y = np.zeros(n, np.complex128)
        lib().fwo_model_schur_sub(self.h, s, _p(vbc), _p(x), _p(y))
        return x, y


def hddm_solve(model, eps=1e-8, max_loop=100000, dom_solver=0, sol_eps=1e-9, ic_shift=1.2, perturb=0,
               threads=1, load=None):
    """Full s-mode HDDM solve; returns dict(resid, x (per-subdomain list),
    u (owned interface values, part-major), stats).  threads > 1: the main
    loop's subdomain Schur steps on OpenMP threads (same sums).  load: a
    per-DOF load (complex, the layout of x_flat) replacing the assembled one."""
    d = model.desc()
    o = FwoOpts(eps, max_loop, dom_solver, sol_eps, ic_shift, 0, perturb, int(threads))
    resid = np.zeros(max_loop + 1)
    nmp = np.asarray(model.nmp, np.int64)
    x = np.zeros(int(nmp.sum()), np.complex128)
    u = np.zeros(int(np.asarray(model.infree).sum()), np.complex128)
    st = FwoStats()
    if load is None:
        rc = lib().fwo_hddm_solve(ctypes.byref(d), ctypes.byref(o), _p(resid), _p(x), _p(u),
                                  ctypes.byref(st))
    else:
        f = np.ascontiguousarray(load, np.complex128)
        if f.shape != x.shape:
            raise ValueError("load must have one entry per subdomain DOF")
        rc = lib().fwo_hddm_solve_load(ctypes.byref(d), ctypes.byref(o), _p(f), _p(resid), _p(x), _p(u),
                                       ctypes.byref(st))
    if rc != 0:
        raise RuntimeError(f"fwo_hddm_solve failed ({rc})")
    offs = np.concatenate([[0], np.cumsum(nmp)])
    xs = [x[offs[s]:offs[s + 1]] for s in range(len(nmp))]
    return dict(resid=resid[: st.loops], x=xs, x_flat=x, u=u,
                stats=dict(loops=st.loops, iters=st.iters, converged=st.converged, r0=st.r0,
                           seconds_loop=st.seconds_loop, sub_solves=st.sub_solves))


def time_iterations(model, n_iter, threads=1):
    """seconds for n_iter interface iterations (subdomain Schur steps on
    `threads` OpenMP threads) and the number of subdomain solves"""
    d = model.desc()
    ns = ctypes.c_long()
    t = lib().fwo_hddm_time_iterations(ctypes.byref(d), n_iter, ctypes.byref(ns), int(threads))
    return t, ns.value


def cocg_icc(rowptr, col, val, b, x0=None, eps=1e-9, ic_shift=1.2, use_icc=True, hist_len=None):
    n = len(rowptr) - 1
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.complex128)
    b = np.ascontiguousarray(b, np.complex128)
    x = np.zeros(n, np.complex128) if x0 is None else np.ascontiguousarray(x0, np.complex128).copy()
    hist = np.zeros(10 * n + 2)
    st = lib().fwo_cocg_icc(n, _p(rowptr), _p(col), _p(val), _p(b), _p(x), eps, ic_shift,
                            int(use_icc), _p(hist))
    return st, x, hist


def icc_make(rowptr, col, val, ic_shift=1.2):
    n = len(rowptr) - 1
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.complex128)
    pc = np.zeros(int(rowptr[n]) + n, np.complex128)
    lib().fwo_icc_make(n, _p(rowptr), _p(col), _p(val), ic_shift, _p(pc))
    return pc
