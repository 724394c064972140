"""ctypes wrapper of the C oracle (oracle/c/oracle.c) -- test infrastructure only."""

from __future__ import annotations

import ctypes
import os

import numpy as np

from . import c_build

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(c_build.OUT):
            c_build.build()
        L = ctypes.CDLL(c_build.OUT)
        vp, i64 = ctypes.c_void_p, ctypes.c_int64
        L.oracle_snh_fun.restype = ctypes.c_double
        L.oracle_snh_fun.argtypes = [i64, vp, vp, vp, vp, vp, vp]
        L.oracle_snh_grad.argtypes = [i64, vp, vp, vp, vp, vp, vp, vp]
        L.oracle_snh_hess_diag.argtypes = [i64, vp, vp, vp, vp, vp, vp, vp]
        L.oracle_snh_hess_quad.restype = ctypes.c_double
        L.oracle_snh_hess_quad.argtypes = [i64, vp, vp, vp, vp, vp, vp, vp]
        L.oracle_snh_pass.restype = ctypes.c_double
        L.oracle_snh_pass.argtypes = [i64, i64, vp, vp, vp, vp, vp, vp, vp, vp]
        L.oracle_set_threads.restype = ctypes.c_int
        L.oracle_set_threads.argtypes = [ctypes.c_int]
        L.oracle_snh_pncg.restype = ctypes.c_int
        L.oracle_snh_pncg.argtypes = [i64, i64, vp, vp, vp, vp, vp, vp, vp, vp, ctypes.c_double, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_double, ctypes.POINTER(ctypes.c_int),
                                      ctypes.POINTER(ctypes.c_double)]
        _lib = L
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


class SnhProblem:
    """Arrays for the C restatement (FP64, reference layout: dhdX (T,4,3))."""

    def __init__(self, cells, dhdX, dV, mu, lam):
        self.cells = _c(cells, np.int32)
        self.dhdX = _c(dhdX, np.float64).reshape(-1, 12)
        self.dV = _c(dV, np.float64)
        self.mu = _c(mu, np.float64)
        self.lam = _c(lam, np.float64)
        self.T = self.cells.shape[0]

    def _args(self):
        return (self.cells.ctypes.data, self.dhdX.ctypes.data, self.dV.ctypes.data, self.mu.ctypes.data,
                self.lam.ctypes.data)

    def fun(self, u):
        u = _c(u, np.float64)
        return lib().oracle_snh_fun(self.T, *self._args(), u.ctypes.data)

    def grad(self, u):
        u = _c(u, np.float64)
        out = np.zeros_like(u)
        lib().oracle_snh_grad(self.T, *self._args(), u.ctypes.data, out.ctypes.data)
        return out

    def hess_diag(self, u):
        u = _c(u, np.float64)
        out = np.zeros_like(u)
        lib().oracle_snh_hess_diag(self.T, *self._args(), u.ctypes.data, out.ctypes.data)
        return out

    def hess_quad(self, u, p):
        u = _c(u, np.float64)
        p = _c(p, np.float64)
        return lib().oracle_snh_hess_quad(self.T, *self._args(), u.ctypes.data, p.ctypes.data)

    def pass_a(self, u, grad, hdiag):
        return lib().oracle_snh_pass(self.T, u.shape[0], *self._args(), u.ctypes.data, grad.ctypes.data,
                                     hdiag.ctypes.data)

    def pncg(self, x, fext, free_dof, rtol=1e-6, max_steps=1500, max_backtracking=8, wolfe_delta=0.1):
        x = _c(x, np.float64).copy()
        fext = _c(fext, np.float64)
        free = _c(free_dof, np.uint8)
        st = ctypes.c_int()
        gn = ctypes.c_double()
        steps = lib().oracle_snh_pncg(self.T, x.shape[0], *self._args(), fext.ctypes.data, free.ctypes.data,
                                      x.ctypes.data, rtol, max_steps, max_backtracking, wolfe_delta,
                                      ctypes.byref(st), ctypes.byref(gn))
        return x, steps, st.value, gn.value


def set_threads(n: int) -> int:
    return lib().oracle_set_threads(int(n))
