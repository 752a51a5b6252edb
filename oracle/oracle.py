"""ctypes wrapper of the CPU oracle (liboracle.so built from nosh_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker of the HIP library. The product
package (nosh_amd) never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


def build() -> None:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-B", "-C", _HERE, "liboracle.so"], check=True,
                   stdout=subprocess.DEVNULL)


def _load():
    if not os.path.exists(_LIB_PATH):
        build()
    lib = C.CDLL(_LIB_PATH)
    vp = C.c_void_p
    sig = {
        "orc_mesh_new": (vp, [_dp, C.c_int64, _ip, C.c_int64, C.c_int]),
        "orc_mesh_free": (None, [vp]),
        "orc_mesh_num_edges": (C.c_int64, [vp]),
        "orc_mesh_edges": (None, [vp, _ip]),
        "orc_mesh_cell_edges": (None, [vp, _ip]),
        "orc_mesh_edge_data": (None, [vp, _dp, _dp]),
        "orc_mesh_control_volumes": (None, [vp, _dp]),
        "orc_cell_edge_coefficients": (None, [vp, _dp]),
        "orc_model_new": (vp, [vp, _dp, _dp, _dp, _dp, C.c_char_p]),
        "orc_model_free": (None, [vp]),
        "orc_model_edge_fields": (None, [vp, _dp, _dp]),
        "orc_set_params": (None, [vp, C.c_double, C.c_double]),
        "orc_set_v_param": (None, [vp, C.c_char_p]),
        "orc_keo_nnz": (C.c_int64, [vp]),
        "orc_crs_get": (None, [vp, C.c_int, _lp, _ip, _dp]),
        "orc_keo_apply": (None, [vp, _dp, _dp]),
        "orc_dkeo_apply": (None, [vp, _dp, _dp]),
        "orc_compute_f": (None, [vp, _dp, _dp]),
        "orc_compute_dfdp": (None, [vp, _dp, C.c_char_p, _dp]),
        "orc_jac_rebuild": (None, [vp, _dp]),
        "orc_jac_apply": (None, [vp, _dp, _dp]),
        "orc_prec_rebuild": (None, [vp, _dp]),
        "orc_prec_op_apply": (None, [vp, _dp, _dp]),
        "orc_set_threads": (C.c_int, [C.c_int]),
        "orc_cheb_apply": (None, [vp, C.c_int, C.c_double, C.c_double, _dp, _dp]),
        "orc_cheb_bound": (C.c_double, [vp]),
        "orc_minres_jac": (C.c_int, [vp, _dp, _dp, C.c_double, C.c_int, C.c_int, C.c_double,
                                     _dp, C.c_void_p]),
        "orc_minres_crs": (C.c_int, [C.c_int64, _lp, _ip, _dp, _dp, _dp, C.c_double,
                                     C.c_int, _dp]),
        "orc_newton": (C.c_int, [vp, _dp, C.c_double, C.c_double, C.c_double, C.c_int,
                                 C.c_double, C.c_int, C.c_int, C.c_double, _dp,
                                 np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_lib = None


def set_threads(n: int) -> int:
    """OpenMP threads of the oracle's parallel loops (0: query); returns the count."""
    return int(lib().orc_set_threads(int(n)))


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


class Mesh:
    """Oracle mesh: edges, cell edges, edge data and control volumes."""

    def __init__(self, coords, cells):
        self.coords = np.ascontiguousarray(coords, dtype=np.float64)
        self.cells = np.ascontiguousarray(cells, dtype=np.int32)
        self.n = self.coords.shape[0]
        self.nv = self.cells.shape[1]
        self.ne = 3 if self.nv == 3 else 6
        self._h = lib().orc_mesh_new(self.coords, self.n, self.cells, self.cells.shape[0],
                                     self.nv)
        self.E = int(lib().orc_mesh_num_edges(self._h))

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.orc_mesh_free(self._h)
            self._h = None

    def edges(self):
        out = np.empty((self.E, 2), dtype=np.int32)
        lib().orc_mesh_edges(self._h, out)
        return out

    def cell_edges(self):
        out = np.empty((self.cells.shape[0], self.ne), dtype=np.int32)
        lib().orc_mesh_cell_edges(self._h, out)
        return out

    def edge_data(self):
        ln = np.empty(self.E)
        cov = np.empty(self.E)
        lib().orc_mesh_edge_data(self._h, ln, cov)
        return ln, cov

    def control_volumes(self):
        out = np.empty(self.n)
        lib().orc_mesh_control_volumes(self._h, out)
        return out

    def cell_edge_coefficients(self):
        out = np.empty((self.cells.shape[0], self.ne))
        lib().orc_cell_edge_coefficients(self._h, out)
        return out


class Model:
    """Oracle NLS model: KEO, dK/dp, F, dF/dp, J, keo_regularized, MINRES, Newton."""

    def __init__(self, mesh: Mesh, A, thickness=None, V=None, dVdp=None,
                 deriv_param: str = "mu"):
        n = mesh.n
        self.mesh = mesh
        self.n = n
        t = np.ones(n) if thickness is None else np.ascontiguousarray(thickness, np.float64)
        v = -np.ones(n) if V is None else np.ascontiguousarray(V, np.float64)
        dv = np.zeros(n) if dVdp is None else np.ascontiguousarray(dVdp, np.float64)
        a = np.ascontiguousarray(A, dtype=np.float64).reshape(-1)
        self._keep = (t, v, dv, a)
        self._h = lib().orc_model_new(mesh._h, t, v, dv, a, deriv_param.encode())

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.orc_model_free(self._h)
            self._h = None

    def edge_fields(self):
        at = np.empty(self.mesh.E)
        al = np.empty(self.mesh.E)
        lib().orc_model_edge_fields(self._h, at, al)
        return at, al

    def set_params(self, g: float, mu: float):
        lib().orc_set_params(self._h, g, mu)

    def _vec_op(self, fn, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        fn(self._h, x, y)
        return y

    def keo(self, x):
        return self._vec_op(lib().orc_keo_apply, x)

    def dkeo(self, x):
        return self._vec_op(lib().orc_dkeo_apply, x)

    def f(self, x):
        return self._vec_op(lib().orc_compute_f, x)

    def dfdp(self, x, param: str):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        lib().orc_compute_dfdp(self._h, x, param.encode(), y)
        return y

    def jac_rebuild(self, psi):
        lib().orc_jac_rebuild(self._h, np.ascontiguousarray(psi, dtype=np.float64))

    def jac(self, x):
        return self._vec_op(lib().orc_jac_apply, x)

    def prec_rebuild(self, psi):
        lib().orc_prec_rebuild(self._h, np.ascontiguousarray(psi, dtype=np.float64))

    def prec_op(self, x):
        return self._vec_op(lib().orc_prec_op_apply, x)

    def cheb(self, r, iters: int, ratio: float, hi: float = 0.0):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.empty_like(r)
        lib().orc_cheb_apply(self._h, iters, ratio, hi, r, z)
        return z

    def cheb_bound(self) -> float:
        return float(lib().orc_cheb_bound(self._h))

    def crs(self, which: int = 0):
        """Real 2n x 2n CRS (ptr, col, val); which: 0 K, 1 dK/dp, 2 P."""
        nnz = int(lib().orc_keo_nnz(self._h))
        ptr = np.empty(2 * self.n + 1, dtype=np.int64)
        col = np.empty(nnz, dtype=np.int32)
        val = np.empty(nnz)
        lib().orc_crs_get(self._h, which, ptr, col, val)
        return ptr, col, val

    def dense(self, which: int = 0):
        ptr, col, val = self.crs(which)
        m = 2 * self.n
        D = np.zeros((m, m))
        rows = np.repeat(np.arange(m), np.diff(ptr))
        D[rows, col] = val
        return D

    def minres(self, b, rtol=1e-10, maxit=1000, cheb_iters=0, ratio=30.0, history=False):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.empty_like(b)
        rr = np.zeros(1)
        hist = np.zeros(maxit) if history else None
        k = lib().orc_minres_jac(self._h, b, x, rtol, maxit, cheb_iters, ratio, rr,
                                 hist.ctypes.data if history else None)
        if history:
            return x, k, float(rr[0]), hist[:k]
        return x, k, float(rr[0])

    def newton(self, psi0, g, mu, tol=1e-10, maxit=30, minres_rtol=1e-10, minres_maxit=1000,
               cheb_iters=0, ratio=30.0):
        psi = np.array(psi0, dtype=np.float64, copy=True)
        nf = np.zeros(maxit + 1)
        mh = np.zeros(maxit + 1, dtype=np.int32)
        it = lib().orc_newton(self._h, psi, g, mu, tol, maxit, minres_rtol, minres_maxit,
                              cheb_iters, ratio, nf, mh)
        k = it if it >= 0 else maxit
        return psi, it, nf[:k + 1], mh[:k]


def minres_crs(ptr, col, val, b, rtol=1e-12, maxit=10000):
    m = b.size
    x = np.empty(m)
    rr = np.zeros(1)
    k = lib().orc_minres_crs(m, np.ascontiguousarray(ptr, np.int64),
                             np.ascontiguousarray(col, np.int32),
                             np.ascontiguousarray(val, np.float64),
                             np.ascontiguousarray(b, np.float64), x, rtol, maxit, rr)
    return x, k, float(rr[0])
