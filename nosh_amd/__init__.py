"""nosh_amd: MI355X-native NLS Newton-Krylov core behind nosh's operator API.

Python mirror of the reference classes over the C ABI of libnosh_amd.so
(include/nosh_amd.h). Names, argument meaning and error behaviour follow the
reference:

=====================================  =========================================
reference (nosh tree)                  here
=====================================  =========================================
nosh::read / nosh::mesh                Mesh(coords, cells)
vector_field::explicit_values          VectorFieldExplicitValues(mesh, A, mu)
scalar_field::constant                 ScalarFieldConstant(mesh, c, name, init)
scalar_field::explicit_values          ScalarFieldExplicitValues(mesh, values)
parameter_matrix::keo                  Keo(mesh, thickness, mvp)
parameter_matrix::DkeoDP               DkeoDP(mesh, thickness, mvp, name)
jacobian_operator                      JacobianOperator(mesh, V, thickness, keo)
keo_regularized                        KeoRegularized(mesh, thickness, mvp)
model_evaluator::nls                   Nls(mesh, mvp, V, g, thickness, x0, name)
=====================================  =========================================

Operators take and return numpy vectors of length 2n (interleaved complex,
reference vertex order) and implement ``apply(X, mode, alpha, beta)`` with the
reference's restrictions: only NO_TRANS, alpha=1, beta=0
(src/jacobian_operator.cpp:48-59); anything else raises ``LogicError``
(std::logic_error). A parameter map lacking a required key raises ``KeyError``
(std::out_of_range from params.at()).

The library must be built (``python -m nosh_amd.build``); there is no CPU
fallback: importing works without it, but every object construction raises
``LibraryNotBuilt``.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, Iterable, Optional

import numpy as np

__all__ = [
    "Mesh", "VectorFieldExplicitValues", "ScalarFieldConstant", "ScalarFieldExplicitValues",
    "Keo", "DkeoDP", "JacobianOperator", "KeoRegularized", "Nls", "NewtonOptions",
    "newton_solve", "minres_solve", "LogicError", "NoshError", "LibraryNotBuilt",
    "NO_TRANS", "TRANS", "CONJ_TRANS", "library_path",
]

NO_TRANS, TRANS, CONJ_TRANS = 0, 1, 2
_HERE = os.path.dirname(os.path.abspath(__file__))
library_path = os.path.join(_HERE, "libnosh_amd.so")


class NoshError(RuntimeError):
    """A failure reported by the library (HIP error, invalid argument ...)."""


class LogicError(NoshError):
    """Unsupported operator mode (the reference throws std::logic_error)."""


class LibraryNotBuilt(NoshError):
    """libnosh_amd.so is missing: build it with ``python -m nosh_amd.build``."""


_OK, _INVALID, _UNSUPPORTED, _MISSING, _HIP, _NOTBUILT, _NOTCONV = range(7)


class NewtonOptions(C.Structure):
    _fields_ = [("tol", C.c_double), ("max_iters", C.c_int), ("minres_rtol", C.c_double),
                ("minres_max_iters", C.c_int), ("cheb_iters", C.c_int),
                ("cheb_ratio", C.c_double), ("precision", C.c_int), ("forcing", C.c_int),
                ("forcing_min", C.c_double), ("forcing_max", C.c_double)]


class NewtonStats(C.Structure):
    _fields_ = [("newton_iters", C.c_int), ("minres_iters_total", C.c_int),
                ("converged", C.c_int), ("final_norm_f", C.c_double), ("seconds", C.c_double),
                ("norm_f_history", C.c_double * 64), ("minres_history", C.c_int * 64)]

    def as_dict(self) -> dict:
        k = min(self.newton_iters, 63)
        return {"newton_iters": self.newton_iters, "minres_iters_total": self.minres_iters_total,
                "converged": bool(self.converged), "final_norm_f": self.final_norm_f,
                "seconds": self.seconds,
                "norm_f_history": list(self.norm_f_history[:k + 1]),
                "minres_history": list(self.minres_history[:k])}


_lib = None


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(library_path):
        raise LibraryNotBuilt(f"{library_path} not found; run `python -m nosh_amd.build`")
    lib = C.CDLL(library_path)
    vp, i32, i64, d = C.c_void_p, C.c_int, C.c_int64, C.c_double
    pvp = C.POINTER(C.c_void_p)
    cpp = C.POINTER(C.c_char_p)
    sig = {
        "nosh_last_error": (C.c_char_p, []),
        "nosh_mesh_create": (i32, [vp, i64, vp, i64, i32, i32, pvp]),
        "nosh_mesh_destroy": (i32, [vp]),
        "nosh_mesh_sizes": (i32, [vp, C.POINTER(i64), C.POINTER(i64), C.POINTER(i64)]),
        "nosh_mesh_edges": (i32, [vp, vp]),
        "nosh_mesh_cell_edges": (i32, [vp, vp]),
        "nosh_mesh_edge_data": (i32, [vp, vp, vp]),
        "nosh_mesh_control_volumes": (i32, [vp, vp]),
        "nosh_mesh_permutation": (i32, [vp, vp]),
        "nosh_mesh_layout": (i32, [vp, C.POINTER(i64), C.POINTER(i64), C.POINTER(i32)]),
        "nosh_mesh_layout_split": (i32, [vp, C.POINTER(i64), C.POINTER(i64)]),
        "nosh_mesh_layout_column_bytes": (i32, [vp, C.POINTER(C.c_int)]),
        "nosh_mesh_to_internal": (i32, [vp, vp, vp, vp]),
        "nosh_mesh_to_reference": (i32, [vp, vp, vp, vp]),
        "nosh_vfield_explicit_values_create": (i32, [vp, vp, d, pvp]),
        "nosh_vfield_edge_projections": (i32, [vp, vp]),
        "nosh_vfield_destroy": (i32, [vp]),
        "nosh_sfield_constant_create": (i32, [vp, d, C.c_char_p, d, pvp]),
        "nosh_sfield_explicit_values_create": (i32, [vp, vp, pvp]),
        "nosh_sfield_destroy": (i32, [vp]),
        "nosh_keo_create": (i32, [vp, vp, vp, pvp]),
        "nosh_dkeo_dp_create": (i32, [vp, vp, vp, C.c_char_p, pvp]),
        "nosh_keo_set_parameters": (i32, [vp, i32, cpp, vp]),
        "nosh_keo_apply": (i32, [vp, vp, vp, i32, i64, i32, d, d]),
        "nosh_keo_apply_device": (i32, [vp, vp, vp, vp]),
        "nosh_keo_apply_internal": (i32, [vp, vp, vp, vp]),
        "nosh_keo_destroy": (i32, [vp]),
        "nosh_jac_create": (i32, [vp, vp, vp, vp, pvp]),
        "nosh_jac_rebuild": (i32, [vp, i32, cpp, vp, vp]),
        "nosh_jac_apply": (i32, [vp, vp, vp, i32, i64, i32, d, d]),
        "nosh_jac_apply_device": (i32, [vp, vp, vp, vp]),
        "nosh_jac_apply_internal": (i32, [vp, vp, vp, vp]),
        "nosh_jac_apply_internal_fp32": (i32, [vp, vp, vp, vp]),
        "nosh_jac_bench_apply": (i32, [vp, vp, vp, i32, i32, i32, vp, C.POINTER(d)]),
        "nosh_jac_destroy": (i32, [vp]),
        "nosh_prec_create": (i32, [vp, vp, vp, pvp]),
        "nosh_prec_set_chebyshev": (i32, [vp, i32, d, d]),
        "nosh_prec_rebuild": (i32, [vp, i32, cpp, vp, vp]),
        "nosh_prec_apply": (i32, [vp, vp, vp, i32, i64, i32, d, d]),
        "nosh_prec_apply_device": (i32, [vp, vp, vp, vp]),
        "nosh_prec_apply_operator": (i32, [vp, vp, vp]),
        "nosh_prec_upper_bound": (i32, [vp, C.POINTER(d)]),
        "nosh_prec_destroy": (i32, [vp]),
        "nosh_nls_create": (i32, [vp, vp, vp, d, vp, vp, C.c_char_p, pvp]),
        "nosh_nls_num_params": (i32, [vp, C.POINTER(i32)]),
        "nosh_nls_param_name": (i32, [vp, i32, C.c_char_p, i32]),
        "nosh_nls_param_nominal": (i32, [vp, i32, C.POINTER(d)]),
        "nosh_nls_eval_f": (i32, [vp, vp, i32, cpp, vp, vp]),
        "nosh_nls_eval_f_device": (i32, [vp, vp, i32, cpp, vp, vp, vp]),
        "nosh_nls_eval_f_internal": (i32, [vp, vp, i32, cpp, vp, vp, vp]),
        "nosh_nls_eval_dfdp": (i32, [vp, vp, i32, cpp, vp, i32, vp]),
        "nosh_nls_create_w_op": (i32, [vp, pvp]),
        "nosh_nls_create_w_prec": (i32, [vp, pvp]),
        "nosh_nls_destroy": (i32, [vp]),
        "nosh_newton_default_options": (i32, [C.POINTER(NewtonOptions)]),
        "nosh_newton_solve": (i32, [vp, vp, i32, cpp, vp, C.POINTER(NewtonOptions),
                                    C.POINTER(NewtonStats)]),
        "nosh_newton_solve_device": (i32, [vp, vp, i32, cpp, vp, C.POINTER(NewtonOptions),
                                           C.POINTER(NewtonStats), vp]),
        "nosh_minres_solve": (i32, [vp, vp, vp, vp, d, i32, C.POINTER(i32), C.POINTER(d)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _check(rc: int) -> None:
    if rc == _OK:
        return
    msg = _lib.nosh_last_error().decode(errors="replace")
    if rc == _UNSUPPORTED:
        raise LogicError(msg)
    if rc == _MISSING:
        raise KeyError(msg)
    if rc == _INVALID:
        raise ValueError(msg)
    if rc == _NOTBUILT:
        raise NoshError("not built: " + msg)
    raise NoshError(f"nosh error {rc}: {msg}")


def _ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


def _vec(x, m: int) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=np.float64)
    if a.ndim != 1 or a.size != m:
        raise ValueError(f"expected a vector of length {m}, got shape {a.shape}")
    return a


class _Params:
    """Marshals a {name: value} map into the (count, names, values) C form."""

    def __init__(self, params: Dict[str, float]):
        items = sorted(params.items())
        self._names = [k.encode() for k, _ in items]
        self.count = len(items)
        self.names = (C.c_char_p * max(1, self.count))(*self._names)
        self.values = np.array([v for _, v in items] or [0.0], dtype=np.float64)

    @property
    def args(self):
        return (self.count, self.names, _ptr(self.values))


class _Handle:
    _destroy = ""

    def __init__(self):
        self._h = C.c_void_p()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value and _lib is not None:
            getattr(_lib, self._destroy)(h)
            self._h = C.c_void_p()


class Mesh(_Handle):
    """Edge/covolume cache built on the GPU (nosh::read + mesh_tri/mesh_tetra).

    coords: (n, 3) float64; cells: (T, 3) triangles or (T, 4) tetrahedra.
    """
    _destroy = "nosh_mesh_destroy"

    def __init__(self, coords, cells, reorder: bool = True):
        super().__init__()
        lib = _load()
        coords = np.ascontiguousarray(coords, dtype=np.float64)
        cells = np.ascontiguousarray(cells, dtype=np.int32)
        if coords.ndim != 2 or coords.shape[1] != 3:
            raise ValueError("coords must be (n, 3)")
        if cells.ndim != 2 or cells.shape[1] not in (3, 4):
            raise ValueError("cells must be (T, 3) or (T, 4)")
        _check(lib.nosh_mesh_create(_ptr(coords), coords.shape[0], _ptr(cells), cells.shape[0],
                                    cells.shape[1], 1 if reorder else 0, C.byref(self._h)))
        n, E, T = C.c_int64(), C.c_int64(), C.c_int64()
        _check(lib.nosh_mesh_sizes(self._h, C.byref(n), C.byref(E), C.byref(T)))
        self.num_vertices, self.num_edges, self.num_cells = n.value, E.value, T.value
        self.vertices_per_cell = cells.shape[1]

    def my_edges(self) -> np.ndarray:
        out = np.empty((self.num_edges, 2), dtype=np.int32)
        _check(_lib.nosh_mesh_edges(self._h, _ptr(out)))
        return out

    def cell_edges(self) -> np.ndarray:
        ne = 3 if self.vertices_per_cell == 3 else 6
        out = np.empty((self.num_cells, ne), dtype=np.int32)
        _check(_lib.nosh_mesh_cell_edges(self._h, _ptr(out)))
        return out

    def get_edge_data(self):
        """(length, covolume) per edge (mesh::get_edge_data)."""
        ln = np.empty(self.num_edges)
        cov = np.empty(self.num_edges)
        _check(_lib.nosh_mesh_edge_data(self._h, _ptr(ln), _ptr(cov)))
        return ln, cov

    def control_volumes(self) -> np.ndarray:
        out = np.empty(self.num_vertices)
        _check(_lib.nosh_mesh_control_volumes(self._h, _ptr(out)))
        return out

    def layout(self) -> dict:
        """Internal sliced-ELLPACK layout sizes (nnz_pad: the full structure;
        nnz_upper / nnz_lower: the split layout the operator kernels read;
        column_bytes: 2 when its columns are 16-bit offsets from the row,
        else 4)."""
        nnz, ns, w = C.c_int64(), C.c_int64(), C.c_int()
        _check(_lib.nosh_mesh_layout(self._h, C.byref(nnz), C.byref(ns), C.byref(w)))
        nu, nl = C.c_int64(), C.c_int64()
        _check(_lib.nosh_mesh_layout_split(self._h, C.byref(nu), C.byref(nl)))
        cb = C.c_int()
        _check(_lib.nosh_mesh_layout_column_bytes(self._h, C.byref(cb)))
        return {"nnz_pad": nnz.value, "nslices": ns.value, "max_width": w.value,
                "nnz_upper": nu.value, "nnz_lower": nl.value, "column_bytes": cb.value}

    def to_internal(self, x_ref_ptr: int, x_int_ptr: int, stream: int = 0):
        _check(_lib.nosh_mesh_to_internal(self._h, C.c_void_p(x_ref_ptr), C.c_void_p(x_int_ptr),
                                          C.c_void_p(stream)))

    def to_reference(self, x_int_ptr: int, x_ref_ptr: int, stream: int = 0):
        _check(_lib.nosh_mesh_to_reference(self._h, C.c_void_p(x_int_ptr), C.c_void_p(x_ref_ptr),
                                           C.c_void_p(stream)))

    def permutation(self) -> np.ndarray:
        out = np.empty(self.num_vertices, dtype=np.int32)
        _check(_lib.nosh_mesh_permutation(self._h, _ptr(out)))
        return out


class VectorFieldExplicitValues(_Handle):
    """Magnetic vector potential from per-vertex values A (n, 3), scaled by mu."""
    _destroy = "nosh_vfield_destroy"

    def __init__(self, mesh: Mesh, A, mu: float):
        super().__init__()
        A = np.ascontiguousarray(A, dtype=np.float64)
        if A.shape != (mesh.num_vertices, 3):
            raise ValueError("A must be (n, 3)")
        self.mesh = mesh
        _check(_load().nosh_vfield_explicit_values_create(mesh._h, _ptr(A), mu,
                                                          C.byref(self._h)))

    def get_scalar_parameters(self):
        return {"mu"}

    def edge_projections(self) -> np.ndarray:
        out = np.empty(self.mesh.num_edges)
        _check(_lib.nosh_vfield_edge_projections(self._h, _ptr(out)))
        return out


class ScalarFieldConstant(_Handle):
    """Constant scalar field c (+ params[param_name] when given)."""
    _destroy = "nosh_sfield_destroy"

    def __init__(self, mesh: Mesh, c: float, param_name: str = "", param_init: float = 0.0):
        super().__init__()
        self.mesh = mesh
        self.param_name = param_name
        _check(_load().nosh_sfield_constant_create(mesh._h, c, param_name.encode(), param_init,
                                                   C.byref(self._h)))


class ScalarFieldExplicitValues(_Handle):
    _destroy = "nosh_sfield_destroy"

    def __init__(self, mesh: Mesh, values):
        super().__init__()
        self.mesh = mesh
        v = np.ascontiguousarray(values, dtype=np.float64)
        if v.shape != (mesh.num_vertices,):
            raise ValueError("values must have one entry per vertex")
        _check(_load().nosh_sfield_explicit_values_create(mesh._h, _ptr(v), C.byref(self._h)))


class _Operator(_Handle):
    _apply = ""

    def _m(self) -> int:
        return 2 * self.mesh.num_vertices

    def apply(self, X, mode: int = NO_TRANS, alpha: float = 1.0, beta: float = 0.0):
        """Y = Op X for a vector (2n,) or column block (2n, k); Y is a new array."""
        X = np.asarray(X, dtype=np.float64)
        m = self._m()
        if X.shape[0] != m:
            raise ValueError(f"expected leading dimension {m}")
        cols = 1 if X.ndim == 1 else X.shape[1]
        Xf = np.ascontiguousarray(X.reshape(m, cols).T)  # columns contiguous
        Yf = np.empty_like(Xf)
        _check(getattr(_lib, self._apply)(self._h, _ptr(Xf), _ptr(Yf), cols, m, mode, alpha,
                                          beta))
        return Yf[0].copy() if X.ndim == 1 else np.ascontiguousarray(Yf.T)

    def __call__(self, x):
        return self.apply(x)


class Keo(_Operator):
    """parameter_matrix::keo: the magnetic kinetic energy operator K."""
    _destroy = "nosh_keo_destroy"
    _apply = "nosh_keo_apply"

    def __init__(self, mesh: Mesh, thickness, mvp: VectorFieldExplicitValues, *, _h=None):
        super().__init__()
        self.mesh, self._keep = mesh, (thickness, mvp)
        if _h is None:
            _check(_load().nosh_keo_create(mesh._h, thickness._h, mvp._h, C.byref(self._h)))
        else:
            self._h = _h

    def set_parameters(self, params: Dict[str, float], vector_params=None):
        p = _Params(params)
        _check(_lib.nosh_keo_set_parameters(self._h, *p.args))

    def apply_device(self, x_ptr: int, y_ptr: int, stream: int = 0):
        _check(_lib.nosh_keo_apply_device(self._h, C.c_void_p(x_ptr), C.c_void_p(y_ptr),
                                          C.c_void_p(stream)))

    def apply_internal(self, x_ptr: int, y_ptr: int, stream: int = 0):
        """y = K x on device vectors in internal order."""
        _check(_lib.nosh_keo_apply_internal(self._h, C.c_void_p(x_ptr), C.c_void_p(y_ptr),
                                            C.c_void_p(stream)))


class DkeoDP(Keo):
    """parameter_matrix::DkeoDP: dK/d(param_name)."""

    def __init__(self, mesh: Mesh, thickness, mvp, param_name: str):
        h = C.c_void_p()
        _check(_load().nosh_dkeo_dp_create(mesh._h, thickness._h, mvp._h, param_name.encode(),
                                           C.byref(h)))
        super().__init__(mesh, thickness, mvp, _h=h)


class JacobianOperator(_Operator):
    """jacobian_operator: J(psi) phi = (K + V + 2g|psi|^2) phi + g psi^2 conj(phi)."""
    _destroy = "nosh_jac_destroy"
    _apply = "nosh_jac_apply"

    def __init__(self, mesh: Mesh, scalar_potential, thickness, keo: Keo, *, _h=None):
        super().__init__()
        self.mesh, self._keep = mesh, (scalar_potential, thickness, keo)
        if _h is None:
            _check(_load().nosh_jac_create(mesh._h, scalar_potential._h, thickness._h, keo._h,
                                           C.byref(self._h)))
        else:
            self._h = _h

    def rebuild(self, params: Dict[str, float], current_x):
        p = _Params(params)
        x = _vec(current_x, self._m())
        _check(_lib.nosh_jac_rebuild(self._h, *p.args, _ptr(x)))

    def apply_device(self, x_ptr: int, y_ptr: int, stream: int = 0):
        _check(_lib.nosh_jac_apply_device(self._h, C.c_void_p(x_ptr), C.c_void_p(y_ptr),
                                          C.c_void_p(stream)))

    def apply_internal(self, x_ptr: int, y_ptr: int, stream: int = 0):
        _check(_lib.nosh_jac_apply_internal(self._h, C.c_void_p(x_ptr), C.c_void_p(y_ptr),
                                            C.c_void_p(stream)))

    def apply_internal_fp32(self, x_ptr: int, y_ptr: int, stream: int = 0):
        _check(_lib.nosh_jac_apply_internal_fp32(self._h, C.c_void_p(x_ptr),
                                                 C.c_void_p(y_ptr), C.c_void_p(stream)))

    def bench_apply(self, x_ptr: int, y_ptr: int, *, fp32: bool = False, cold: bool = False,
                    reps: int = 20, stream: int = 0) -> float:
        """Average kernel duration (seconds) of the internal-order apply,
        stamped on the kernel dispatch itself; cold: a 1 GiB read before
        each launch evicts the caches (measurement helper)."""
        out = C.c_double()
        _check(_lib.nosh_jac_bench_apply(self._h, C.c_void_p(x_ptr), C.c_void_p(y_ptr),
                                         int(fp32), int(cold), int(reps), C.c_void_p(stream),
                                         C.byref(out)))
        return out.value


class KeoRegularized(_Operator):
    """keo_regularized: P = K + g-blocks of psi; apply() is the approximate
    inverse (Chebyshev iteration with 2x2 block-Jacobi scaling)."""
    _destroy = "nosh_prec_destroy"
    _apply = "nosh_prec_apply"

    def __init__(self, mesh: Mesh, thickness, mvp, *, _h=None):
        super().__init__()
        self.mesh, self._keep = mesh, (thickness, mvp)
        if _h is None:
            _check(_load().nosh_prec_create(mesh._h, thickness._h, mvp._h, C.byref(self._h)))
        else:
            self._h = _h

    def set_chebyshev(self, iters: int = 4, ratio: float = 30.0, upper: float = 0.0):
        _check(_lib.nosh_prec_set_chebyshev(self._h, iters, ratio, upper))

    def rebuild(self, params: Dict[str, float], current_x):
        p = _Params(params)
        x = _vec(current_x, self._m())
        _check(_lib.nosh_prec_rebuild(self._h, *p.args, _ptr(x)))

    def apply_operator(self, x) -> np.ndarray:
        """y = P x (the regularized KEO itself, not its inverse)."""
        x = _vec(x, self._m())
        y = np.empty_like(x)
        _check(_lib.nosh_prec_apply_operator(self._h, _ptr(x), _ptr(y)))
        return y

    def upper_bound(self) -> float:
        v = C.c_double()
        _check(_lib.nosh_prec_upper_bound(self._h, C.byref(v)))
        return v.value


class Nls(_Handle):
    """model_evaluator::nls: F(psi) = (K + V + g|psi|^2) psi and its derivatives."""
    _destroy = "nosh_nls_destroy"

    def __init__(self, mesh: Mesh, mvp, scalar_potential, g: float, thickness,
                 initial_x=None, deriv_parameter: str = "mu"):
        super().__init__()
        self.mesh, self._keep = mesh, (mvp, scalar_potential, thickness)
        x0 = None if initial_x is None else _vec(initial_x, 2 * mesh.num_vertices)
        self._x0 = x0
        _check(_load().nosh_nls_create(mesh._h, mvp._h, scalar_potential._h, g, thickness._h,
                                       None if x0 is None else _ptr(x0),
                                       deriv_parameter.encode(), C.byref(self._h)))

    def get_p_names(self):
        n = C.c_int()
        _check(_lib.nosh_nls_num_params(self._h, C.byref(n)))
        out = []
        buf = C.create_string_buffer(256)
        for k in range(n.value):
            _check(_lib.nosh_nls_param_name(self._h, k, buf, 256))
            out.append(buf.value.decode())
        return out

    def get_nominal_values(self) -> Dict[str, float]:
        out = {}
        for k, name in enumerate(self.get_p_names()):
            v = C.c_double()
            _check(_lib.nosh_nls_param_nominal(self._h, k, C.byref(v)))
            out[name] = v.value
        return out

    def eval_f(self, x, params: Dict[str, float]) -> np.ndarray:
        m = 2 * self.mesh.num_vertices
        x = _vec(x, m)
        f = np.empty(m)
        p = _Params(params)
        _check(_lib.nosh_nls_eval_f(self._h, _ptr(x), *p.args, _ptr(f)))
        return f

    def eval_f_internal(self, x_ptr: int, params: Dict[str, float], f_ptr: int, stream: int = 0):
        """F on device vectors in internal order (no permutation)."""
        p = _Params(params)
        _check(_lib.nosh_nls_eval_f_internal(self._h, C.c_void_p(x_ptr), *p.args,
                                             C.c_void_p(f_ptr), C.c_void_p(stream)))

    def eval_dfdp(self, x, params: Dict[str, float], column: int) -> np.ndarray:
        m = 2 * self.mesh.num_vertices
        x = _vec(x, m)
        out = np.empty(m)
        p = _Params(params)
        _check(_lib.nosh_nls_eval_dfdp(self._h, _ptr(x), *p.args, column, _ptr(out)))
        return out

    def create_W_op(self) -> JacobianOperator:
        h = C.c_void_p()
        _check(_lib.nosh_nls_create_w_op(self._h, C.byref(h)))
        op = JacobianOperator(self.mesh, None, None, None, _h=h)
        op._keep = (self,)  # shares this model's keo
        return op

    def create_W_prec(self) -> KeoRegularized:
        h = C.c_void_p()
        _check(_lib.nosh_nls_create_w_prec(self._h, C.byref(h)))
        op = KeoRegularized(self.mesh, None, None, _h=h)
        op._keep = (self,)
        return op


def default_newton_options(**kw) -> NewtonOptions:
    o = NewtonOptions()
    _check(_load().nosh_newton_default_options(C.byref(o)))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise AttributeError(k)
        setattr(o, k, v)
    return o


def newton_solve(model: Nls, psi0, params: Dict[str, float], options: Optional[NewtonOptions] = None,
                 **kw):
    """Newton-MINRES on F(psi) = 0; returns (psi, stats dict)."""
    o = options if options is not None else default_newton_options(**kw)
    psi = np.array(_vec(psi0, 2 * model.mesh.num_vertices), copy=True)
    st = NewtonStats()
    p = _Params(params)
    _check(_lib.nosh_newton_solve(model._h, _ptr(psi), *p.args, C.byref(o), C.byref(st)))
    return psi, st.as_dict()


def minres_solve(jac: JacobianOperator, prec: Optional[KeoRegularized], b, rtol: float = 1e-10,
                 max_iters: int = 1000):
    """MINRES on J x = b from x = 0; returns (x, iterations, relative residual)."""
    b = _vec(b, 2 * jac.mesh.num_vertices)
    x = np.empty_like(b)
    it, rr = C.c_int(), C.c_double()
    _check(_lib.nosh_minres_solve(jac._h, None if prec is None else prec._h, _ptr(b), _ptr(x),
                                  rtol, max_iters, C.byref(it), C.byref(rr)))
    return x, it.value, rr.value


def newton_solve_device(model: Nls, psi_ptr: int, params: Dict[str, float],
                        options: Optional[NewtonOptions] = None, stream: int = 0, **kw) -> dict:
    """Newton-MINRES on a device vector (2n doubles, reference order) in place."""
    o = options if options is not None else default_newton_options(**kw)
    st = NewtonStats()
    p = _Params(params)
    _check(_lib.nosh_newton_solve_device(model._h, C.c_void_p(psi_ptr), *p.args, C.byref(o),
                                         C.byref(st), C.c_void_p(stream)))
    return st.as_dict()


class ContinuationOptions(C.Structure):
    _fields_ = [("parameter", C.c_char_p), ("start", C.c_double), ("stop", C.c_double),
                ("steps", C.c_int), ("predictor", C.c_int), ("max_halvings", C.c_int),
                ("newton", NewtonOptions)]


class ContinuationStep(C.Structure):
    _fields_ = [("param", C.c_double), ("newton_iters", C.c_int), ("minres_iters", C.c_int),
                ("halvings", C.c_int), ("converged", C.c_int), ("norm_f", C.c_double),
                ("gibbs", C.c_double), ("seconds", C.c_double),
                ("predictor_used", C.c_int), ("fallback", C.c_int),
                ("pred_error", C.c_double), ("tan_error", C.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


def _load_more():
    lib = _load()
    if getattr(lib, "_nosh_more", False):
        return lib
    vp, i32, d = C.c_void_p, C.c_int, C.c_double
    cpp = C.POINTER(C.c_char_p)
    sig = {
        "nosh_nls_gibbs_energy": (i32, [vp, vp, C.POINTER(d)]),
        "nosh_nls_inner_product": (i32, [vp, vp, vp, C.POINTER(d)]),
        "nosh_nls_gibbs_energy_noalloc": (i32, [vp, vp, vp, vp]),
        "nosh_bench_stream": (i32, [i32, C.c_int64, i32, i32, i32, C.POINTER(d)]),
        "nosh_continuation_default_options": (i32, [C.POINTER(ContinuationOptions)]),
        "nosh_continuation_run": (i32, [vp, vp, i32, cpp, vp, C.POINTER(ContinuationOptions),
                                        C.POINTER(ContinuationStep), C.POINTER(i32)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    lib._nosh_more = True
    return lib


def gibbs_energy(model: Nls, psi) -> float:
    """-sum c |psi|^4 / sum c (the value nls::gibbs_energy computes)."""
    lib = _load_more()
    x = _vec(psi, 2 * model.mesh.num_vertices)
    v = C.c_double()
    _check(lib.nosh_nls_gibbs_energy(model._h, _ptr(x), C.byref(v)))
    return v.value


def gibbs_energy_noalloc(model: Nls, psi_ptr: int, energy_ptr: int, stream: int = 0) -> None:
    """The Gibbs energy with no allocation: psi (2n doubles, reference order)
    and energy (1 double) are device pointers; the value is written on
    `stream`. Meshes above 4,194,304 vertices are refused; calls that overlap
    on one GPU race on the library's scratch array."""
    lib = _load_more()
    _check(lib.nosh_nls_gibbs_energy_noalloc(model._h, C.c_void_p(psi_ptr),
                                             C.c_void_p(energy_ptr), C.c_void_p(stream)))


def bench_stream(kind: str, nbytes: int, reps: int = 20, blocks_per_cu: int = 8,
                 nontemporal: bool = True) -> float:
    """Seconds per launch of the library's streaming kernel ("read", "write"
    or "copy") over nbytes-byte buffers (copy moves 2 nbytes per launch)."""
    lib = _load_more()
    k = {"read": 0, "write": 1, "copy": 2}[kind]
    t = C.c_double()
    _check(lib.nosh_bench_stream(k, int(nbytes), reps, blocks_per_cu, int(nontemporal),
                                 C.byref(t)))
    return t.value


def inner_product(model: Nls, phi, psi) -> float:
    """sum c Re(conj(phi) psi) / sum c (the value nls::inner_product computes)."""
    lib = _load_more()
    m = 2 * model.mesh.num_vertices
    a, b = _vec(phi, m), _vec(psi, m)
    v = C.c_double()
    _check(lib.nosh_nls_inner_product(model._h, _ptr(a), _ptr(b), C.byref(v)))
    return v.value


def continuation(model: Nls, psi0, params: Dict[str, float], parameter: str = "mu",
                 start: float = 0.0, stop: float = 1.0, steps: int = 10, predictor: int = 1,
                 max_halvings: int = 4, newton: Optional[NewtonOptions] = None, **newton_kw):
    """Natural-parameter continuation; returns (final psi, list of step records).

    Raises NoshError(NOT_CONVERGED) only through the records: a failed step
    ends the run early and is reported with converged=0.
    """
    lib = _load_more()
    o = ContinuationOptions()
    _check(lib.nosh_continuation_default_options(C.byref(o)))
    pname = parameter.encode()
    o.parameter, o.start, o.stop, o.steps = pname, start, stop, steps
    o.predictor, o.max_halvings = predictor, max_halvings
    o.newton = newton if newton is not None else default_newton_options(**newton_kw)
    psi = np.array(_vec(psi0, 2 * model.mesh.num_vertices), copy=True)
    rec = (ContinuationStep * (steps + 1))()
    done = C.c_int()
    p = _Params(params)
    rc = lib.nosh_continuation_run(model._h, _ptr(psi), *p.args, C.byref(o), rec, C.byref(done))
    if rc not in (_OK, _NOTCONV):
        _check(rc)
    out = [rec[i].as_dict() for i in range(done.value)]
    if rc == _NOTCONV and done.value <= steps:
        out.append(rec[done.value].as_dict())
    return psi, out


class BatchInstanceStats(C.Structure):
    _fields_ = [("newton_iters", C.c_int), ("minres_iters", C.c_int), ("converged", C.c_int),
                ("final_norm_f", C.c_double)]


def _load_batch():
    lib = _load_more()
    if getattr(lib, "_nosh_batch", False):
        return lib
    vp, i32 = C.c_void_p, C.c_int
    sig = {
        "nosh_batch_create": (i32, [i32, vp, vp, vp, vp, C.POINTER(vp)]),
        "nosh_batch_set_tuning": (i32, [vp, i32, i32, i32, i32, i32]),
        "nosh_batch_autotune": (i32, [vp, vp, C.POINTER(NewtonOptions), i32, C.POINTER(i32),
                                      C.POINTER(i32), C.POINTER(i32)]),
        "nosh_batch_solve": (i32, [vp, vp, C.POINTER(NewtonOptions),
                                   C.POINTER(BatchInstanceStats)]),
        "nosh_batch_solve_device": (i32, [vp, vp, C.POINTER(NewtonOptions),
                                          C.POINTER(BatchInstanceStats), vp]),
        "nosh_batch_destroy": (i32, [vp]),
        "nosh_batch_time_cheb_step": (i32, [vp, i32, C.POINTER(C.c_double),
                                            C.POINTER(C.c_double)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    lib._nosh_batch = True
    return lib


class Batch(_Handle):
    """Many Newton-MINRES solves on meshes of uneven size, advanced together.

    groups: list of (model, count); mu, g: one value per instance, group-major.
    """
    _destroy = "nosh_batch_destroy"

    def __init__(self, groups, mu, g):
        super().__init__()
        lib = _load_batch()
        self.models = [m for m, _ in groups]
        self.counts = [int(c) for _, c in groups]
        self.sizes = [2 * m.mesh.num_vertices for m in self.models]
        self.ninst = sum(self.counts)
        self.mu = np.ascontiguousarray(mu, dtype=np.float64)
        self.g = np.ascontiguousarray(g, dtype=np.float64)
        if self.mu.size != self.ninst or self.g.size != self.ninst:
            raise ValueError("mu and g need one value per instance")
        hs = (C.c_void_p * len(groups))(*[m._h for m in self.models])
        cnt = np.array(self.counts, dtype=np.int32)
        _check(lib.nosh_batch_create(len(groups), hs, _ptr(cnt), _ptr(self.mu), _ptr(self.g),
                                     C.byref(self._h)))

    @property
    def total_length(self) -> int:
        return sum(c * s for c, s in zip(self.counts, self.sizes))

    def time_cheb_step(self, reps: int = 50):
        """(us per launch, algorithmic bytes per launch) of the batched
        Chebyshev step, all instances active, at the current states."""
        us, by = C.c_double(), C.c_double()
        _check(_load_batch().nosh_batch_time_cheb_step(self._h, reps, C.byref(us), C.byref(by)))
        return us.value, by.value

    def split(self, psi):
        """Instance vectors of a concatenated state."""
        out, off = [], 0
        for c, s in zip(self.counts, self.sizes):
            for _ in range(c):
                out.append(psi[off:off + s])
                off += s
        return out

    def set_tuning(self, block_threads=256, mode=0, graph_iters=4, poll=16, fold=1):
        """mode 0: hipGraph iterations; 1: persistent cooperative kernel (one or two
        workgroups per CU striding over the row tiles, grid barriers with a bounded
        wait); 2: hipGraph with two lanes per thread in the gathering kernels (every
        group at least 2 lanes). All modes give the same results bit for bit; an
        unsupported mode raises NoshError and leaves the batch in mode 0."""
        _check(_lib.nosh_batch_set_tuning(self._h, block_threads, fold, mode, graph_iters,
                                          poll))

    def autotune(self, psi, options: Optional[NewtonOptions] = None, use_cache=True, **kw):
        """Returns (block_threads, mode, fold) of the fastest candidate."""
        o = options if options is not None else default_newton_options(**kw)
        x = _vec(psi, self.total_length)
        t, f, m = C.c_int(), C.c_int(), C.c_int()
        _check(_lib.nosh_batch_autotune(self._h, _ptr(x), C.byref(o), 1 if use_cache else 0,
                                        C.byref(t), C.byref(f), C.byref(m)))
        return t.value, m.value, f.value

    def solve(self, psi0, options: Optional[NewtonOptions] = None, **kw):
        o = options if options is not None else default_newton_options(**kw)
        psi = np.array(_vec(psi0, self.total_length), copy=True)
        st = (BatchInstanceStats * self.ninst)()
        _check(_lib.nosh_batch_solve(self._h, _ptr(psi), C.byref(o), st))
        return psi, [{k: getattr(st[i], k) for k, _ in BatchInstanceStats._fields_}
                     for i in range(self.ninst)]

    def solve_device(self, psi_ptr: int, options: Optional[NewtonOptions] = None, stream=0, **kw):
        o = options if options is not None else default_newton_options(**kw)
        st = (BatchInstanceStats * self.ninst)()
        _check(_lib.nosh_batch_solve_device(self._h, C.c_void_p(psi_ptr), C.byref(o), st,
                                            C.c_void_p(stream)))
        return [{k: getattr(st[i], k) for k, _ in BatchInstanceStats._fields_}
                for i in range(self.ninst)]

    def continuation(self, psi0, span: float, steps: int, surrogate: Optional[dict] = None,
                     options: Optional[NewtonOptions] = None, predictor: str = "auto",
                     **kw):
        """Steps every instance (a branch) from its mu_0 to mu_0 + span in
        ``steps`` equal steps, all branches solved together, each step from
        the ``predictor`` guess ("tangent": psi_k + h dpsi/dmu from one
        lockstep solve of all branches; "secant": through the last two
        states; "auto": the tangent on the first step and after a step that
        a branch did not pass from its first guess, else the secant). With ``surrogate`` (SurrogateOptions fields; ``{}`` for the
        defaults) the MLP seeds Newton once trained. Returns (psi, step
        records, surrogate stats or None)."""
        pred = {"tangent": 1, "secant": 2, "auto": 3}[predictor]
        lib = _load_mlp()
        if not getattr(lib, "_nosh_bcont", False):
            _load_surrogate(lib)
            lib.nosh_batch_continuation_run.restype = C.c_int
            lib.nosh_batch_continuation_run.argtypes = [
                C.c_void_p, C.c_void_p, C.c_double, C.c_int, C.c_int, C.POINTER(NewtonOptions),
                C.POINTER(SurrogateOptions), C.POINTER(BatchContinuationStep), C.POINTER(C.c_int),
                C.POINTER(SurrogateStats)]
            lib._nosh_bcont = True
        o = options if options is not None else default_newton_options(**kw)
        psi = np.array(_vec(psi0, self.total_length), copy=True)
        so = None
        if surrogate is not None:
            so = SurrogateOptions()
            _check(lib.nosh_surrogate_default_options(C.byref(so)))
            for k, v in surrogate.items():
                setattr(so, k, v)
        rec = (BatchContinuationStep * (steps + 1))()
        done = C.c_int()
        st = SurrogateStats()
        rc = lib.nosh_batch_continuation_run(self._h, _ptr(psi), span, steps, pred,
                                             C.byref(o),
                                             C.byref(so) if so is not None else None, rec,
                                             C.byref(done), C.byref(st))
        if rc not in (_OK, _NOTCONV):
            _check(rc)
        out = [{k: getattr(rec[i], k) for k, _ in BatchContinuationStep._fields_}
               for i in range(done.value)]
        stats = {k: getattr(st, k) for k, _ in SurrogateStats._fields_} if so is not None else None
        return psi, out, stats


class BatchContinuationStep(C.Structure):
    _fields_ = [("param", C.c_double), ("newton_iters", C.c_int), ("minres_iters", C.c_int),
                ("converged", C.c_int), ("seeded", C.c_int), ("fallbacks", C.c_int),
                ("max_norm_f", C.c_double), ("seconds", C.c_double), ("pred_error", C.c_double),
                ("sec_error", C.c_double), ("halvings", C.c_int), ("retries", C.c_int),
                ("serial_fallbacks", C.c_int)]


def _load_mlp():
    lib = _load_batch()
    if getattr(lib, "_nosh_mlp", False):
        return lib
    vp, i32, i64 = C.c_void_p, C.c_int, C.c_int64
    sig = {
        "nosh_mlp_create": (i32, [i32, i32, i32, i32, C.c_uint64, C.POINTER(vp)]),
        "nosh_mlp_num_params": (i32, [vp, C.POINTER(i64), C.POINTER(i64)]),
        "nosh_mlp_get_params": (i32, [vp, vp]),
        "nosh_mlp_set_params": (i32, [vp, vp]),
        "nosh_mlp_forward": (i32, [vp, vp, i64, vp]),
        "nosh_mlp_forward_device": (i32, [vp, vp, i64, vp, vp]),
        "nosh_mlp_infer": (i32, [vp, vp, i64, vp]),
        "nosh_mlp_infer_device": (i32, [vp, vp, i64, vp, vp]),
        "nosh_mlp_gradients": (i32, [vp, vp, vp, i64, vp, C.POINTER(C.c_double)]),
        "nosh_mlp_train_step_device": (i32, [vp, vp, vp, i64, C.c_float,
                                             C.POINTER(C.c_double), vp]),
        "nosh_mlp_destroy": (i32, [vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    lib._nosh_mlp = True
    return lib


class Mlp(_Handle):
    """Per-vertex MLP (BF16 MFMA kernels, FP32 accumulation and masters)."""
    _destroy = "nosh_mlp_destroy"

    def __init__(self, n_in=7, width=256, depth=4, n_out=2, seed=0):
        super().__init__()
        lib = _load_mlp()
        _check(lib.nosh_mlp_create(n_in, width, depth, n_out, seed, C.byref(self._h)))
        self.n_in, self.n_out = n_in, n_out
        self.dims = [n_in] + [width] * depth + [n_out]
        npar, nw = C.c_int64(), C.c_int64()
        _check(lib.nosh_mlp_num_params(self._h, C.byref(npar), C.byref(nw)))
        self.nparams, self.nweights = npar.value, nw.value

    def get_params(self) -> np.ndarray:
        p = np.empty(self.nparams, np.float32)
        _check(_lib.nosh_mlp_get_params(self._h, _ptr(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(_lib.nosh_mlp_set_params(self._h, _ptr(p)))

    def forward(self, X) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.empty((X.shape[0], self.n_out), np.float32)
        _check(_lib.nosh_mlp_forward(self._h, _ptr(X), X.shape[0], _ptr(Y)))
        return Y

    def infer(self, X) -> np.ndarray:
        """Same outputs as forward() from the fused inference kernel."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.empty((X.shape[0], self.n_out), np.float32)
        _check(_lib.nosh_mlp_infer(self._h, _ptr(X), X.shape[0], _ptr(Y)))
        return Y

    def infer_device(self, x_ptr: int, M: int, y_ptr: int, stream: int = 0):
        _check(_lib.nosh_mlp_infer_device(self._h, C.c_void_p(x_ptr), M, C.c_void_p(y_ptr),
                                          C.c_void_p(stream)))

    def gradients(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.empty(self.nparams, np.float32)
        loss = C.c_double()
        _check(_lib.nosh_mlp_gradients(self._h, _ptr(X), _ptr(T), X.shape[0], _ptr(g),
                                       C.byref(loss)))
        return loss.value, g

    def train_step_device(self, x_ptr: int, t_ptr: int, M: int, lr: float, stream: int = 0):
        loss = C.c_double()
        _check(_lib.nosh_mlp_train_step_device(self._h, C.c_void_p(x_ptr), C.c_void_p(t_ptr),
                                               M, lr, C.byref(loss), C.c_void_p(stream)))
        return loss.value


class SurrogateOptions(C.Structure):
    _fields_ = [("width", C.c_int), ("depth", C.c_int), ("train_after", C.c_int),
                ("train_steps", C.c_int), ("batch", C.c_int), ("lr", C.c_float),
                ("newton_budget", C.c_int), ("energy_tol", C.c_double), ("seed", C.c_uint64),
                ("retrain_steps", C.c_int), ("smooth_sweeps", C.c_int),
                ("accept_ratio", C.c_double)]


class SurrogateStats(C.Structure):
    _fields_ = [("trained", C.c_int), ("train_seconds", C.c_double), ("final_loss", C.c_double),
                ("steps_seeded", C.c_int), ("fallbacks", C.c_int)]


def _load_surrogate(lib):
    if not getattr(lib, "_nosh_sur", False):
        lib.nosh_surrogate_default_options.restype = C.c_int
        lib.nosh_surrogate_default_options.argtypes = [C.POINTER(SurrogateOptions)]
        lib.nosh_continuation_run_surrogate.restype = C.c_int
        lib.nosh_continuation_run_surrogate.argtypes = [
            C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_char_p), C.c_void_p,
            C.POINTER(ContinuationOptions), C.POINTER(SurrogateOptions),
            C.POINTER(ContinuationStep), C.POINTER(C.c_int), C.POINTER(SurrogateStats)]
        lib._nosh_sur = True
    return lib


def continuation_surrogate(model: Nls, psi0, params: Dict[str, float], parameter: str = "mu",
                           start: float = 0.0, stop: float = 1.0, steps: int = 10,
                           newton: Optional[NewtonOptions] = None, surrogate: Optional[dict] = None,
                           **newton_kw):
    """Surrogate-seeded continuation; returns (psi, step records, surrogate stats)."""
    lib = _load_surrogate(_load_mlp())
    o = ContinuationOptions()
    _check(lib.nosh_continuation_default_options(C.byref(o)))
    pname = parameter.encode()
    o.parameter, o.start, o.stop, o.steps, o.predictor = pname, start, stop, steps, 1
    o.newton = newton if newton is not None else default_newton_options(**newton_kw)
    so = SurrogateOptions()
    _check(lib.nosh_surrogate_default_options(C.byref(so)))
    for k, v in (surrogate or {}).items():
        setattr(so, k, v)
    psi = np.array(_vec(psi0, 2 * model.mesh.num_vertices), copy=True)
    rec = (ContinuationStep * (steps + 1))()
    done = C.c_int()
    st = SurrogateStats()
    p = _Params(params)
    rc = lib.nosh_continuation_run_surrogate(model._h, _ptr(psi), *p.args, C.byref(o),
                                             C.byref(so), rec, C.byref(done), C.byref(st))
    if rc not in (_OK, _NOTCONV):
        _check(rc)
    out = [rec[i].as_dict() for i in range(done.value)]
    return psi, out, {k: getattr(st, k) for k, _ in SurrogateStats._fields_}


# ---------------------------------------------------------------- stability eigenvalues
class EigenOptions(C.Structure):
    _fields_ = [("nev", C.c_int), ("block_size", C.c_int), ("max_iters", C.c_int),
                ("tol", C.c_double), ("which", C.c_int), ("deflate_gauge", C.c_int),
                ("use_prec", C.c_int), ("cheb_iters", C.c_int), ("cheb_ratio", C.c_double),
                ("seed", C.c_uint64)]


class EigenStats(C.Structure):
    _fields_ = [("iters", C.c_int), ("converged", C.c_int), ("applies", C.c_int),
                ("max_residual", C.c_double), ("scale", C.c_double), ("seconds", C.c_double)]


class ArclengthOptions(C.Structure):
    _fields_ = [("parameter", C.c_char_p), ("start", C.c_double), ("min_value", C.c_double),
                ("max_value", C.c_double), ("max_steps", C.c_int), ("direction", C.c_int),
                ("initial_step", C.c_double), ("min_step", C.c_double),
                ("max_step", C.c_double), ("aggressiveness", C.c_double),
                ("predictor", C.c_int), ("border_rtol", C.c_double), ("eigenvalues", C.c_int),
                ("newton", NewtonOptions)]


class ArclengthStep(C.Structure):
    _fields_ = [("param", C.c_double), ("step", C.c_double), ("param_dot", C.c_double),
                ("newton_iters", C.c_int), ("minres_iters", C.c_int), ("failures", C.c_int),
                ("converged", C.c_int), ("norm_f", C.c_double), ("gibbs", C.c_double),
                ("seconds", C.c_double), ("num_eigs", C.c_int), ("num_negative", C.c_int),
                ("eigs", C.c_double * 8)]

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "eigs"}
        d["eigs"] = [self.eigs[i] for i in range(self.num_eigs)]
        return d


def _load_eigen():
    lib = _load()
    if getattr(lib, "_nosh_eig", False):
        return lib
    vp, i32 = C.c_void_p, C.c_int
    cpp = C.POINTER(C.c_char_p)
    sig = {
        "nosh_eigen_default_options": (i32, [C.POINTER(EigenOptions)]),
        "nosh_nls_jacobian_eigenpairs": (i32, [vp, vp, i32, cpp, vp, C.POINTER(EigenOptions),
                                               vp, vp, C.POINTER(EigenStats)]),
        "nosh_arclength_default_options": (i32, [C.POINTER(ArclengthOptions)]),
        "nosh_arclength_run": (i32, [vp, vp, i32, cpp, vp, C.POINTER(ArclengthOptions),
                                     C.POINTER(ArclengthStep), vp, C.POINTER(i32)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    lib._nosh_eig = True
    return lib


def jacobian_eigenpairs(model: Nls, psi, params: Dict[str, float], nev: int = 4,
                        vectors: bool = False, **kw):
    """Extreme eigenpairs of J(psi) (nosh-eig): returns (evals, evecs or None, stats).

    Options as nosh_eigen_options (block_size, max_iters, tol, which,
    deflate_gauge, use_prec, cheb_iters, cheb_ratio, seed). Raises
    NoshError when fewer than nev pairs converge.
    """
    lib = _load_eigen()
    o = EigenOptions()
    _check(lib.nosh_eigen_default_options(C.byref(o)))
    o.nev = nev
    o.block_size = kw.pop("block_size", min(16, nev + 2))
    for k, v in kw.items():
        if k not in dict(EigenOptions._fields_):
            raise AttributeError(k)
        setattr(o, k, v)
    m = 2 * model.mesh.num_vertices
    x = _vec(psi, m)
    ev = np.zeros(nev)
    vec = np.zeros((nev, m)) if vectors else None
    st = EigenStats()
    p = _Params(params)
    _check(lib.nosh_nls_jacobian_eigenpairs(model._h, _ptr(x), *p.args, C.byref(o), _ptr(ev),
                                            _ptr(vec) if vectors else None, C.byref(st)))
    return ev, vec, {k: getattr(st, k) for k, _ in EigenStats._fields_}


def arclength(model: Nls, psi0, params: Dict[str, float], parameter: str = "mu",
              start: float = 0.0, max_steps: int = 20, states: bool = False,
              newton: Optional[NewtonOptions] = None, newton_kw: Optional[dict] = None, **kw):
    """Pseudo-arclength continuation (LOCA's "Arc Length" stepper).

    Returns (final psi, step records[, states (steps_done x 2n)]). Keyword
    options as nosh_arclength_options (min_value, max_value, direction,
    initial_step, min_step, max_step, aggressiveness, predictor,
    border_rtol, eigenvalues). A run that ends because the step size fell
    below min_step returns the accepted records plus the failed one.
    """
    lib = _load_eigen()
    o = ArclengthOptions()
    _check(lib.nosh_arclength_default_options(C.byref(o)))
    pname = parameter.encode()
    o.parameter, o.start, o.max_steps = pname, start, max_steps
    for k, v in kw.items():
        if k not in dict(ArclengthOptions._fields_):
            raise AttributeError(k)
        setattr(o, k, v)
    o.newton = newton if newton is not None else default_newton_options(**(newton_kw or {}))
    m = 2 * model.mesh.num_vertices
    psi = np.array(_vec(psi0, m), copy=True)
    rec = (ArclengthStep * (max_steps + 1))()
    done = C.c_int()
    st = np.zeros((max_steps + 1, m)) if states else None
    p = _Params(params)
    rc = lib.nosh_arclength_run(model._h, _ptr(psi), *p.args, C.byref(o), rec,
                                _ptr(st) if states else None, C.byref(done))
    if rc not in (_OK, _NOTCONV):
        _check(rc)
    out = [rec[i].as_dict() for i in range(done.value)]
    if rc == _NOTCONV and done.value <= max_steps:
        out.append(rec[done.value].as_dict())
    if states:
        return psi, out, st[:done.value]
    return psi, out
