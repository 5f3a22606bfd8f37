"""ctypes binding of libfdfd (include/fdfd.h).

The library is the product path: there is no Python or CPU fallback.  If the
shared object is missing, importing this module raises immediately.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(HERE, "..", "lib", "libfdfd.so"))

FDFD_TM, FDFD_TE = 1, 2
FDFD_FP64, FDFD_FP32_IR = 0, 1
FDFD_WHICH_LM, FDFD_WHICH_LR = 0, 1
(FDFD_INJECT_NONE, FDFD_INJECT_PIVOT_FALLBACK, FDFD_INJECT_SINGULAR, FDFD_INJECT_NOCONV,
 FDFD_INJECT_OOM) = range(5)

STATUS = {
    0: "FDFD_OK", -1: "FDFD_ERR_INVALID", -2: "FDFD_ERR_OOM", -3: "FDFD_ERR_SINGULAR",
    -4: "FDFD_ERR_NOCONV", -5: "FDFD_ERR_HIP", -6: "FDFD_ERR_NOTFOUND", -7: "FDFD_ERR_STATE",
}


class fdfd_complex(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


class fdfd_stats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int), ("restarts", C.c_int), ("residual", C.c_double),
        ("t_factor_ms", C.c_double), ("t_solve_ms", C.c_double),
        ("pivot_growth", C.c_double), ("singular_line", C.c_int), ("nconv", C.c_int),
        ("pivot_fallback", C.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class fdfd_mode_problem(C.Structure):
    _fields_ = [("n", C.c_int), ("pol", C.c_int), ("omega", C.c_double), ("dh", C.c_double),
                ("neff", fdfd_complex), ("eps", C.c_void_p)]


class fdfd_flux_plane(C.Structure):
    _fields_ = [("normal", C.c_int), ("index", C.c_int), ("lo", C.c_int), ("hi", C.c_int)]


class FDFDError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_PI = C.POINTER(C.c_int)
_PD = C.POINTER(C.c_double)
_PS = C.POINTER(fdfd_stats)

# name -> (restype, argtypes)
PROTOTYPES = {
    "fdfd_version": (_I, [_PI, _PI, _PI]),
    "fdfd_last_error": (C.c_char_p, [_P]),
    "fdfd_arch": (C.c_char_p, []),
    "fdfd_create": (_I, [_I, C.POINTER(_P)]),
    "fdfd_destroy": (_I, [_P]),
    "fdfd_set_grid": (_I, [_P, _I, _I, _D, _D, _I, _I, _D]),
    "fdfd_set_eps": (_I, [_P, _P]),
    "fdfd_set_eps_d": (_I, [_P, _P]),
    "fdfd_factor": (_I, [_P, _I, _D, _I, _PI, _PS]),
    "fdfd_release_factor": (_I, [_P, _I]),
    "fdfd_solve": (_I, [_P, _I, _I, _P, _P, _D, _PS]),
    "fdfd_solve_d": (_I, [_P, _I, _I, _P, _P, _D, _PS]),
    "fdfd_linsolve": (_I, [_P, _I, _I, _I, _P, _P, _D, _PS]),
    "fdfd_apply": (_I, [_P, _I, _I, _P, _P]),
    "fdfd_residual": (_I, [_P, _I, _P, _P, _PD]),
    "fdfd_adjoint_grad": (_I, [_P, _I, _P, _P, fdfd_complex, _P]),
    "fdfd_eigenfrequency": (_I, [_P, _I, _D, _I, _I, _I, _I, _D, _I, _P, _P, _PS]),
    "fdfd_eigenmode": (_I, [_P, _I, C.POINTER(fdfd_mode_problem), _I, _I, _D, _D, _I, _P, _P,
                            _PI, _PS]),
    "fdfd_modulated_solve": (_I, [_P, _D, _D, _I, _I, _P, _P, _I, _D, _P, _PS]),
    "fdfd_poynting": (_I, [_P, _I, _I, _I, _P, _P]),
    "fdfd_poynting_d": (_I, [_P, _I, _I, _I, _P, _P]),
    "fdfd_flux": (_I, [_P, _I, _I, _I, _D, _D, _P, _I, C.POINTER(fdfd_flux_plane), _PD]),
    "fdfd_flux_d": (_I, [_P, _I, _I, _I, _D, _D, _P, _I, C.POINTER(fdfd_flux_plane), _PD]),
    "fdfd_flux_d_noalloc": (_I, [_P, _I, _I, _I, _D, _D, _P, _I, C.POINTER(fdfd_flux_plane),
                                 _PD]),
    "fdfd_vector_eigenmode": (_I, [_P, _D, fdfd_complex, _I, _I, _I, _D, _I, _P, _P, _PS]),
    "fdfd_batch_create": (_I, [_P, C.POINTER(_P)]),
    "fdfd_batch_add": (_I, [_P, _I, _I, _D, _D, _I, _I, _D, _I, _P, _I, _P, _I, _P, _PI]),
    "fdfd_batch_run": (_I, [_P, _I, _D, _PS]),
    "fdfd_batch_result": (_I, [_P, _I, _I, _I, _P]),
    "fdfd_batch_destroy": (_I, [_P]),
    "fdfd_profile_enable": (_I, [_P, _I]),
    "fdfd_profile_read": (_I, [_P, _I, _PD, C.POINTER(C.c_longlong), _PD, _PD]),
    "fdfd_set_tuning": (_I, [_P, _I, _I]),
    "fdfd_debug_inject": (_I, [_P, _I, _I]),
    "fdfd_autotune": (_I, [_P, _I, _D, _I, _PS]),
    "fdfd_get_tuning": (_I, [_P, _PI, _I]),
    "fdfd_solve_pullback": (_I, [_P, _I, _P, _P, _D, _P, _PS]),
    "fdfd_solve_pullback_d": (_I, [_P, _I, _P, _P, _D, _P, _PS]),
}

_lib = None


def load(path: str | None = None):
    """Load libfdfd and attach prototypes.  Raises if the library is absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ImportError(
            f"libfdfd not built ({p} missing); run python fdfd.jl_amd/build_lib.py")
    lib = C.CDLL(p)
    for name, (res, args) in PROTOTYPES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(lib, handle, rc: int):
    if rc != 0:
        msg = lib.fdfd_last_error(handle)
        raise FDFDError(rc, msg.decode() if msg else "")
    return rc
