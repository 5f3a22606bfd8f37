"""Driven solve, linear solves and eigenfrequencies through libfdfd.

Mirrors the reference entry points:
  solve(d::Device, pol=TM)                   src/solver/driven.jl:4-59
  eigenfrequency(d, pol, nev; which=:LM)     src/solver/eigen.jl:69-115
The FDFD_SOLVER environment switch of src/solver/solver.jl:10 is not
consulted: this package has one solver, the GPU block solver.
"""
from __future__ import annotations

import ctypes as C
import logging
import time

import numpy as np

from . import _lib
from ._lib import FDFD_FP32_IR, FDFD_FP64, FDFD_WHICH_LM, FDFD_WHICH_LR, check, fdfd_stats
from .core import TE, TM, Device, FieldTE, FieldTM, Grid

log = logging.getLogger("fdfd")

PRECISIONS = {"fp64": FDFD_FP64, "mixed": FDFD_FP32_IR, "fp32": FDFD_FP32_IR,
              FDFD_FP64: FDFD_FP64, FDFD_FP32_IR: FDFD_FP32_IR}


def _flat(a: np.ndarray) -> np.ndarray:
    """Nx x Ny (x fastest) -> contiguous complex128 vector."""
    return np.ascontiguousarray(np.asarray(a, dtype=np.complex128).ravel(order="F"))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class Engine:
    """One libfdfd handle (one HIP stream) bound to a grid."""

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        h = C.c_void_p()
        rc = self.lib.fdfd_create(device, C.byref(h))
        if rc != 0:
            raise _lib.FDFDError(rc, "fdfd_create failed (no GPU visible?)")
        self.h = h
        self.grid = None

    def close(self):
        if self.h:
            self.lib.fdfd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        return check(self.lib, self.h, rc)

    # -- problem -----------------------------------------------------------
    def set_grid(self, g: Grid):
        self._chk(self.lib.fdfd_set_grid(self.h, g.N[0], g.N[1], g.dx, g.dy, g.Npml[0],
                                         g.Npml[1], g.L0))
        self.grid = g

    def set_eps(self, eps_r: np.ndarray):
        e = _flat(eps_r)
        assert e.size == len(self.grid)
        self._chk(self.lib.fdfd_set_eps(self.h, _ptr(e)))

    def factor(self, pol: int, omega: float, precision="fp64"):
        fid = C.c_int()
        st = fdfd_stats()
        self._chk(self.lib.fdfd_factor(self.h, int(pol), float(omega), PRECISIONS[precision],
                                       C.byref(fid), C.byref(st)))
        return fid.value, st.as_dict()

    def release(self, fid: int):
        self._chk(self.lib.fdfd_release_factor(self.h, fid))

    # -- solves ------------------------------------------------------------
    def solve_fields(self, fid: int, srcs: np.ndarray, tol: float = 1e-10):
        """srcs: Nx x Ny (x fastest) or Nx x Ny x nrhs -> Nx x Ny x 3 x nrhs."""
        g = self.grid
        srcs = np.asarray(srcs, dtype=np.complex128)
        if srcs.ndim == 2:
            srcs = srcs[:, :, None]
        nrhs = srcs.shape[2]
        s = np.ascontiguousarray(np.stack([_flat(srcs[:, :, r]) for r in range(nrhs)]))
        out = np.empty((nrhs, 3, len(g)), dtype=np.complex128)
        st = fdfd_stats()
        self._chk(self.lib.fdfd_solve(self.h, fid, nrhs, _ptr(s), _ptr(out), tol, C.byref(st)))
        fields = out.reshape(nrhs, 3, g.N[1], g.N[0]).transpose(3, 2, 1, 0)
        return fields, st.as_dict()

    def linsolve(self, fid: int, b: np.ndarray, transpose: bool = False, tol: float = 1e-12):
        """x = A \\ b or A^T \\ b with the reference-scaled operator A."""
        b = np.ascontiguousarray(np.asarray(b, dtype=np.complex128).reshape(-1))
        x = np.empty_like(b)
        st = fdfd_stats()
        self._chk(self.lib.fdfd_linsolve(self.h, fid, int(transpose), 1, _ptr(b), _ptr(x), tol,
                                         C.byref(st)))
        return x, st.as_dict()

    def apply(self, fid: int, x: np.ndarray, transpose: bool = False):
        x = np.ascontiguousarray(np.asarray(x, dtype=np.complex128).reshape(-1))
        y = np.empty_like(x)
        self._chk(self.lib.fdfd_apply(self.h, fid, int(transpose), _ptr(x), _ptr(y)))
        return y

    def residual(self, fid: int, x: np.ndarray, b: np.ndarray) -> float:
        x = np.ascontiguousarray(np.asarray(x, dtype=np.complex128).reshape(-1))
        b = np.ascontiguousarray(np.asarray(b, dtype=np.complex128).reshape(-1))
        rel = C.c_double()
        self._chk(self.lib.fdfd_residual(self.h, fid, _ptr(x), _ptr(b), C.byref(rel)))
        return rel.value

    def adjoint_grad(self, fid: int, e: np.ndarray, lam: np.ndarray, weight: complex):
        e = np.ascontiguousarray(np.asarray(e, dtype=np.complex128).reshape(-1))
        lam = np.ascontiguousarray(np.asarray(lam, dtype=np.complex128).reshape(-1))
        grad = np.empty(e.size, dtype=np.float64)
        w = _lib.fdfd_complex(weight.real, weight.imag)
        self._chk(self.lib.fdfd_adjoint_grad(self.h, fid, _ptr(e), _ptr(lam), w, _ptr(grad)))
        return grad

    def eigenfrequency(self, pol: int, omega0: float, nev: int, ncv: int = 0, which="LM",
                       precision="fp64", tol: float = 1e-12, maxiter: int = 300,
                       want_fields: bool = True):
        g = self.grid
        om = np.empty(nev, dtype=np.complex128)
        fl = np.empty((nev, 3, len(g)), dtype=np.complex128) if want_fields else None
        st = fdfd_stats()
        wh = FDFD_WHICH_LM if which in ("LM", FDFD_WHICH_LM) else FDFD_WHICH_LR
        self._chk(self.lib.fdfd_eigenfrequency(
            self.h, int(pol), float(omega0), int(nev), int(ncv), wh, PRECISIONS[precision], tol,
            maxiter, _ptr(om), _ptr(fl) if fl is not None else None, C.byref(st)))
        if fl is not None:
            fl = fl.reshape(nev, 3, g.N[1], g.N[0]).transpose(0, 3, 2, 1)
        return om, fl, st.as_dict()

    def batch_solve(self, jobs, precision="mixed", tol: float = 1e-10, keep: bool = True,
                    fetch=None):
        """Batched driven solves.  jobs: list of (grid, pol, eps_r, omegas,
        srcs[Nx, Ny, nsrc]).  Returns (results, stats) where results[j] is an
        array [nomega, nsrc, Nx, Ny, 3] (or None when keep=False).  With
        fetch = [(job, iomega, isrc), ...] only those fields are copied back,
        and results is a dict keyed by those triples.
        Resets the engine's grid (and releases its factors)."""
        b = C.c_void_p()
        self._chk(self.lib.fdfd_batch_create(self.h, C.byref(b)))
        try:
            ids = []
            for (g, pol, eps, omegas, srcs) in jobs:
                e = _flat(eps)
                om = np.ascontiguousarray(np.asarray(omegas, dtype=np.float64))
                srcs = np.asarray(srcs, dtype=np.complex128)
                if srcs.ndim == 2:
                    srcs = srcs[:, :, None]
                ss = np.ascontiguousarray(np.stack([_flat(srcs[:, :, r])
                                                    for r in range(srcs.shape[2])]))
                jid = C.c_int()
                self._chk(self.lib.fdfd_batch_add(
                    b, g.N[0], g.N[1], g.dx, g.dy, g.Npml[0], g.Npml[1], g.L0, int(pol),
                    _ptr(e), om.size, _ptr(om), ss.shape[0], _ptr(ss), C.byref(jid)))
                ids.append((jid.value, g, om.size, ss.shape[0]))
            st = fdfd_stats()
            self._chk(self.lib.fdfd_batch_run(b, PRECISIONS[precision], tol, C.byref(st)))
            if fetch is not None:
                sel = {}
                for (j, o, r) in fetch:
                    jid, g, _, _ = ids[j]
                    buf = np.empty(3 * len(g), dtype=np.complex128)
                    self._chk(self.lib.fdfd_batch_result(b, jid, o, r, _ptr(buf)))
                    sel[(j, o, r)] = buf.reshape(3, g.N[1], g.N[0]).transpose(2, 1, 0)
                self.grid = None
                return sel, st.as_dict()
            out = []
            for jid, g, nom, ns in ids:
                if not keep:
                    out.append(None)
                    continue
                res = np.empty((nom, ns, g.N[0], g.N[1], 3), dtype=np.complex128)
                buf = np.empty(3 * len(g), dtype=np.complex128)
                for o in range(nom):
                    for r in range(ns):
                        self._chk(self.lib.fdfd_batch_result(b, jid, o, r, _ptr(buf)))
                        res[o, r] = buf.reshape(3, g.N[1], g.N[0]).transpose(2, 1, 0)
                out.append(res)
            self.grid = None
            return out, st.as_dict()
        finally:
            self.lib.fdfd_batch_destroy(b)

    def profile(self, enable: bool = True):
        self._chk(self.lib.fdfd_profile_enable(self.h, int(enable)))

    def profile_read(self, kernel: int):
        ms, n, fl, by = C.c_double(), C.c_longlong(), C.c_double(), C.c_double()
        self._chk(self.lib.fdfd_profile_read(self.h, kernel, C.byref(ms), C.byref(n),
                                             C.byref(fl), C.byref(by)))
        return {"ms": ms.value, "launches": n.value, "flops": fl.value, "bytes": by.value}

    def set_tuning(self, panel_width: int = 0, force_full_search: bool = False):
        self._chk(self.lib.fdfd_set_tuning(self.h, panel_width, int(force_full_search)))

    def autotune(self, pol: int, omega: float, precision="mixed"):
        """Time every factorization variant on the current grid and eps and keep
        the fastest (fdfd_autotune); recorded in FDFD_TUNE_CACHE if set.
        Returns (knobs, winning step ms)."""
        st = fdfd_stats()
        self._chk(self.lib.fdfd_autotune(self.h, int(pol), float(omega), PRECISIONS[precision],
                                         C.byref(st)))
        return self.tuning(), st.t_factor_ms

    def tuning(self):
        """The handle's factorization variants: gemm_b3, trail_h3, fuse_panels,
        gemm_waves, gemm3m."""
        k = (C.c_int * 5)()
        self._chk(self.lib.fdfd_get_tuning(self.h, k, 5))
        return dict(zip(("gemm_b3", "trail_h3", "fuse_panels", "gemm_waves", "gemm3m"), list(k)))

    def debug_inject(self, what: int, arg: int = 0):
        """One-shot failure injection (fdfd_debug_inject) for error-path tests."""
        self._chk(self.lib.fdfd_debug_inject(self.h, int(what), int(arg)))


_default_engine = None


def engine() -> Engine:
    global _default_engine
    if _default_engine is None:
        _default_engine = Engine(0)
    return _default_engine


def solve(d: Device, pol: int = TM, precision="fp64", tol: float = 1e-10, eng: Engine = None,
          return_stats: bool = False):
    from .modulated import ModulatedDevice, solve_modulated
    if isinstance(d, ModulatedDevice):  # solve(d::ModulatedDevice), modulation.jl:35
        return solve_modulated(d, precision=precision, tol=tol, eng=eng,
                               return_stats=return_stats)
    """solve(d::Device, pol::Polarization=TM), src/solver/driven.jl:4-59.

    One factorization per frequency (driven.jl:11); returns a FieldTM/FieldTE,
    or a list when d.omega has several entries (driven.jl:57-58).  With modes
    present the source is rebuilt from them at every frequency
    (driven.jl:15-19, modes.setup_modes_all)."""
    eng = eng or engine()
    eng.set_grid(d.grid)
    eng.set_eps(d.eps_r)
    fields, stats = [], []
    for i, omega in enumerate(d.omega):
        log.info("Frequency: %d/%d", i + 1, len(d.omega))
        t0 = time.time()
        if d.modes:
            from .modes import setup_modes_all
            setup_modes_all(d, float(omega), eng=eng)
        fid, fst = eng.factor(pol, float(omega), precision)
        try:
            f, sst = eng.solve_fields(fid, d.src, tol)
        finally:
            eng.release(fid)
        log.info("Unknowns: %d  Solve time: %.3f s", len(d.grid), time.time() - t0)
        cls = FieldTM if pol == TM else FieldTE
        fields.append(cls(d.grid, omega, f[:, :, :, 0]))
        stats.append({**fst, **{k: v for k, v in sst.items() if k != "t_factor_ms"}})
    out = fields[0] if len(fields) == 1 else fields
    if return_stats:
        return out, (stats[0] if len(stats) == 1 else stats)
    return out


def eigenfrequency(d: Device, pol: int, nev: int, which: str = "LM", precision="fp64",
                   tol: float = 1e-12, ncv: int = 0, eng: Engine = None):
    """eigenfrequency(d, pol, nev; which=:LM), src/solver/eigen.jl:69-115.
    Returns (omega::Vector{Complex}, fields)."""
    eng = eng or engine()
    eng.set_grid(d.grid)
    eng.set_eps(d.eps_r)
    om, fl, st = eng.eigenfrequency(pol, float(d.omega[0]), nev, ncv, which, precision, tol)
    cls = FieldTM if pol == TM else FieldTE
    return om, [cls(d.grid, om[i], fl[i]) for i in range(nev)]
