"""Reverse-mode derivative of the driven solve with respect to eps_r.

The Python mirror of the Julia package's ChainRulesCore rule
(``ChainRulesCore.rrule(::typeof(solve), d::Device, pol)`` in FDFDAMD.jl):

    field, pullback = solve_rrule(d, TM)
    deps = pullback(dfield)        # dfield: cotangent of field.data (Nx x Ny x 3)

The reference (src/solver/driven.jl:4-59) has no derivative.  The forward pass
is the reference's solve (one factorization per frequency).  The factor of
every frequency is kept until the pullback has run, so the pullback costs one
transposed FGMRES solve on the same factor plus one gradient kernel per
frequency (fdfd_solve_pullback): for a real loss L with cotangent
dfield = dL/dRe F + i dL/dIm F,

    dL/deps_k = -Re(lam_k w^2 eps0' Ez_k),  A^T lam = conj(bar Ez + M^H bar H),

with M the H recovery of driven.jl:40-41 and dA/deps_k = w^2 eps0' e_k e_k^T
(driven.jl:21,35).  TM only; the source is held fixed (a mode source whose
slice crosses the design region would also depend on eps_r).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .core import TM, Device, FieldTM
from .solve import _flat, _ptr, engine


def _field_cotangent(df, shape) -> np.ndarray:
    """Cotangent of FieldTM.data (Nx x Ny x 3), or of Ez alone (Nx x Ny), as
    the component-major vector of fdfd_solve's layout."""
    Nx, Ny = shape
    df = np.asarray(df, dtype=np.complex128)
    if df.shape == (Nx, Ny):
        full = np.zeros((Nx, Ny, 3), dtype=np.complex128)
        full[:, :, 0] = df
        df = full
    if df.shape != (Nx, Ny, 3):
        raise ValueError(f"field cotangent must be {Nx}x{Ny}x3 or {Nx}x{Ny}, got {df.shape}")
    return np.ascontiguousarray(df.transpose(2, 1, 0)).ravel()


class _Pullback:
    def __init__(self, eng, fids, ez, shape, tol):
        self.eng, self.fids, self.ez, self.shape, self.tol = eng, fids, ez, shape, tol
        self.stats = []

    def __call__(self, dfield):
        """dfield: cotangent of the field (or a list, one per frequency).
        Returns dL/deps_r (real, Nx x Ny), summed over the frequencies.  The
        engine must still hold the forward factors: a set_grid on it (or a
        solve of another device) in between frees them (FDFD_ERR_NOTFOUND)."""
        if self.fids is None:
            raise RuntimeError("pullback already released")
        many = len(self.fids) > 1
        dfs = list(dfield) if many else [dfield]
        if len(dfs) != len(self.fids):
            raise ValueError("one cotangent per frequency is required")
        Nx, Ny = self.shape
        total = np.zeros(Nx * Ny)
        g = np.empty(Nx * Ny)
        self.stats = []
        for fid, ez, df in zip(self.fids, self.ez, dfs):
            if df is None:
                continue
            v = _field_cotangent(df, self.shape)
            st = _lib.fdfd_stats()
            self.eng._chk(self.eng.lib.fdfd_solve_pullback(
                self.eng.h, fid, _ptr(ez), _ptr(v), float(self.tol), _ptr(g), C.byref(st)))
            total += g
            self.stats.append(st.as_dict())
        return total.reshape((Nx, Ny), order="F")

    def release(self):
        """Free the factors.  A factor already freed by the engine (set_grid,
        another solve's grid, batch runs) is skipped."""
        if self.fids is not None and self.eng.h:
            for fid in self.fids:
                try:
                    self.eng.release(fid)
                except _lib.FDFDError as e:
                    if e.code != -6:  # FDFD_ERR_NOTFOUND
                        raise
        self.fids = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def solve_rrule(d: Device, pol: int = TM, precision="fp64", tol: float = 1e-12, eng=None):
    """rrule of solve(d::Device, pol) with respect to d.eps_r.

    Returns (field, pullback) where field is what ``fdfd.solve.solve`` returns
    (a FieldTM, or a list over d.omega) and ``pullback(dfield)`` returns
    dL/deps_r.  Call ``pullback.release()`` (or drop it) to free the factors."""
    if pol != TM:
        raise NotImplementedError("the solve pullback is implemented for TM")
    if d.modes:
        raise NotImplementedError("mode sources depend on eps_r; set the source explicitly")
    eng = eng or engine()
    eng.set_grid(d.grid)
    eng.set_eps(d.eps_r)
    fields, fids, ezs = [], [], []
    try:
        for omega in d.omega:
            fid, _ = eng.factor(pol, float(omega), precision)
            fids.append(fid)
            f, _ = eng.solve_fields(fid, d.src, tol)
            fields.append(FieldTM(d.grid, omega, f[:, :, :, 0]))
            ezs.append(_flat(f[:, :, 0, 0]))
    except Exception:
        for fid in fids:
            eng.release(fid)
        raise
    pb = _Pullback(eng, fids, ezs, d.grid.N, tol)
    return (fields[0] if len(fields) == 1 else fields), pb


def objective_and_gradient(d: Device, m: np.ndarray, precision="fp64", tol: float = 1e-12,
                           eng=None):
    """J = |m^T Ez|^2 and dJ/deps_r through solve_rrule (one frequency): the
    mode-overlap objective of an inverse-design step.  m: Nx x Ny weights."""
    f, pb = solve_rrule(d, TM, precision, tol, eng)
    try:
        ez = f.data[:, :, 0]
        s = np.sum(np.asarray(m) * ez)
        # dJ/dRe ez + i dJ/dIm ez = 2 s conj(m)
        grad = pb(2 * s * np.conj(m))
    finally:
        pb.release()
    return float(abs(s) ** 2), grad, f
