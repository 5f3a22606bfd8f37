"""Modulated devices (MF-FDFD): ModulatedDevice, setup_deps and their solve.

Mirrors
  ModulatedDevice(grid, w, Omega, nsidebands; sharedpml)   src/solver/modulation.jl:4-27
  setup_Δϵᵣ!(d, region, value)                             src/solver/modulation.jl:29-33
  solve(d::ModulatedDevice)                                src/solver/modulation.jl:35-119
The coupled (2K+1)-sideband system is solved in libfdfd (fdfd_modulated_solve):
every sideband operator is factored on the GPU and FP64 FGMRES on the coupled
operator, preconditioned by them, reaches the requested residual.
"""
from __future__ import annotations

import ctypes as C
import logging
from dataclasses import dataclass, field
from typing import Callable

import numpy as np

from . import _lib
from .core import TM, FieldTM, Grid, _mask_values

log = logging.getLogger("fdfd")


@dataclass
class ModulatedDevice:
    """src/solver/modulation.jl:4-27: eps_r, deps_r (Delta eps_r), src are
    complex Nx x Ny; omega is the list of carrier frequencies."""

    grid: Grid
    omega: np.ndarray
    Omega: float
    nsidebands: int
    sharedpml: bool = True
    eps_r: np.ndarray = None
    deps_r: np.ndarray = None
    src: np.ndarray = None
    modes: list = field(default_factory=list)

    def __post_init__(self):
        self.omega = np.atleast_1d(np.asarray(self.omega, dtype=np.float64))
        if self.eps_r is None:
            self.eps_r = np.ones(self.grid.N, dtype=np.complex128)
        if self.deps_r is None:
            self.deps_r = np.zeros(self.grid.N, dtype=np.complex128)
        if self.src is None:
            self.src = np.zeros(self.grid.N, dtype=np.complex128)


def setup_deps(d: ModulatedDevice, region: Callable, value):
    """setup_Δϵᵣ!(d, region, value): src/solver/modulation.jl:32-33 (region
    and value evaluated at cell centres)."""
    _mask_values(d.deps_r, d.grid, region, value)


setup_Δϵᵣ = setup_deps


def solve_modulated(d: ModulatedDevice, precision="fp64", tol: float = 1e-10, eng=None,
                    return_stats: bool = False):
    """solve(d::ModulatedDevice) (modulation.jl:35-119): a list over carriers
    of lists over sidebands j = -K..K of FieldTM at wn = w + j*Omega.  Mode
    sources rebuild src at every carrier (TM, as the reference does)."""
    from .modes import setup_modes_all
    from .solve import PRECISIONS, _flat, _ptr, engine
    eng = eng or engine()
    eng.set_grid(d.grid)
    eng.set_eps(d.eps_r)
    g = d.grid
    N = len(g)
    K = int(d.nsidebands)
    nf = 2 * K + 1
    deps = _flat(d.deps_r)
    fields, stats = [], []
    for i, omega in enumerate(d.omega):
        log.info("Frequency: %d/%d", i + 1, len(d.omega))
        if d.modes:
            setup_modes_all(d, float(omega), eng=eng)
        src = _flat(d.src)
        out = np.empty((nf, 3, N), dtype=np.complex128)
        st = _lib.fdfd_stats()
        eng._chk(eng.lib.fdfd_modulated_solve(
            eng.h, float(omega), float(d.Omega), K, int(bool(d.sharedpml)), _ptr(deps), _ptr(src),
            PRECISIONS[precision], float(tol), _ptr(out), C.byref(st)))
        data = out.reshape(nf, 3, g.N[1], g.N[0]).transpose(0, 3, 2, 1)
        wn = float(omega) + d.Omega * np.arange(-K, K + 1)
        fields.append([FieldTM(g, wn[j], data[j]) for j in range(nf)])
        stats.append(st.as_dict())
    return (fields, stats) if return_stats else fields
