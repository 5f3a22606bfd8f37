"""fdfd: host package of the MI355X FDFD engine (libfdfd).

Mirrors the FDFD.jl driven-solve / eigenfrequency path: Grid, Device,
setup_eps / setup_src, mode sources, solve, eigenfrequency, FieldTM / FieldTE,
flux.
All numerics run in libfdfd on the GPU.
"""
from .core import (C0, DEFAULT_L0, EPS0, ETA0, MU0, TE, TM, XHAT, YHAT, ZHAT, Device, Field,
                   FieldAll, FieldTE, FieldTM, Grid, Point, normalize_parameters, setup_eps,
                   setup_src)
from .solve import Engine, eigenfrequency, engine, solve
from .flux import (flux_plane, flux_surface_integral, flux_surface_integrals, poynting,
                   probe_field, scattering_parameters)
from .modes import (Mode, add_mode, eigenmode, eigenmode_batch, get_modes, setup_mode,
                    vector_eigenmode)
from .modulated import ModulatedDevice, setup_deps, solve_modulated

__all__ = [
    "C0", "DEFAULT_L0", "EPS0", "ETA0", "MU0", "TE", "TM", "XHAT", "YHAT", "ZHAT", "Device",
    "Field", "FieldTE", "FieldTM", "Grid", "Point", "normalize_parameters", "setup_eps",
    "setup_src", "Engine", "eigenfrequency", "engine", "solve", "flux_surface_integral",
    "poynting", "probe_field", "flux_plane", "flux_surface_integrals", "scattering_parameters",
    "Mode", "add_mode", "eigenmode", "eigenmode_batch", "get_modes",
    "setup_mode", "vector_eigenmode", "FieldAll", "ModulatedDevice", "setup_deps", "solve_modulated",
]
