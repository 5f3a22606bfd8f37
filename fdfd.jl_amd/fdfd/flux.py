"""Post-processing of solved fields (host FP64): src/flux.jl.

These run on the host over arrays already returned by the GPU solve; they
are O(Nx + Ny) reductions used for reporting, not part of the hot path.
"""
from __future__ import annotations

import numpy as np

from .core import XHAT, YHAT, Field, FieldTE, FieldTM, Point


def _avg(a: np.ndarray, axis: int) -> np.ndarray:
    """grid_average (src/grid.jl:157-162): (a + circshift(a, 1))/2."""
    return (a + np.roll(a, 1, axis=axis)) / 2


def probe_field(f: Field, component: str, pt: Point):
    """src/flux.jl:6-12 (nearest cell within half a cell)."""
    if component not in f.components:
        raise ValueError(f"{component} is invalid for this polarization; valid: {f.components}")
    g = f.grid
    xi = np.nonzero(np.abs(g.xc() - pt.x) <= g.dx / 2)[0]
    yi = np.nonzero(np.abs(g.yc() - pt.y) <= g.dy / 2)[0]
    return f[component][xi[0], yi[0]]


def poynting(f: Field):
    """src/flux.jl:15-30.  TM: Sx = -1/2 Re(avg_x(Ez) conj(Hy)),
    Sy = 1/2 Re(avg_y(Ez) conj(Hx)).  TE (the reference branch reads an
    undefined field, flux.jl:26): Sx = 1/2 Re(Ey conj(avg_x Hz)),
    Sy = -1/2 Re(Ex conj(avg_y Hz))."""
    if isinstance(f, FieldTM):
        Ez = f["Ez"]
        return (-0.5 * np.real(_avg(Ez, 0) * np.conj(f["Hy"])),
                0.5 * np.real(_avg(Ez, 1) * np.conj(f["Hx"])))
    Hz = f["Hz"]
    return (0.5 * np.real(f["Ey"] * np.conj(_avg(Hz, 0))),
            -0.5 * np.real(f["Ex"] * np.conj(_avg(Hz, 1))))


def flux_surface_integral(f: Field, center: Point, width: float, normal: int):
    """src/flux.jl:37-66: flux through the segment of half-width `width`
    centred at `center` with normal x.  TM follows flux.jl:39-47; TE (broken
    in the reference, flux.jl:49-54) uses the TE analogue of the same
    formula.  The y-normal case is unimplemented in the reference too."""
    g = f.grid
    if normal != XHAT:
        raise NotImplementedError("ŷ normal flux calculation not implemented (flux.jl:58-65)")
    xs, ys = g.xc(), g.yc()
    xi = int(np.nonzero(np.abs(xs - center.x) <= g.dx / 2)[0][0])
    yi = np.nonzero((ys >= center.y - width) & (ys <= center.y + width))[0]
    if isinstance(f, FieldTM):
        Ez = f["Ez"][xi:xi + 2, yi]
        Ezx = (Ez[0] + Ez[1]) / 2
        Hy = f["Hy"][xi, yi]
        return float(np.sum(-0.5 * np.real(Ezx * np.conj(Hy))) * g.dy)
    Hz = f["Hz"][xi:xi + 2, yi]
    Hzx = (Hz[0] + Hz[1]) / 2
    Ey = f["Ey"][xi, yi]
    return float(np.sum(0.5 * np.real(Ey * np.conj(Hzx))) * g.dy)
