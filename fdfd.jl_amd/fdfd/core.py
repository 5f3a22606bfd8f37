"""Host-side mirror of the FDFD.jl problem description (types, Grid, Device,
FieldTM/FieldTE).  Only bookkeeping lives here: painting eps_r and src,
coordinates and index maps.  All operator arithmetic runs in libfdfd on the
GPU (see solve.py).

Reference: src/types.jl, src/grid.jl:1-125, src/device.jl, src/data.jl.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, Sequence

import numpy as np

# src/types.jl:6-9
EPS0 = 8.85418782e-12
MU0 = 1.25663706e-6
C0 = math.sqrt(1 / EPS0 / MU0)
ETA0 = math.sqrt(MU0 / EPS0)
DEFAULT_L0 = 1e-6  # src/grid.jl:5

# src/types.jl:18-20
TM, TE = 1, 2
XHAT, YHAT, ZHAT = 1, 2, 3


def normalize_parameters(L0: float = DEFAULT_L0):
    """(eps0*L0, mu0*L0, c0/L0): src/device.jl:40-43."""
    return (EPS0 * L0, MU0 * L0, C0 / L0)


@dataclass(frozen=True)
class Point:
    """src/types.jl:23-26."""
    x: float
    y: float


@dataclass
class Grid:
    """2-D grid, src/grid.jl:7-33.  N = round(L/dh) with ties to even."""

    L: tuple
    L0: float
    N: tuple
    Npml: tuple
    bounds: tuple

    def __init__(self, dh: float, npml: Sequence[int], xrange: Sequence[float],
                 yrange: Sequence[float], L0: float = DEFAULT_L0):
        L = (float(xrange[1] - xrange[0]), float(yrange[1] - yrange[0]))
        self.L = L
        self.L0 = float(L0)
        self.N = (int(round(L[0] / dh)), int(round(L[1] / dh)))
        npml = tuple(int(v) for v in np.ravel(npml))
        self.Npml = (npml[0], npml[1] if len(npml) > 1 else npml[0])
        self.bounds = ((float(xrange[0]), float(yrange[0])), (float(xrange[1]), float(yrange[1])))

    def __len__(self):
        return self.N[0] * self.N[1]

    @property
    def shape(self):
        return self.N

    def size(self):
        return self.N

    @property
    def dx(self) -> float:  # src/grid.jl:68-70 (span / N)
        return (self.bounds[1][0] - self.bounds[0][0]) / self.N[0]

    @property
    def dy(self) -> float:  # src/grid.jl:72-74
        return (self.bounds[1][1] - self.bounds[0][1]) / self.N[1]

    def xc(self) -> np.ndarray:  # src/grid.jl:76-78
        return self.bounds[0][0] + self.dx * (np.arange(self.N[0]) + 0.5)

    def yc(self) -> np.ndarray:  # src/grid.jl:80-82
        return self.bounds[0][1] + self.dy * (np.arange(self.N[1]) + 0.5)

    def xe(self) -> np.ndarray:
        return self.bounds[0][0] + self.dx * np.arange(self.N[0] + 1)

    def ye(self) -> np.ndarray:
        return self.bounds[0][1] + self.dy * np.arange(self.N[1] + 1)

    def x2ind(self, x: float) -> int:  # src/grid.jl:112-117 (1-based)
        ind = int(round((x - self.bounds[0][0]) / self.L[0] * self.N[0])) + 1
        return min(max(ind, 1), self.N[0])

    def y2ind(self, y: float) -> int:  # src/grid.jl:120-125
        ind = int(round((y - self.bounds[0][1]) / self.L[1] * self.N[1])) + 1
        return min(max(ind, 1), self.N[1])

    def coord2ind(self, pt: Point):  # src/grid.jl:103-109
        return (self.x2ind(pt.x), self.y2ind(pt.y))


@dataclass
class Device:
    """src/device.jl:19-35: eps_r and src are complex Nx x Ny arrays."""

    grid: Grid
    omega: np.ndarray
    eps_r: np.ndarray = None
    src: np.ndarray = None
    modes: list = field(default_factory=list)  # src/device.jl:24 (see modes.py)

    def __post_init__(self):
        self.omega = np.atleast_1d(np.asarray(self.omega, dtype=np.float64))
        if self.eps_r is None:
            self.eps_r = np.ones(self.grid.N, dtype=np.complex128)
        if self.src is None:
            self.src = np.zeros(self.grid.N, dtype=np.complex128)


def _mask_values(pixels: np.ndarray, g: Grid, region: Callable, value):
    """src/device.jl:63-83: region and value evaluated at cell centres."""
    X, Y = np.meshgrid(g.xc(), g.yc(), indexing="ij")
    mask = np.vectorize(region, otypes=[bool])(X, Y)
    if callable(value):
        pixels[mask] = np.vectorize(value, otypes=[np.complex128])(X, Y)[mask]
    else:
        pixels[mask] = value


def setup_eps(d: Device, region: Callable, value):
    """setup_ϵᵣ!(d, region, value): src/device.jl:89."""
    _mask_values(d.eps_r, d.grid, region, value)


def setup_src(d: Device, region, value=None, normal: int | None = None):
    """setup_src!: region/value (device.jl:92), Point (:95-98, writes 1im),
    Point + normal direction (:101-108, writes 1im along a full row/column)."""
    if isinstance(region, Point):
        ix, iy = d.grid.coord2ind(region)
        if normal is None:
            d.src[ix - 1, iy - 1] = 1j
        elif normal == XHAT:
            d.src[ix - 1, :] = 1j
        elif normal == YHAT:
            d.src[:, iy - 1] = 1j
        return
    _mask_values(d.src, d.grid, region, value)


# aliases with the reference spelling (Python normalises identifiers to NFKC)
setup_ϵᵣ = setup_eps


@dataclass
class Field:
    """FieldTM / FieldTE (src/data.jl:50-78): data is Nx x Ny x 3 complex with
    axes (xc, yc, component)."""

    grid: Grid
    omega: complex
    data: np.ndarray
    components: tuple = ()

    def __getitem__(self, key):
        if isinstance(key, str):
            return self.data[:, :, self.components.index(key)]
        return self.data[key]

    @property
    def x(self):
        return self.grid.xc()

    @property
    def y(self):
        return self.grid.yc()


class FieldTM(Field):
    def __init__(self, grid, omega, data):
        super().__init__(grid, complex(omega), data, ("Ez", "Hx", "Hy"))


class FieldTE(Field):
    def __init__(self, grid, omega, data):
        super().__init__(grid, complex(omega), data, ("Hz", "Ex", "Ey"))


class FieldAll(Field):
    """src/data.jl:80-92: all six components (full-vector modes)."""

    def __init__(self, grid, omega, data):
        super().__init__(grid, complex(omega), data, ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"))
