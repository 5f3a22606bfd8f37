"""Poynting vector, flux through planes, probes and S-parameters (src/flux.jl).

  probe_field(field, component, pt)                 flux.jl:6-12  (host indexing)
  poynting(field)                                   flux.jl:15-30 (fdfd_poynting)
  flux_surface_integral(field, center, width, nrm)  flux.jl:37-66 (fdfd_flux)
  scattering_parameters(fields, d)                  flux.jl:68-91 (commented out in
                                                    the reference; mode overlaps)
The reference's TE branches (flux.jl:26-27, 49-54) read undefined fields and
its y-normal flux is unimplemented (:58-65); here they are completed with the
TE analogue / the same formula along the other axis.  Plane indices are
resolved on the host; the sums run in libfdfd on the GPU.
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import _lib
from .core import XHAT, YHAT, Field, FieldTE, FieldTM, Grid, Point


def probe_field(f: Field, component: str, pt: Point):
    """src/flux.jl:6-12 (nearest cell within half a cell)."""
    if component not in f.components:
        raise ValueError(f"{component} is invalid for this polarization; valid: {f.components}")
    g = f.grid
    xi = np.nonzero(np.abs(g.xc() - pt.x) <= g.dx / 2)[0]
    yi = np.nonzero(np.abs(g.yc() - pt.y) <= g.dy / 2)[0]
    return f[component][xi[0], yi[0]]


def flux_plane(g: Grid, center: Point, width: float, normal: int):
    """The plane of flux_surface_integral (flux.jl:40-42): the line nearest the
    centre (atvalue, tolerance h/2) and the cells whose centres lie in
    [c - width, c + width] along it.  Returns (normal, index, lo, hi), 0-based."""
    xs, ys = g.xc(), g.yc()
    if normal == XHAT:
        ci = np.nonzero(np.abs(xs - center.x) <= g.dx / 2)[0]
        line = np.nonzero((ys >= center.y - width) & (ys <= center.y + width))[0]
    elif normal == YHAT:
        ci = np.nonzero(np.abs(ys - center.y) <= g.dy / 2)[0]
        line = np.nonzero((xs >= center.x - width) & (xs <= center.x + width))[0]
    else:
        raise ValueError("normal must be XHAT or YHAT")
    if ci.size == 0 or line.size == 0:
        raise ValueError("flux plane outside the grid")
    return (int(normal), int(ci[0]), int(line[0]), int(line[-1]))


def _pol(f: Field) -> int:
    return 1 if isinstance(f, FieldTM) else 2


def _fields_flat(f: Field) -> np.ndarray:
    """Nx x Ny x 3 -> component-major, x fastest (the fdfd_solve layout)."""
    return np.ascontiguousarray(np.asarray(f.data, dtype=np.complex128).transpose(2, 1, 0)).ravel()


def poynting(f: Field, eng=None):
    """poynting(field) (flux.jl:15-30): (Sx, Sy), Nx x Ny each.
    TM: Sx = -1/2 Re(avg_x(Ez) conj(Hy)), Sy = 1/2 Re(avg_y(Ez) conj(Hx));
    TE: Sx = 1/2 Re(Ey conj(avg_x Hz)), Sy = -1/2 Re(Ex conj(avg_y Hz))."""
    from .solve import engine
    eng = eng or engine()
    g = f.grid
    N = len(g)
    data = _fields_flat(f)
    S = np.empty(2 * N, dtype=np.float64)
    eng._chk(eng.lib.fdfd_poynting(eng.h, _pol(f), g.N[0], g.N[1],
                                   data.ctypes.data_as(C.c_void_p), S.ctypes.data_as(C.c_void_p)))
    S = S.reshape(2, g.N[1], g.N[0])
    return S[0].T.copy(), S[1].T.copy()


def flux_surface_integrals(f: Field, planes: Sequence, eng=None) -> np.ndarray:
    """Fluxes through several planes in one launch; planes are
    (center, width, normal) triples or resolved (normal, index, lo, hi)."""
    from .solve import engine
    eng = eng or engine()
    g = f.grid
    res = [p if len(p) == 4 else flux_plane(g, *p) for p in planes]
    arr = (_lib.fdfd_flux_plane * len(res))(*[_lib.fdfd_flux_plane(*p) for p in res])
    data = _fields_flat(f)
    out = np.empty(len(res), dtype=np.float64)
    eng._chk(eng.lib.fdfd_flux(eng.h, _pol(f), g.N[0], g.N[1], g.dx, g.dy,
                               data.ctypes.data_as(C.c_void_p), len(res), arr,
                               out.ctypes.data_as(C.POINTER(C.c_double))))
    return out


def flux_surface_integral(f: Field, center: Point, width: float, normal: int, eng=None) -> float:
    """flux_surface_integral(field, center, width, normal) (flux.jl:37-66)."""
    return float(flux_surface_integrals(f, [(center, width, normal)], eng=eng)[0])


def scattering_parameters(fields: Sequence[Field], d, eng=None) -> np.ndarray:
    """The mode-overlap S-parameters sketched (commented out) in
    flux.jl:68-91: S[i, j] = sum((Ec - Ej) conj(Ej)) / sum(|Ej|^2) for the
    first mode (the input port, incident mode subtracted) and
    sum(Ec conj(Ej)) / sum(|Ej|^2) for the others, with Ec the computed first
    field component on mode j's slice and Ej the mode at omega_i (unit norm,
    largest entry real positive).  All (omega, mode) eigenproblems go to one
    batched fdfd_eigenmode call."""
    from .modes import eigenmode_batch, slice_indices
    if len(fields) != len(d.omega):
        raise ValueError("one field per frequency of the device is required")
    g = d.grid
    idx, slices, dhs, pols, oms, neffs = [], [], [], [], [], []
    for omega in d.omega:
        for mode in d.modes:
            ix, iy, dh = slice_indices(g, mode.pt, mode.dir, mode.width)
            idx.append((ix, iy))
            slices.append(d.eps_r[np.ix_(ix - 1, iy - 1)].ravel())
            dhs.append(dh)
            pols.append(mode.pol)
            oms.append(float(omega))
            neffs.append(mode.neff)
    S = np.zeros((len(d.omega), len(d.modes)), dtype=np.complex128)
    if not slices:
        return S
    _, vecs, _ = eigenmode_batch(slices, dhs, pols, oms, neffs, 1, g.L0, eng=eng)
    q = 0
    for i, f in enumerate(fields):
        for j in range(len(d.modes)):
            ix, iy = idx[q]
            Ej = vecs[q][:, 0]
            Ec = np.asarray(f.data)[np.ix_(ix - 1, iy - 1)][:, :, 0].ravel()
            numer = (Ec - Ej) * np.conj(Ej) if j == 0 else Ec * np.conj(Ej)
            S[i, j] = np.sum(numer) / np.sum(Ej * np.conj(Ej))
            q += 1
    return S
