"""Mode sources: Mode, add_mode, get_modes, setup_mode and the 1-D eigenmode.

Mirrors
  struct Mode / add_mode!(d, mode)                  src/device.jl:5-11,113-115
  setup_mode!(d, pol, w, neff, pt, normal, width)   src/device.jl:118-121
  get_modes(d, pol, w, neff, nmodes, pt, normal, width)  src/device.jl:124-149
  eigenmode(d::Device{1}, pol, neff, nev)           src/solver/eigen.jl:6-29
The slice bookkeeping is host work; every eigenproblem is solved by
fdfd_eigenmode in libfdfd (one GPU workgroup per slice, batched).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Sequence

import numpy as np

from . import _lib
from .core import DEFAULT_L0, XHAT, YHAT, Device, Grid, Point


@dataclass
class Mode:
    """src/device.jl:5-11."""
    pol: int
    dir: int
    neff: complex
    pt: Point
    width: float


def add_mode(d: Device, mode: Mode):
    """add_mode!(d, mode): src/device.jl:113-115."""
    d.modes.append(mode)


def eigenmode_batch(slices: Sequence[np.ndarray], dh, pol, omega, neff, nev: int = 1,
                    L0: float = DEFAULT_L0, ncv: int = 0, tol: float = 0.0,
                    maxiter: int = 0, eng=None):
    """Eigenmodes of several 1-D slices in one launch.  dh, pol, omega, neff
    are scalars or one value per slice.  Returns (beta[nslice, nev],
    [vectors n_p x nev], stats)."""
    from .solve import engine
    eng = eng or engine()
    k = len(slices)

    def per(v):
        return list(v) if np.ndim(v) else [v] * k
    dhs, pols, oms, nes = per(dh), per(pol), per(omega), per(neff)
    eps = [np.ascontiguousarray(np.asarray(s, dtype=np.complex128).ravel()) for s in slices]
    probs = (_lib.fdfd_mode_problem * k)()
    for p in range(k):
        ne = complex(nes[p])
        probs[p] = _lib.fdfd_mode_problem(eps[p].size, int(pols[p]), float(oms[p]), float(dhs[p]),
                                          _lib.fdfd_complex(ne.real, ne.imag),
                                          eps[p].ctypes.data_as(C.c_void_p))
    ntot = sum(e.size for e in eps)
    beta = np.empty(k * nev, dtype=np.complex128)
    vecs = np.empty(ntot * nev, dtype=np.complex128)
    nconv = (C.c_int * k)()
    st = _lib.fdfd_stats()
    eng._chk(eng.lib.fdfd_eigenmode(eng.h, k, probs, int(nev), int(ncv), float(L0), float(tol),
                                    int(maxiter), beta.ctypes.data_as(C.c_void_p),
                                    vecs.ctypes.data_as(C.c_void_p), nconv, C.byref(st)))
    out, o = [], 0
    for e in eps:
        out.append(vecs[o:o + e.size * nev].reshape(nev, e.size).T)
        o += e.size * nev
    return beta.reshape(k, nev), out, st.as_dict()


def eigenmode(eps1d: np.ndarray, dh: float, pol: int, omega: float, neff, nev: int = 1,
              L0: float = DEFAULT_L0, eng=None):
    """eigenmode(d::Device{1}, pol, neff, nev) (src/solver/eigen.jl:6-29) on
    the 1-D grid Grid(n, [0 n*dh]) with eps1d.  Returns (beta[nev], vectors
    n x nev): the nev eigenvalues beta^2 nearest (omega/c0 neff)^2, nearest
    first, beta = sqrt(beta^2)."""
    beta, vecs, _ = eigenmode_batch([eps1d], dh, pol, omega, neff, nev, L0, eng=eng)
    return beta[0], vecs[0]


def vector_eigenmode(d: Device, neff, nev: int = 1, precision="fp64", tol: float = 0.0,
                     ncv: int = 0, maxiter: int = 0, eng=None):
    """eigenmode(d::AbstractDevice{2}, neff, nev) (src/solver/eigen.jl:31-66):
    full-vector modes of the cross-section d.eps_r at d.omega[0] (periodic,
    no PML).  Returns (beta[nev], [FieldAll]) nearest (omega/c0 neff)^2 first."""
    from .core import FieldAll
    from .solve import PRECISIONS, engine
    eng = eng or engine()
    eng.set_grid(d.grid)
    eng.set_eps(d.eps_r)
    g = d.grid
    N = len(g)
    beta = np.empty(nev, dtype=np.complex128)
    fl = np.empty((nev, 6, N), dtype=np.complex128)
    st = _lib.fdfd_stats()
    ne = complex(neff)
    eng._chk(eng.lib.fdfd_vector_eigenmode(
        eng.h, float(d.omega[0]), _lib.fdfd_complex(ne.real, ne.imag), int(nev), int(ncv),
        PRECISIONS[precision], float(tol), int(maxiter), beta.ctypes.data_as(C.c_void_p),
        fl.ctypes.data_as(C.c_void_p), C.byref(st)))
    data = fl.reshape(nev, 6, g.N[1], g.N[0]).transpose(0, 3, 2, 1)
    omega = float(d.omega[0])
    return beta, [FieldAll(g, omega, data[i]) for i in range(nev)]


def slice_indices(g: Grid, pt: Point, slicenormal: int, slicewidth: float):
    """src/device.jl:125-142: 1-based index arrays of the odd-length slice
    through pt and its cell size.  An x-normal slice runs along y
    (round(width/dy) points), a y-normal one along x; even counts get one
    more point."""
    indx, indy = g.coord2ind(pt)
    if slicenormal == XHAT:
        srcpoints = int(round(slicewidth / g.dy))
    elif slicenormal == YHAT:
        srcpoints = int(round(slicewidth / g.dx))
    else:
        raise ValueError("slice normal must be XHAT or YHAT")
    if srcpoints % 2 == 0:
        srcpoints += 1
    M = (srcpoints - 1) // 2
    if slicenormal == XHAT:
        ix, iy, dh = np.array([indx]), indy + np.arange(-M, M + 1), g.dy
    else:
        ix, iy, dh = indx + np.arange(-M, M + 1), np.array([indy]), g.dx
    if ix.min() < 1 or iy.min() < 1 or ix.max() > g.N[0] or iy.max() > g.N[1]:
        raise IndexError(f"mode slice {ix.min()}:{ix.max()} x {iy.min()}:{iy.max()} leaves the "
                         f"{g.N[0]} x {g.N[1]} grid")
    return ix, iy, dh


def get_modes(d: Device, pol: int, omega: float, neff, nmodes: int, pt: Point,
              slicenormal: int, slicewidth: float, eng=None):
    """get_modes (src/device.jl:124-149): (beta, vectors, indx, indy), with
    1-based index arrays as in the reference."""
    ix, iy, dh = slice_indices(d.grid, pt, slicenormal, slicewidth)
    eps1d = d.eps_r[np.ix_(ix - 1, iy - 1)].ravel()
    beta, vecs = eigenmode(eps1d, dh, pol, omega, neff, nmodes, d.grid.L0, eng=eng)
    return beta, vecs, ix, iy


def setup_mode(d: Device, pol: int, omega: float, neff, pt: Point, srcnormal: int,
               srcwidth: float, eng=None):
    """setup_mode! (src/device.jl:118-121): adds normalize(abs.(v)) of the
    mode nearest the guess to src on the slice."""
    _, v, ix, iy = get_modes(d, pol, omega, neff, 1, pt, srcnormal, srcwidth, eng=eng)
    a = np.abs(v[:, 0])
    d.src[np.ix_(ix - 1, iy - 1)] += (a / np.linalg.norm(a)).reshape(ix.size, iy.size)


def setup_modes_all(d: Device, omega: float, eng=None):
    """The source of driven.jl:15-19 at one frequency: src reset, then every
    mode of d added with TM polarization (the reference passes TM whatever
    mode.pol says).  All slices are solved in one batched call."""
    d.src = np.zeros(d.grid.N, dtype=np.complex128)
    if not d.modes:
        return
    idx, slices, dhs, neffs = [], [], [], []
    for mode in d.modes:
        ix, iy, dh = slice_indices(d.grid, mode.pt, mode.dir, mode.width)
        idx.append((ix, iy))
        slices.append(d.eps_r[np.ix_(ix - 1, iy - 1)].ravel())
        dhs.append(dh)
        neffs.append(mode.neff)
    from .core import TM
    _, vecs, _ = eigenmode_batch(slices, dhs, TM, omega, neffs, 1, d.grid.L0, eng=eng)
    for (ix, iy), v in zip(idx, vecs):
        a = np.abs(v[:, 0])
        d.src[np.ix_(ix - 1, iy - 1)] += (a / np.linalg.norm(a)).reshape(ix.size, iy.size)
