"""CPU oracle for the FDFD driven-solve / eigenfrequency path.

TEST INFRASTRUCTURE ONLY.  This module is the checker: it restates the
reference algorithm (FDFD.jl, Julia) in numpy/scipy so that the HIP engine's
results can be compared against it.  Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import it; the product path
(``fdfd.jl_amd``) never does.

Every function cites the reference file:line it follows.  The sparse matrices
are built in the reference's own order of operations (``kron`` of 1-D periodic
difference matrices, left-multiplied by diagonal inverse stretch factors,
products formed as sparse-sparse products) so that the assembled ``A`` and
``b`` are the reference's ``A`` and ``b`` up to floating-point rounding of the
same operations.

Parity pinning (what checks each function against something other than itself;
see tests/test_oracle.py):
  constants / normalisation ........ literal values, c0*sqrt(eps0*mu0)==1
  Grid, dx, xc, x2ind .............. hand-computed cases, banker's rounding ties
  delta (periodic differences) ..... dense brute force; Dxf^T == -Dxb; FFT symbol
  create_sfactor / S_create ........ design attenuation 2 eta0 int sigma = -lnR; grading
                                     exponent m = 3.5 from sigma at two depths (ratio
                                     (l1/l2)^m); literal sigma at the deepest cell
  grid_average ..................... np.roll restatement + hand case
  TM/TE operator ................... dense kron brute force; FFT eigenvalues for
                                     npml=0; Kronecker-sum separability with PML
  solve (SuperLU + FP64 refinement). residual certificate; FFT closed form for
                                     npml=0 uniform eps; dense zgesv tiny case
  field recovery ................... discrete Ampere (TM) and Faraday (TE) identities of
                                     the solved fields; the TE one fails for a swapped
                                     pairing of the averaged permittivities
  eigenfrequency ................... FFT (npml=0) and Kronecker-sum (uniform eps, PML)
                                     closed forms, TM and TE; TE eigenfield identity
  workload generators (config_c*) .. hand-derived region bounds, seeds, source rows
  modulated (MF-FDFD) system ....... per-Fourier-mode closed form (uniform eps and
                                     deps, npml=0); A^T = W^-1 A W with W = Sxb Syb
  2-D full-vector eigenmode ........ closed-form spectrum (uniform eps); slabs uniform
                                     along x or y reduce exactly to the 1-D TM mode
  1-D eigenmode (mode sources) ..... closed-form spectrum (uniform eps, TM and TE);
                                     lattice slab dispersion relations (TM and TE)
                                     derived independently; hand-computed slices
  flux_surface_integral ............ lossless energy conservation between planes
                                     (TM and TE, x and y normals)
  scattering_parameters ............ modal propagation in a straight guide: |S|
                                     constant and phase advancing by the lattice
                                     wavenumber arccos(1 - b^2 dx^2/2)/dx
  adjoint gradient ................. central finite differences
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field as dc_field
from typing import Callable, Sequence

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

# ---------------------------------------------------------------------------
# constants: src/types.jl:6-9 ; enums src/types.jl:18-21
# ---------------------------------------------------------------------------
EPS0 = 8.85418782e-12
MU0 = 1.25663706e-6
C0 = math.sqrt(1 / EPS0 / MU0)
ETA0 = math.sqrt(MU0 / EPS0)
DEFAULT_L0 = 1e-6  # src/grid.jl:5

TM = 1  # src/types.jl:20
TE = 2
XHAT, YHAT = 1, 2  # src/types.jl:18
FORWARD, BACKWARD = 1, 2  # src/types.jl:19


def normalize_parameters(L0: float = DEFAULT_L0):
    """src/device.jl:40 -- (eps0*L0, mu0*L0, c0/L0)."""
    return (EPS0 * L0, MU0 * L0, C0 / L0)


# ---------------------------------------------------------------------------
# Grid: src/grid.jl:7-125
# ---------------------------------------------------------------------------
@dataclass
class Grid:
    """2-D grid (src/grid.jl:7-33).  ``N = round(L/dh)`` with ties-to-even
    (Julia ``round`` default, src/grid.jl:28; Python ``round`` is also
    ties-to-even)."""

    L: tuple
    L0: float
    N: tuple
    Npml: tuple
    bounds: tuple  # ((x1, y1), (x2, y2))

    @classmethod
    def create(cls, dh: float, npml: Sequence[int], xrange: Sequence[float],
               yrange: Sequence[float], L0: float = DEFAULT_L0) -> "Grid":
        L = (float(xrange[1] - xrange[0]), float(yrange[1] - yrange[0]))
        N = (int(round(L[0] / dh)), int(round(L[1] / dh)))
        return cls(L, L0, N, (int(npml[0]), int(npml[1])),
                   ((float(xrange[0]), float(yrange[0])),
                    (float(xrange[1]), float(yrange[1]))))

    @property
    def shape(self):
        return self.N

    def __len__(self):
        return self.N[0] * self.N[1]

    # src/grid.jl:68-74 : span / N, not dh
    @property
    def dx(self) -> float:
        return (self.bounds[1][0] - self.bounds[0][0]) / self.N[0]

    @property
    def dy(self) -> float:
        return (self.bounds[1][1] - self.bounds[0][1]) / self.N[1]

    def dh(self, w: int) -> float:
        return self.dx if w == XHAT else self.dy

    # src/grid.jl:76-82 (Julia range arithmetic is TwicePrecision; plain
    # float64 here -- masks are specified by integer index in the workloads)
    def xc(self) -> np.ndarray:
        return self.bounds[0][0] + self.dx * (np.arange(self.N[0]) + 0.5)

    def yc(self) -> np.ndarray:
        return self.bounds[0][1] + self.dy * (np.arange(self.N[1]) + 0.5)

    # src/grid.jl:112-125
    def x2ind(self, x: float) -> int:
        ind = int(round((x - self.bounds[0][0]) / self.L[0] * self.N[0])) + 1
        return min(max(ind, 1), self.N[0])

    def y2ind(self, y: float) -> int:
        ind = int(round((y - self.bounds[0][1]) / self.L[1] * self.N[1])) + 1
        return min(max(ind, 1), self.N[1])

    def coord2ind(self, x: float, y: float):
        """src/grid.jl:103-109 (1-based indices)."""
        return (self.x2ind(x), self.y2ind(y))


# ---------------------------------------------------------------------------
# derivative operators: src/grid.jl:128-154
# ---------------------------------------------------------------------------
def _d1(n: int, h: float, s: int) -> sp.csc_matrix:
    """1-D periodic difference; forward: spdiagm(1=>1, 0=>-1, -n+1=>[1])/h
    (src/grid.jl:138), backward: spdiagm(-1=>-1, 0=>1, n-1=>[-1])/h (:141)."""
    inv = 1.0 / h
    rows, cols, vals = [], [], []
    for i in range(n):
        if s == FORWARD:
            rows += [i, i]
            cols += [i, (i + 1) % n]
            vals += [-1.0, 1.0]
        else:
            rows += [i, i]
            cols += [i, (i - 1) % n]
            vals += [1.0, -1.0]
    if n == 1:  # the +1 and -1 land on the same entry and cancel
        return sp.csc_matrix((1, 1), dtype=np.float64)
    m = sp.csc_matrix((np.array(vals) * inv, (rows, cols)), shape=(n, n))
    m.sum_duplicates()
    return m


def delta(w: int, s: int, g: Grid) -> sp.csc_matrix:
    """src/grid.jl:128-154: kron(I_Ny, D_x) for x, kron(D_y, I_Nx) for y
    (column-major, x fastest)."""
    Nx, Ny = g.N
    if w == XHAT:
        return sp.kron(sp.identity(Ny, format="csc"), _d1(Nx, g.dx, s), format="csc")
    return sp.kron(_d1(Ny, g.dy, s), sp.identity(Nx, format="csc"), format="csc")


def grid_average(a: np.ndarray, w: int) -> np.ndarray:
    """src/grid.jl:157-162: (a + circshift(a, 1 along w))/2, periodic."""
    if a.ndim == 1:
        return (a + np.roll(a, 1)) / 2
    if w == XHAT:
        return (a + np.roll(a, 1, axis=0)) / 2
    if w == YHAT:
        return (a + np.roll(a, 1, axis=1)) / 2
    return a


# ---------------------------------------------------------------------------
# SC-PML: src/pml.jl:1-63
# ---------------------------------------------------------------------------
def create_sfactor(w: int, s: int, g: Grid, omega: float, m: float = 3.5,
                   lnR: float = -12.0) -> np.ndarray:
    """src/pml.jl:1-31.  s(l) = 1 - i*sigma(l)/(omega*eps0'),
    sigma(l) = sigma_max*(l/T)^m, sigma_max = -(m+1)lnR/(2*eta0*T)
    (eta0 unnormalised, eps0 normalised)."""
    eps0, _, _ = normalize_parameters(g.L0)
    dw = g.dh(w)
    nw_pml = g.Npml[w - 1]
    Tw = nw_pml * dw
    Nw = g.N[w - 1]
    out = np.ones(Nw, dtype=np.complex128)
    if nw_pml == 0:
        return out
    smax = -(m + 1) * lnR / (2 * ETA0 * Tw)

    def S(l):
        sig = smax * (l / Tw) ** m
        return 1 - 1j * sig / (omega * eps0)

    for i in range(1, Nw + 1):
        if s == FORWARD:
            if i <= nw_pml:
                out[i - 1] = S(dw * (nw_pml - i + 0.5))
            elif i > Nw - nw_pml:
                out[i - 1] = S(dw * (i - (Nw - nw_pml) - 0.5))
        else:
            if i <= nw_pml:
                out[i - 1] = S(dw * (nw_pml - i + 1))
            elif i > Nw - nw_pml:
                out[i - 1] = S(dw * (i - (Nw - nw_pml) - 1))
    return out


def sfactor_vectors(g: Grid, omega: float):
    """The four inverse stretch vectors 1/s of src/pml.jl:33-54 (x-forward,
    x-backward, y-forward, y-backward), kept 1-D."""
    return tuple(1.0 / create_sfactor(w, s, g, omega)
                 for (w, s) in ((XHAT, FORWARD), (XHAT, BACKWARD),
                                (YHAT, FORWARD), (YHAT, BACKWARD)))


def S_create(g: Grid, omega: float):
    """src/pml.jl:33-63: diagonal matrices of 1/s broadcast to 2-D."""
    Nx, Ny = g.N
    sxf, sxb, syf, syb = sfactor_vectors(g, omega)
    Sxf = np.repeat(sxf[:, None], Ny, axis=1)
    Sxb = np.repeat(sxb[:, None], Ny, axis=1)
    Syf = np.repeat(syf[None, :], Nx, axis=0)
    Syb = np.repeat(syb[None, :], Nx, axis=0)
    f = lambda a: sp.diags(a.ravel(order="F"), format="csc")
    return (f(Sxf), f(Sxb), f(Syf), f(Syb))


# ---------------------------------------------------------------------------
# Device: src/device.jl:19-108
# ---------------------------------------------------------------------------
@dataclass
class Device:
    """src/device.jl:19-35; eps_r and src are complex Nx x Ny (x fastest)."""

    grid: Grid
    omega: np.ndarray
    eps_r: np.ndarray = None
    src: np.ndarray = None
    modes: list = dc_field(default_factory=list)  # src/device.jl:24

    def __post_init__(self):
        self.omega = np.atleast_1d(np.asarray(self.omega, dtype=np.float64))
        if self.eps_r is None:
            self.eps_r = np.ones(self.grid.N, dtype=np.complex128)
        if self.src is None:
            self.src = np.zeros(self.grid.N, dtype=np.complex128)

    def setup_eps_region(self, region: Callable, value):
        """src/device.jl:63-86: region evaluated at cell centres."""
        _mask_values(self.eps_r, self.grid, region, value)

    def setup_src_region(self, region: Callable, value):
        """src/device.jl:92."""
        _mask_values(self.src, self.grid, region, value)

    def setup_src_point(self, x: float, y: float):
        """src/device.jl:95-98: writes 1im at coord2ind."""
        ix, iy = self.grid.coord2ind(x, y)
        self.src[ix - 1, iy - 1] = 1j

    def setup_src_line(self, x: float, y: float, normal: int):
        """src/device.jl:101-108."""
        ix, iy = self.grid.coord2ind(x, y)
        if normal == XHAT:
            self.src[ix - 1, :] = 1j
        elif normal == YHAT:
            self.src[:, iy - 1] = 1j


def _mask_values(pixels, g: Grid, region, value):
    X, Y = np.meshgrid(g.xc(), g.yc(), indexing="ij")
    mask = np.vectorize(region)(X, Y).astype(bool)
    if callable(value):
        pixels[mask] = np.vectorize(value)(X, Y)[mask]
    else:
        pixels[mask] = value


# ---------------------------------------------------------------------------
# operator assembly: src/solver/driven.jl:21-53
# ---------------------------------------------------------------------------
def stretched_deltas(g: Grid, omega: float):
    """src/solver/driven.jl:25-31: S*delta for the four derivatives."""
    Sxf, Sxb, Syf, Syb = S_create(g, omega)
    dxb = (Sxb @ delta(XHAT, BACKWARD, g)).tocsc()
    dxf = (Sxf @ delta(XHAT, FORWARD, g)).tocsc()
    dyb = (Syb @ delta(YHAT, BACKWARD, g)).tocsc()
    dyf = (Syf @ delta(YHAT, FORWARD, g)).tocsc()
    return dxf, dxb, dyf, dyb


def assemble(g: Grid, eps_r: np.ndarray, omega: float, pol: int):
    """TM: A = dxf*mu0^-1*dxb + dyf*mu0^-1*dyb + w^2*T_eps (driven.jl:35).
    TE: A = dxf*Texi*dxb + dyf*Teyi*dyb + w^2*mu0*I (driven.jl:45)."""
    eps0, mu0, _ = normalize_parameters(g.L0)
    dxf, dxb, dyf, dyb = stretched_deltas(g, omega)
    N = len(g)
    if pol == TM:
        Teps = sp.diags(eps0 * eps_r.ravel(order="F"), format="csc")
        A = (dxf * (mu0 ** -1)) @ dxb + (dyf * (mu0 ** -1)) @ dyb + (omega ** 2) * Teps
    else:
        Texi = sp.diags((1 / grid_average(eps0 * eps_r, XHAT)).ravel(order="F"), format="csc")
        Teyi = sp.diags((1 / grid_average(eps0 * eps_r, YHAT)).ravel(order="F"), format="csc")
        A = dxf @ Texi @ dxb + dyf @ Teyi @ dyb + (omega ** 2 * mu0) * sp.identity(N, format="csc")
    return A.tocsc()


def rhs(omega: float, src: np.ndarray) -> np.ndarray:
    """b = 1im*omega*src[:] (driven.jl:36,46)."""
    return (1j * omega) * src.ravel(order="F").astype(np.complex128)


# ---------------------------------------------------------------------------
# the reference's `\`: src/solver/solver.jl:4-41 (lu(A)\b)
# ---------------------------------------------------------------------------
def rel_residual(A, x, b) -> float:
    r = b - A @ x
    return float(np.linalg.norm(r) / np.linalg.norm(b))


def linear_solve(A, b, refine_tol: float = 1e-14, max_refine: int = 8,
                 lu=None, trans: str = "N"):
    """FP64 sparse direct solve (SuperLU, the scipy stand-in for UMFPACK's
    lu(A)\\b at solver.jl:35) followed by FP64 iterative refinement so the
    relative residual certifies the solution.  ``trans='T'`` solves A^T x=b."""
    if lu is None:
        lu = spla.splu(A.tocsc(), permc_spec="COLAMD")
    Aop = A if trans == "N" else A.T
    x = lu.solve(b, trans=trans)
    for _ in range(max_refine):
        r = b - Aop @ x
        if np.linalg.norm(r) <= refine_tol * np.linalg.norm(b):
            break
        x = x + lu.solve(r, trans=trans)
    return x, lu


@dataclass
class FieldTM:
    """src/data.jl:50-63: data is Nx x Ny x 3 (Ez, Hx, Hy)."""
    grid: Grid
    omega: complex
    data: np.ndarray


@dataclass
class FieldTE:
    """src/data.jl:65-78: data is Nx x Ny x 3 (Hz, Ex, Ey)."""
    grid: Grid
    omega: complex
    data: np.ndarray


def _stack(g: Grid, *vs):
    sz = g.N
    return np.stack([v.reshape(sz, order="F") for v in vs], axis=2)


def solve(d: Device, pol: int = TM, return_system: bool = False):
    """src/solver/driven.jl:4-59.  With modes present the source is reset and
    rebuilt from the TM modes at every frequency (driven.jl:15-19; the mode
    polarization is always TM there)."""
    g = d.grid
    eps0, mu0, _ = normalize_parameters(g.L0)
    out = []
    systems = []
    for omega in d.omega:
        omega = float(omega)
        if d.modes:
            d.src = np.zeros(g.N, dtype=np.complex128)
            for mode in d.modes:
                setup_mode(d, TM, omega, mode.neff, mode.pt, mode.dir, mode.width)
        A = assemble(g, d.eps_r, omega, pol)
        b = rhs(omega, d.src)
        x, lu = linear_solve(A, b)
        dxf, dxb, dyf, dyb = stretched_deltas(g, omega)
        if pol == TM:
            ez = x
            hx = (-1 / 1j / omega / mu0) * (dyb @ ez)   # driven.jl:40
            hy = (1 / 1j / omega / mu0) * (dxb @ ez)    # driven.jl:41
            out.append(FieldTM(g, complex(omega), _stack(g, ez, hx, hy)))
        else:
            Texi = sp.diags((1 / grid_average(eps0 * d.eps_r, XHAT)).ravel(order="F"))
            Teyi = sp.diags((1 / grid_average(eps0 * d.eps_r, YHAT)).ravel(order="F"))
            hz = x
            ex = (1 / 1j / omega) * (Teyi @ (dyb @ hz))     # driven.jl:50
            ey = (1 / 1j / omega) * (Texi @ (-(dxb @ hz)))  # driven.jl:51
            out.append(FieldTE(g, complex(omega), _stack(g, hz, ex, ey)))
        systems.append((A, b, lu))
    res = out[0] if len(out) == 1 else out
    return (res, systems) if return_system else res


# ---------------------------------------------------------------------------
# modulated device (MF-FDFD): src/solver/modulation.jl:4-119
# ---------------------------------------------------------------------------
@dataclass
class ModulatedDevice:
    """src/solver/modulation.jl:4-27: eps_r, deps_r (the modulation amplitude
    Delta eps_r), src; omega (carrier list), Omega, nsidebands, sharedpml."""

    grid: Grid
    omega: np.ndarray
    Omega: float
    nsidebands: int
    sharedpml: bool = True
    eps_r: np.ndarray = None
    deps_r: np.ndarray = None
    src: np.ndarray = None
    modes: list = dc_field(default_factory=list)

    def __post_init__(self):
        self.omega = np.atleast_1d(np.asarray(self.omega, dtype=np.float64))
        if self.eps_r is None:
            self.eps_r = np.ones(self.grid.N, dtype=np.complex128)
        if self.deps_r is None:
            self.deps_r = np.zeros(self.grid.N, dtype=np.complex128)
        if self.src is None:
            self.src = np.zeros(self.grid.N, dtype=np.complex128)


def sideband_operator(g: Grid, eps_r: np.ndarray, omega: float, omega_pml: float):
    """modulation.jl:80-89: Sxb dxb (1/mu0) Sxf dxf + Syb dyb (1/mu0) Syf dyf
    + w^2 T_eps, stretch factors at omega_pml (note the order: backward
    differences outside, forward inside, unlike driven.jl:35)."""
    eps0, mu0, _ = normalize_parameters(g.L0)
    Sxf, Sxb, Syf, Syb = S_create(g, omega_pml)
    dxf, dxb = delta(XHAT, FORWARD, g), delta(XHAT, BACKWARD, g)
    dyf, dyb = delta(YHAT, FORWARD, g), delta(YHAT, BACKWARD, g)
    Te = sp.diags(eps0 * eps_r.ravel(order="F"))
    L = Sxb @ dxb * (1 / mu0) @ Sxf @ dxf + Syb @ dyb * (1 / mu0) @ Syf @ dyf
    return (L + omega ** 2 * Te).tocsc(), (Sxf, Sxb, Syf, Syb)


def modulated_system(d: ModulatedDevice, omega: float):
    """modulation.jl:44-101 for one carrier: the (2K+1) N matrix, b, the
    sideband frequencies and stretch factors."""
    g = d.grid
    eps0, mu0, _ = normalize_parameters(g.L0)
    K = d.nsidebands
    nf = 2 * K + 1
    wn = omega + d.Omega * np.arange(-K, K + 1)
    N = len(g)
    As, Ss = [], []
    for j in range(nf):
        A, S = sideband_operator(g, d.eps_r, wn[j], omega if d.sharedpml else wn[j])
        As.append(A)
        Ss.append(S)
    A = sp.block_diag(As, format="csc")
    if K > 0:
        Td = sp.diags(eps0 * d.deps_r.ravel(order="F"))
        Cp = sp.kron(sp.diags([0.5 * wn[:-1] ** 2], [1]), Td.conj())   # :95-96
        Cm = sp.kron(sp.diags([0.5 * wn[1:] ** 2], [-1]), Td)          # :97-98
        A = (A + Cp + Cm).tocsc()
    b = np.zeros(nf * N, dtype=np.complex128)
    b[K * N:(K + 1) * N] = 1j * omega * d.src.ravel(order="F")      # :67-69
    return A, b, wn, Ss


def solve_modulated(d: ModulatedDevice):
    """solve(d::ModulatedDevice), modulation.jl:35-119: fields[i][j] is the
    FieldTM of carrier i at sideband j (wn = w + (j-K) Omega), H from the
    forward stretched differences (:107-109).  Modes rebuild src at every
    carrier with TM modes (:52-56)."""
    g = d.grid
    _, mu0, _ = normalize_parameters(g.L0)
    N = len(g)
    out = []
    for omega in d.omega:
        omega = float(omega)
        if d.modes:
            d.src = np.zeros(g.N, dtype=np.complex128)
            for mode in d.modes:
                setup_mode(d, TM, omega, mode.neff, mode.pt, mode.dir, mode.width)
        A, b, wn, Ss = modulated_system(d, omega)
        x, _ = linear_solve(A, b)
        dyf, dxf = delta(YHAT, FORWARD, g), delta(XHAT, FORWARD, g)
        row = []
        for j in range(len(wn)):
            Sxf, _, Syf, _ = Ss[j]
            ez = x[j * N:(j + 1) * N]
            hx = -1 / 1j / wn[j] / mu0 * (Syf @ (dyf @ ez))
            hy = 1 / 1j / wn[j] / mu0 * (Sxf @ (dxf @ ez))
            row.append(FieldTM(g, complex(wn[j]), _stack(g, ez, hx, hy)))
        out.append(row)
    return out


# ---------------------------------------------------------------------------
# mode sources: src/device.jl:5-11,113-149 and the 1-D eigenmode of
# src/solver/eigen.jl:6-29
# ---------------------------------------------------------------------------
@dataclass
class Mode:
    """src/device.jl:5-11.  pt is an (x, y) pair."""
    pol: int
    dir: int
    neff: complex
    pt: tuple
    width: float


def add_mode(d: Device, mode: Mode):
    """src/device.jl:113-115."""
    d.modes.append(mode)


def mode_operator_1d(eps1d: np.ndarray, dh: float, omega: float, pol: int,
                     L0: float = DEFAULT_L0):
    """eigen.jl:7-21 on the 1-D device of get_modes (device.jl:144-146):
    Grid(n, [0 n*dh]) has dx = (n*dh - 0)/n (grid.jl:53-58,68-70) and no PML;
    periodic differences (grid.jl:138,141).
      TM: A = w^2 mu0 Te + dxf dxb
      TE: A = w^2 mu0 Te + Te dxf Texinv dxb
    with Te = diag(eps0 eps), Texinv = diag(1/(eps0 grid_average(eps)))."""
    eps1d = np.asarray(eps1d, dtype=np.complex128)
    n = eps1d.size
    eps0, mu0, _ = normalize_parameters(L0)
    dx = (n * dh - 0.0) / n
    dxf, dxb = _d1(n, dx, FORWARD), _d1(n, dx, BACKWARD)
    Te = sp.diags(eps0 * eps1d)
    if pol == TM:
        return omega ** 2 * mu0 * Te + dxf @ dxb
    Texinv = sp.diags((eps0 * grid_average(eps1d, XHAT)) ** -1)
    return omega ** 2 * mu0 * Te + Te @ dxf @ Texinv @ dxb


def eigenmode_1d(eps1d: np.ndarray, dh: float, omega: float, pol: int, neff, nev: int,
                 L0: float = DEFAULT_L0):
    """eigen.jl:6-29: (beta^2, vectors) = eigs(A, nev, sigma=(w/c0 neff)^2),
    beta = sqrt(beta^2).  Shift-invert ARPACK with which=:LM returns the nev
    eigenvalues nearest sigma; here they come from a dense eigendecomposition
    (exact selection, the slices are small), nearest first.  Vectors have unit
    2-norm with their largest-magnitude entry real and positive (ARPACK's
    phase is arbitrary)."""
    _, _, c0 = normalize_parameters(L0)
    A = mode_operator_1d(eps1d, dh, omega, pol, L0).toarray()
    sigma = (omega / c0 * complex(neff)) ** 2
    w, V = np.linalg.eig(A)
    order = np.argsort(np.abs(w - sigma), kind="stable")[:nev]
    V = V[:, order]
    for k in range(nev):
        v = V[:, k] / np.linalg.norm(V[:, k])
        i = int(np.argmax(np.abs(v)))
        V[:, k] = v * (np.conj(v[i]) / abs(v[i]))
    return np.sqrt(w[order].astype(np.complex128)), V, sigma


def slice_indices(g: Grid, pt, slicenormal: int, slicewidth: float):
    """device.jl:125-142: 1-based indices of the odd-length slice through pt
    and its cell size.  x-normal slices run along y (width / dy points), y-normal
    slices along x; an even count is made odd by adding one."""
    indx, indy = g.coord2ind(*pt)
    srcpoints = int(round(slicewidth / (g.dy if slicenormal == XHAT else g.dx)))
    if srcpoints % 2 == 0:
        srcpoints += 1
    M = (srcpoints - 1) // 2
    if slicenormal == XHAT:
        return np.array([indx]), indy + np.arange(-M, M + 1), g.dy
    return indx + np.arange(-M, M + 1), np.array([indy]), g.dx


def get_modes(d: Device, pol: int, omega: float, neff, nmodes: int, pt, slicenormal: int,
              slicewidth: float):
    """device.jl:124-149.  Returns (beta, vectors, indx, indy) with 1-based
    index arrays; a slice leaving the grid raises (Julia BoundsError)."""
    ix, iy, dh = slice_indices(d.grid, pt, slicenormal, slicewidth)
    if ix.min() < 1 or iy.min() < 1 or ix.max() > d.grid.N[0] or iy.max() > d.grid.N[1]:
        raise IndexError("mode slice leaves the grid")
    eps1d = d.eps_r[np.ix_(ix - 1, iy - 1)].ravel()
    beta, vecs, _ = eigenmode_1d(eps1d, dh, omega, pol, neff, nmodes, d.grid.L0)
    return beta, vecs, ix, iy


def setup_mode(d: Device, pol: int, omega: float, neff, pt, srcnormal: int, srcwidth: float):
    """device.jl:118-121: src[slice] += normalize(abs.(v)) of the mode nearest
    the guess."""
    _, v, ix, iy = get_modes(d, pol, omega, neff, 1, pt, srcnormal, srcwidth)
    a = np.abs(v[:, 0])
    d.src[np.ix_(ix - 1, iy - 1)] += (a / np.linalg.norm(a)).reshape(ix.size, iy.size)


# ---------------------------------------------------------------------------
# full-vector waveguide modes: src/solver/eigen.jl:31-66
# ---------------------------------------------------------------------------
@dataclass
class FieldAll:
    """src/data.jl:80-92: data is Nx x Ny x 6 (Ex, Ey, Ez, Hx, Hy, Hz)."""
    grid: Grid
    omega: complex
    data: np.ndarray


def vector_mode_operator(g: Grid, eps_r: np.ndarray, omega: float):
    """eigen.jl:35-45: the (Hx, Hy) operator
      w^2 mu0 blockdiag(Tey, Tex) + blockdiag(Tey, Tex) [-dyf; dxf] Tezinv [-dyb, dxb]
      + [dxb; dyb] [dxf, dyf]
    with Tex = eps0 avg_x eps, Tey = eps0 avg_y eps, Tezinv = 1/(eps0 avg_y avg_x eps);
    plain periodic differences (no PML)."""
    eps0, mu0, _ = normalize_parameters(g.L0)
    Tex = sp.diags(eps0 * grid_average(eps_r, XHAT).ravel(order="F"))
    Tey = sp.diags(eps0 * grid_average(eps_r, YHAT).ravel(order="F"))
    Tezi = sp.diags((eps0 * grid_average(grid_average(eps_r, XHAT), YHAT).ravel(order="F")) ** -1)
    dxb, dxf = delta(XHAT, BACKWARD, g), delta(XHAT, FORWARD, g)
    dyb, dyf = delta(YHAT, BACKWARD, g), delta(YHAT, FORWARD, g)
    B = sp.block_diag([Tey, Tex])
    return (omega ** 2 * mu0 * B + B @ sp.vstack([-dyf, dxf]) @ Tezi @ sp.hstack([-dyb, dxb])
            + sp.vstack([dxb, dyb]) @ sp.hstack([dxf, dyf])).tocsc()


def vector_eigenmode(g: Grid, eps_r: np.ndarray, omega: float, neff, nev: int):
    """eigenmode(d::AbstractDevice{2}, neff, nev), eigen.jl:31-66: ARPACK
    shift-invert (eigs with sigma = (w/c0 neff)^2), nearest first; fields of
    eigen.jl:54-61 from (Hx, Hy) of unit norm with the largest entry real
    positive (ARPACK's phase is arbitrary)."""
    eps0, mu0, c0 = normalize_parameters(g.L0)
    A = vector_mode_operator(g, eps_r, omega)
    sigma = (omega / c0 * complex(neff)) ** 2
    N = len(g)
    lam, vec = spla.eigs(A, k=nev, sigma=sigma, which="LM", tol=1e-14)
    order = np.argsort(np.abs(lam - sigma), kind="stable")
    lam, vec = lam[order], vec[:, order]
    beta = np.sqrt(lam.astype(np.complex128))
    Texi = sp.diags((eps0 * grid_average(eps_r, XHAT).ravel(order="F")) ** -1)
    Teyi = sp.diags((eps0 * grid_average(eps_r, YHAT).ravel(order="F")) ** -1)
    Tezi = sp.diags((eps0 * grid_average(grid_average(eps_r, XHAT), YHAT).ravel(order="F")) ** -1)
    dxb, dxf = delta(XHAT, BACKWARD, g), delta(XHAT, FORWARD, g)
    dyb, dyf = delta(YHAT, BACKWARD, g), delta(YHAT, FORWARD, g)
    fields = []
    for i in range(nev):
        v = vec[:, i] / np.linalg.norm(vec[:, i])
        k = int(np.argmax(np.abs(v)))
        v = v * (np.conj(v[k]) / abs(v[k]))
        hx, hy = v[:N], v[N:]
        b = beta[i]
        hz = 1 / (1j * b) * (dxf @ hx + dyf @ hy)
        ex = 1 / (1j * omega) * (Texi @ (dyb @ hz + 1j * b * hy))
        ey = 1 / (1j * omega) * (Teyi @ (-1j * b * hx - dxb @ hz))
        ez = 1 / (1j * omega) * (Tezi @ (dxb @ hy - dyb @ hx))
        data = np.stack([c.reshape(g.N, order="F") for c in (ex, ey, ez, hx, hy, hz)], axis=2)
        fields.append(FieldAll(g, complex(omega), data))
    return beta, fields, lam, sigma


# ---------------------------------------------------------------------------
# eigenfrequency: src/solver/eigen.jl:69-115
# ---------------------------------------------------------------------------
def eigen_operator(g: Grid, eps_r: np.ndarray, omega0: float, pol: int):
    """TM: A = diag(1/eps_r) (dxf dxb + dyf dyb), sigma = -w0^2 mu0 eps0
    (eigen.jl:83-86).  TE: A = dxf Tex^-1 dxb + dyf Tey^-1 dyb,
    sigma = -w0^2 mu0 (eigen.jl:100-104)."""
    eps0, mu0, _ = normalize_parameters(g.L0)
    dxf, dxb, dyf, dyb = stretched_deltas(g, omega0)
    if pol == TM:
        Tinv = sp.diags(1 / eps_r.ravel(order="F"), format="csc")
        A = Tinv @ dxf @ dxb + Tinv @ dyf @ dyb
        sigma = -omega0 ** 2 * mu0 * eps0
    else:
        Texi = sp.diags((1 / grid_average(eps0 * eps_r, XHAT)).ravel(order="F"))
        Teyi = sp.diags((1 / grid_average(eps0 * eps_r, YHAT)).ravel(order="F"))
        A = dxf @ Texi @ dxb + dyf @ Teyi @ dyb
        sigma = -omega0 ** 2 * mu0
    return A.tocsc(), sigma


def eigenfrequency(d: Device, pol: int, nev: int, ncv: int | None = None,
                   tol: float = 0.0, which: str = "LM"):
    """Shift-invert Arnoldi via ARPACK (scipy.sparse.linalg.eigs with sigma),
    the same ARPACK family Arpack.jl wraps (eigen.jl:86,104); `which` applies
    to theta = 1/(lambda - sigma) in both.  Eigenvalues are returned nearest
    sigma first (LM) or largest Re theta first (LR); fields use the forward
    stretched differences (eigen.jl:90-91, 108-109)."""
    g = d.grid
    eps0, mu0, _ = normalize_parameters(g.L0)
    omega0 = float(d.omega[0])
    A, sigma = eigen_operator(g, d.eps_r, omega0, pol)
    if ncv is None:
        ncv = max(20, 2 * nev + 1)
    lam, vec = spla.eigs(A, k=nev, sigma=sigma, which=which, ncv=ncv, tol=tol)
    if which == "LR":
        order = np.argsort(-np.real(1 / (lam - sigma)))
    else:
        order = np.argsort(np.abs(lam - sigma))
    lam, vec = lam[order], vec[:, order]
    dxf, dxb, dyf, dyb = stretched_deltas(g, omega0)
    fields = []
    if pol == TM:
        w = np.sqrt(-lam / mu0 / eps0)
        for i in range(nev):
            ez = vec[:, i]
            hx = (-1 / 1j / w[i] / mu0) * (dyf @ ez)
            hy = (1 / 1j / w[i] / mu0) * (dxf @ ez)
            fields.append(FieldTM(g, w[i], _stack(g, ez, hx, hy)))
    else:
        w = np.sqrt(-lam / mu0)
        Texi = sp.diags((1 / grid_average(eps0 * d.eps_r, XHAT)).ravel(order="F"))
        Teyi = sp.diags((1 / grid_average(eps0 * d.eps_r, YHAT)).ravel(order="F"))
        for i in range(nev):
            hz = vec[:, i]
            ex = (1 / 1j / w[i]) * (Texi @ (dyb @ hz))
            ey = (1 / 1j / w[i]) * (Teyi @ (-(dxb @ hz)))
            fields.append(FieldTE(g, w[i], _stack(g, hz, ex, ey)))
    return w, fields, lam, sigma


# ---------------------------------------------------------------------------
# flux: src/flux.jl:37-47 (TM, x-normal)
# ---------------------------------------------------------------------------
def flux_surface_integral_tm_x(f: FieldTM, xcen: float, ycen: float, width: float) -> float:
    """-1/2 sum Re(avg_x(Ez) conj(Hy)) dy over y in [yc-w, yc+w] at the column
    nearest xcen (atvalue tolerance dx/2), flux.jl:39-47."""
    g = f.grid
    xs, ys = g.xc(), g.yc()
    cand = np.nonzero(np.abs(xs - xcen) <= g.dx / 2)[0]
    if cand.size == 0:
        raise ValueError("no column within dx/2")
    xi = int(cand[0])
    yi = np.nonzero((ys >= ycen - width) & (ys <= ycen + width))[0]
    Ez = f.data[xi:xi + 2, yi, 0]
    # grid_average over the 2-row view, first row: circshift((1,0)) of a
    # 2-row array puts row 2 into row 1, so the average is (Ez[x]+Ez[x+1])/2
    Ezx = (Ez[0, :] + Ez[1, :]) / 2
    Hy = f.data[xi, yi, 2]
    return float(np.sum(-0.5 * np.real(Ezx * np.conj(Hy))) * g.dy)


def poynting(f):
    """flux.jl:15-30.  TM: Sx = -1/2 Re(avg_x(Ez) conj(Hy)),
    Sy = 1/2 Re(avg_y(Ez) conj(Hx)), avg = grid_average (periodic, i and
    i-1).  The TE branch of the reference reads an undefined field (:26-27);
    here it is the TE analogue, Sx = 1/2 Re(Ey conj(avg_x Hz)),
    Sy = -1/2 Re(Ex conj(avg_y Hz)) (the x, y parts of 1/2 Re(E x H*))."""
    u = f.data[:, :, 0]
    if isinstance(f, FieldTM):
        return (-0.5 * np.real(grid_average(u, XHAT) * np.conj(f.data[:, :, 2])),
                0.5 * np.real(grid_average(u, YHAT) * np.conj(f.data[:, :, 1])))
    return (0.5 * np.real(f.data[:, :, 2] * np.conj(grid_average(u, XHAT))),
            -0.5 * np.real(f.data[:, :, 1] * np.conj(grid_average(u, YHAT))))


def flux_surface_integral(f, xcen: float, ycen: float, width: float, normal: int) -> float:
    """flux.jl:37-66.  x-normal TM is the reference's formula
    (flux_surface_integral_tm_x).  The reference's x-normal TE branch is
    broken (:49-54: undefined pt, :Hy for TE) and its y-normal branches are
    unimplemented (:58-65); here they are completed the same way: the line
    nearest the centre (tolerance h/2), cells within [c - w, c + w] along it,
    the first field averaged with the next line (index + 1), times the cell
    size along the line.
      TE x: 1/2 Re(Ey conj(avg Hz))     TM y: 1/2 Re(avg Ez conj(Hx))
      TE y: -1/2 Re(Ex conj(avg Hz))"""
    g = f.grid
    xs, ys = g.xc(), g.yc()
    if normal == XHAT:
        ci = np.nonzero(np.abs(xs - xcen) <= g.dx / 2)[0]
        if ci.size == 0:
            raise ValueError("no column within dx/2")
        i = int(ci[0])
        line = np.nonzero((ys >= ycen - width) & (ys <= ycen + width))[0]
        u = (f.data[i, line, 0] + f.data[i + 1, line, 0]) / 2
        if isinstance(f, FieldTM):
            return float(np.sum(-0.5 * np.real(u * np.conj(f.data[i, line, 2]))) * g.dy)
        return float(np.sum(0.5 * np.real(f.data[i, line, 2] * np.conj(u))) * g.dy)
    cj = np.nonzero(np.abs(ys - ycen) <= g.dy / 2)[0]
    if cj.size == 0:
        raise ValueError("no row within dy/2")
    j = int(cj[0])
    line = np.nonzero((xs >= xcen - width) & (xs <= xcen + width))[0]
    u = (f.data[line, j, 0] + f.data[line, j + 1, 0]) / 2
    if isinstance(f, FieldTM):
        return float(np.sum(0.5 * np.real(u * np.conj(f.data[line, j, 1]))) * g.dx)
    return float(np.sum(-0.5 * np.real(f.data[line, j, 1] * np.conj(u))) * g.dx)


def scattering_parameters(fields, d):
    """flux.jl:68-91 (commented out in the reference, with `mode.coor` for
    `mode.pt`): S[i, j] = sum((Ec - Ej) conj(Ej)) / sum(|Ej|^2) for the first
    mode, sum(Ec conj(Ej)) / sum(|Ej|^2) for the others, Ec the computed
    first field component on mode j's slice and Ej the mode (unit norm,
    largest entry real positive)."""
    S = np.zeros((len(d.omega), len(d.modes)), dtype=np.complex128)
    for i, omega in enumerate(d.omega):
        for j, mode in enumerate(d.modes):
            _, Ej, ix, iy = get_modes(d, mode.pol, float(omega), mode.neff, 1, mode.pt, mode.dir,
                                      mode.width)
            Ej = Ej[:, 0]
            Ec = fields[i].data[np.ix_(ix - 1, iy - 1)][:, :, 0].ravel()
            numer = (Ec - Ej) * np.conj(Ej) if j == 0 else Ec * np.conj(Ej)
            S[i, j] = np.sum(numer) / np.sum(Ej * np.conj(Ej))
    return S


# ---------------------------------------------------------------------------
# workload generators (integer-index geometry; x fastest; 0-based arrays)
# ---------------------------------------------------------------------------
def omega_of_wavelength(lam_um: float, L0: float = DEFAULT_L0) -> float:
    _, _, c0 = normalize_parameters(L0)
    return 2 * math.pi * c0 / lam_um


def config_c1():
    """200x120 cells, dl=0.02, 15-cell PML, eps=12.25 guide on iy 54:68
    (1-based), point source 1im at (51,61), lambda=1.55."""
    g = Grid.create(0.02, (15, 15), (0.0, 4.0), (-1.2, 1.2))
    d = Device(g, omega_of_wavelength(1.55))
    d.eps_r[:, 53:68] = 12.25
    d.src[50, 60] = 1j
    return d


def random_pixel_device(Nx: int, Ny: int, npml: int, lam_um: float, seed,
                        design: int | None = None, pixel: int = 8,
                        guide: int = 15, src_len: int = 25, src_ix: int = 51,
                        dl: float = 0.02):
    """Random-pixel coupler: centred design region of pixel x pixel blocks
    with eps in {1, 12.25} (Bernoulli 0.5), two 15-cell guides through the
    PML on the centre row, and a 25-cell line source (value 1.0) at
    ix=src_ix.  At 1024x1024 the design is ix,iy in 213:812 (1-based)."""
    g = Grid.create(dl, (npml, npml), (-Nx * dl / 2, Nx * dl / 2), (-Ny * dl / 2, Ny * dl / 2))
    assert g.N == (Nx, Ny), g.N
    d = Device(g, omega_of_wavelength(lam_um))
    if design is None:
        design = int(round(600 / 1024 * min(Nx, Ny) / pixel)) * pixel
    npx = design // pixel
    rng = np.random.Generator(np.random.PCG64(seed))
    pix = rng.random((npx, npx)) < 0.5
    x0 = (Nx - design) // 2
    y0 = (Ny - design) // 2
    eps = np.ones((Nx, Ny))
    blk = np.kron(np.where(pix, 12.25, 1.0), np.ones((pixel, pixel)))
    eps[x0:x0 + design, y0:y0 + design] = blk
    # 1024: iy 506:520 (1-based) -> 0-based 505..519
    yg0 = Ny // 2 - guide // 2
    eps[:x0, yg0:yg0 + guide] = 12.25
    eps[x0 + design:, yg0:yg0 + guide] = 12.25
    d.eps_r[:] = eps
    # 1024: iy 501:525 (1-based) -> 0-based 500..524
    ys0 = Ny // 2 - src_len // 2
    d.src[src_ix - 1, ys0:ys0 + src_len] = 1.0
    return d


def config_c2(n: int = 1024, npml: int = 20, seed: int = 2):
    return random_pixel_device(n, n, npml, 1.55, seed)


def config_c3():
    """Batched multi-frequency, multi-source TM workload: grids (Nx, Ny) in
    {(256, 256), (384, 512), (512, 512)}, 20-cell PML, the random-pixel
    permittivity of config_c2 scaled to each grid (design side
    round(600/1024 * min(Nx, Ny)/8)*8 cells, 8x8 pixels, seed (3, grid
    index)), 15-cell guides through the centre row, 64 wavelengths uniform
    in [1.50, 1.60] um and 16 unit point sources (1im, device.jl:97) per grid
    at ix ~ U{21..floor(0.2 Nx)}, iy ~ U{21..Ny-20} (1-based, seed (30, grid
    index)), shared by every wavelength.  The reference would run one
    solve() per (grid, source) with a factorization per wavelength
    (driven.jl:11-55): 192 factorizations and 3072 right-hand sides.

    Returns a list of (Device, omegas[64], srcs[Nx, Ny, 16])."""
    lams = np.linspace(1.50, 1.60, 64)
    oms = np.array([omega_of_wavelength(l) for l in lams])
    out = []
    for gi, (Nx, Ny) in enumerate(((256, 256), (384, 512), (512, 512))):
        d = random_pixel_device(Nx, Ny, 20, 1.55, [3, gi])
        d.src[:] = 0
        rng = np.random.Generator(np.random.PCG64([30, gi]))
        ix = rng.integers(21, int(0.2 * Nx) + 1, size=16)
        iy = rng.integers(21, Ny - 20 + 1, size=16)
        srcs = np.zeros((Nx, Ny, 16), dtype=np.complex128)
        for r in range(16):
            srcs[ix[r] - 1, iy[r] - 1, r] = 1j
        out.append((d, oms, srcs))
    return out


def config_c4(Nx: int = 1024, Ny: int = 768, npml: int = 20, design: int = 256,
              seed: int = 4, obs_from_pml: int = 50, waist_um: float = 0.25):
    """Adjoint-gradient workload: Nx x Ny, 20-cell PML, centred design region
    of design x design cells (1024 x 768: ix 385:640, iy 257:512, 1-based)
    with eps_r ~ U[1, 12.25] per cell (seed 4, drawn x fastest), 15-cell
    guides through the centre row left and right of it, the 25-cell line
    source of config_c2 at ix = 51, lambda = 1.55 um.  The objective is
    J = |sum_iy Ez[ix_obs, iy] conj(E_mode[iy])|^2 on the column
    ix_obs = Nx - npml - 50 (1-based; 954 at Nx = 1024, 1 um before the right
    PML) with E_mode(y) = exp(-((y - y_guide)/0.25 um)^2) at the cell centres
    and y_guide the centre of the guide cells.  The 32 finite-difference
    cells are rng(seed=40).choice(design^2, 32) over the design cells, x
    fastest.

    Returns (Device, m[N] (J = |m^T e|^2), design_cells[design^2] (0-based
    linear indices, x fastest), fd_cells[32] (linear indices))."""
    d = random_pixel_device(Nx, Ny, npml, 1.55, 0, design=8)
    d.eps_r[:] = 1.0
    guide = 15
    x0, y0 = (Nx - design) // 2, (Ny - design) // 2
    rng = np.random.Generator(np.random.PCG64(seed))
    vals = rng.uniform(1.0, 12.25, design * design).reshape((design, design), order="F")
    d.eps_r[x0:x0 + design, y0:y0 + design] = vals
    yg0 = Ny // 2 - guide // 2
    d.eps_r[:x0, yg0:yg0 + guide] = 12.25
    d.eps_r[x0 + design:, yg0:yg0 + guide] = 12.25
    g = d.grid
    yc = g.yc()
    y_guide = float(np.mean(yc[yg0:yg0 + guide]))
    emode = np.exp(-((yc - y_guide) / waist_um) ** 2)
    ix_obs = Nx - npml - obs_from_pml - 1  # 0-based
    mfield = np.zeros((Nx, Ny), dtype=np.complex128)
    mfield[ix_obs, :] = np.conj(emode)
    m = mfield.ravel(order="F")
    X, Y = np.meshgrid(np.arange(x0, x0 + design), np.arange(y0, y0 + design), indexing="ij")
    cells = (X + Y * Nx).ravel(order="F")
    pick = np.random.Generator(np.random.PCG64(40)).choice(design * design, 32, replace=False)
    return d, m, cells, cells[pick]


def config_c5(n: int = 768, npml: int = 20, a: float = 0.5, r_over_a: float = 0.2,
              eps_rod: float = 8.9, dl: float = 0.02):
    """Eigenfrequency workload: n x n cells (dl 0.02 um, 20-cell PML, domain
    centred on the origin), a triangular lattice (a = 0.5 um) of dielectric
    rods eps_r = 8.9, radius 0.2a, centres (a(i + j/2), a j sqrt(3)/2) for
    all integers i, j, with the rod at the origin removed.  A cell belongs to
    a rod when its centre satisfies (xc - x0)^2 + (yc - y0)^2 <= r^2
    (device.jl:63-67 evaluates regions at cell centres).  TM, omega0 =
    2 pi c0 / 1.55 um; eigenfrequency nev = 8 nearest, ncv = 20.
    Returns the Device (omega = omega0)."""
    half = n * dl / 2
    g = Grid.create(dl, (npml, npml), (-half, half), (-half, half))
    assert g.N == (n, n), g.N
    d = Device(g, omega_of_wavelength(1.55))
    xc, yc = g.xc(), g.yc()
    X, Y = np.meshgrid(xc, yc, indexing="ij")
    r = r_over_a * a
    eps = np.ones((n, n))
    jmax = int(math.ceil(half / (a * math.sqrt(3) / 2))) + 1
    for j in range(-jmax, jmax + 1):
        y0 = a * j * math.sqrt(3) / 2
        imax = int(math.ceil(half / a + abs(j) / 2)) + 1
        for i in range(-imax, imax + 1):
            x0 = a * (i + j / 2)
            if i == 0 and j == 0:
                continue
            if abs(x0) > half + r or abs(y0) > half + r:
                continue
            eps[(X - x0) ** 2 + (Y - y0) ** 2 <= r * r] = eps_rod
    d.eps_r[:] = eps
    return d
