"""GPU parity of the Poynting / flux kernels (fdfd_poynting, fdfd_flux,
fdfd_flux_d through the C ABI) and of the S-parameters (mode overlaps on GPU
eigenmodes) against the oracle's restatement of src/flux.jl.

Bar (floating point, FP64 reductions in a different order): Poynting maps to
1e-14 relative to their largest entry, fluxes to 1e-12 relative, S to 1e-9.
"""
import numpy as np
import pytest

import fdfd_oracle as O
from fdfd import core, flux, modes
from fdfd.solve import solve as gpu_solve

pytestmark = pytest.mark.gpu


def _fields(pol, seed=4, Nx=50, Ny=40):
    rng = np.random.default_rng(seed)
    g = core.Grid(0.02, (0, 0), (0, Nx * 0.02), (0, Ny * 0.02))
    og = O.Grid.create(0.02, (0, 0), (0, Nx * 0.02), (0, Ny * 0.02))
    data = rng.standard_normal((Nx, Ny, 3)) + 1j * rng.standard_normal((Nx, Ny, 3))
    if pol == O.TM:
        return core.FieldTM(g, 1.0, data), O.FieldTM(og, 1.0, data)
    return core.FieldTE(g, 1.0, data), O.FieldTE(og, 1.0, data)


@pytest.mark.parametrize("pol", [O.TM, O.TE])
def test_poynting_matches_oracle(engine, pol):
    f, fo = _fields(pol)
    sx, sy = flux.poynting(f, eng=engine)
    ox, oy = O.poynting(fo)
    for a, b in ((sx, ox), (sy, oy)):
        assert np.abs(a - b).max() <= 1e-14 * np.abs(b).max()


@pytest.mark.parametrize("pol", [O.TM, O.TE])
def test_flux_planes_match_oracle(engine, pol):
    f, fo = _fields(pol)
    planes = [((0.31, 0.4), 0.2, core.XHAT), ((0.55, 0.21), 0.1, core.XHAT),
              ((0.5, 0.41), 0.3, core.YHAT), ((0.011, 0.75), 0.05, core.YHAT),
              ((0.97, 0.4), 0.5, core.XHAT)]
    got = flux.flux_surface_integrals(f, [(core.Point(*c), w, n) for c, w, n in planes], eng=engine)
    ref = np.array([O.flux_surface_integral(fo, c[0], c[1], w, n) for c, w, n in planes])
    print(f"flux pol={pol}: {got} vs {ref}", flush=True)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_flux_device_pointer_entry(engine):
    """fdfd_flux_d on a device-resident field array gives fdfd_flux's values."""
    import ctypes as C
    from fdfd import _lib
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    f, _ = _fields(O.TM, seed=9)
    g = f.grid
    flat = flux._fields_flat(f)
    dev = C.c_void_p()
    assert hip.hipMalloc(C.byref(dev), flat.nbytes) == 0
    assert hip.hipMemcpy(dev, flat.ctypes.data_as(C.c_void_p), flat.nbytes, 1) == 0
    planes = [flux.flux_plane(g, core.Point(0.31, 0.4), 0.2, core.XHAT),
              flux.flux_plane(g, core.Point(0.5, 0.41), 0.3, core.YHAT)]
    arr = (_lib.fdfd_flux_plane * 2)(*[_lib.fdfd_flux_plane(*p) for p in planes])
    out = np.empty(2)
    try:
        engine._chk(engine.lib.fdfd_flux_d(engine.h, 1, g.N[0], g.N[1], g.dx, g.dy, dev, 2, arr,
                                           out.ctypes.data_as(C.POINTER(C.c_double))))
    finally:
        hip.hipFree(dev)
    ref = flux.flux_surface_integrals(f, planes, eng=engine)
    assert np.abs(out - ref).max() <= 1e-14 * np.abs(ref).max()


def test_scattering_parameters_match_oracle_and_propagate(engine):
    """A straight guide driven by its mode (solved on the GPU): S from GPU
    eigenmodes matches the oracle's S on the oracle's fields, and along the
    guide |S| is constant and its phase advances by -k D with the lattice
    wavenumber k = arccos(1 - b^2 dx^2/2)/dx."""
    g = core.Grid(0.02, (15, 15), (0.0, 6.0), (-1.0, 1.0))
    d = core.Device(g, O.omega_of_wavelength(1.55))
    d.eps_r[:, 43:58] = 12.25
    w = 1.6
    modes.add_mode(d, modes.Mode(O.TM, core.XHAT, 3.2, core.Point(1.0, 0.0), w))
    f = gpu_solve(d, O.TM, tol=1e-13, eng=engine)
    og = O.Grid.create(0.02, (15, 15), (0.0, 6.0), (-1.0, 1.0))
    o = O.Device(og, d.omega)
    o.eps_r[:] = d.eps_r
    O.add_mode(o, O.Mode(O.TM, O.XHAT, 3.2, (1.0, 0.0), w))
    fo = O.solve(o, O.TM)
    for x in (3.0, 4.0):
        modes.add_mode(d, modes.Mode(O.TM, core.XHAT, 3.2, core.Point(x, 0.0), w))
        O.add_mode(o, O.Mode(O.TM, O.XHAT, 3.2, (x, 0.0), w))
    S = flux.scattering_parameters([f], d, eng=engine)
    So = O.scattering_parameters([fo], o)
    print(f"S gpu {S} oracle {So}", flush=True)
    assert np.abs(S - So).max() <= 1e-9 * np.abs(So).max()
    beta, _ = modes.eigenmode(d.eps_r[150, 10:91], g.dy, O.TM, float(d.omega[0]), 3.2, 1, eng=engine)
    k = np.arccos(1 - (beta[0].real * g.dx) ** 2 / 2) / g.dx
    D = (g.x2ind(4.0) - g.x2ind(3.0)) * g.dx
    assert abs(abs(S[0, 2]) / abs(S[0, 1]) - 1) < 1e-5
    assert abs(np.angle(np.exp(1j * (np.angle(S[0, 2] / S[0, 1]) + k * D)))) < 1e-5
