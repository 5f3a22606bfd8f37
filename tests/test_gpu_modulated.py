"""GPU parity of the modulated (MF-FDFD) solve, fdfd_modulated_solve through
the C ABI, against the oracle's restatement of solve(d::ModulatedDevice)
(src/solver/modulation.jl:35-119): the (2K+1) N coupled system assembled as
the reference does and solved by refined sparse LU.

Bar (floating point): every sideband's fields within 1e-10 of the oracle's,
relative to the largest sideband norm, at a requested residual of 1e-13; the
reported residual at or below the request.
"""
import numpy as np
import pytest

import fdfd_oracle as O
from fdfd import core, modes
from fdfd._lib import FDFDError
from fdfd.modulated import ModulatedDevice, setup_deps, solve_modulated

pytestmark = pytest.mark.gpu

FIELD_TOL = 1e-10


def _pair(Nx, Ny, npml, K, sharedpml, seed):
    rng = np.random.default_rng(seed)
    dl = 0.02
    g = core.Grid(dl, (npml, npml), (0.0, Nx * dl), (0.0, Ny * dl))
    og = O.Grid.create(dl, (npml, npml), (0.0, Nx * dl), (0.0, Ny * dl))
    om = O.omega_of_wavelength(1.55)
    Om = 0.06 * om
    d = ModulatedDevice(g, om, Om, K, sharedpml)
    o = O.ModulatedDevice(og, om, Om, K, sharedpml)
    eps = 1 + 11.25 * rng.random(g.N)
    deps = np.zeros(g.N, dtype=complex)
    reg = (slice(Nx // 4, 3 * Nx // 4), slice(Ny // 3, 2 * Ny // 3))
    shape = deps[reg].shape
    deps[reg] = 0.8 * rng.random(shape) * np.exp(2j * np.pi * rng.random(shape))
    src = np.zeros(g.N, dtype=complex)
    src[npml + 3, Ny // 2 - 3:Ny // 2 + 4] = 1.0
    for x in (d, o):
        x.eps_r[:] = eps
        x.deps_r[:] = deps
        x.src[:] = src
    return d, o


def _check(fg, fo, tag):
    scale = max(np.linalg.norm(f.data) for f in fo)
    errs = [np.linalg.norm(a.data - b.data) / scale for a, b in zip(fg, fo)]
    print(f"{tag}: sideband errors {['%.1e' % e for e in errs]}", flush=True)
    assert max(errs) <= FIELD_TOL
    for a, b in zip(fg, fo):
        assert abs(a.omega - b.omega) <= 1e-15 * abs(b.omega)


@pytest.mark.parametrize("K", [0, 1, 2])
@pytest.mark.parametrize("sharedpml", [True, False])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_modulated_matches_oracle(engine, K, sharedpml, precision):
    d, o = _pair(48, 40, 4, K, sharedpml, seed=10 + K)
    fg, st = solve_modulated(d, precision=precision, tol=1e-13, eng=engine, return_stats=True)
    fo = O.solve_modulated(o)
    print(f"K={K} shared={sharedpml} {precision}: its={st[0]['iterations']} "
          f"res={st[0]['residual']:.1e}", flush=True)
    assert st[0]["residual"] <= 1e-13
    _check(fg[0], fo[0], f"K={K} shared={sharedpml} {precision}")


@pytest.fixture(scope="module")
def modwg():
    """The modulated waveguide of the reference's tests (runtests.jl:40-62):
    Grid(0.01, [15 15], [0 5], [-1 1]) (500 x 200), carrier 2 pi 1.939e14,
    Omega = 4.541e14, one sideband each side, guide eps 12.25 for |y| <= a/2,
    deps = exp(i q x) on 1 <= x <= 4, -a/2 <= y <= 0, a TM mode source
    Mode(TM, x, 3.5, [0.2, 0], 4a).  (No cell centre lies on a region edge.)"""
    a, q, L = 0.2202, 2.9263, 5.0
    g = core.Grid(0.01, (15, 15), (0.0, L), (-1.0, 1.0))
    og = O.Grid.create(0.01, (15, 15), (0.0, L), (-1.0, 1.0))
    d = ModulatedDevice(g, 2 * np.pi * 1.939e14, 4.541e14, 1)
    o = O.ModulatedDevice(og, 2 * np.pi * 1.939e14, 4.541e14, 1)
    wg = lambda x, y: -a / 2 <= y <= a / 2  # noqa: E731
    mod = lambda x, y: 1 <= x <= L - 1 and -a / 2 <= y <= 0  # noqa: E731
    core.setup_eps(d, wg, 12.25)
    setup_deps(d, mod, lambda x, y: np.exp(1j * q * x))
    o.eps_r[:] = d.eps_r
    o.deps_r[:] = d.deps_r
    modes.add_mode(d, modes.Mode(O.TM, core.XHAT, 3.5, core.Point(0.2, 0.0), 4 * a))
    O.add_mode(o, O.Mode(O.TM, O.XHAT, 3.5, (0.2, 0.0), 4 * a))
    fo = O.solve_modulated(o)
    return d, o, fo


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_reference_modulated_waveguide(engine, modwg, precision):
    d, o, fo = modwg
    fg, st = solve_modulated(d, precision=precision, tol=1e-13, eng=engine, return_stats=True)
    print(f"modwg {precision}: its={st[0]['iterations']} res={st[0]['residual']:.1e} "
          f"factor {st[0]['t_factor_ms']:.1f} ms total {st[0]['t_solve_ms']:.1f} ms", flush=True)
    assert np.abs(d.src - o.src).max() <= 1e-9  # mode source rebuilt identically
    _check(fg[0], fo[0], f"modwg {precision}")
    # the modulation moves power into the sidebands
    assert np.linalg.norm(fg[0][0].data) > 1e-3 * np.linalg.norm(fg[0][1].data)


def test_modulated_invalid_arguments(engine):
    d, _ = _pair(24, 20, 2, 1, True, seed=1)
    d.Omega = 2 * float(d.omega[0])  # w - Omega < 0
    with pytest.raises(FDFDError):
        solve_modulated(d, eng=engine)
    d.Omega = 1e13
    d.nsidebands = 17
    with pytest.raises(FDFDError):
        solve_modulated(d, eng=engine)
