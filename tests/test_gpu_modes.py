"""GPU parity of the mode-source path (fdfd_eigenmode through the C ABI)
against the oracle's restatement of eigenmode / get_modes / setup_mode!
(src/solver/eigen.jl:6-29, src/device.jl:118-149) and of the driven solve
with modes (src/solver/driven.jl:15-19).

Bar (floating point): propagation constants to 1e-10 relative; eigenvectors
(unit norm, largest entry real positive on both sides) to 1e-8 where the
eigenvalue is isolated; mode-source profiles to 1e-9; driven fields with
mode sources to 1e-10 relative.
"""
import numpy as np
import pytest

import fdfd_oracle as O
from fdfd import core, modes
from fdfd._lib import FDFDError
from fdfd.solve import solve as gpu_solve

pytestmark = pytest.mark.gpu

BETA_TOL = 1e-10
VEC_TOL = 1e-8
OM = O.omega_of_wavelength(1.55)


def _slab(n, core_cells, ec, ecl, loss=0.0, seed=None):
    e = np.full(n, ecl, dtype=np.complex128)
    c0 = (n - core_cells) // 2
    e[c0:c0 + core_cells] = ec
    if loss:
        e += 1j * loss
    if seed is not None:
        e += 0.05 * np.random.default_rng(seed).random(n)
    return e


def _compare(beta, vecs, eps, dh, pol, neff, nev):
    ob, ov, _ = O.eigenmode_1d(eps, dh, OM, pol, neff, nev)
    rel = np.abs(beta - ob) / np.abs(ob)
    print(f"  beta gpu {beta} oracle {ob} rel {rel.max():.2e}", flush=True)
    assert rel.max() <= BETA_TOL
    lam = ob ** 2
    full = np.linalg.eigvals(O.mode_operator_1d(eps, dh, OM, pol).toarray())
    for k in range(nev):
        # eigenvectors are compared only where the eigenvalue is isolated in
        # the whole spectrum (relative gap >= 1e-4)
        gap = np.sort(np.abs(full - lam[k]))[1] / abs(lam[k])
        if gap < 1e-4:
            continue
        err = np.abs(vecs[:, k] - ov[:, k]).max()
        print(f"  vector {k}: max err {err:.2e} (gap {gap:.1e})", flush=True)
        assert err <= VEC_TOL


@pytest.mark.parametrize("pol", [O.TM, O.TE])
@pytest.mark.parametrize("n,core_cells,nev", [(41, 15, 1), (101, 15, 3), (315, 15, 2), (64, 20, 4)])
def test_eigenmode_matches_oracle(engine, pol, n, core_cells, nev):
    eps = _slab(n, core_cells, 12.25, 2.25, loss=0.01, seed=n)
    neff = 3.2
    print(f"eigenmode pol={pol} n={n} nev={nev}", flush=True)
    beta, vecs = modes.eigenmode(eps, 0.02, pol, OM, neff, nev, eng=engine)
    _compare(beta, vecs, eps, 0.02, pol, neff, nev)


@pytest.mark.parametrize("pol", [O.TM, O.TE])
def test_eigenmode_uniform_closed_form(engine, pol):
    """Uniform eps: beta^2 = k0^2 eps - (4/dx^2) sin^2(pi p/n); with the shift
    at p = 0 (neff = sqrt(eps)) the nearest eigenvalue is k0^2 eps exactly
    and its eigenvector is constant."""
    n, eps = 33, 4.0 + 0.2j
    beta, vecs = modes.eigenmode(np.full(n, eps), 0.02, pol, OM, np.sqrt(eps), 1, eng=engine)
    k2 = OM ** 2 * O.EPS0 * O.DEFAULT_L0 * O.MU0 * O.DEFAULT_L0
    assert abs(beta[0] ** 2 - k2 * eps) <= BETA_TOL * abs(k2 * eps)
    assert np.abs(vecs[:, 0] - 1 / np.sqrt(n)).max() <= VEC_TOL


def test_eigenmode_batch_equals_single_calls(engine):
    """Several slices of different lengths, polarizations and frequencies in
    one launch give the same results as one call each."""
    slices = [_slab(41, 15, 12.25, 2.25), _slab(77, 9, 12.0, 1.0, seed=3),
              _slab(200, 30, 8.0, 2.0, loss=0.02)]
    pols = [O.TM, O.TE, O.TM]
    oms = [OM, OM * 1.03, OM * 0.97]
    neffs = [3.0, 2.8, 2.5 + 0.01j]
    beta, vecs, st = modes.eigenmode_batch(slices, 0.02, pols, oms, neffs, 2, eng=engine)
    print(f"batch stats {st}", flush=True)
    for p in range(3):
        b1, v1 = modes.eigenmode(slices[p], 0.02, pols[p], oms[p], neffs[p], 2, eng=engine)
        assert np.abs(beta[p] - b1).max() <= 1e-13 * np.abs(b1).max()
        assert np.abs(vecs[p] - v1).max() <= 1e-12


def test_eigenmode_invalid_arguments(engine):
    with pytest.raises(FDFDError):
        modes.eigenmode(np.ones(2), 0.02, O.TM, OM, 1.0, 1, eng=engine)  # n < 3
    with pytest.raises(FDFDError):
        modes.eigenmode(np.ones(5), 0.02, O.TM, OM, 1.0, 5, eng=engine)  # nev >= n
    with pytest.raises(FDFDError):
        modes.eigenmode(np.ones(9), 0.02, 3, OM, 1.0, 1, eng=engine)  # pol


def _wg_devices():
    """The waveguide case of the reference's tests (runtests.jl:23-27):
    Grid(0.02, [15 15], [0 10], [-1 1]), a guide of eps 12 through the
    whole length, Mode(TM, x, 3.5, Point(1.0, 0), 0.8), f = 200 THz.  The
    guide is painted on rows 44..58 (1-based; |y| <= 0.15 falls on cell-centre
    ties, so integer rows are used)."""
    g = core.Grid(0.02, (15, 15), (0.0, 10.0), (-1.0, 1.0))
    om = 2 * np.pi * 200e12
    d = core.Device(g, om)
    d.eps_r[:, 43:58] = 12.0
    modes.add_mode(d, modes.Mode(O.TM, core.XHAT, 3.5, core.Point(1.0, 0.0), 0.8))
    o = O.Device(O.Grid.create(0.02, (15, 15), (0.0, 10.0), (-1.0, 1.0)), om)
    o.eps_r[:, 43:58] = 12.0
    O.add_mode(o, O.Mode(O.TM, O.XHAT, 3.5, (1.0, 0.0), 0.8))
    return d, o


def test_setup_mode_matches_oracle(engine):
    d, o = _wg_devices()
    om = float(d.omega[0])
    modes.setup_mode(d, O.TM, om, 3.5, core.Point(1.0, 0.0), core.XHAT, 0.8, eng=engine)
    O.setup_mode(o, O.TM, om, 3.5, (1.0, 0.0), O.XHAT, 0.8)
    err = np.abs(d.src - o.src).max()
    print(f"setup_mode profile err {err:.2e}", flush=True)
    assert err <= 1e-9
    assert abs(np.linalg.norm(d.src) - 1) <= 1e-12


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_solve_with_mode_source_matches_oracle(engine, precision):
    """driven.jl:15-19: the source is rebuilt from the TM mode at every
    frequency, then solved; two frequencies."""
    d, o = _wg_devices()
    d.omega = np.array([2 * np.pi * 200e12, 2 * np.pi * 190e12])
    o.omega = d.omega.copy()
    f_gpu = gpu_solve(d, O.TM, precision=precision, tol=1e-13, eng=engine)
    f_or = O.solve(o, O.TM)
    for fg, fo in zip(f_gpu, f_or):
        err = np.linalg.norm(fg.data - fo.data) / np.linalg.norm(fo.data)
        print(f"mode-source solve {precision}: field err {err:.2e}", flush=True)
        assert err <= 1e-10
    assert np.abs(d.src - o.src).max() <= 1e-9  # the last frequency's source
