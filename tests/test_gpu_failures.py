"""Error paths and less common options on the GPU, through the C ABI.

Failures are forced from the host with fdfd_debug_inject (one-shot), so every
kernel still runs normally: the windowed-pivot fallback (pass 0 rejected, full
search redone from the operator), a zero pivot reported on a chosen line, an
FGMRES refinement cut short, and a failed factor allocation.  After each the
handle keeps working.  Also: argument validation of fdfd_eigenfrequency, and
which = LR (Arpack's largest real part of 1/(lambda - sigma), the mode the
reference notebook's ring-resonator example uses) against ARPACK."""
import ctypes as C

import numpy as np
import pytest

import fdfd_oracle as O
from fdfd import _lib, core
from fdfd._lib import (FDFD_INJECT_NOCONV, FDFD_INJECT_OOM, FDFD_INJECT_PIVOT_FALLBACK,
                       FDFD_INJECT_SINGULAR, FDFDError)

pytestmark = pytest.mark.gpu

EPS0n, MU0n, C0n = O.normalize_parameters()


def _cgrid(g):
    return core.Grid(g.dx, g.Npml, (g.bounds[0][0], g.bounds[1][0]),
                     (g.bounds[0][1], g.bounds[1][1]), g.L0)


@pytest.fixture(scope="module")
def small():
    d = O.random_pixel_device(96, 80, 10, 1.55, 7)
    return d, float(d.omega[0]), O.solve(d, O.TM)


def _setup(engine, d):
    engine.set_grid(_cgrid(d.grid))
    engine.set_eps(d.eps_r)


def _ok_solve(engine, d, om, ref, prec="fp64"):
    fid, st = engine.factor(O.TM, om, prec)
    try:
        f, sst = engine.solve_fields(fid, d.src, tol=1e-12)
    finally:
        engine.release(fid)
    err = np.linalg.norm(f[:, :, :, 0] - ref.data) / np.linalg.norm(ref.data)
    assert err <= 1e-10, err
    return st


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_injected_pivot_fallback_redoes_factorization(engine, small, precision):
    """Pass 0 (candidate-window pivots) is rejected, so the whole reduction is
    redone with the full-column search from the operator's coefficients: the
    factor reports pivot_fallback = 1 and solves to the oracle's field."""
    d, om, ref = small
    _setup(engine, d)
    engine.debug_inject(FDFD_INJECT_PIVOT_FALLBACK)
    st = _ok_solve(engine, d, om, ref, precision)
    assert st["pivot_fallback"] == 1
    st = _ok_solve(engine, d, om, ref, precision)   # one-shot
    assert st["pivot_fallback"] == 0


def test_injected_fallback_rebuilds_one_factor_of_a_batch(engine, small):
    """In a batched build only the factor that failed its window verification
    is redone (nested full-search build swapped in); all results still match
    the oracle and the batch reports the fallback."""
    d, om, ref = small
    oms = np.array([O.omega_of_wavelength(l) for l in (1.50, 1.55, 1.60)])
    srcs = d.src[:, :, None]
    engine.debug_inject(FDFD_INJECT_PIVOT_FALLBACK)
    res, st = engine.batch_solve([(_cgrid(d.grid), O.TM, d.eps_r, oms, srcs)],
                                 precision="mixed", tol=1e-12)
    assert st["pivot_fallback"] == 1 and st["residual"] <= 1e-12
    A0 = None
    for o, w in enumerate(oms):
        A0 = O.assemble(d.grid, d.eps_r, float(w), O.TM)
        x, _ = O.linear_solve(A0, O.rhs(float(w), d.src))
        ez = res[0][o, 0, :, :, 0].ravel(order="F")
        assert np.linalg.norm(ez - x) / np.linalg.norm(x) <= 1e-10, o


def test_injected_singular_pivot(engine, small):
    d, om, ref = small
    _setup(engine, d)
    engine.debug_inject(FDFD_INJECT_SINGULAR, 7)
    with pytest.raises(FDFDError) as ei:
        engine.factor(O.TM, om, "fp64")
    assert ei.value.code == -3 and "line block 7" in str(ei.value)
    _ok_solve(engine, d, om, ref)


def test_injected_refinement_failure(engine, small):
    """Refinement cut to one FGMRES iteration with complex64 factors cannot
    reach 1e-12: FDFD_ERR_NOCONV with the residual reached in the message."""
    d, om, ref = small
    _setup(engine, d)
    fid, _ = engine.factor(O.TM, om, "mixed")
    try:
        engine.debug_inject(FDFD_INJECT_NOCONV, 1)
        with pytest.raises(FDFDError) as ei:
            engine.solve_fields(fid, d.src, tol=1e-12)
        assert ei.value.code == -4 and "did not reach the tolerance" in str(ei.value)
        f, st = engine.solve_fields(fid, d.src, tol=1e-12)
        assert st["residual"] <= 1e-12
    finally:
        engine.release(fid)


def test_injected_allocation_failure(engine, small):
    d, om, ref = small
    _setup(engine, d)
    engine.debug_inject(FDFD_INJECT_OOM)
    with pytest.raises(FDFDError) as ei:
        engine.factor(O.TM, om, "mixed")
    assert ei.value.code == -2
    _ok_solve(engine, d, om, ref, "mixed")


def test_inject_rejects_unknown_kind(engine):
    with pytest.raises(FDFDError) as ei:
        engine.debug_inject(99)
    assert ei.value.code == -1


def test_eigenfrequency_validates_precision_and_which(engine, small):
    d, om, _ = small
    _setup(engine, d)
    lib, h = engine.lib, engine.h
    w = np.empty(2, dtype=np.complex128)
    st = _lib.fdfd_stats()
    args = lambda which, prec: (h, O.TM, om, 2, 0, which, prec, 1e-10, 10,  # noqa: E731
                                w.ctypes.data_as(C.c_void_p), None, C.byref(st))
    assert lib.fdfd_eigenfrequency(*args(0, 5)) == -1
    assert b"precision" in lib.fdfd_last_error(h)
    assert lib.fdfd_eigenfrequency(*args(7, 0)) == -1
    assert b"which" in lib.fdfd_last_error(h)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_which_lr_ring_resonator_matches_arpack(engine, precision):
    """The reference notebook's ring resonator (Example_simulations.ipynb cell
    31): Grid(0.01, [15 15], [-2 2], [-2 2]) (400 x 400), omega0 = 2 pi 200e12,
    eps 12.25 for 0.8 < r <= 1.0 (the inner cylinder, listed first, wins
    inside r <= 0.8), TE, nev = 4, which = :LR.  Eigenvalues against the
    oracle's ARPACK with which='LR' and the same shift: 1e-9, in LR order."""
    g = O.Grid.create(0.01, (15, 15), (-2.0, 2.0), (-2.0, 2.0))
    d = O.Device(g, 2 * np.pi * 200e12)
    X, Y = np.meshgrid(g.xc(), g.yc(), indexing="ij")
    r = np.hypot(X, Y)
    d.eps_r[(r > 0.8) & (r <= 1.0)] = 12.25
    engine.set_grid(_cgrid(g))
    engine.set_eps(d.eps_r)
    wo, _, lam_o, sigma = O.eigenfrequency(d, O.TE, 4, which="LR")
    w, fl, st = engine.eigenfrequency(O.TE, float(d.omega[0]), 4, which="LR",
                                      precision=precision, tol=1e-12)
    print(f"ring LR {precision}: gpu {w} arpack {wo}", flush=True)
    theta = np.real(1 / (-w ** 2 * MU0n - sigma))
    assert np.all(np.diff(theta) <= 1e-12 * np.abs(theta).max())   # LR order
    for k in range(4):
        assert np.min(np.abs(wo - w[k])) / abs(w[k]) <= 1e-9
