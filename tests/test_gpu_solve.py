"""GPU parity of the driven-solve path against the CPU oracle.

Bar (floating point): the field error relative to the oracle's FP64 solution
is <= 1e-10 when both are solved to a relative residual far below it, the
operator apply agrees with the assembled reference matrix to <= 1e-13, and
every reported residual is <= the requested tolerance.
"""
import numpy as np
import pytest

import fdfd_oracle as O
from fdfd import core
from fdfd.solve import Engine

pytestmark = pytest.mark.gpu

APPLY_TOL = 1e-13
FIELD_TOL = 1e-10


def _device(Nx, Ny, npml, seed, lam=1.55, dl=0.02, complex_eps=False):
    rng = np.random.default_rng(seed)
    g = O.Grid.create(dl, (npml, npml), (0.0, Nx * dl), (0.0, Ny * dl))
    assert g.N == (Nx, Ny)
    d = O.Device(g, O.omega_of_wavelength(lam))
    d.eps_r[:] = 1 + 11.25 * rng.random((Nx, Ny))
    if complex_eps:
        d.eps_r[:] += 0.3j * rng.random((Nx, Ny))
    d.src[:] = rng.standard_normal((Nx, Ny)) + 1j * rng.standard_normal((Nx, Ny))
    return d


def _engine_grid(eng, d):
    g = d.grid
    eng.set_grid(core.Grid(0.02, g.Npml, (g.bounds[0][0], g.bounds[1][0]),
                           (g.bounds[0][1], g.bounds[1][1]), g.L0))
    assert eng.grid.N == g.N
    eng.set_eps(d.eps_r)


CASES = [(24, 20, 4), (20, 24, 4), (16, 16, 0), (13, 9, 3), (9, 13, 2), (40, 8, 4)]


@pytest.mark.parametrize("Nx,Ny,npml", CASES)
@pytest.mark.parametrize("pol", [O.TM, O.TE])
def test_apply_matches_reference_matrix(engine, Nx, Ny, npml, pol):
    d = _device(Nx, Ny, npml, seed=Nx * 100 + Ny, complex_eps=True)
    _engine_grid(engine, d)
    om = float(d.omega[0])
    fid, _ = engine.factor(pol, om, "fp64")
    try:
        A = O.assemble(d.grid, d.eps_r, om, pol)
        x = O.rhs(om, d.src)
        for tr in (False, True):
            y = engine.apply(fid, x, transpose=tr)
            ref = (A.T if tr else A) @ x
            err = np.linalg.norm(y - ref) / np.linalg.norm(ref)
            print(f"apply {Nx}x{Ny} npml={npml} pol={pol} tr={tr} rel={err:.2e}", flush=True)
            assert err <= APPLY_TOL
    finally:
        engine.release(fid)


@pytest.mark.parametrize("Nx,Ny,npml", CASES)
@pytest.mark.parametrize("pol", [O.TM, O.TE])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_linsolve_matches_oracle(engine, Nx, Ny, npml, pol, precision):
    d = _device(Nx, Ny, npml, seed=7 + Nx + Ny)
    _engine_grid(engine, d)
    om = float(d.omega[0])
    A = O.assemble(d.grid, d.eps_r, om, pol)
    b = O.rhs(om, d.src)
    xo, _ = O.linear_solve(A, b)
    xt, _ = O.linear_solve(A, b, trans="T")
    fid, fst = engine.factor(pol, om, precision)
    try:
        for tr, ref in ((False, xo), (True, xt)):
            x, st = engine.linsolve(fid, b, transpose=tr, tol=1e-13)
            err = np.linalg.norm(x - ref) / np.linalg.norm(ref)
            Aop = A.T if tr else A
            res = np.linalg.norm(b - Aop @ x) / np.linalg.norm(b)
            print(f"linsolve {Nx}x{Ny} pol={pol} {precision} tr={tr} err={err:.2e} "
                  f"res={res:.2e} its={st['iterations']}", flush=True)
            assert res <= 1e-12
            assert err <= FIELD_TOL
    finally:
        engine.release(fid)


@pytest.mark.parametrize("gemm", ["fp32", "fp16x3"])
@pytest.mark.parametrize("Nx,Ny,npml", [(96, 80, 8), (72, 100, 6)])
def test_mixed_gemm_modes_match_oracle(monkeypatch, gemm, Nx, Ny, npml):
    """Both GEMM kernels of the mixed path (complex64 factors with FP32 MFMA
    products, or fp16x3 products) give the oracle's solution through the FP64
    refinement, on grids deep enough for several dense reduction levels."""
    monkeypatch.setenv("FDFD_MIXED_GEMM", gemm)  # read when the handle is created
    eng = Engine(0)
    try:
        d = _device(Nx, Ny, npml, seed=3 + Nx)
        _engine_grid(eng, d)
        om = float(d.omega[0])
        A = O.assemble(d.grid, d.eps_r, om, O.TM)
        b = O.rhs(om, d.src)
        xo, _ = O.linear_solve(A, b)
        fid, _ = eng.factor(O.TM, om, "mixed")
        x, st = eng.linsolve(fid, b, tol=1e-13)
        eng.release(fid)
        err = np.linalg.norm(x - xo) / np.linalg.norm(xo)
        res = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
        print(f"mixed gemm={gemm} {Nx}x{Ny}: err={err:.2e} res={res:.2e} its={st['iterations']}",
              flush=True)
        assert res <= 1e-12
        assert err <= FIELD_TOL
    finally:
        eng.close()


@pytest.mark.parametrize("Nx,Ny,npml", [(37, 90, 6), (90, 50, 6)])
def test_fp16x3_gemm_tracks_fp32_gemm(monkeypatch, Nx, Ny, npml):
    """On line lengths that are not multiples of the 16 x 32 operand tiles
    (37, 50), the fp16x3 products stay close enough to the FP32 ones that the
    refinement needs about as many iterations: a wrong or stale operand tile
    would leave FGMRES converging, but only with a much weaker preconditioner."""
    d = _device(Nx, Ny, npml, seed=5 + Ny)
    om = float(d.omega[0])
    A = O.assemble(d.grid, d.eps_r, om, O.TM)
    b = O.rhs(om, d.src)
    its = {}
    for gemm in ("fp32", "fp16x3"):
        monkeypatch.setenv("FDFD_MIXED_GEMM", gemm)  # read when the handle is created
        eng = Engine(0)
        try:
            _engine_grid(eng, d)
            fid, _ = eng.factor(O.TM, om, "mixed")
            x, st = eng.linsolve(fid, b, tol=1e-13)
            eng.release(fid)
        finally:
            eng.close()
        res = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
        its[gemm] = st["iterations"]
        print(f"{Nx}x{Ny} gemm={gemm}: res={res:.2e} its={its[gemm]}", flush=True)
        assert res <= 1e-12
    assert its["fp16x3"] <= its["fp32"] + 4


@pytest.mark.parametrize("Nx,Ny,npml", [(100, 84, 6), (76, 120, 6), (300, 264, 8)])
def test_fp16x3_trailing_tracks_fp32_trailing(monkeypatch, Nx, Ny, npml):
    """The Gauss-Jordan trailing update on fp16x3 products (line lengths
    n % 4 == 0; here 84, 76 and 264, not multiples of the 64 x 64 tiles, with
    a last panel pair shorter than 64 columns; at 264 a 64-row strip spans
    two workgroups of 4 and 1 tiles) needs about as many refinement
    iterations as the FP32 one and reaches the same residual."""
    d = _device(Nx, Ny, npml, seed=7 + Nx)
    om = float(d.omega[0])
    A = O.assemble(d.grid, d.eps_r, om, O.TM)
    b = O.rhs(om, d.src)
    xo, _ = O.linear_solve(A, b)
    its = {}
    for trail in ("fp32", "fp16x3"):
        monkeypatch.setenv("FDFD_MIXED_TRAIL", trail)  # read when the handle is created
        eng = Engine(0)
        try:
            _engine_grid(eng, d)
            fid, _ = eng.factor(O.TM, om, "mixed")
            x, st = eng.linsolve(fid, b, tol=1e-13)
            eng.release(fid)
        finally:
            eng.close()
        res = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
        err = np.linalg.norm(x - xo) / np.linalg.norm(xo)
        its[trail] = st["iterations"]
        print(f"{Nx}x{Ny} trail={trail}: res={res:.2e} err={err:.2e} its={its[trail]}", flush=True)
        assert res <= 1e-12
        assert err <= FIELD_TOL
    assert its["fp16x3"] <= its["fp32"] + 3


@pytest.mark.parametrize("switch", ["FDFD_FUSE_PANELS=0", "FDFD_GEMM3M=0", "FDFD_GEMM_WAVES=4",
                                    "FDFD_GRAPHS=0", "FDFD_MIXED_TRAIL=fp32"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_handle_switches_match_oracle(monkeypatch, switch, precision):
    """Every kernel variant selectable at handle creation (one panel per
    trailing pass, four-multiplication GEMM, four-wave GEMM workgroups, no
    graph replay) gives the oracle's solution, plain and transposed."""
    key, val = switch.split("=")
    monkeypatch.setenv(key, val)  # read when the handle is created
    eng = Engine(0)
    try:
        d = _device(72, 100, 6, seed=11)
        _engine_grid(eng, d)
        om = float(d.omega[0])
        A = O.assemble(d.grid, d.eps_r, om, O.TM)
        b = O.rhs(om, d.src)
        fid, _ = eng.factor(O.TM, om, precision)
        try:
            for tr in (False, True):
                ref, _ = O.linear_solve(A, b, trans="T" if tr else "N")
                x, st = eng.linsolve(fid, b, transpose=tr, tol=1e-13)
                Aop = A.T if tr else A
                err = np.linalg.norm(x - ref) / np.linalg.norm(ref)
                res = np.linalg.norm(b - Aop @ x) / np.linalg.norm(b)
                print(f"{switch} {precision} tr={tr}: err={err:.2e} res={res:.2e} "
                      f"its={st['iterations']}", flush=True)
                assert res <= 1e-12
                assert err <= FIELD_TOL
        finally:
            eng.release(fid)
    finally:
        eng.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("pol", [O.TM, O.TE])
def test_config_c1_fields(engine, precision, pol):
    """200x120 guide + point source (the CPU-runnable workload), all 3 fields."""
    d = O.config_c1()
    _engine_grid(engine, d)
    fo = O.solve(d, pol)
    fid, fst = engine.factor(pol, float(d.omega[0]), precision)
    try:
        f, st = engine.solve_fields(fid, d.src, tol=1e-13)
    finally:
        engine.release(fid)
    f = f[:, :, :, 0]
    for c in range(3):
        err = np.linalg.norm(f[:, :, c] - fo.data[:, :, c]) / np.linalg.norm(fo.data[:, :, c])
        print(f"C1 pol={pol} {precision} comp={c} err={err:.2e} res={st['residual']:.2e} "
              f"its={st['iterations']} growth={fst['pivot_growth']:.2e}", flush=True)
        assert err <= FIELD_TOL
    assert st["residual"] <= 1e-13


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_window_pivot_search_matches_full_search(engine, precision):
    """The windowed pivot search (every pivot verified against all unpivoted
    rows with the threshold test |pivot| >= 0.1 max) and the full-column
    search give the same solution.  On this workload the largest row outside
    a window is at most 1.03 times the window pivot (deepest reduction level,
    measured on the CPU), so the verification passes and no fallback
    refactorization happens."""
    d = O.random_pixel_device(160, 144, 12, 1.55, 5)
    _engine_grid(engine, d)
    om = float(d.omega[0])
    A = O.assemble(d.grid, d.eps_r, om, O.TM)
    b = O.rhs(om, d.src)
    xs = {}
    for full in (False, True):
        engine.set_tuning(0, force_full_search=full)
        try:
            fid, fst = engine.factor(O.TM, om, precision)
            x, st = engine.linsolve(fid, b, tol=1e-13)
            engine.release(fid)
        finally:
            engine.set_tuning(0, force_full_search=False)
        assert fst["pivot_fallback"] == (1 if full else 0)
        xs[full] = x
        print(f"{precision} full={full}: growth={fst['pivot_growth']:.3e} its={st['iterations']}",
              flush=True)
    err = np.linalg.norm(xs[True] - xs[False]) / np.linalg.norm(xs[True])
    assert err <= FIELD_TOL
    res = np.linalg.norm(b - A @ xs[False]) / np.linalg.norm(b)
    assert res <= 1e-12
