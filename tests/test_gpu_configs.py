"""The benchmark configurations at their full sizes on the GPU, through the
C ABI, checked against the oracle's FP64 operators and solves.

  C2  TM and TE at 1024^2, plain FP64 and mixed factors: FP64 residual of the
      returned field on the oracle's assembled A (a sparse matvec on the host),
      H/E recovery against the oracle's differences of the returned field,
      mixed against plain, and the TM plain field against the oracle's refined
      SuperLU solve.
  C3  the whole batched sweep (3 grids x 64 wavelengths x 16 sources) in one
      fdfd_batch_run; sampled results against oracle solves / host residuals.
  C4  dJ/deps on all 65,536 design cells of the 1024 x 768 coupler through the
      solve pullback, against central finite differences (h = 1e-4) on the 32
      seeded cells, and mixed against plain.
  C5  the 8 eigenfrequencies nearest 2 pi c0/1.55 um of the 768^2 rod lattice,
      plain and mixed factors: eigen-residual on the oracle's operator, the two
      precisions against each other (omega and Q), and the same lattice at
      160^2 against ARPACK.
"""
import math

import numpy as np
import pytest

import fdfd_oracle as O
from fdfd import adjoint, core

pytestmark = pytest.mark.gpu

EPS0n, MU0n, C0n = O.normalize_parameters()


def _cgrid(g):
    return core.Grid(g.dx, g.Npml, (g.bounds[0][0], g.bounds[1][0]),
                     (g.bounds[0][1], g.bounds[1][1]), g.L0)


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


# ---------------------------------------------------------------------------
# C2: TM and TE at 1024^2
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c2(engine):
    d = O.config_c2()
    om = float(d.omega[0])
    engine.set_grid(_cgrid(d.grid))
    engine.set_eps(d.eps_r)
    out = {}
    for pol in (O.TM, O.TE):
        for prec in ("fp64", "mixed"):
            print(f"C2 1024^2 pol={pol} {prec}: factor + solve", flush=True)
            fid, fst = engine.factor(pol, om, prec)
            try:
                f, st = engine.solve_fields(fid, d.src, tol=1e-12)
            finally:
                engine.release(fid)
            out[(pol, prec)] = (f[:, :, :, 0], {**fst, **st})
            print(f"   factor {fst['t_factor_ms']:.1f} ms, solve {st['t_solve_ms']:.1f} ms, "
                  f"{st['iterations']} its, residual {st['residual']:.2e}", flush=True)
    return d, om, out


@pytest.mark.parametrize("pol", [O.TM, O.TE])
def test_c2_residual_on_oracle_operator(c2, pol):
    """||b - A x|| / ||b|| <= 1e-11 in FP64 on the reference operator
    (driven.jl:35-36 / 45-46) for the returned field, both precisions."""
    d, om, out = c2
    A = O.assemble(d.grid, d.eps_r, om, pol)
    b = O.rhs(om, d.src)
    for prec in ("fp64", "mixed"):
        f, st = out[(pol, prec)]
        x = f[:, :, 0].ravel(order="F")
        r = O.rel_residual(A, x, b)
        print(f"C2 pol={pol} {prec}: host residual {r:.2e} (library {st['residual']:.2e})",
              flush=True)
        assert r <= 1e-11, (pol, prec, r)


@pytest.mark.parametrize("pol", [O.TM, O.TE])
def test_c2_field_recovery(c2, pol):
    """H (TM, driven.jl:40-41) and E (TE, driven.jl:50-51) from the returned
    Ez / Hz by the oracle's stretched differences: 1e-12 relative."""
    d, om, out = c2
    g = d.grid
    dxf, dxb, dyf, dyb = O.stretched_deltas(g, om)
    f, _ = out[(pol, "fp64")]
    u = f[:, :, 0].ravel(order="F")
    if pol == O.TM:
        c1 = (-1 / 1j / om / MU0n) * (dyb @ u)
        c2_ = (1 / 1j / om / MU0n) * (dxb @ u)
    else:
        Texi = 1 / O.grid_average(EPS0n * d.eps_r, O.XHAT).ravel(order="F")
        Teyi = 1 / O.grid_average(EPS0n * d.eps_r, O.YHAT).ravel(order="F")
        c1 = (1 / 1j / om) * (Teyi * (dyb @ u))
        c2_ = (1 / 1j / om) * (Texi * (-(dxb @ u)))
    assert _rel(f[:, :, 1].ravel(order="F"), c1) <= 1e-12
    assert _rel(f[:, :, 2].ravel(order="F"), c2_) <= 1e-12


@pytest.mark.parametrize("pol", [O.TM, O.TE])
def test_c2_mixed_matches_plain(c2, pol):
    """Mixed (complex64 factors + FP64 FGMRES) against plain FP64 factors,
    both refined to 1e-12: every field component within 1e-9 relative."""
    d, om, out = c2
    fp, _ = out[(pol, "fp64")]
    fm, _ = out[(pol, "mixed")]
    errs = [_rel(fm[:, :, c], fp[:, :, c]) for c in range(3)]
    print(f"C2 pol={pol}: mixed vs plain field errors {['%.1e' % e for e in errs]}", flush=True)
    assert max(errs) <= 1e-9


@pytest.mark.slow
def test_c2_tm_plain_matches_oracle_solve(c2):
    """TM plain FP64 field against the oracle's refined sparse LU solve of the
    same 1,048,576-unknown system (the reference's lu(A)\\b): 1e-10."""
    d, om, out = c2
    ref = O.solve(d, O.TM)
    for prec in ("fp64", "mixed"):
        f, _ = out[(O.TM, prec)]
        errs = [_rel(f[:, :, c], ref.data[:, :, c]) for c in range(3)]
        print(f"C2 TM {prec} vs oracle: {['%.1e' % e for e in errs]}", flush=True)
        assert max(errs) <= 1e-10


# ---------------------------------------------------------------------------
# C3: the batched sweep
# ---------------------------------------------------------------------------
def test_c3_batched_sweep(engine):
    """192 factorizations and 3,072 right-hand sides in one fdfd_batch_run
    (complex64 factors + FP64 refinement to 1e-10).  The library certifies
    every solve's FP64 residual; sampled fields are checked against the
    oracle's solve (256^2) and the oracle's operator (the other grids)."""
    jobs = O.config_c3()
    spec = [(_cgrid(d.grid), O.TM, d.eps_r, oms, srcs) for d, oms, srcs in jobs]
    fetch = [(0, 0, 0), (0, 31, 7), (0, 63, 15), (1, 0, 3), (1, 40, 12), (2, 17, 5), (2, 63, 0)]
    print("C3: 3 grids x 64 wavelengths x 16 sources, one fdfd_batch_run", flush=True)
    res, st = engine.batch_solve(spec, precision="mixed", tol=1e-10, fetch=fetch)
    nsolve = sum(len(oms) * srcs.shape[2] for _, oms, srcs in jobs)
    t = (st["t_factor_ms"] + st["t_solve_ms"]) / 1e3
    print(f"C3: {nsolve} solves, factor {st['t_factor_ms']:.0f} ms, solve {st['t_solve_ms']:.0f} ms"
          f" ({nsolve / t:.1f} solves/s device time), {st['iterations']} FGMRES its, "
          f"worst residual {st['residual']:.2e}", flush=True)
    assert nsolve == 3072
    assert st["residual"] <= 1e-10
    for (j, o, r), f in res.items():
        d, oms, srcs = jobs[j]
        A = O.assemble(d.grid, d.eps_r, float(oms[o]), O.TM)
        b = O.rhs(float(oms[o]), srcs[:, :, r])
        x = f[:, :, 0].ravel(order="F")
        res_h = O.rel_residual(A, x, b)
        assert res_h <= 1e-10, (j, o, r, res_h)
        if j == 0:
            xo, _ = O.linear_solve(A, b)
            err = _rel(x, xo)
            assert err <= 1e-8, (j, o, r, err)


# ---------------------------------------------------------------------------
# C4: adjoint gradient through the solve pullback
# ---------------------------------------------------------------------------
def _c4_fd(engine, d, m, e, om, cells, h=1e-4):
    """Central differences of J = |m^T e|^2 in eps_k, each side solved on its
    own factorization of A(eps +- h e_k) to 1e-12.  The perturbed field is
    e + delta with A(eps +- h) delta = b - A(eps +- h) e, so
    J(eps +- h) - J(eps) = 2 Re(conj(s) m^T delta) + |m^T delta|^2 is formed
    without cancellation against J itself."""
    g = d.grid
    A = O.assemble(g, d.eps_r, om, O.TM)
    b = O.rhs(om, d.src)
    r0 = b - A @ e
    s0 = m @ e
    c = om ** 2 * EPS0n
    out = []
    for k in cells:
        dj = []
        for sgn in (1, -1):
            eps = d.eps_r.ravel(order="F").copy()
            eps[k] += sgn * h
            engine.set_eps(eps.reshape(g.N, order="F"))
            fid, _ = engine.factor(O.TM, om, "mixed")
            try:
                rhs = r0.copy()
                rhs[k] -= sgn * h * c * e[k]
                delta, st = engine.linsolve(fid, rhs, tol=1e-12)
            finally:
                engine.release(fid)
            ds = m @ delta
            dj.append(2 * np.real(np.conj(s0) * ds) + abs(ds) ** 2)
        out.append((dj[0] - dj[1]) / (2 * h))
    engine.set_eps(d.eps_r)
    return np.array(out)


def test_c4_gradient_matches_finite_differences(engine):
    d, m, cells, fd_cells = O.config_c4()
    om = float(d.omega[0])
    dev = core.Device(_cgrid(d.grid), om)
    dev.eps_r[:] = d.eps_r
    dev.src[:] = d.src
    mf = m.reshape(d.grid.N, order="F")
    print("C4 1024x768: objective and gradient (fp64, mixed)", flush=True)
    J, grad, f = adjoint.objective_and_gradient(dev, mf, precision="fp64", tol=1e-12, eng=engine)
    Jm, gradm, _ = adjoint.objective_and_gradient(dev, mf, precision="mixed", tol=1e-12,
                                                  eng=engine)
    g64 = grad.ravel(order="F")[cells]
    g32 = gradm.ravel(order="F")[cells]
    print(f"C4: J = {J:.6e}, |grad| max {np.abs(g64).max():.3e}, mixed vs plain "
          f"{_rel(g32, g64):.1e}", flush=True)
    assert abs(Jm - J) <= 1e-10 * J
    assert _rel(g32, g64) <= 1e-8
    # outside the design region the gradient is still defined; its norm is not checked
    e = f.data[:, :, 0].ravel(order="F")
    engine.set_grid(_cgrid(d.grid))
    engine.set_eps(d.eps_r)
    fd = _c4_fd(engine, d, m, e, om, fd_cells)
    ga = grad.ravel(order="F")[fd_cells]
    scale = np.abs(g64).max()
    rel = np.abs(ga - fd) / np.maximum(np.abs(fd), 1e-3 * scale)
    print(f"C4: adjoint vs central FD on 32 cells: max rel {rel.max():.2e}", flush=True)
    assert rel.max() <= 1e-6


# ---------------------------------------------------------------------------
# C5: eigenfrequencies of the rod lattice
# ---------------------------------------------------------------------------
def _match(a, b):
    """Greedy nearest matching of two eigenvalue lists; returns b reordered."""
    b = list(b)
    out = []
    for x in a:
        i = int(np.argmin([abs(x - y) for y in b]))
        out.append(b.pop(i))
    return np.array(out)


def test_c5_eigenfrequencies(engine):
    d = O.config_c5()
    om0 = float(d.omega[0])
    engine.set_grid(_cgrid(d.grid))
    engine.set_eps(d.eps_r)
    A, sigma = O.eigen_operator(d.grid, d.eps_r, om0, O.TM)
    res = {}
    for prec in ("fp64", "mixed"):
        print(f"C5 768^2 eigenfrequency nev=8 {prec}", flush=True)
        w, fl, st = engine.eigenfrequency(O.TM, om0, 8, ncv=20, precision=prec, tol=1e-11)
        print(f"   factor {st['t_factor_ms']:.0f} ms, {st['iterations']} FGMRES its, "
              f"{st['restarts']} restarts, nconv {st['nconv']}", flush=True)
        lam = -w ** 2 * MU0n * EPS0n
        worst = 0.0
        for q in range(8):
            v = fl[q][:, :, 0].ravel(order="F")
            r = np.linalg.norm(A @ v - lam[q] * v) / (abs(lam[q]) * np.linalg.norm(v))
            worst = max(worst, r)
        Q = w.real / (2 * np.abs(w.imag))
        print(f"   omega/omega0 {np.round(w.real / om0, 6)}, Q {np.round(Q, 1)}, "
              f"eigen-residual {worst:.1e}", flush=True)
        assert worst <= 1e-7
        res[prec] = w
    wp, wm = res["fp64"], _match(res["fp64"], res["mixed"])
    assert np.max(np.abs(wm - wp) / np.abs(wp)) <= 1e-10
    Qp = wp.real / (2 * np.abs(wp.imag))
    Qm = wm.real / (2 * np.abs(wm.imag))
    qerr = np.max(np.abs(Qm - Qp) / Qp)
    print(f"C5: mixed vs plain: omega {np.max(np.abs(wm - wp) / np.abs(wp)):.1e}, Q {qerr:.1e}",
          flush=True)
    assert qerr <= 1e-5


def test_c5_lattice_matches_arpack(engine):
    """The same lattice on 160^2 cells against the oracle's ARPACK
    shift-invert eigs (eigen.jl:86): omega to 1e-9, eigenvector overlap."""
    d = O.config_c5(n=160)
    om0 = float(d.omega[0])
    engine.set_grid(_cgrid(d.grid))
    engine.set_eps(d.eps_r)
    wo, fo, lam_o, sigma = O.eigenfrequency(d, O.TM, 8, ncv=20)
    for prec in ("fp64", "mixed"):
        w, fl, st = engine.eigenfrequency(O.TM, om0, 8, ncv=20, precision=prec, tol=1e-12)
        wm = _match(wo, w)
        err = np.max(np.abs(wm - wo) / np.abs(wo))
        print(f"C5 160^2 {prec}: omega vs ARPACK {err:.1e}", flush=True)
        assert err <= 1e-9


# ---------------------------------------------------------------------------
# the reference notebook's modulated waveguide (Example 3)
# ---------------------------------------------------------------------------
def test_notebook_modulated_waveguide_photon_numbers(engine):
    """Example_simulations.ipynb cells 17-25: Grid(0.01, [15 10], [0 15],
    [-1 1]) (1500 x 200), carrier 2 pi 1.939e14, Omega = 4.541e14, one
    sideband each side, a = 0.2202, q = 2.9263, guide eps 12.25 for
    |y| <= a/2, deps = exp(i q x) on 1.5 <= x <= 11.7, -a/2 <= y <= 0, TM
    mode source Mode(TM, x, 3.5, [0.2, 0], 4a).  The notebook prints
    Nin = 1.3625216010889075e-20, Nout = 1.3618731014650896e-20 and
    Nout/Nin = 0.9995240445191477 (photon-number fluxes through x = 1.25 and
    x = 11.95).  No cell centre lies on a region edge.  Bar: 1e-12 relative."""
    from fdfd import flux, modes
    from fdfd.modulated import ModulatedDevice, setup_deps, solve_modulated
    a, q = 0.2202, 2.9263
    w, Om = 2 * math.pi * 1.939e14, 4.541e14
    g = core.Grid(0.01, (15, 10), (0.0, 15.0), (-1.0, 1.0))
    assert g.N == (1500, 200)
    d = ModulatedDevice(g, w, Om, 1)
    core.setup_eps(d, lambda x, y: -a / 2 <= y <= a / 2, 12.25)
    setup_deps(d, lambda x, y: 1.5 <= x <= 11.7 and -a / 2 <= y <= 0,
               lambda x, y: np.exp(1j * q * x))
    modes.add_mode(d, modes.Mode(O.TM, core.XHAT, 3.5, core.Point(0.2, 0.0), 4 * a))
    for prec in ("fp64", "mixed"):
        fs, st = solve_modulated(d, precision=prec, tol=1e-12, eng=engine, return_stats=True)
        f = fs[0]
        nin = flux.flux_surface_integral(f[1], core.Point(1.25, 0.0), np.inf, core.XHAT,
                                         eng=engine) / w
        nout = (flux.flux_surface_integral(f[2], core.Point(11.95, 0.0), np.inf, core.XHAT,
                                           eng=engine) / (Om + w)
                + flux.flux_surface_integral(f[1], core.Point(11.95, 0.0), np.inf, core.XHAT,
                                             eng=engine) / w
                + flux.flux_surface_integral(f[0], core.Point(11.95, 0.0), np.inf, core.XHAT,
                                             eng=engine) / (w - Om))
        ratio = nout / nin
        print(f"notebook modulated waveguide {prec}: {st[0]['iterations']} its, Nin {nin!r}, "
              f"Nout {nout!r}, Nout/Nin {ratio!r}", flush=True)
        # measured agreement: 2e-14 (Nin), 5e-15 (Nout), 3e-14 (ratio)
        assert nin == pytest.approx(1.3625216010889075e-20, rel=1e-12)
        assert nout == pytest.approx(1.3618731014650896e-20, rel=1e-12)
        assert ratio == pytest.approx(0.9995240445191477, rel=1e-12)
