"""Several sources on one factor (fdfd_solve with nrhs > 1): the lockstep
multi-source FGMRES (one pass over the factor per iteration for all sources)
against one-at-a-time solves and the oracle, TM and TE, plain and mixed,
including more sources than one lockstep group (16) and a zero source."""
import numpy as np
import pytest

import fdfd_oracle as O
from fdfd import core

pytestmark = pytest.mark.gpu


def _cgrid(g):
    return core.Grid(g.dx, g.Npml, (g.bounds[0][0], g.bounds[1][0]),
                     (g.bounds[0][1], g.bounds[1][1]), g.L0)


@pytest.mark.parametrize("pol", [O.TM, O.TE])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_multi_source_solve_matches_single(engine, pol, precision):
    d = O.random_pixel_device(120, 104, 10, 1.55, 9)
    om = float(d.omega[0])
    Nx, Ny = d.grid.N
    rng = np.random.default_rng(13)
    ns = 19
    srcs = np.zeros((Nx, Ny, ns), dtype=complex)
    for r in range(ns):
        if r == 5:
            continue  # an all-zero source: zero field
        srcs[rng.integers(12, Nx - 12), rng.integers(12, Ny - 12), r] = 1j
        srcs[:, :, r] += (r % 3 == 0) * 0.1 * d.src
    engine.set_grid(_cgrid(d.grid))
    engine.set_eps(d.eps_r)
    fid, _ = engine.factor(pol, om, precision)
    try:
        fm, st = engine.solve_fields(fid, srcs, tol=1e-12)
        print(f"multi pol={pol} {precision}: {ns} sources, {st['iterations']} its, "
              f"residual {st['residual']:.1e}, {st['t_solve_ms']:.1f} ms", flush=True)
        assert st["residual"] <= 1e-12
        for r in range(ns):
            fs, _ = engine.solve_fields(fid, srcs[:, :, r], tol=1e-12)
            scale = max(np.linalg.norm(fs[:, :, :, 0]), 1e-300)
            err = np.linalg.norm(fm[:, :, :, r] - fs[:, :, :, 0]) / scale
            if r == 5:
                assert not np.any(fm[:, :, :, r])
            else:
                assert err <= 1e-10, (r, err)
    finally:
        engine.release(fid)
    A = O.assemble(d.grid, d.eps_r, om, pol)
    for r in (0, 17):
        x, _ = O.linear_solve(A, O.rhs(om, srcs[:, :, r]))
        ez = fm[:, :, 0, r].ravel(order="F")
        assert np.linalg.norm(ez - x) / np.linalg.norm(x) <= 1e-10


@pytest.mark.parametrize("shape", [(120, 104), (72, 136)])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_mfma_and_vector_products_agree(monkeypatch, shape, precision):
    """The reduced solve's multi-source matrix-vector step on the matrix
    cores (default) and on the vector ALUs (FDFD_MV_MFMA=0) give the same
    fields, on a grid whose lines are rows (120 x 104: n = 104, strided
    vectors) and one whose lines are columns (72 x 136: n = 72); neither n is
    a multiple of the 64-row tile or the 64-column stage."""
    from fdfd.solve import Engine
    Nx, Ny = shape
    d = O.random_pixel_device(Nx, Ny, 10, 1.55, 21)
    om = float(d.omega[0])
    rng = np.random.default_rng(5)
    ns = 21  # a full chunk of 16 and a partial one
    srcs = np.zeros((Nx, Ny, ns), dtype=complex)
    for r in range(ns):
        srcs[rng.integers(12, Nx - 12), rng.integers(12, Ny - 12), r] = 1 + 1j * r
    out = {}
    for mf in ("1", "0"):
        monkeypatch.setenv("FDFD_MV_MFMA", mf)
        eng = Engine(0)
        try:
            eng.set_grid(_cgrid(d.grid))
            eng.set_eps(d.eps_r)
            fid, _ = eng.factor(O.TM, om, precision)
            f, st = eng.solve_fields(fid, srcs, tol=1e-12)
            eng.release(fid)
        finally:
            eng.close()
        print(f"{shape} {precision} mfma={mf}: {st['iterations']} its, residual "
              f"{st['residual']:.1e}", flush=True)
        assert st["residual"] <= 1e-12
        out[mf] = f
    for r in range(ns):
        a, b = out["1"][:, :, :, r], out["0"][:, :, :, r]
        assert np.linalg.norm(a - b) / np.linalg.norm(b) <= 1e-10, r
