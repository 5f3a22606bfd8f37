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
