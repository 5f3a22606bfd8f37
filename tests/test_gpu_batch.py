"""Batched driven solves (several grids x frequencies x sources in one call)
agree with one-at-a-time solves and with the oracle."""
import numpy as np
import pytest

import fdfd_oracle as O
from fdfd import core

pytestmark = pytest.mark.gpu


def _job(Nx, Ny, seed, nsrc):
    d = O.random_pixel_device(Nx, Ny, 6, 1.55, seed, design=8 * (min(Nx, Ny) // 16), src_ix=10)
    rng = np.random.default_rng(seed)
    srcs = np.zeros((Nx, Ny, nsrc), dtype=complex)
    for r in range(nsrc):
        srcs[rng.integers(7, Nx // 4), rng.integers(7, Ny - 7), r] = 1j
    g = d.grid
    cg = core.Grid(g.dx, g.Npml, (g.bounds[0][0], g.bounds[1][0]),
                   (g.bounds[0][1], g.bounds[1][1]), g.L0)
    return d, cg, srcs


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_matches_single_solves(engine, precision):
    lams = np.linspace(1.50, 1.60, 3)
    oms = np.array([O.omega_of_wavelength(l) for l in lams])
    specs = [_job(48, 40, 3, 2), _job(40, 56, 4, 2)]
    jobs = [(cg, O.TM, d.eps_r, oms, srcs) for (d, cg, srcs) in specs]
    res, st = engine.batch_solve(jobs, precision=precision, tol=1e-12)
    print(f"batch {precision}: its={st['iterations']} res={st['residual']:.2e} "
          f"factor={st['t_factor_ms']:.1f} ms solve={st['t_solve_ms']:.1f} ms", flush=True)
    assert st["residual"] <= 1e-12
    for (d, cg, srcs), r in zip(specs, res):
        for o, om in enumerate(oms):
            A = O.assemble(d.grid, d.eps_r, om, O.TM)
            for k in range(srcs.shape[2]):
                b = O.rhs(om, srcs[:, :, k])
                x, _ = O.linear_solve(A, b)
                ez = r[o, k, :, :, 0].ravel(order="F")
                err = np.linalg.norm(ez - x) / np.linalg.norm(x)
                assert err <= 1e-9, (o, k, err)


def test_pipelined_batch_matches_sequential_lanes(monkeypatch):
    """The two-lane pipeline (next group factored on a second stream while the
    current one is solved) gives the same fields, bit for bit, as running the
    lanes in sequence (FDFD_BATCH_LANES=1), over uneven grids that need
    several groups each (FDFD_BATCH_MEMORY_GB=0.02 allows one frequency per
    group), mixed and FP64."""
    from fdfd.solve import Engine
    monkeypatch.setenv("FDFD_BATCH_MEMORY_GB", "0.02")
    lams = np.linspace(1.50, 1.60, 4)
    oms = np.array([O.omega_of_wavelength(l) for l in lams])
    specs = [_job(96, 80, 5, 3), _job(64, 88, 6, 2)]
    jobs = [(cg, O.TM, d.eps_r, oms, srcs) for (d, cg, srcs) in specs]
    for precision in ("mixed", "fp64"):
        out = {}
        for lanes in ("2", "1"):
            monkeypatch.setenv("FDFD_BATCH_LANES", lanes)
            eng = Engine(0)
            try:
                res, st = eng.batch_solve(jobs, precision=precision, tol=1e-11)
            finally:
                eng.close()
            print(f"{precision} lanes={lanes}: its={st['iterations']} res={st['residual']:.2e}",
                  flush=True)
            assert st["residual"] <= 1e-11
            out[lanes] = res
        for a, b in zip(out["2"], out["1"]):
            assert np.array_equal(a, b), precision
