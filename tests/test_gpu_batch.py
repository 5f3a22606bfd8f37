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
