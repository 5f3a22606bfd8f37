"""fdfd_autotune: every candidate variant solves the problem, the chosen one
is kept on the handle and recorded in FDFD_TUNE_CACHE, and a new handle picks
it up from the file at factorization time; the tuned factors still solve to
the oracle's field."""
import os

import numpy as np
import pytest

import fdfd_oracle as O
from fdfd import core
from fdfd.solve import Engine

pytestmark = pytest.mark.gpu


def _cgrid(g):
    return core.Grid(g.dx, g.Npml, (g.bounds[0][0], g.bounds[1][0]),
                     (g.bounds[0][1], g.bounds[1][1]), g.L0)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_autotune_records_and_reuses(tmp_path, precision):
    d = O.random_pixel_device(128, 96, 10, 1.55, 21)
    om = float(d.omega[0])
    cache = tmp_path / "tune.txt"
    old = os.environ.get("FDFD_TUNE_CACHE")
    os.environ["FDFD_TUNE_CACHE"] = str(cache)
    try:
        e1 = Engine(0)
        e1.set_grid(_cgrid(d.grid))
        e1.set_eps(d.eps_r)
        knobs, ms = e1.autotune(O.TM, om, precision)
        print(f"autotune {precision}: {knobs}, {ms:.2f} ms per step", flush=True)
        e1.close()
        lines = [l.split() for l in cache.read_text().splitlines() if not l.startswith("#")]
        assert len(lines) == 1
        n, m, prec, pol = map(int, lines[0][:4])
        assert (n, m, pol) == (96, 128, 1) and prec == (0 if precision == "fp64" else 1)
        rec = list(map(int, lines[0][4:9]))
        assert rec == [knobs[k] for k in ("gemm_b3", "trail_h3", "fuse_panels", "gemm_waves",
                                          "gemm3m")]
        # a fresh handle takes the cached variants when it factors this shape
        e2 = Engine(0)
        e2.set_grid(_cgrid(d.grid))
        e2.set_eps(d.eps_r)
        fid, _ = e2.factor(O.TM, om, precision)
        assert e2.tuning() == knobs
        f, st = e2.solve_fields(fid, d.src, tol=1e-12)
        e2.release(fid)
        e2.close()
        ref = O.solve(d, O.TM)
        err = np.linalg.norm(f[:, :, :, 0] - ref.data) / np.linalg.norm(ref.data)
        assert err <= 1e-10
    finally:
        if old is None:
            del os.environ["FDFD_TUNE_CACHE"]
        else:
            os.environ["FDFD_TUNE_CACHE"] = old
