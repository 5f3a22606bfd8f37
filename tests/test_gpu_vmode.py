"""GPU parity of the full-vector mode solver (fdfd_vector_eigenmode through
the C ABI) against the oracle's restatement of eigenmode(d::AbstractDevice{2},
neff, nev) (src/solver/eigen.jl:31-66): ARPACK shift-invert on the assembled
(Hx, Hy) operator.

Bar (floating point): beta to 1e-10 relative; the six field components of
isolated modes (unit-norm (Hx, Hy), largest entry real positive on both
sides) to 1e-7 relative to their largest entry.
"""
import numpy as np
import pytest

import fdfd_oracle as O
from fdfd import core
from fdfd.modes import eigenmode, vector_eigenmode

pytestmark = pytest.mark.gpu

OM = O.omega_of_wavelength(1.55)


def _strip(dl=0.04, w=0.52, t=0.24, clad=2.1):
    """A silicon strip (eps 12.25) in oxide, centred in a 3 x 2 um window."""
    g = core.Grid(dl, (0, 0), (-1.5, 1.5), (-1.0, 1.0))
    og = O.Grid.create(dl, (0, 0), (-1.5, 1.5), (-1.0, 1.0))
    d = core.Device(g, OM)
    d.eps_r[:] = clad
    X, Y = np.meshgrid(g.xc(), g.yc(), indexing="ij")
    d.eps_r[(np.abs(X) < w / 2) & (np.abs(Y) < t / 2)] = 12.25
    return d, og


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_vector_modes_match_oracle(engine, precision):
    d, og = _strip()
    beta, fields = vector_eigenmode(d, 2.4, 3, precision=precision, eng=engine)
    ob, ofs, lam, _ = O.vector_eigenmode(og, d.eps_r, OM, 2.4, 3)
    print(f"{precision}: beta {beta} oracle {ob}", flush=True)
    assert np.abs(beta - ob).max() <= 1e-10 * np.abs(ob).max()
    full = np.sort(np.abs(lam))
    for k in range(3):
        gap = np.sort(np.abs(lam - lam[k]))[1] / abs(lam[k])
        if gap < 1e-4:
            continue
        a, b = fields[k].data, ofs[k].data
        err = np.abs(a - b).max() / np.abs(b).max()
        print(f"  mode {k}: field err {err:.2e} (gap {gap:.1e})", flush=True)
        assert err <= 1e-7


def test_vector_mode_slab_equals_1d_tm(engine):
    """A slab uniform along x: the fundamental vector mode is the 1-D TM mode
    of the eps(y) profile (computed by the 1-D mode kernel)."""
    prof = np.concatenate([np.full(20, 2.25), np.full(9, 12.25), np.full(20, 2.25)]) + 0j
    g = core.Grid(0.02, (0, 0), (0, 0.12), (0, 0.98))
    d = core.Device(g, OM)
    d.eps_r[:] = np.tile(prof, (g.N[0], 1))
    b1, _ = eigenmode(prof, 0.02, O.TM, OM, 3.0, 1, eng=engine)
    k0 = OM / O.normalize_parameters()[2]
    # a guess near (not at) the eigenvalue: the shift itself must not be singular
    beta, fields = vector_eigenmode(d, b1[0] / k0 * (1 + 1e-3), 1, eng=engine)
    print(f"slab: vector {beta[0]} 1-D TM {b1[0]}", flush=True)
    assert abs(beta[0] - b1[0]) <= 1e-10 * abs(b1[0])
    assert np.abs(fields[0]["Hx"]).max() <= 1e-8 * np.abs(fields[0]["Hy"]).max()
