"""The reverse-mode rule of the TM driven solve (fdfd.adjoint.solve_rrule,
the Python mirror of FDFDAMD's ChainRulesCore.rrule, on fdfd_solve_pullback)
against central finite differences of the oracle's FP64 solves, for a loss
that reads all three field components, and for several frequencies at once."""
import numpy as np
import pytest

import fdfd_oracle as O
from fdfd import adjoint, core
from fdfd._lib import FDFDError

pytestmark = pytest.mark.gpu

EPS0n, MU0n, C0n = O.normalize_parameters()


def _device(oms):
    g = O.Grid.create(0.02, (5, 5), (0, 0.8), (0, 0.64))
    d = O.Device(g, oms)
    rng = np.random.default_rng(17)
    d.eps_r[:] = 1 + 11.25 * rng.random(g.N)
    d.src[8, 20] = 1j
    d.src[9, 10:14] = 0.5
    cg = core.Grid(g.dx, g.Npml, (g.bounds[0][0], g.bounds[1][0]),
                   (g.bounds[0][1], g.bounds[1][1]), g.L0)
    dev = core.Device(cg, oms)
    dev.eps_r[:] = d.eps_r
    dev.src[:] = d.src
    return d, dev


def _weights(shape, seed):
    rng = np.random.default_rng(seed)
    mz = np.zeros(shape, dtype=complex)
    mz[30, 8:24] = rng.standard_normal(16) + 1j * rng.standard_normal(16)
    wx = np.zeros(shape)
    wx[25:35, 5:12] = rng.random((10, 7))
    my = np.zeros(shape, dtype=complex)
    my[28, 20:30] = rng.standard_normal(10)
    return mz, wx, my


def _loss(F, w):
    """L = |sum mz Ez|^2 + sum |wx Hx|^2 + |sum my Hy|^2 and its cotangent
    dL/dRe F + i dL/dIm F."""
    mz, wx, my = w
    ez, hx, hy = F[:, :, 0], F[:, :, 1], F[:, :, 2]
    sz, sy = np.sum(mz * ez), np.sum(my * hy)
    L = abs(sz) ** 2 + np.sum(np.abs(wx * hx) ** 2) + abs(sy) ** 2
    bar = np.stack([2 * sz * np.conj(mz), 2 * wx ** 2 * hx, 2 * sy * np.conj(my)], axis=2)
    return L, bar


def _fd(d, w, cells, h=1e-4):
    out = []
    for (i, j) in cells:
        vals = []
        for sgn in (1, -1):
            e = d.eps_r.copy()
            e[i, j] += sgn * h
            dd = O.Device(d.grid, d.omega)
            dd.eps_r[:] = e
            dd.src[:] = d.src
            fs = O.solve(dd, O.TM)
            fs = fs if isinstance(fs, list) else [fs]
            vals.append(sum(_loss(f.data, w)[0] for f in fs))
        out.append((vals[0] - vals[1]) / (2 * h))
    return np.array(out)


CELLS = [(30, 16), (22, 10), (33, 25), (12, 20), (6, 6)]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_rrule_three_component_loss_matches_fd(engine, precision):
    d, dev = _device(O.omega_of_wavelength(1.55))
    w = _weights(d.grid.N, 3)
    f, pb = adjoint.solve_rrule(dev, precision=precision, tol=1e-13, eng=engine)
    try:
        L, bar = _loss(f.data, w)
        g = pb(bar)
    finally:
        pb.release()
    fd = _fd(d, w, CELLS)
    ga = np.array([g[c] for c in CELLS])
    rel = np.abs(ga - fd) / np.maximum(np.abs(fd), 1e-3 * np.abs(g).max())
    print(f"rrule {precision}: L={L:.4e} rel err vs FD {rel.max():.1e}", flush=True)
    assert rel.max() <= 1e-6


def test_rrule_sums_over_frequencies(engine):
    oms = [O.omega_of_wavelength(l) for l in (1.50, 1.55, 1.62)]
    d, dev = _device(oms)
    w = _weights(d.grid.N, 5)
    fs, pb = adjoint.solve_rrule(dev, tol=1e-13, eng=engine)
    try:
        assert len(fs) == 3
        g = pb([_loss(f.data, w)[1] for f in fs])
    finally:
        pb.release()
    # one frequency at a time adds up to the same gradient
    parts = []
    for k, om in enumerate(oms):
        _, dk = _device(om)
        fk, pk = adjoint.solve_rrule(dk, tol=1e-13, eng=engine)
        try:
            parts.append(pk(_loss(fk.data, w)[1]))
        finally:
            pk.release()
        assert np.allclose(fk.data, fs[k].data, rtol=0, atol=1e-12 * np.abs(fk.data).max())
    assert np.allclose(g, sum(parts), rtol=1e-10, atol=1e-12 * np.abs(g).max())
    fd = _fd(d, w, CELLS[:3])
    ga = np.array([g[c] for c in CELLS[:3]])
    assert np.max(np.abs(ga - fd) / np.maximum(np.abs(fd), 1e-3 * np.abs(g).max())) <= 1e-6


def test_pullback_rejects_te(engine):
    d, dev = _device(O.omega_of_wavelength(1.55))
    with pytest.raises(NotImplementedError):
        adjoint.solve_rrule(dev, pol=core.TE, eng=engine)
    engine.set_grid(dev.grid)
    engine.set_eps(dev.eps_r)
    fid, _ = engine.factor(core.TE, float(dev.omega[0]), "fp64")
    try:
        z = np.zeros(len(dev.grid), dtype=np.complex128)
        z3 = np.zeros(3 * len(dev.grid), dtype=np.complex128)
        g = np.empty(len(dev.grid))
        import ctypes as C
        from fdfd import _lib
        st = _lib.fdfd_stats()
        rc = engine.lib.fdfd_solve_pullback(engine.h, fid, z.ctypes.data_as(C.c_void_p),
                                            z3.ctypes.data_as(C.c_void_p), 1e-12,
                                            g.ctypes.data_as(C.c_void_p), C.byref(st))
        assert rc == -1
        with pytest.raises(FDFDError):
            _lib.check(engine.lib, engine.h, rc)
    finally:
        engine.release(fid)
