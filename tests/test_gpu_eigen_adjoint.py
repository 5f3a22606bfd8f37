"""GPU parity of fdfd_eigenfrequency (shift-invert Arnoldi on the block
factors) against the oracle's ARPACK shift-invert, and of the adjoint
gradient against the oracle's adjoint (itself pinned by finite differences)
and against central differences of GPU solves.

Bars: eigenfrequencies to 1e-9 relative, eigenvector overlap |<u, v>| to
1 - 1e-8 for simple eigenvalues, gradients to 1e-8 relative (oracle) and
1e-6 relative (central differences, h = 1e-4)."""
import numpy as np
import pytest

import fdfd_oracle as O
from fdfd import core

pytestmark = pytest.mark.gpu

EPS0n, MU0n, C0n = O.normalize_parameters()


def _engine_grid(eng, d):
    g = d.grid
    eng.set_grid(core.Grid(g.dx, g.Npml, (g.bounds[0][0], g.bounds[1][0]),
                           (g.bounds[0][1], g.bounds[1][1]), g.L0))
    assert eng.grid.N == g.N
    eng.set_eps(d.eps_r)


def _cavity(Nx=44, Ny=40, npml=5, pol=O.TM):
    g = O.Grid.create(0.05, (npml, npml), (0, Nx * 0.05), (0, Ny * 0.05))
    d = O.Device(g, 2 * np.pi * C0n / 1.6)
    d.eps_r[14:30, 12:28] = 9.0 if pol == O.TM else 6.0
    return d


@pytest.mark.parametrize("pol", [O.TM, O.TE])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_eigenfrequency_matches_arpack(engine, pol, precision):
    d = _cavity(pol=pol)
    _engine_grid(engine, d)
    nev = 4
    w_ref, f_ref, lam_ref, sigma = O.eigenfrequency(d, pol, nev, tol=1e-14)
    om, fl, st = engine.eigenfrequency(pol, float(d.omega[0]), nev, precision=precision,
                                       tol=1e-12)
    print(f"eig pol={pol} {precision}: gpu {om} ref {w_ref} restarts={st['restarts']} "
          f"its={st['iterations']}", flush=True)
    assert st["nconv"] >= nev
    for k in range(nev):
        err = np.min(np.abs(w_ref - om[k])) / np.abs(w_ref[k])
        assert err <= 1e-9
    # eigenvectors of simple eigenvalues (compare the primary field)
    for k in range(nev):
        j = int(np.argmin(np.abs(w_ref - om[k])))
        if np.sum(np.abs(w_ref - w_ref[j]) < 1e-6 * abs(w_ref[j])) > 1:
            continue
        u = fl[k][:, :, 0].ravel()
        v = f_ref[j].data[:, :, 0].ravel()
        ov = abs(np.vdot(u, v)) / (np.linalg.norm(u) * np.linalg.norm(v))
        assert ov >= 1 - 1e-8


def _objective_setup(seed=21):
    g = O.Grid.create(0.02, (8, 8), (0, 1.2), (0, 0.8))  # 60 x 40
    d = O.Device(g, O.omega_of_wavelength(1.55))
    rng = np.random.default_rng(seed)
    d.eps_r[20:40, 10:30] = 1 + 11.25 * rng.random((20, 20))
    d.eps_r[:20, 18:22] = 12.25
    d.eps_r[40:, 18:22] = 12.25
    d.src[12, 16:24] = 1.0
    mvec = np.zeros(g.N, dtype=complex)
    yy = np.arange(g.N[1])
    mvec[48, :] = np.conj(np.exp(-((yy - 20) / 5.0) ** 2))
    return d, mvec.ravel(order="F")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_adjoint_gradient(engine, precision):
    d, m = _objective_setup()
    _engine_grid(engine, d)
    om = float(d.omega[0])
    b = O.rhs(om, d.src)
    # oracle gradient (the oracle's adjoint formula is pinned by FD in test_oracle)
    A = O.assemble(d.grid, d.eps_r, om, O.TM)
    e_o, _ = O.linear_solve(A, b)
    lam_o, _ = O.linear_solve(A, m, trans="T")
    s_o = m @ e_o
    g_ref = -np.real(2 * np.conj(s_o) * lam_o * om ** 2 * EPS0n * e_o)
    fid, _ = engine.factor(O.TM, om, precision)
    try:
        e, _ = engine.linsolve(fid, b, tol=1e-13)
        lam, _ = engine.linsolve(fid, m, transpose=True, tol=1e-13)
        s = m @ e
        grad = engine.adjoint_grad(fid, e, lam, 2 * np.conj(s))
    finally:
        engine.release(fid)
    rel = np.linalg.norm(grad - g_ref) / np.linalg.norm(g_ref)
    print(f"adjoint {precision}: |g - g_ref|/|g_ref| = {rel:.2e}", flush=True)
    assert rel <= 1e-8

    # central differences of GPU solves on a few design cells
    def J(eps):
        engine.set_eps(eps)
        f2, _ = engine.factor(O.TM, om, "fp64")
        try:
            x, _ = engine.linsolve(f2, b, tol=1e-14)
        finally:
            engine.release(f2)
        return abs(m @ x) ** 2

    h = 1e-4
    for cell in [(25, 15), (33, 22), (21, 28)]:
        k = cell[0] + cell[1] * d.grid.N[0]
        ep, em = d.eps_r.copy(), d.eps_r.copy()
        ep[cell] += h
        em[cell] -= h
        fd = (J(ep) - J(em)) / (2 * h)
        print(f"  cell {cell}: adjoint {grad[k]:.6e} fd {fd:.6e}", flush=True)
        assert abs(grad[k] - fd) <= 1e-6 * abs(fd) + 1e-9 * np.abs(grad).max()
    engine.set_eps(d.eps_r)
