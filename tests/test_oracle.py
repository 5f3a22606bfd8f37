"""Pins of the CPU oracle (oracle/fdfd_oracle.py) against facts that do not
come from the oracle's own code: literal constants, hand-computed grid
cases, independent loop-built operators, FFT / Kronecker-sum closed forms,
the PML's design attenuation, a discrete Maxwell identity and finite
differences.  CPU only."""
import math

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import fdfd_oracle as O

EPS0n, MU0n, C0n = O.normalize_parameters()


# --- constants and normalisation (src/types.jl:6-9, src/device.jl:40) ------
def test_constants():
    assert O.EPS0 == 8.85418782e-12 and O.MU0 == 1.25663706e-6
    assert abs(O.C0 * math.sqrt(O.EPS0 * O.MU0) - 1) < 1e-15
    assert abs(O.ETA0 - 376.730313) < 1e-5
    assert EPS0n == 8.85418782e-18 and MU0n == pytest.approx(1.25663706e-12, rel=1e-15)
    assert C0n == pytest.approx(2.9979245813099606e14, rel=1e-14)


# --- Grid (src/grid.jl:26-33, 68-82, 112-125) --------------------------------
def test_grid_c1():
    g = O.Grid.create(0.02, (15, 15), (0.0, 4.0), (-1.2, 1.2))
    assert g.N == (200, 120)
    assert g.dx == pytest.approx(0.02) and g.dy == pytest.approx(0.02)
    assert g.coord2ind(1.0, 0.0) == (51, 61)
    assert g.xc()[0] == pytest.approx(0.01) and g.yc()[-1] == pytest.approx(1.19)


def test_grid_rounding_ties_to_even():
    # N = round(L/dh): 2.5 -> 2, 3.5 -> 4 (Julia round is ties-to-even)
    assert O.Grid.create(1.0, (0, 0), (0, 2.5), (0, 3.5)).N == (2, 4)
    g = O.Grid.create(1.0, (0, 0), (0, 4.0), (0, 4.0))
    # (x - x1)/L*N = 1.5 -> round = 2 -> index 3 ; 2.5 -> 2 -> index 3
    assert g.x2ind(1.5) == 3 and g.x2ind(2.5) == 3 and g.x2ind(0.5) == 1
    assert g.x2ind(-10) == 1 and g.x2ind(99) == 4  # clamped


# --- periodic differences (src/grid.jl:128-154) ------------------------------
def _dense_1d(n, h, fwd):
    D = np.zeros((n, n))
    for i in range(n):
        if fwd:
            D[i, i] -= 1 / h
            D[i, (i + 1) % n] += 1 / h
        else:
            D[i, i] += 1 / h
            D[i, (i - 1) % n] -= 1 / h
    return D


@pytest.mark.parametrize("n", [1, 2, 3, 7])
def test_delta_1d_and_transpose(n):
    g = O.Grid.create(0.5, (0, 0), (0, 0.5 * n), (0, 0.5 * 3))
    dxf = O.delta(O.XHAT, O.FORWARD, g).toarray()
    dxb = O.delta(O.XHAT, O.BACKWARD, g).toarray()
    assert np.allclose(dxf[:n, :n], _dense_1d(n, 0.5, True))
    assert np.allclose(dxf, -dxb.T)  # Dxf^T = -Dxb for the periodic operators


def test_delta_2d_layout():
    """x fastest: y-difference couples k and k+Nx, periodic in y."""
    Nx, Ny = 4, 3
    g = O.Grid.create(1.0, (0, 0), (0, Nx), (0, Ny))
    dyf = O.delta(O.YHAT, O.FORWARD, g).toarray()
    f = np.arange(Nx * Ny, dtype=float) ** 2
    F = f.reshape((Nx, Ny), order="F")
    ref = (np.roll(F, -1, axis=1) - F).ravel(order="F")
    assert np.allclose(dyf @ f, ref)


def test_grid_average():
    a = np.array([[1.0, 2.0], [3.0, 5.0], [7.0, 11.0]])
    assert np.allclose(O.grid_average(a, O.XHAT), [[4.0, 6.5], [2.0, 3.5], [5.0, 8.0]])
    assert np.allclose(O.grid_average(a, O.YHAT), [[1.5, 1.5], [4.0, 4.0], [9.0, 9.0]])


# --- SC-PML (src/pml.jl) -----------------------------------------------------
def test_sfactor_design_attenuation():
    """The conductivity integrates to the design reflection: for a graded
    profile, 2 eta0 * int_0^T sigma dl = -lnR = 12.  The forward factors sit
    at cell centres, so the midpoint rule gives the integral."""
    npml, dl = 40, 0.01
    g = O.Grid.create(dl, (npml, npml), (0, 2.0), (0, 2.0))
    om = O.omega_of_wavelength(1.55)
    s = O.create_sfactor(O.XHAT, O.FORWARD, g, om)
    sigma = -s.imag * om * EPS0n
    integral = np.sum(sigma[:npml]) * dl
    assert 2 * O.ETA0 * integral == pytest.approx(12.0, rel=2e-3)
    assert np.allclose(s[npml:-npml], 1.0)
    assert np.allclose(s, s[::-1])                      # symmetric profile
    assert np.all(np.diff(sigma[:npml]) < 0)           # grows toward the wall
    sb = O.create_sfactor(O.XHAT, O.BACKWARD, g, om)
    assert sb[-npml] == 1.0 and sb[0].imag < s[0].imag  # edge sample is deeper


def test_sfactor_no_pml():
    g = O.Grid.create(0.02, (0, 0), (0, 1.0), (0, 1.0))
    for w in (O.XHAT, O.YHAT):
        for sd in (O.FORWARD, O.BACKWARD):
            assert np.all(O.create_sfactor(w, sd, g, 1e15) == 1.0)


# --- operator: FFT closed form (npml = 0) ------------------------------------
@pytest.mark.parametrize("pol", [O.TM, O.TE])
def test_operator_fft_spectrum(pol):
    Nx, Ny, dx = 10, 7, 0.03
    g = O.Grid.create(dx, (0, 0), (0, Nx * dx), (0, Ny * dx))
    om = O.omega_of_wavelength(1.3)
    eps = 2.5
    d = O.Device(g, om)
    d.eps_r[:] = eps
    A = O.assemble(g, d.eps_r, om, pol)
    X, Y = np.meshgrid(np.arange(Nx), np.arange(Ny), indexing="ij")
    for p, q in [(0, 0), (1, 2), (3, 6), (5, 1)]:
        mode = np.exp(2j * np.pi * (p * X / Nx + q * Y / Ny)).ravel(order="F")
        lap = -4 * (np.sin(np.pi * p / Nx) ** 2 / g.dx ** 2 + np.sin(np.pi * q / Ny) ** 2 / g.dy ** 2)
        if pol == O.TM:
            lam = lap / MU0n + om ** 2 * EPS0n * eps
        else:
            lam = lap / (EPS0n * eps) + om ** 2 * MU0n
        assert np.allclose(A @ mode, lam * mode, rtol=1e-12, atol=1e-12 * abs(lam))


def test_solve_fft_closed_form():
    Nx, Ny, dx = 16, 12, 0.025
    g = O.Grid.create(dx, (0, 0), (0, Nx * dx), (0, Ny * dx))
    om = O.omega_of_wavelength(1.55)
    d = O.Device(g, om)
    d.eps_r[:] = 3.0
    rng = np.random.default_rng(0)
    d.src[:] = rng.standard_normal((Nx, Ny)) + 1j * rng.standard_normal((Nx, Ny))
    f = O.solve(d, O.TM)
    b = (1j * om) * d.src
    p = np.fft.fftfreq(Nx) * Nx
    q = np.fft.fftfreq(Ny) * Ny
    P, Q = np.meshgrid(p, q, indexing="ij")
    lam = (-4 * (np.sin(np.pi * P / Nx) ** 2 / g.dx ** 2 + np.sin(np.pi * Q / Ny) ** 2 / g.dy ** 2)
           / MU0n + om ** 2 * EPS0n * 3.0)
    ez = np.fft.ifft2(np.fft.fft2(b) / lam)
    assert np.linalg.norm(f.data[:, :, 0] - ez) / np.linalg.norm(ez) < 1e-12


# --- operator with PML: Kronecker sum of independently built 1-D operators --
def _op1d(n, h, sf, sb):
    """s_f[i] (s_b[i+1](f[i+1]-f[i]) - s_b[i](f[i]-f[i-1])) / h^2, periodic."""
    M = np.zeros((n, n), dtype=complex)
    for i in range(n):
        ip, im = (i + 1) % n, (i - 1) % n
        M[i, ip] += sf[i] * sb[ip] / h ** 2
        M[i, i] -= sf[i] * (sb[ip] + sb[i]) / h ** 2
        M[i, im] += sf[i] * sb[i] / h ** 2
    return M


def test_tm_operator_kronecker_sum_with_pml():
    Nx, Ny, dx = 11, 9, 0.05
    g = O.Grid.create(dx, (3, 2), (0, Nx * dx), (0, Ny * dx))
    om = O.omega_of_wavelength(1.55)
    eps = 4.0
    d = O.Device(g, om)
    d.eps_r[:] = eps
    A = O.assemble(g, d.eps_r, om, O.TM).toarray()
    sxf, sxb, syf, syb = O.sfactor_vectors(g, om)
    Ax = _op1d(Nx, g.dx, sxf, sxb) / MU0n
    Ay = _op1d(Ny, g.dy, syf, syb) / MU0n
    K = np.kron(np.eye(Ny), Ax) + np.kron(Ay, np.eye(Nx)) + om ** 2 * EPS0n * eps * np.eye(Nx * Ny)
    assert np.abs(A - K).max() <= 1e-12 * np.abs(K).max()
    # spectrum is the sum of the 1-D spectra
    ev = np.sort_complex(np.linalg.eigvals(A))
    ex = np.sort_complex((np.linalg.eigvals(Ax)[:, None] + np.linalg.eigvals(Ay)[None, :]).ravel()
                         + om ** 2 * EPS0n * eps)
    assert np.allclose(ev, ex, rtol=1e-9, atol=1e-9 * np.abs(ex).max())


def test_te_operator_loop_restatement():
    """TE with non-uniform eps against a per-cell loop of driven.jl:45."""
    Nx, Ny, dx = 6, 5, 0.05
    g = O.Grid.create(dx, (2, 1), (0, Nx * dx), (0, Ny * dx))
    om = O.omega_of_wavelength(1.4)
    rng = np.random.default_rng(3)
    eps = 1 + 5 * rng.random((Nx, Ny)) + 0.1j * rng.random((Nx, Ny))
    A = O.assemble(g, eps, om, O.TE).toarray()
    sxf, sxb, syf, syb = O.sfactor_vectors(g, om)
    tx = lambda i, j: 1 / (EPS0n * (eps[i % Nx, j] + eps[(i - 1) % Nx, j]) / 2)
    ty = lambda i, j: 1 / (EPS0n * (eps[i, j % Ny] + eps[i, (j - 1) % Ny]) / 2)
    K = np.zeros_like(A)
    idx = lambda i, j: (i % Nx) + (j % Ny) * Nx
    for i in range(Nx):
        for j in range(Ny):
            k = idx(i, j)
            ip = (i + 1) % Nx
            K[k, idx(i + 1, j)] += sxf[i] * tx(i + 1, j) * sxb[ip] / dx ** 2
            K[k, k] -= sxf[i] * (tx(i + 1, j) * sxb[ip] + tx(i, j) * sxb[i]) / dx ** 2
            K[k, idx(i - 1, j)] += sxf[i] * tx(i, j) * sxb[i] / dx ** 2
            jp = (j + 1) % Ny
            K[k, idx(i, j + 1)] += syf[j] * ty(i, j + 1) * syb[jp] / dx ** 2
            K[k, k] -= syf[j] * (ty(i, j + 1) * syb[jp] + ty(i, j) * syb[j]) / dx ** 2
            K[k, idx(i, j - 1)] += syf[j] * ty(i, j) * syb[j] / dx ** 2
            K[k, k] += om ** 2 * MU0n
    assert np.abs(A - K).max() <= 1e-12 * np.abs(K).max()


# --- solve, fields, flux -----------------------------------------------------
def test_solve_residual_and_dense_agreement():
    g = O.Grid.create(0.02, (4, 4), (0, 0.4), (0, 0.36))
    d = O.Device(g, O.omega_of_wavelength(1.55))
    rng = np.random.default_rng(5)
    d.eps_r[:] = 1 + 11.25 * (rng.random(g.N) < 0.5)
    d.src[10, 9] = 1j
    f, systems = O.solve(d, O.TM, return_system=True)
    A, b, _ = systems[0]
    ez = f.data[:, :, 0].ravel(order="F")
    assert O.rel_residual(A, ez, b) < 1e-14
    xd = np.linalg.solve(A.toarray(), b)
    assert np.linalg.norm(ez - xd) / np.linalg.norm(xd) < 1e-11


def test_tm_fields_satisfy_ampere():
    """Recovered H satisfies dxf Hy - dyf Hx - i w eps0 eps Ez = src, which
    follows from A ez = i w src and the definitions of Hx, Hy; checks the
    H recovery (driven.jl:40-41) through the physics, not the formula."""
    d = O.config_c1()
    f = O.solve(d, O.TM)
    g = d.grid
    om = float(d.omega[0])
    dxf, dxb, dyf, dyb = O.stretched_deltas(g, om)
    ez, hx, hy = (f.data[:, :, c].ravel(order="F") for c in range(3))
    lhs = dxf @ hy - dyf @ hx - 1j * om * EPS0n * d.eps_r.ravel(order="F") * ez
    src = d.src.ravel(order="F")
    assert np.linalg.norm(lhs - src) / np.linalg.norm(src) < 1e-9


def test_flux_indexing_hand_case():
    g = O.Grid.create(1.0, (0, 0), (0, 4.0), (0, 6.0))
    data = np.zeros((4, 6, 3), dtype=complex)
    data[:, :, 0] = 2.0   # Ez
    data[:, :, 2] = 0.5j  # Hy
    data[2, :, 0] = 4.0   # column 3 (x = 2.5), averaged with column 2
    f = O.FieldTM(g, 1.0, data)
    # centre x=1.5 -> column index 1 (0-based); rows with |y - 3| <= 1: y = 2.5, 3.5
    val = O.flux_surface_integral_tm_x(f, 1.5, 3.0, 1.0)
    ez_avg = (2.0 + 4.0) / 2
    assert val == pytest.approx(2 * (-0.5 * np.real(ez_avg * np.conj(0.5j))) * 1.0)


def test_flux_conservation_lossless_slab():
    """Npml_y = 0 (periodic, lossless in y): the power through two vertical
    planes on the same side of the source agrees up to the half-cell
    staggering of the reference's flux formula."""
    g = O.Grid.create(0.02, (20, 0), (0, 4.0), (0, 1.0))
    d = O.Device(g, O.omega_of_wavelength(1.55))
    d.eps_r[:, 20:31] = 12.25
    d.src[50, :] = 1.0
    f = O.solve(d, O.TM)
    p1 = O.flux_surface_integral_tm_x(f, 2.01, 0.5, 0.6)
    p2 = O.flux_surface_integral_tm_x(f, 3.01, 0.5, 0.6)
    assert p1 > 0 and abs(p1 - p2) / abs(p1) < 2e-2


# --- eigenfrequency (src/solver/eigen.jl) ------------------------------------
def test_eigenfrequency_fft_closed_form():
    Nx, Ny, dx = 12, 10, 0.05
    g = O.Grid.create(dx, (0, 0), (0, Nx * dx), (0, Ny * dx))
    om0 = 2 * np.pi * C0n / 1.2
    d = O.Device(g, om0)
    eps = 2.0
    d.eps_r[:] = eps
    w, fields, lam, sigma = O.eigenfrequency(d, O.TM, 4)
    P, Q = np.meshgrid(np.arange(Nx), np.arange(Ny), indexing="ij")
    lam_all = (-4 * (np.sin(np.pi * P / Nx) ** 2 / dx ** 2 + np.sin(np.pi * Q / Ny) ** 2 / dx ** 2)
               / eps).ravel()
    near = lam_all[np.argsort(np.abs(lam_all - sigma))]
    # degenerate clusters: compare the sorted distances
    assert np.allclose(np.sort(np.abs(lam - sigma)), np.sort(np.abs(near[:4] - sigma)), rtol=1e-9)
    # omega = sqrt(-lambda/(mu0 eps0)) (eigen.jl:87)
    assert np.allclose(w ** 2, -lam / MU0n / EPS0n, rtol=1e-12,
                       atol=1e-12 * np.abs(lam).max() / MU0n / EPS0n)


def test_eigenfrequency_kronecker_with_pml():
    Nx, Ny, dx = 14, 12, 0.05
    g = O.Grid.create(dx, (3, 3), (0, Nx * dx), (0, Ny * dx))
    om0 = 2 * np.pi * C0n / 1.0
    d = O.Device(g, om0)
    d.eps_r[:] = 1.0
    w, fields, lam, sigma = O.eigenfrequency(d, O.TM, 5)
    sxf, sxb, syf, syb = O.sfactor_vectors(g, om0)
    a = np.linalg.eigvals(_op1d(Nx, dx, sxf, sxb))
    b = np.linalg.eigvals(_op1d(Ny, dx, syf, syb))
    allv = (a[:, None] + b[None, :]).ravel()
    near = allv[np.argsort(np.abs(allv - sigma))][:5]
    assert np.allclose(np.sort(np.abs(lam - sigma)), np.sort(np.abs(near - sigma)), rtol=1e-8)


# --- adjoint gradient (new; checked by central finite differences) ----------
def test_adjoint_gradient_matches_fd():
    g = O.Grid.create(0.02, (4, 4), (0, 0.5), (0, 0.4))
    om = O.omega_of_wavelength(1.55)
    rng = np.random.default_rng(11)
    eps = 1 + 11.25 * rng.random(g.N)
    src = np.zeros(g.N, dtype=complex)
    src[6, 10] = 1j
    mvec = np.zeros(g.N, dtype=complex)
    mvec[18, 5:15] = np.exp(-((np.arange(5, 15) - 10) / 3.0) ** 2)
    m = mvec.ravel(order="F")

    def J(e_r):
        A = O.assemble(g, e_r, om, O.TM)
        e, _ = O.linear_solve(A, O.rhs(om, src))
        s = m @ e
        return abs(s) ** 2, A, e, s

    J0, A, e, s = J(eps)
    lam, _ = O.linear_solve(A, m, trans="T")
    grad = -np.real(2 * np.conj(s) * lam * om ** 2 * EPS0n * e)
    h = 1e-4
    for k in [(6, 12), (12, 8), (18, 10), (3, 3)]:
        kk = k[0] + k[1] * g.N[0]
        ep, em = eps.copy(), eps.copy()
        ep[k] += h
        em[k] -= h
        fd = (J(ep)[0] - J(em)[0]) / (2 * h)
        assert abs(grad[kk] - fd) <= 1e-6 * max(abs(fd), 1e-3 * np.abs(grad).max())


# --- mode sources: 1-D eigenmode (eigen.jl:6-29), get_modes / setup_mode! ----
L0 = O.DEFAULT_L0
K0SQ = lambda om: om ** 2 * O.EPS0 * L0 * O.MU0 * L0  # noqa: E731


@pytest.mark.parametrize("pol", [O.TM, O.TE])
@pytest.mark.parametrize("n", [37, 40])
def test_mode_operator_uniform_spectrum(pol, n):
    """Uniform eps: A is circulant with eigenvalues
    k0^2 eps - (4/dx^2) sin^2(pi p / n), p = 0..n-1, for TM and TE alike
    (the TE factors eps/avg(eps) are 1).  Complex eps included."""
    om = O.omega_of_wavelength(1.55)
    eps = 4.0 + 0.3j
    dh = 0.02
    A = O.mode_operator_1d(np.full(n, eps), dh, om, pol).toarray()
    dx = (n * dh) / n
    p = np.arange(n)
    ref = K0SQ(om) * eps - 4.0 / dx ** 2 * np.sin(np.pi * p / n) ** 2
    got = np.linalg.eigvals(A)
    err = np.abs(np.sort_complex(got) - np.sort_complex(ref)).max() / np.abs(ref).max()
    assert err <= 1e-12, err


def _slab(a: int, clad: int, ec: float, ecl: float):
    """2a+1 core cells (eps ec) centred in 2*clad cladding cells (eps ecl)."""
    return np.concatenate([np.full(clad, ecl), np.full(2 * a + 1, ec), np.full(clad, ecl)])


def _slab_roots(g, lo, hi, nscan=4000):
    xs = np.linspace(lo, hi, nscan + 2)[1:-1]
    vals = np.array([g(x) for x in xs])
    roots = []
    for i in range(len(xs) - 1):
        if np.sign(vals[i]) != np.sign(vals[i + 1]):
            from scipy.optimize import brentq
            roots.append(brentq(g, xs[i], xs[i + 1], xtol=1e-15 * hi, rtol=1e-15))
    return roots


@pytest.mark.parametrize("pol", [O.TM, O.TE])
def test_mode_slab_lattice_dispersion(pol):
    """Even guided modes of a symmetric slab on the lattice, derived without
    the matrix: core f_i = cos(theta i) with 2cos(theta) - 2 = dx^2 (b2 - k0^2 ec),
    cladding f_{a+k} = D r^k with r + 1/r - 2 = dx^2 (b2 - k0^2 ecl), |r| < 1.
    The two rows at the interface fix b2:
      TM: r cos(theta a) = cos(theta (a+1))
      TE: with em = (ec+ecl)/2 (grid_average), rho = ecl/em,
          D = rho f_a / (1 + rho r - r),
          (ec/em)(D r - f_a) - (f_a - f_{a-1}) - (2cos(theta) - 2) f_a = 0.
    Every root must be an eigenvalue of the restated operator, and
    eigenmode_1d with a guess near it must return it.  The cladding is wide
    enough that the periodic wrap is below round-off."""
    om = O.omega_of_wavelength(1.55)
    k2 = K0SQ(om)
    a, clad, ec, ecl, dh = 7, 150, 12.25, 2.25, 0.02
    dx = dh
    em = 0.5 * (ec + ecl)
    rho = ecl / em

    def g(b2):
        c = 1 + dx ** 2 * (b2 - k2 * ec) / 2
        th = np.arccos(c)
        q = 1 + dx ** 2 * (b2 - k2 * ecl) / 2
        r = q - np.sqrt(q * q - 1)
        fa, fa1, fp1 = np.cos(th * a), np.cos(th * (a - 1)), np.cos(th * (a + 1))
        if pol == O.TM:
            return r * fa - fp1
        D = rho * fa / (1 + rho * r - r)
        return (ec / em) * (D * r - fa) - (fa - fa1) - (2 * c - 2) * fa

    roots = _slab_roots(g, k2 * ecl, k2 * ec)
    assert len(roots) >= 1
    eps1d = _slab(a, clad, ec, ecl)
    lam = np.linalg.eigvals(O.mode_operator_1d(eps1d, dh, om, pol).toarray())
    for b2 in roots:
        assert np.abs(lam - b2).min() <= 1e-11 * b2
        neff = np.sqrt(b2) / np.sqrt(k2) * (1 + 1e-4)
        beta, v, _ = O.eigenmode_1d(eps1d, dh, om, pol, neff, 1)
        assert abs(beta[0] ** 2 - b2) <= 1e-11 * b2
        # the guided mode is even about the core centre
        vv = np.abs(v[:, 0])
        assert np.abs(vv - vv[::-1]).max() <= 1e-9


def test_mode_slices_and_setup_mode():
    """Slice indices of device.jl:125-142 by hand for the waveguide case of the
    reference's tests (runtests.jl:23-27): Grid(0.02, [15 15], [0 10], [-1 1])
    is 500 x 100; Point(1.0, 0) -> (51, 51); width 0.8 -> 40 points -> odd 41,
    rows 31..71 of column 51.  y-normal: width 0.3 -> 15 points (odd), columns
    44..58 of row 51.  setup_mode! adds a unit-norm, non-negative profile that
    is even for a centred guide, and leaves every other cell alone."""
    g = O.Grid.create(0.02, (15, 15), (0.0, 10.0), (-1.0, 1.0))
    assert g.N == (500, 100)
    ix, iy, dh = O.slice_indices(g, (1.0, 0.0), O.XHAT, 0.8)
    assert list(ix) == [51] and list(iy) == list(range(31, 72)) and dh == g.dy
    ix, iy, dh = O.slice_indices(g, (1.0, 0.0), O.YHAT, 0.3)
    assert list(ix) == list(range(44, 59)) and list(iy) == [51] and dh == g.dx
    d = O.Device(g, O.omega_of_wavelength(1.55))
    d.eps_r[:, 43:58] = 12.0   # guide rows 44..58 (1-based), centred on row 51
    d.src[3, 3] = 5.0
    O.setup_mode(d, O.TM, float(d.omega[0]), 3.5, (1.0, 0.0), O.XHAT, 0.8)
    prof = d.src[50, 30:71]
    assert np.all(prof.real >= 0) and np.all(prof.imag == 0)
    assert abs(np.linalg.norm(prof) - 1) <= 1e-14
    assert np.abs(prof - prof[::-1]).max() <= 1e-9
    rest = d.src.copy()
    rest[50, 30:71] = 0
    rest[3, 3] -= 5.0
    assert not rest.any()
    with pytest.raises(IndexError):
        O.get_modes(d, O.TM, float(d.omega[0]), 3.5, 1, (1.0, -0.95), O.XHAT, 0.8)


# --- modulated device, MF-FDFD (modulation.jl:35-119) ------------------------
def test_modulated_fourier_closed_form():
    """Uniform eps and deps, npml = 0: every 2-D Fourier mode (p, q) of the
    coupled system is an independent (2K+1) x (2K+1) tridiagonal system
      M[j][j]   = -(4/mu0)(sin^2(pi p/Nx)/dx^2 + sin^2(pi q/Ny)/dy^2) + wn_j^2 eps0 eps
      M[j][j+1] = 0.5 wn_j^2 eps0 conj(deps),  M[j+1][j] = 0.5 wn_{j+1}^2 eps0 deps
    (the backward-forward product has the same symbol as forward-backward).
    Solving mode by mode through the FFT must reproduce the oracle's sparse
    solve; a swapped conj, a wrong sideband index or a wrong omega in the
    coupling fails by O(1)."""
    Nx, Ny, K = 12, 10, 2
    g = O.Grid.create(0.05, (0, 0), (0, Nx * 0.05), (0, Ny * 0.05))
    om, Om = O.omega_of_wavelength(1.55), 1.7e14
    d = O.ModulatedDevice(g, om, Om, K, sharedpml=True)
    d.eps_r[:] = 3.0 + 0.1j
    dep = 0.7 + 0.4j
    d.deps_r[:] = dep
    rng = np.random.default_rng(5)
    d.src[:] = rng.standard_normal(g.N) + 1j * rng.standard_normal(g.N)
    A, b, wn, _ = O.modulated_system(d, om)
    x, _ = O.linear_solve(A, b)
    eps0, mu0, _ = O.normalize_parameters(g.L0)
    nf = 2 * K + 1
    N = len(g)
    bhat = np.fft.fft2(d.src * 1j * om)   # Nx x Ny (x fastest in the flat layout)
    p = np.arange(Nx)[:, None]
    q = np.arange(Ny)[None, :]
    lap = -(4 / mu0) * (np.sin(np.pi * p / Nx) ** 2 / g.dx ** 2 + np.sin(np.pi * q / Ny) ** 2 / g.dy ** 2)
    xhat = np.zeros((nf, Nx, Ny), dtype=complex)
    for ip in range(Nx):
        for iq in range(Ny):
            M = np.zeros((nf, nf), dtype=complex)
            for j in range(nf):
                M[j, j] = lap[ip, iq] + wn[j] ** 2 * eps0 * d.eps_r[0, 0]
                if j + 1 < nf:
                    M[j, j + 1] = 0.5 * wn[j] ** 2 * eps0 * np.conj(dep)
                    M[j + 1, j] = 0.5 * wn[j + 1] ** 2 * eps0 * dep
            rhs = np.zeros(nf, dtype=complex)
            rhs[K] = bhat[ip, iq]
            xhat[:, ip, iq] = np.linalg.solve(M, rhs)
    for j in range(nf):
        ref = np.fft.ifft2(xhat[j]).ravel(order="F")
        got = x[j * N:(j + 1) * N]
        assert np.linalg.norm(got - ref) <= 1e-11 * np.linalg.norm(x)


def test_sideband_operator_transpose_identity_and_cell_loop():
    """With PML the sideband operator Sxb dxb Sxf dxf + Syb dyb Syf dyf + w^2 T
    satisfies A^T = W^-1 A W for W = diag(sxb_x syb_y) (dxf^T = -dxb), which
    the driven ordering (W = sxf syf) does not; and it equals a per-cell loop
    built from the difference definitions."""
    g = O.Grid.create(0.05, (4, 3), (0, 0.8), (0, 0.6))
    om = O.omega_of_wavelength(1.55)
    eps = 1 + 5 * np.random.default_rng(2).random(g.N)
    A, (Sxf, Sxb, Syf, Syb) = O.sideband_operator(g, eps, om, om * 1.1)
    sxb = np.diag(Sxb.toarray()).reshape(g.N, order="F")
    syb = np.diag(Syb.toarray()).reshape(g.N, order="F")
    W = sp_diag((sxb * syb).ravel(order="F"))
    lhs = A.T.toarray()
    rhs = (np.linalg.inv(W) @ A.toarray() @ W)
    assert np.abs(lhs - rhs).max() <= 1e-12 * np.abs(lhs).max()
    # the driven-form W does not work for this operator
    sxf = np.diag(Sxf.toarray()).reshape(g.N, order="F")
    syf = np.diag(Syf.toarray()).reshape(g.N, order="F")
    Wf = sp_diag((sxf * syf).ravel(order="F"))
    assert np.abs(lhs - np.linalg.inv(Wf) @ A.toarray() @ Wf).max() > 1e-6 * np.abs(lhs).max()
    # per-cell loop
    eps0, mu0, _ = O.normalize_parameters(g.L0)
    sx_f, sx_b = sxf[:, 0], sxb[:, 0]
    sy_f, sy_b = syf[0, :], syb[0, :]
    Nx, Ny = g.N
    e = np.random.default_rng(3).standard_normal(g.N) + 0j
    y = np.zeros(g.N, dtype=complex)
    for i in range(Nx):
        for j in range(Ny):
            ip, im, jp, jm = (i + 1) % Nx, (i - 1) % Nx, (j + 1) % Ny, (j - 1) % Ny
            lx = sx_b[i] * (sx_f[i] * (e[ip, j] - e[i, j]) - sx_f[im] * (e[i, j] - e[im, j])) / g.dx ** 2
            ly = sy_b[j] * (sy_f[j] * (e[i, jp] - e[i, j]) - sy_f[jm] * (e[i, j] - e[i, jm])) / g.dy ** 2
            y[i, j] = (lx + ly) / mu0 + om ** 2 * eps0 * eps[i, j] * e[i, j]
    got = (A @ e.ravel(order="F")).reshape(g.N, order="F")
    assert np.abs(got - y).max() <= 1e-12 * np.abs(y).max()


def sp_diag(v):
    return np.diag(v)


# --- flux, Poynting, S-parameters (flux.jl:15-91), TE and y-normal completed --
def test_flux_hand_cases_te_and_y_normal():
    """Hand values: a 4 x 6 grid of unit cells, constant fields except one
    line, so the averaged neighbour (index + 1) and the conjugated factor are
    visible in the result."""
    g = O.Grid.create(1.0, (0, 0), (0, 4.0), (0, 6.0))
    # TE, x normal: 1/2 Re(Ey conj(avg Hz)); column 1 averaged with column 2
    d = np.zeros((4, 6, 3), dtype=complex)
    d[:, :, 0] = 2.0          # Hz
    d[2, :, 0] = 4.0 + 2.0j
    d[:, :, 2] = 1.0 + 1.0j   # Ey
    v = O.flux_surface_integral(O.FieldTE(g, 1.0, d), 1.5, 3.0, 1.0, O.XHAT)
    # avg Hz = 3 + 1j; Re((1+1j)(3-1j)) = 4; two rows (y = 2.5, 3.5); dy = 1
    assert v == pytest.approx(2 * 0.5 * 4.0)
    # TM, y normal: 1/2 Re(avg Ez conj(Hx)); row 2 averaged with row 3
    d = np.zeros((4, 6, 3), dtype=complex)
    d[:, :, 0] = 1.0j         # Ez
    d[:, 3, 0] = 3.0j
    d[:, :, 1] = 2.0j         # Hx
    v = O.flux_surface_integral(O.FieldTM(g, 1.0, d), 2.0, 2.5, 1.0, O.YHAT)
    # avg Ez = 2j; Re(2j * conj(2j)) = 4; columns x = 1.5, 2.5; dx = 1
    assert v == pytest.approx(2 * 0.5 * 4.0)
    # TE, y normal: -1/2 Re(Ex conj(avg Hz))
    d = np.zeros((4, 6, 3), dtype=complex)
    d[:, :, 0] = 1.0
    d[:, 3, 0] = -3.0         # avg Hz over rows 2, 3 = -1
    d[:, :, 1] = 5.0          # Ex
    v = O.flux_surface_integral(O.FieldTE(g, 1.0, d), 2.0, 2.5, 1.0, O.YHAT)
    assert v == pytest.approx(2 * (-0.5) * (5.0 * -1.0))
    # Poynting maps: TE with constant fields, Sx = 1/2 Re(Ey conj Hz)
    d = np.zeros((4, 6, 3), dtype=complex)
    d[:, :, 0], d[:, :, 1], d[:, :, 2] = 2.0, 1.0j, 3.0
    sx, sy = O.poynting(O.FieldTE(g, 1.0, d))
    assert np.allclose(sx, 3.0) and np.allclose(sy, -0.5 * np.real(1j * 2.0))


@pytest.mark.parametrize("pol", [O.TM, O.TE])
def test_flux_conservation_te_and_y_normal(pol):
    """Lossless, source-free region between two planes: the same power
    crosses both (up to the half-cell staggering of the flux formula), for
    x-normal planes and, on the transposed device, y-normal planes."""
    g = O.Grid.create(0.02, (20, 0), (0, 4.0), (0, 1.0))
    d = O.Device(g, O.omega_of_wavelength(1.55))
    d.eps_r[:, 20:31] = 12.25
    d.src[50, :] = 1.0
    f = O.solve(d, pol)
    p1 = O.flux_surface_integral(f, 2.01, 0.5, 0.6, O.XHAT)
    p2 = O.flux_surface_integral(f, 3.01, 0.5, 0.6, O.XHAT)
    assert p1 > 0 and abs(p1 - p2) / abs(p1) < 2e-2
    gt = O.Grid.create(0.02, (0, 20), (0, 1.0), (0, 4.0))
    dt = O.Device(gt, d.omega)
    dt.eps_r[:] = d.eps_r.T
    dt.src[:] = d.src.T
    ft = O.solve(dt, pol)
    q1 = O.flux_surface_integral(ft, 0.5, 2.01, 0.6, O.YHAT)
    q2 = O.flux_surface_integral(ft, 0.5, 3.01, 0.6, O.YHAT)
    assert q1 > 0 and abs(q1 - q2) / abs(q1) < 2e-2
    # x <-> y transposition maps Ez, Hy to Ez, -Hx (TM) and Hz, Ey to Hz, -Ex
    # (TE): the y-normal power is the x-normal power of the original device
    assert abs(q1 - p1) / abs(p1) < 1e-10


def test_scattering_parameters_modal_propagation():
    """A straight lossless guide excited by its own mode: the overlap of the
    computed field with the mode keeps its magnitude along the guide, and its
    phase advances by the lattice wavenumber k = arccos(1 - b^2 dx^2 / 2)/dx
    of the mode's beta (the exact dispersion of the x differences), to 1e-5
    (PML reflections and the slice's periodic wrap are below that)."""
    g = O.Grid.create(0.02, (15, 15), (0.0, 6.0), (-1.0, 1.0))
    d = O.Device(g, O.omega_of_wavelength(1.55))
    d.eps_r[:, 43:58] = 12.25   # rows 44..58 (1-based), centred on row 51
    w = 1.6
    O.add_mode(d, O.Mode(O.TM, O.XHAT, 3.2, (1.0, 0.0), w))
    f = O.solve(d, O.TM)          # driven by the input mode only
    # ports: the input and two output planes (modes are sources in solve, so
    # the output ports are added after it)
    O.add_mode(d, O.Mode(O.TM, O.XHAT, 3.2, (3.0, 0.0), w))
    O.add_mode(d, O.Mode(O.TM, O.XHAT, 3.2, (4.0, 0.0), w))
    S = O.scattering_parameters([f], d)
    beta, _, ix, _ = O.get_modes(d, O.TM, float(d.omega[0]), 3.2, 1, (3.0, 0.0), O.XHAT, w)
    k = np.arccos(1 - (beta[0].real * g.dx) ** 2 / 2) / g.dx
    _, _, ix2, _ = O.get_modes(d, O.TM, float(d.omega[0]), 3.2, 1, (4.0, 0.0), O.XHAT, w)
    D = (ix2[0] - ix[0]) * g.dx
    s1, s2 = S[0, 1], S[0, 2]
    assert abs(abs(s2) / abs(s1) - 1) < 1e-5
    # the solved fields go as exp(-i k x) downstream of the source
    dphi = np.angle(s2 / s1)
    assert abs(np.angle(np.exp(1j * (dphi + k * D)))) < 1e-5


# --- full-vector waveguide modes (eigen.jl:31-66) -----------------------------
def test_vector_mode_uniform_spectrum():
    """Uniform eps: the polarization coupling vanishes (the periodic
    differences commute) and both blocks are Laplacians, so the spectrum is
    k0^2 eps - (4/dx^2) sin^2(pi p/Nx) - (4/dy^2) sin^2(pi q/Ny), each twice."""
    g = O.Grid.create(0.05, (0, 0), (0, 0.6), (0, 0.5))
    om = O.omega_of_wavelength(1.55)
    eps = 2.5 + 0.1j
    A = O.vector_mode_operator(g, np.full(g.N, eps), om).toarray()
    Nx, Ny = g.N
    P, Q = np.meshgrid(np.arange(Nx), np.arange(Ny), indexing="ij")
    base = (K0SQ(om) * eps - 4 / g.dx ** 2 * np.sin(np.pi * P / Nx) ** 2
            - 4 / g.dy ** 2 * np.sin(np.pi * Q / Ny) ** 2).ravel()
    ref = np.sort_complex(np.concatenate([base, base]))
    got = np.sort_complex(np.linalg.eigvals(A))
    assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref).max()


@pytest.mark.parametrize("axis", ["y", "x"])
def test_vector_mode_slab_reduces_to_1d_tm(axis):
    """A slab uniform along x (eps varies with y only): (Hx, Hy) = (0, h(y))
    uniform in x is an exact eigenvector, with h the 1-D TM mode of the eps(y)
    slice (A_yy = k0^2 avg_x eps + dyb dyf, and avg_x eps = eps; A_xy kills
    x-uniform fields).  Likewise (h(x), 0) for a slab uniform along y.  The
    fundamental vector mode's beta must equal the 1-D TM beta."""
    om = O.omega_of_wavelength(1.55)
    prof = np.concatenate([np.full(20, 2.25), np.full(9, 12.25), np.full(20, 2.25)]) + 0j
    if axis == "y":
        g = O.Grid.create(0.02, (0, 0), (0, 0.12), (0, 0.98))
        eps = np.tile(prof, (g.N[0], 1))
    else:
        g = O.Grid.create(0.02, (0, 0), (0, 0.98), (0, 0.12))
        eps = np.tile(prof[:, None], (1, g.N[1]))
    b1, v1, _ = O.eigenmode_1d(prof, 0.02, om, O.TM, 3.0, 1)
    beta, fields, _, _ = O.vector_eigenmode(g, eps, om, b1[0] / np.sqrt(K0SQ(om)) * (1 + 1e-3), 1)
    assert abs(beta[0] - b1[0]) <= 1e-10 * abs(b1[0])
    f = fields[0].data
    h = f[:, :, 4] if axis == "y" else f[:, :, 3]   # Hy or Hx carries the mode
    other = f[:, :, 3] if axis == "y" else f[:, :, 4]
    assert np.abs(other).max() <= 1e-8 * np.abs(h).max()
    line = h[0, :] if axis == "y" else h[:, 0]
    line = line / np.linalg.norm(line)
    assert np.abs(np.abs(line) - np.abs(v1[:, 0])).max() <= 1e-8


# --- PML grading exponent (src/pml.jl:1-10) ----------------------------------
def test_sfactor_grading_exponent():
    """sigma(l) = sigma_max (l/T)^m with m = 3.5: the ratio of the
    conductivities at two depths is (l1/l2)^m whatever sigma_max is, so this
    pins m, which the integral test cannot (there sigma_max ~ (m + 1) cancels).
    Forward factor of cell 1 sits at l = dw (Npml - 1 + 0.5) = 0.39 um,
    backward factor of cell 10 at l = dw (Npml - 10 + 1) = 0.22 um
    (Npml = 20, dw = 0.02 um, T = 0.4 um).  (0.39/0.22)^3.5 = 7.417315...,
    whereas m = 3 would give 5.5709 and m = 4 would give 9.8757."""
    g = O.Grid.create(0.02, (20, 20), (0, 4.0), (0, 4.0))
    om = O.omega_of_wavelength(1.55)
    sf = O.create_sfactor(O.XHAT, O.FORWARD, g, om)
    sb = O.create_sfactor(O.XHAT, O.BACKWARD, g, om)
    sig = lambda s: -s.imag * om * EPS0n  # noqa: E731
    assert sig(sf[0]) / sig(sb[9]) == pytest.approx(7.4173153, rel=1e-7)
    # absolute value at the deepest forward sample: sigma_max = 4.5*12/(2 eta0 T)
    # = 0.17917326 with eta0 = 376.7303132 ohm; times (0.975)^3.5 = 0.91520030
    assert sig(sf[0]) == pytest.approx(0.17917326 * 0.91520030, rel=1e-7)
    # right edge mirrors the left: cell N (1-based) sits at l = dw (N - (N - 20) - 0.5)
    assert sf[-1] == pytest.approx(sf[0], rel=1e-14)
    assert sb[-1] == pytest.approx(sb[1], rel=1e-14)  # dw (20 - 1) both


# --- TE field recovery: the discrete Faraday identity -----------------------
def test_te_fields_satisfy_faraday():
    """With A_TE hz = i w src (driven.jl:45-46) and the driven field recovery
    Ex = Tey^-1 dyb hz/(i w), Ey = -Tex^-1 dxb hz/(i w) (driven.jl:50-51),
    dyf Ex - dxf Ey - i w mu0' Hz = src holds exactly.  It holds only with
    that pairing of the averaged permittivities (y-average for Ex, x-average
    for Ey); on a random non-uniform eps a swapped pairing, a missing average
    or a wrong sign of either component fails it."""
    g = O.Grid.create(0.02, (6, 5), (0, 0.8), (0, 0.7))
    d = O.Device(g, O.omega_of_wavelength(1.55))
    rng = np.random.default_rng(21)
    d.eps_r[:] = 1 + 11.25 * rng.random(g.N)
    d.src[13, 17] = 1j
    d.src[30, 9] = 0.5
    f = O.solve(d, O.TE)
    om = float(d.omega[0])
    dxf, dxb, dyf, dyb = O.stretched_deltas(g, om)
    hz, ex, ey = (f.data[:, :, c].ravel(order="F") for c in range(3))
    lhs = dyf @ ex - dxf @ ey - 1j * om * MU0n * hz
    src = d.src.ravel(order="F")
    assert np.linalg.norm(lhs - src) / np.linalg.norm(src) < 1e-9
    # the swapped pairing does not satisfy it (the test can fail)
    Texi = 1 / O.grid_average(EPS0n * d.eps_r, O.XHAT).ravel(order="F")
    Teyi = 1 / O.grid_average(EPS0n * d.eps_r, O.YHAT).ravel(order="F")
    ex_sw = Texi * (dyb @ hz) / (1j * om)
    ey_sw = -Teyi * (dxb @ hz) / (1j * om)
    bad = dyf @ ex_sw - dxf @ ey_sw - 1j * om * MU0n * hz
    assert np.linalg.norm(bad - src) / np.linalg.norm(src) > 1e-2


# --- TE eigenfrequency (src/solver/eigen.jl:97-114) ---------------------------
def test_te_eigenfrequency_fft_closed_form():
    """npml = 0, uniform eps: the TE eigen operator dxf Tex^-1 dxb +
    dyf Tey^-1 dyb is (1/(eps0' eps)) times the periodic Laplacian, so
    lambda = -(4/(eps0' eps)) (sin^2(pi p/Nx)/dx^2 + sin^2(pi q/Ny)/dy^2),
    sigma = -w0^2 mu0' and omega = sqrt(-lambda/mu0') (eigen.jl:104-105)."""
    Nx, Ny, dx = 12, 10, 0.05
    g = O.Grid.create(dx, (0, 0), (0, Nx * dx), (0, Ny * dx))
    om0 = 2 * np.pi * C0n / 1.2
    d = O.Device(g, om0)
    eps = 2.0
    d.eps_r[:] = eps
    w, fields, lam, sigma = O.eigenfrequency(d, O.TE, 4)
    assert sigma == pytest.approx(-om0 ** 2 * MU0n, rel=1e-15)
    P, Q = np.meshgrid(np.arange(Nx), np.arange(Ny), indexing="ij")
    lam_all = (-4 * (np.sin(np.pi * P / Nx) ** 2 / dx ** 2 + np.sin(np.pi * Q / Ny) ** 2 / dx ** 2)
               / (EPS0n * eps)).ravel()
    near = lam_all[np.argsort(np.abs(lam_all - sigma))]
    assert np.allclose(np.sort(np.abs(lam - sigma)), np.sort(np.abs(near[:4] - sigma)), rtol=1e-9)
    assert np.allclose(w ** 2, -lam / MU0n, rtol=1e-12, atol=1e-12 * np.abs(lam).max() / MU0n)


def test_te_eigenfrequency_kronecker_with_pml():
    """Uniform eps with PML: the TE eigen operator is (1/(eps0' eps)) times the
    Kronecker sum of the 1-D stretched operators, so its eigenvalues near
    sigma are sums of 1-D eigenvalues scaled by 1/(eps0' eps)."""
    Nx, Ny, dx = 14, 12, 0.05
    g = O.Grid.create(dx, (3, 3), (0, Nx * dx), (0, Ny * dx))
    om0 = 2 * np.pi * C0n / 1.0
    d = O.Device(g, om0)
    eps = 1.7
    d.eps_r[:] = eps
    w, fields, lam, sigma = O.eigenfrequency(d, O.TE, 5)
    sxf, sxb, syf, syb = O.sfactor_vectors(g, om0)
    a = np.linalg.eigvals(_op1d(Nx, dx, sxf, sxb))
    b = np.linalg.eigvals(_op1d(Ny, dx, syf, syb))
    allv = (a[:, None] + b[None, :]).ravel() / (EPS0n * eps)
    near = allv[np.argsort(np.abs(allv - sigma))][:5]
    assert np.allclose(np.sort(np.abs(lam - sigma)), np.sort(np.abs(near - sigma)), rtol=1e-8)
    # eigenfields (eigen.jl:106-109) of uniform eps: dyf Ex - dxf Ey = i w mu0' Hz
    # (the near-null mode of the periodic operator, |lambda| ~ 1e-14 |sigma|,
    # has no usable identity and is skipped)
    dxf, dxb, dyf, dyb = O.stretched_deltas(g, om0)
    checked = 0
    for f, wq, lq in zip(fields, w, lam):
        if abs(lq) < 1e-6 * abs(sigma):
            continue
        checked += 1
        hz, ex, ey = (f.data[:, :, c].ravel(order="F") for c in range(3))
        lhs = dyf @ ex - dxf @ ey
        assert np.linalg.norm(lhs - 1j * wq * MU0n * hz) <= 1e-6 * np.linalg.norm(lhs)
    assert checked >= 4
