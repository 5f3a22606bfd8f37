"""The workload generators of the five benchmark configurations (CPU):
sizes, region bounds in 1-based cell indices, seeds and source placement are
checked against hand-derived values, so a GPU test that uses them runs the
stated workload."""
import math

import numpy as np
import pytest

import fdfd_oracle as O


def test_c1_geometry():
    d = O.config_c1()
    assert d.grid.N == (200, 120)
    rows = np.nonzero(d.eps_r[0].real > 1)[0] + 1
    assert (rows.min(), rows.max()) == (54, 68)
    assert np.count_nonzero(d.src) == 1 and d.src[50, 60] == 1j


def test_c2_geometry():
    d = O.config_c2()
    assert d.grid.N == (1024, 1024) and d.grid.Npml == (20, 20)
    e = d.eps_r.real
    design = np.zeros_like(e, dtype=bool)
    design[212:812, 212:812] = True
    assert set(np.unique(e[design])) == {1.0, 12.25}
    # 8x8 pixels: constant on each block
    blk = e[212:812, 212:812].reshape(75, 8, 75, 8)
    assert np.all(blk == blk[:, :1, :, :1])
    rows = np.nonzero(e[0] > 1)[0] + 1
    assert (rows.min(), rows.max()) == (506, 520)
    assert np.all(e[812:, 505:520] == 12.25)
    ys = np.nonzero(d.src[50])[0] + 1
    assert (ys.min(), ys.max(), ys.size) == (501, 525, 25) and np.count_nonzero(d.src) == 25


def test_c3_geometry():
    jobs = O.config_c3()
    assert [d.grid.N for d, _, _ in jobs] == [(256, 256), (384, 512), (512, 512)]
    lam = 2 * math.pi * O.normalize_parameters()[2] / jobs[0][1]
    assert lam[0] == pytest.approx(1.50) and lam[-1] == pytest.approx(1.60) and lam.size == 64
    # design side round(600/1024 * min(N)/8)*8: 152, 224, 304 (37.5 rounds to even 38)
    for (d, oms, srcs), side in zip(jobs, (152, 224, 304)):
        Nx, Ny = d.grid.N
        e = d.eps_r.real
        x0, y0 = (Nx - side) // 2, (Ny - side) // 2
        assert set(np.unique(e[x0:x0 + side, y0:y0 + side])) == {1.0, 12.25}
        assert np.all(e[x0 - 1, : Ny // 2 - 7] == 1.0)  # left of the design, off the guide
        assert np.all(e[:x0, Ny // 2 - 7:Ny // 2 + 8] == 12.25)
        assert not d.src.any()
        assert srcs.shape == (Nx, Ny, 16)
        for r in range(16):
            ix, iy = np.nonzero(srcs[:, :, r])
            assert ix.size == 1 and srcs[ix[0], iy[0], r] == 1j
            assert 21 <= ix[0] + 1 <= int(0.2 * Nx) and 21 <= iy[0] + 1 <= Ny - 20
    # the permittivity seeds differ per grid
    assert not np.array_equal(jobs[1][0].eps_r[:256, :256], jobs[2][0].eps_r[:256, :256])


def test_c4_geometry():
    d, m, cells, fd = O.config_c4()
    Nx, Ny = d.grid.N
    assert (Nx, Ny) == (1024, 768)
    ix, iy = cells % Nx + 1, cells // Nx + 1
    assert (ix.min(), ix.max(), iy.min(), iy.max()) == (385, 640, 257, 512)
    vals = d.eps_r.ravel(order="F")[cells].real
    assert vals.min() >= 1.0 and vals.max() <= 12.25 and np.std(vals) > 3.0
    assert len(set(fd.tolist())) == 32 and set(fd.tolist()) <= set(cells.tolist())
    rows = np.nonzero(d.eps_r[0].real > 1)[0] + 1
    assert (rows.min(), rows.max()) == (378, 392)
    ys = np.nonzero(d.src[50])[0] + 1
    assert (ys.min(), ys.max()) == (373, 397)
    mf = m.reshape((Nx, Ny), order="F")
    cols = np.nonzero(np.abs(mf).sum(axis=1))[0] + 1
    assert list(cols) == [954]
    # Gaussian waist 0.25 um centred on the guide (cell centres 377..391, 0-based)
    yc = d.grid.yc()
    peak = int(np.argmax(np.abs(mf[953])))
    assert peak in (383, 384)
    k = 384 + int(round(0.25 / 0.02))
    assert abs(mf[953, k]) == pytest.approx(math.exp(-((yc[k] - yc[384]) / 0.25) ** 2), rel=1e-12)


def test_c5_geometry():
    d = O.config_c5()
    g = d.grid
    assert g.N == (768, 768)
    e = d.eps_r.real
    xc, yc = g.xc(), g.yc()
    cell = lambda x, y: (int(np.argmin(np.abs(xc - x))), int(np.argmin(np.abs(yc - y))))  # noqa: E731
    assert e[cell(0.0, 0.0)] == 1.0                        # centre rod removed
    a = 0.5
    for (x, y) in [(a, 0), (-a, 0), (a / 2, a * math.sqrt(3) / 2), (-a / 2, -a * math.sqrt(3) / 2)]:
        assert e[cell(x, y)] == 8.9
    assert e[cell(a / 2, 0.0)] == 1.0                      # between rods
    assert set(np.unique(e)) == {1.0, 8.9}
    # around the rod at (3a/2, a sqrt(3)/2): every centre within r = 0.1 um is
    # rod, every centre between r and a - r is background
    X, Y = np.meshgrid(xc, yc, indexing="ij")
    dist = np.hypot(X - 1.5 * a, Y - a * math.sqrt(3) / 2)
    assert np.all(e[dist <= 0.1 - 1e-9] == 8.9) and np.all(e[(dist > 0.1 + 1e-9) & (dist < 0.4)] == 1.0)
    assert 70 < np.count_nonzero(dist <= 0.1) < 90        # ~ pi r^2 / dl^2 = 78.5 cells
    assert np.array_equal(e, e[::-1, ::-1])                # inversion symmetric
