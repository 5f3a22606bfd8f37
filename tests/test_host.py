"""Host-side mirror of the reference problem description (CPU only):
Grid / index maps, Device painting, FieldTM indexing, flux post-processing."""
import numpy as np
import pytest

import fdfd_oracle as O
from fdfd import core, flux


@pytest.mark.parametrize("dh,npml,xr,yr", [
    (0.02, (15, 15), (0.0, 4.0), (-1.2, 1.2)),
    (0.02, (20, 20), (-10.24, 10.24), (-10.24, 10.24)),
    (1.0, (0, 0), (0.0, 2.5), (0.0, 3.5)),
    (0.05, (3, 5), (-1.0, 1.35), (0.2, 1.0)),
])
def test_grid_matches_oracle(dh, npml, xr, yr):
    g = core.Grid(dh, npml, xr, yr)
    o = O.Grid.create(dh, npml, xr, yr)
    assert g.N == o.N and g.Npml == o.Npml
    assert g.dx == o.dx and g.dy == o.dy
    assert np.array_equal(g.xc(), o.xc()) and np.array_equal(g.yc(), o.yc())
    for x in np.linspace(xr[0] - 0.3, xr[1] + 0.3, 37):
        assert g.x2ind(x) == o.x2ind(x)
    for y in np.linspace(yr[0] - 0.3, yr[1] + 0.3, 29):
        assert g.y2ind(y) == o.y2ind(y)


def test_setup_src_point_and_line():
    g = core.Grid(0.02, (15, 15), (0.0, 4.0), (-1.2, 1.2))
    d = core.Device(g, 1e15)
    core.setup_src(d, core.Point(1.0, 0.0))
    assert d.src[50, 60] == 1j and np.count_nonzero(d.src) == 1  # device.jl:97 writes 1im
    d2 = core.Device(g, 1e15)
    core.setup_src(d2, core.Point(1.0, 0.0), normal=core.XHAT)
    assert np.all(d2.src[50, :] == 1j) and np.count_nonzero(d2.src) == g.N[1]


def test_setup_eps_region_and_callable():
    g = core.Grid(0.1, (0, 0), (0.0, 1.0), (0.0, 1.0))
    d = core.Device(g, 1e15)
    core.setup_eps(d, lambda x, y: x < 0.5, 4.0)
    assert np.all(d.eps_r[:5, :] == 4.0) and np.all(d.eps_r[5:, :] == 1.0)
    core.setup_eps(d, lambda x, y: y > 0.8, lambda x, y: 2.0 + x)
    assert d.eps_r[3, 9] == pytest.approx(2.35)


def test_field_indexing_layout():
    g = core.Grid(1.0, (0, 0), (0, 3.0), (0, 2.0))
    data = np.arange(3 * 2 * 3).reshape(3, 2, 3).astype(complex)
    f = core.FieldTM(g, 2.0, data)
    assert np.array_equal(f["Hx"], data[:, :, 1])
    assert f.components == ("Ez", "Hx", "Hy")


@pytest.mark.parametrize("center,width,normal", [
    ((0.31, 0.4), 0.2, core.XHAT), ((0.55, 0.21), 0.1, core.XHAT), ((0.5, 0.41), 0.3, core.YHAT),
    ((0.011, 0.75), 0.05, core.YHAT),
])
def test_flux_plane_matches_oracle_sum(center, width, normal):
    """flux_plane resolves the reference's line and interval (flux.jl:40-42):
    summing the oracle's flux formula over the resolved cells gives the
    oracle's flux_surface_integral."""
    rng = np.random.default_rng(2)
    g = core.Grid(0.02, (0, 0), (0, 1.0), (0, 0.8))
    og = O.Grid.create(0.02, (0, 0), (0, 1.0), (0, 0.8))
    data = rng.standard_normal((50, 40, 3)) + 1j * rng.standard_normal((50, 40, 3))
    nrm, idx, lo, hi = flux.flux_plane(g, core.Point(*center), width, normal)
    fo = O.FieldTM(og, 1.0, data)
    if normal == core.XHAT:
        u = (data[idx, lo:hi + 1, 0] + data[idx + 1, lo:hi + 1, 0]) / 2
        ref = np.sum(-0.5 * np.real(u * np.conj(data[idx, lo:hi + 1, 2]))) * g.dy
    else:
        u = (data[lo:hi + 1, idx, 0] + data[lo:hi + 1, idx + 1, 0]) / 2
        ref = np.sum(0.5 * np.real(u * np.conj(data[lo:hi + 1, idx, 1]))) * g.dx
    assert O.flux_surface_integral(fo, center[0], center[1], width, normal) == pytest.approx(ref, rel=1e-13)
    with pytest.raises(ValueError):
        flux.flux_plane(g, core.Point(5.0, 0.4), 0.2, core.XHAT)


@pytest.mark.parametrize("pt,normal,width", [
    ((1.0, 0.0), core.XHAT, 0.8), ((1.0, 0.0), core.YHAT, 0.3), ((2.5, 0.31), core.XHAT, 0.61),
    ((7.77, -0.5), core.YHAT, 1.0), ((1.0, 0.0), core.XHAT, 0.001),
])
def test_mode_slices_match_oracle(pt, normal, width):
    """Slice indices of get_modes (src/device.jl:125-142): host mirror and
    oracle agree, odd counts, the slice centred on coord2ind(pt)."""
    from fdfd import modes
    g = core.Grid(0.02, (15, 15), (0.0, 10.0), (-1.0, 1.0))
    o = O.Grid.create(0.02, (15, 15), (0.0, 10.0), (-1.0, 1.0))
    ix, iy, dh = modes.slice_indices(g, core.Point(*pt), normal, width)
    oix, oiy, odh = O.slice_indices(o, pt, normal, width)
    assert np.array_equal(ix, oix) and np.array_equal(iy, oiy) and dh == odh
    assert (ix.size * iy.size) % 2 == 1


def test_mode_slice_outside_grid_raises():
    from fdfd import modes
    g = core.Grid(0.02, (15, 15), (0.0, 10.0), (-1.0, 1.0))
    with pytest.raises(IndexError):
        modes.slice_indices(g, core.Point(1.0, -0.95), core.XHAT, 0.8)
