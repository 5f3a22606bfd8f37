"""fdfd_flux_d_noalloc (the no-allocation variant of fdfd_flux_d) against
fdfd_flux_d through the C ABI: two calls in a row, the second on a smaller
field with fewer planes, bit-identical sums; the plane limit is enforced."""
import ctypes as C

import numpy as np
import pytest

import fdfd_oracle as O
from fdfd import _lib, core
from fdfd._lib import FDFDError

pytestmark = pytest.mark.gpu


def _hip():
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    return hip


def _field(Nx, Ny, seed):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((3, Ny, Nx)) + 1j * rng.standard_normal((3, Ny, Nx))
    return np.ascontiguousarray(f.astype(np.complex128)).ravel()


def _planes(Nx, Ny, count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for p in range(count):
        if p % 2 == 0:
            idx = int(rng.integers(0, Nx - 1))
            lo = int(rng.integers(0, Ny - 1))
            hi = int(rng.integers(lo, Ny))
            out.append((1, idx, lo, hi))
        else:
            idx = int(rng.integers(0, Ny - 1))
            lo = int(rng.integers(0, Nx - 1))
            hi = int(rng.integers(lo, Nx))
            out.append((2, idx, lo, hi))
    return (_lib.fdfd_flux_plane * count)(*[_lib.fdfd_flux_plane(*p) for p in out])


def _both(engine, hip, pol, Nx, Ny, nplanes, seed):
    flat = _field(Nx, Ny, seed)
    dev = C.c_void_p()
    assert hip.hipMalloc(C.byref(dev), flat.nbytes) == 0
    try:
        assert hip.hipMemcpy(dev, flat.ctypes.data_as(C.c_void_p), flat.nbytes, 1) == 0
        arr = _planes(Nx, Ny, nplanes, seed + 1)
        ref = np.empty(nplanes)
        got = np.full(nplanes, np.nan)
        pd = C.POINTER(C.c_double)
        engine._chk(engine.lib.fdfd_flux_d(engine.h, pol, Nx, Ny, 0.02, 0.03, dev, nplanes, arr,
                                           ref.ctypes.data_as(pd)))
        engine._chk(engine.lib.fdfd_flux_d_noalloc(engine.h, pol, Nx, Ny, 0.02, 0.03, dev,
                                                   nplanes, arr, got.ctypes.data_as(pd)))
        return ref, got
    finally:
        hip.hipFree(dev)


@pytest.mark.parametrize("pol", [1, 2])
def test_noalloc_flux_matches_flux_d_twice(engine, pol):
    hip = _hip()
    print(f"noalloc flux pol={pol}: 300 x 260 field, 4096 planes", flush=True)
    ref, got = _both(engine, hip, pol, 300, 260, 4096, 3)
    assert np.array_equal(ref, got)
    print(f"noalloc flux pol={pol}: 64 x 48 field, 37 planes", flush=True)
    ref2, got2 = _both(engine, hip, pol, 64, 48, 37, 11)
    assert np.array_equal(ref2, got2)
    assert not np.array_equal(ref[:37], ref2)


def test_noalloc_flux_rejects_too_many_planes(engine):
    hip = _hip()
    Nx, Ny = 16, 12
    flat = _field(Nx, Ny, 0)
    dev = C.c_void_p()
    assert hip.hipMalloc(C.byref(dev), flat.nbytes) == 0
    try:
        assert hip.hipMemcpy(dev, flat.ctypes.data_as(C.c_void_p), flat.nbytes, 1) == 0
        arr = _planes(Nx, Ny, 4097, 2)
        out = np.empty(4097)
        rc = engine.lib.fdfd_flux_d_noalloc(engine.h, 1, Nx, Ny, 0.02, 0.02, dev, 4097, arr,
                                            out.ctypes.data_as(C.POINTER(C.c_double)))
        assert rc == -1
        assert b"at most 4096 planes" in engine.lib.fdfd_last_error(engine.h)
    finally:
        hip.hipFree(dev)
