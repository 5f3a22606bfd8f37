"""The C-ABI library loads and exports every function include/fdfd.h declares
(CPU only: no kernel is launched)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fdfd.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(fdfd_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_api():
    names = declared_functions()
    for must in ("fdfd_create", "fdfd_set_grid", "fdfd_set_eps", "fdfd_factor", "fdfd_solve",
                 "fdfd_solve_d", "fdfd_linsolve", "fdfd_eigenfrequency", "fdfd_adjoint_grad",
                 "fdfd_batch_run"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from fdfd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libfdfd not built (run __graft_entry__.build())")
    lib = _lib.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the ctypes prototypes cover the whole header
    assert set(declared_functions()) == set(_lib.PROTOTYPES)
    assert lib.fdfd_arch() == b"gfx950"
    ma, mi, pa = (_lib.C.c_int(), _lib.C.c_int(), _lib.C.c_int())
    assert lib.fdfd_version(_lib.C.byref(ma), _lib.C.byref(mi), _lib.C.byref(pa)) == 0


def test_missing_library_fails_loudly(tmp_path):
    from fdfd import _lib

    with pytest.raises(ImportError):
        _lib.load(str(tmp_path / "nope.so"))


def test_invalid_handle_calls_return_errors():
    """Argument validation runs before any device call."""
    from fdfd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libfdfd not built")
    lib = _lib.load()
    assert lib.fdfd_destroy(None) == -1
    assert lib.fdfd_set_grid(None, 4, 4, 0.1, 0.1, 0, 0, 1e-6) == -1
    assert lib.fdfd_factor(None, 1, 1.0, 0, None, None) == -1
    assert lib.fdfd_create(0, None) == -1
