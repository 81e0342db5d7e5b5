"""The C-ABI library loads on the host and exports every entry point of include/lex_capi.h."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lex_capi.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lex_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_symbols()
    assert len(names) >= 30
    for must in ("lex_adv_scalar", "lex_acoustic_substep", "lex_bicgstab", "lex_rk3_step", "lex_smag_momentum"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from lex_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("liblexhip.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    from lex_amd import _lib
    assert set(declared_symbols()) <= set(_lib.exported_symbols())


def test_no_compute_without_gpu_raises_cleanly():
    """Calling the binding with CPU tensors raises instead of silently computing elsewhere."""
    import torch
    from lex_amd import _lib
    with pytest.raises(_lib.LexError):
        _lib.ptr(torch.zeros(3))


def test_abi_version_and_error_reporting():
    from lex_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("liblexhip.so not built")
    lib = _lib.load()
    assert lib.lex_abi_version() == 2
    # an argument error is reported without touching the GPU
    g = _lib.grid(2, 2, 2, 1.0, 1.0, 1.0)
    rc = lib.lex_div(g, None, None, None, None, None)
    assert rc < 0
    assert b"invalid grid" in lib.lex_last_error() or b"null" in lib.lex_last_error()
