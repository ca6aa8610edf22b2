"""The C-ABI library loads and exports every symbol include/macau_amd.h
declares (no compute calls: CPU only)."""
import ctypes
import os
import re

from macau_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    with open(os.path.join(ROOT, "include", "macau_amd.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(macau_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = _declared()
    assert "macau_run" in names and "macau_sample_latents" in names and "macau_solve_blockcg" in names
    assert len(names) >= 40


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(_declared()) == set(_lib.SIGNATURES)


def test_version_and_error_string():
    L = _lib.lib()
    assert L.macau_version().decode() == "0.1.0"
    assert isinstance(_lib.last_error(), str)


def test_probit_factory():
    """make_macau_probit (macau.cpp:207-211): a handle for valid sizes (the
    constructor makes no device calls), NULL with an error otherwise."""
    L = _lib.lib()
    h = L.macau_make_probit(2, 4)
    assert h
    L.macau_free(h)
    assert not L.macau_make_probit(2, 0)
    assert "num_latent" in _lib.last_error()
