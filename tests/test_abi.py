"""The C-ABI library loads and exports every symbol include/tmrx.h declares
(no GPU work)."""
from tmr_amd import _lib


def test_exports_every_declared_symbol(native):
    names = _lib.declared_symbols()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(native, n)]
    assert missing == []


def test_ctypes_signatures_cover_header():
    assert set(_lib.declared_symbols()) <= set(_lib._SIGS)


def test_version(native):
    assert _lib.lib().tmrx_version().decode().startswith("tmrx")
