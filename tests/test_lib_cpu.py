"""CPU-side checks of the C ABI: the library is built, loads, and exports every entry
point include/donof_hip.h declares (no GPU calls)."""
import ctypes
import os

from donof_amd import _lib


def test_header_declares_entry_points():
    names = _lib.library_symbols()
    for required in ("dn_ctx_create", "dn_sys_create", "dn_jkh_mo", "dn_lagrange2",
                     "dn_last_error"):
        assert required in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(_lib.LIB_PATH), "libdonofhip.so not built"
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in _lib.library_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_header():
    assert set(_lib.library_symbols()) == set(_lib._SIGS)


def test_last_error_without_gpu():
    lib = _lib.load_library()
    assert isinstance(lib.dn_last_error(), bytes)
