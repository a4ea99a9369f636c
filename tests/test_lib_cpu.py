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


def test_split_footprint_below_fp64():
    """A split-precision system keeps b only as its FP16 hi/lo planes (half the bytes of
    the FP64 b), so it needs less HBM than the FP64 system of the same sizes (at the
    benchmark sizes; tiny systems are dominated by the 128-row padding of the planes)."""
    from donof_amd import api, synthetic
    for nbf, naux, nbf5 in ((400, 1600, 120), (800, 3200, 240), (200, 800, 60)):
        p = synthetic.synthetic_param(nbf, naux, nbf5)
        f64 = api.footprint(p, _lib.DN_FP64)
        f16 = api.footprint(p, _lib.DN_SPLIT_F16)
        assert f16 < f64, (nbf, naux, nbf5, f16, f64)
