"""dn_jkh_mo_noalloc (the J/K-MO build on caller-owned device arrays, intermediates in the
library's static scratch array) against dn_jkh_mo on the same inputs and against the
oracle: two calls in a row, the second with a smaller system, then a size past the
scratch limit (a clear DN_ERR_ARG, nothing launched).  FP64 bar: max|err| <= 1e-10."""
import ctypes as ct

import numpy as np
import pytest

from donof_amd import _lib, api, synthetic
from donof_amd._lib import DN_ERR_ARG, DN_OK

pytestmark = pytest.mark.gpu

TOL = 1e-10
H2D, D2H = 1, 2


class _Hip:
    """hipMalloc/hipMemcpy of the HIP runtime libdonofhip.so itself links (the test's own
    device arrays; the entry point under test allocates nothing)."""

    def __init__(self):
        _lib.load_library()
        self.rt = ct.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [ct.POINTER(ct.c_void_p), ct.c_size_t]
        self.rt.hipFree.argtypes = [ct.c_void_p]
        self.rt.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
        self.rt.hipSetDevice.argtypes = [ct.c_int]
        self.rt.hipStreamSynchronize.argtypes = [ct.c_void_p]
        self.bufs = []
        assert self.rt.hipSetDevice(0) == 0

    def put(self, a):
        a = np.asfortranarray(a, dtype=np.float64)
        p = ct.c_void_p()
        assert self.rt.hipMalloc(ct.byref(p), max(a.nbytes, 8)) == 0
        self.bufs.append(p)
        assert self.rt.hipMemcpy(p, a.ctypes.data_as(ct.c_void_p), a.nbytes, H2D) == 0
        return p

    def zeros(self, shape):
        return self.put(np.full(shape, np.nan, order="F"))

    def get(self, p, shape):
        out = np.empty(shape, order="F")
        assert self.rt.hipMemcpy(out.ctypes.data_as(ct.c_void_p), p, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def _noalloc(hip, nbf, naux, nbf5, b, C, H):
    lib = _lib.load_library()
    db, dC, dH = hip.put(b), hip.put(C), hip.put(H)
    dJ, dK, dHc = hip.zeros((nbf5, nbf5)), hip.zeros((nbf5, nbf5)), hip.zeros((nbf5,))
    rc = lib.dn_jkh_mo_noalloc(0, nbf, naux, nbf5, db, dC, dH, dJ, dK, dHc, None)
    assert rc == DN_OK, lib.dn_last_error()
    assert hip.rt.hipStreamSynchronize(None) == 0
    return hip.get(dJ, (nbf5, nbf5)), hip.get(dK, (nbf5, nbf5)), hip.get(dHc, (nbf5,))


def test_noalloc_matches_dn_jkh_mo_two_calls(oracle):
    hip = _Hip()
    try:
        # first call larger (even nbf), second smaller (odd nbf, ragged tiles)
        for nbf, naux, nbf5, seed in ((120, 480, 41, 11), (37, 90, 13, 12)):
            print("case noalloc", nbf, naux, nbf5, flush=True)
            s = synthetic.SyntheticSystem(nbf, naux, nbf5, seed)
            b = synthetic.make_b(nbf, naux, seed)
            J, K, Hc = _noalloc(hip, nbf, naux, nbf5, b, s.C, s.H)
            dev = api.System(s.p)
            dev.upload_b(b)
            dev.set_symmetry(False)
            Je, Ke, He = api.computeJKH_MO(s.C, s.H, dev, s.p)
            dev.close()
            assert np.abs(J - Je).max() <= TOL
            assert np.abs(K - Ke).max() <= TOL
            assert np.abs(Hc - He).max() <= TOL
            Jr, Kr, Hr = oracle.jkh_mo(s.C, s.H, b, nbf5)
            assert np.abs(J - Jr).max() <= TOL
            assert np.abs(K - Kr).max() <= TOL
            assert np.abs(Hc - Hr).max() <= TOL
    finally:
        hip.free()


def test_noalloc_rejects_sizes_past_the_scratch():
    lib = _lib.load_library()
    one = ct.c_void_p(16)  # never dereferenced: the size check comes first
    rc = lib.dn_jkh_mo_noalloc(0, 400, 1600, 120, one, one, one, one, one, one, None)
    assert rc == DN_ERR_ARG
    assert b"scratch" in lib.dn_last_error()
