"""Host-side mirror of DoNOF.jl's hot-path interface, running on libdonofhip.

Names and argument meaning follow the reference so calling code reads the same:

  computeJKH_MO(C, H, I, p)                    integrals.jl:131-140
  JKH_MO(C, H, b, nbf, nbf5, nbfaux)           integrals.jl:144-172
  compute_Lagrange2(C, n, H, b, cj12, ck12, nalpha, nbeta)   utils.jl:1-59

`I`/`b` is a `System`: the device-resident three-index tensor of one solve, the analogue
of the CuArray the reference's GPU extension dispatches on (ext/DoNOFGPUExt.jl:52,63,150).
Host arrays are float64; they are passed column-major, as Julia stores them.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import _lib
from ._lib import DN_FP64, DnError, check, dptr, dn_dims


class Context:
    """One GPU with an HBM budget (bytes; 0 = all free memory)."""

    def __init__(self, device=0, hbm_budget=0):
        self.lib = _lib.load_library()
        h = ct.c_void_p()
        check(self.lib.dn_ctx_create(int(device), ct.c_uint64(int(hbm_budget)), ct.byref(h)))
        self.handle = h
        self.device = device

    @property
    def hbm_used(self):
        u = ct.c_uint64()
        check(self.lib.dn_ctx_hbm_used(self.handle, ct.byref(u)))
        return u.value

    def close(self):
        if self.handle:
            self.lib.dn_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def dims_of(p):
    d = dn_dims()
    d.nbf, d.nbfaux, d.nbf5 = p.nbf, p.nbfaux, p.nbf5
    d.no1, d.ndoc, d.nsoc = p.no1, p.ndoc, p.nsoc
    d.nbeta, d.nalpha, d.ncwo, d.nv = p.nbeta, p.nalpha, p.ncwo, p.nv
    return d


def footprint(p, precision=DN_FP64):
    """HBM bytes a System of these sizes reserves."""
    lib = _lib.load_library()
    out = ct.c_uint64()
    d = dims_of(p)
    check(lib.dn_sys_footprint(ct.byref(d), int(precision), ct.byref(out)))
    return out.value


class System:
    """Device-resident state of one solve: b (nbf x nbf x nbfaux), C, H, occupations,
    outputs and workspaces, on its own HIP stream."""

    def __init__(self, p, ctx=None, precision=DN_FP64):
        self.ctx = ctx if ctx is not None else default_context()
        self.lib = self.ctx.lib
        self.p = p
        self.nbf, self.nbfaux, self.nbf5 = p.nbf, p.nbfaux, p.nbf5
        self.dims = dims_of(p)
        h = ct.c_void_p()
        check(self.lib.dn_sys_create(self.ctx.handle, ct.byref(self.dims), int(precision),
                                     ct.byref(h)))
        self.handle = h

    # -- b -------------------------------------------------------------------------
    def upload_b(self, b):
        b = np.asarray(b, dtype=np.float64)
        if b.shape != (self.nbf, self.nbf, self.nbfaux):
            raise ValueError("b must have shape (nbf, nbf, nbfaux)")
        b = np.asfortranarray(b)
        check(self.lib.dn_sys_upload_b(self.handle, dptr(b)))
        return self

    def generate_b(self, seed):
        check(self.lib.dn_sys_generate_b(self.handle, ct.c_uint64(int(seed))))
        return self

    def download_b(self):
        out = np.empty((self.nbf, self.nbf, self.nbfaux), order="F")
        check(self.lib.dn_sys_download_b(self.handle, dptr(out)))
        return out

    # -- resident J/K build -------------------------------------------------------------
    def set_orbitals(self, C):
        C = _colmajor(C, (self.nbf, self.nbf), "C")
        check(self.lib.dn_sys_set_orbitals(self.handle, dptr(C)))
        self.sync()

    def set_hcore(self, H):
        H = _colmajor(H, (self.nbf, self.nbf), "H")
        check(self.lib.dn_sys_set_hcore(self.handle, dptr(H)))
        self.sync()

    def jkh_mo_resident(self):
        check(self.lib.dn_jkh_mo_resident(self.handle))

    def get_jkh(self):
        n5 = self.nbf5
        J = np.empty((n5, n5), order="F")
        K = np.empty((n5, n5), order="F")
        Hc = np.empty(n5)
        check(self.lib.dn_sys_get_jkh(self.handle, dptr(J), dptr(K), dptr(Hc)))
        return J, K, Hc

    def sync(self):
        check(self.lib.dn_sys_sync(self.handle))

    def set_timing(self, enable):
        check(self.lib.dn_sys_set_timing(self.handle, 1 if enable else 0))

    def get_timing(self):
        ms = ct.c_double()
        n = ct.c_long()
        check(self.lib.dn_sys_get_timing(self.handle, ct.byref(ms), ct.byref(n)))
        return ms.value, n.value

    def close(self):
        if getattr(self, "handle", None):
            self.lib.dn_sys_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _colmajor(a, shape, name):
    a = np.asarray(a, dtype=np.float64)
    if a.shape != shape:
        raise ValueError("%s must have shape %s, got %s" % (name, shape, a.shape))
    return np.asfortranarray(a)


def _vec(a, n, name):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if a.shape != (n,):
        raise ValueError("%s must have shape (%d,)" % (name, n))
    return a


def _require_system(b):
    if not isinstance(b, System):
        raise TypeError("the three-index tensor must be a device-resident donof_amd.System "
                        "(create one with System(p) and upload_b / generate_b)")
    return b


# --------------------------------------------------------------------------------------
# integrals.jl:131-172
# --------------------------------------------------------------------------------------


def JKH_MO(C, H, b, nbf, nbf5, nbfaux):
    """(J_MO, K_MO, H_core) of the RI tensor b for orbitals C (integrals.jl:144-172)."""
    s = _require_system(b)
    if (nbf, nbf5, nbfaux) != (s.nbf, s.nbf5, s.nbfaux):
        raise ValueError("sizes do not match the resident system")
    C = _colmajor(C, (nbf, nbf), "C")
    H = _colmajor(H, (nbf, nbf), "H")
    J = np.empty((nbf5, nbf5), order="F")
    K = np.empty((nbf5, nbf5), order="F")
    Hc = np.empty(nbf5)
    check(s.lib.dn_jkh_mo(s.handle, dptr(C), dptr(H), dptr(J), dptr(K), dptr(Hc)))
    return J, K, Hc


def computeJKH_MO(C, H, I, p):
    """integrals.jl:131-140 (RI branch; the 4-index branch is out of scope)."""
    if not p.RI:
        raise NotImplementedError("only the RI path (p.RI = true) is implemented")
    return JKH_MO(C, H, I, p.nbf, p.nbf5, p.nbfaux)


# --------------------------------------------------------------------------------------
# utils.jl:1-59, 216-227
# --------------------------------------------------------------------------------------


def compute_Lagrange2(C, n, H, b, cj12, ck12, nalpha, nbeta, return_energy=False):
    """(elag, H_mat) of utils.jl:1-59.  With return_energy=True also returns
    computeE_elec(H_mat, n, elag) evaluated on the device (utils.jl:216-227)."""
    s = _require_system(b)
    nbf, nbf5 = s.nbf, s.nbf5
    C = _colmajor(C, (nbf, nbf), "C")
    H = _colmajor(H, (nbf, nbf), "H")
    n = _vec(n, nbf5, "n")
    cj12 = _colmajor(cj12, (nbf5, nbf5), "cj12")
    ck12 = _colmajor(ck12, (nbf5, nbf5), "ck12")
    elag = np.empty((nbf, nbf), order="F")
    Hmat = np.empty((nbf, nbf5), order="F")
    E = ct.c_double()
    check(s.lib.dn_lagrange2(s.handle, dptr(C), dptr(n), dptr(H), dptr(cj12), dptr(ck12),
                             int(nalpha), int(nbeta), dptr(elag), dptr(Hmat), ct.byref(E)))
    if return_energy:
        return elag, Hmat, E.value
    return elag, Hmat


__all__ = ["Context", "System", "default_context", "footprint", "JKH_MO", "computeJKH_MO",
           "compute_Lagrange2", "DnError"]
