"""ctypes binding of libdonofhip.so (C ABI declared in include/donof_hip.h)."""
from __future__ import annotations

import ctypes as ct
import os
import re

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libdonofhip.so")
HEADER = os.path.join(os.path.dirname(PKG), "include", "donof_hip.h")

DN_OK, DN_ERR_ARG, DN_ERR_HIP, DN_ERR_NOMEM, DN_ERR_BUDGET, DN_ERR_NONFINITE, DN_ERR_STATE = range(7)
DN_FP64, DN_SPLIT_F16 = 0, 1
ERROR_NAMES = {1: "DN_ERR_ARG", 2: "DN_ERR_HIP", 3: "DN_ERR_NOMEM", 4: "DN_ERR_BUDGET",
               5: "DN_ERR_NONFINITE", 6: "DN_ERR_STATE"}


class DnError(RuntimeError):
    """A libdonofhip call returned a nonzero status."""

    def __init__(self, code, msg):
        super().__init__("%s: %s" % (ERROR_NAMES.get(code, code), msg))
        self.code = code


class dn_dims(ct.Structure):
    _fields_ = [(f, ct.c_int32) for f in
                ("nbf", "nbfaux", "nbf5", "no1", "ndoc", "nsoc", "nbeta", "nalpha", "ncwo", "nv")]


class dn_functional(ct.Structure):
    _fields_ = [("ipnof", ct.c_int32), ("ista", ct.c_int32), ("occ_method", ct.c_int32),
                ("highspin", ct.c_int32), ("mspin", ct.c_int32), ("h_cut", ct.c_double)]


class dn_orbopt_state(ct.Structure):
    _fields_ = [("method", ct.c_int32), ("maxloop", ct.c_int32), ("alpha", ct.c_double),
                ("threshgorb", ct.c_double), ("nit", ct.c_int32), ("success", ct.c_int32),
                ("E", ct.c_double)]


DN_ADAM, DN_ADABELIEF, DN_YOGI, DN_DEMON = range(4)
DN_OCC_SOFTMAX, DN_OCC_TRIGONOMETRIC = range(2)

_P = ct.c_void_p
_D = ct.POINTER(ct.c_double)
_I = ct.c_int
_SIGS = {
    "dn_last_error": (ct.c_char_p, []),
    "dn_device_count": (_I, [ct.POINTER(ct.c_int)]),
    "dn_ctx_create": (_I, [_I, ct.c_uint64, ct.POINTER(_P)]),
    "dn_ctx_destroy": (_I, [_P]),
    "dn_ctx_hbm_used": (_I, [_P, ct.POINTER(ct.c_uint64)]),
    "dn_ctx_hbm_budget": (_I, [_P, ct.POINTER(ct.c_uint64)]),
    "dn_sys_create": (_I, [_P, ct.POINTER(dn_dims), _I, ct.POINTER(_P)]),
    "dn_sys_destroy": (_I, [_P]),
    "dn_sys_footprint": (_I, [ct.POINTER(dn_dims), _I, ct.POINTER(ct.c_uint64)]),
    "dn_sys_upload_b": (_I, [_P, _D]),
    "dn_sys_set_ri_factor": (_I, [_P, _D]),
    "dn_grad_create": (_I, [_P, _D, _D, _D, _D, _D, ct.POINTER(_P)]),
    "dn_grad_atom": (_I, [_P, _P, _P, _P, _P, _P, _D]),
    "dn_grad_get_densities": (_I, [_P, _D, _D, _D, _D]),
    "dn_grad_destroy": (_I, [_P]),
    "dn_sys_download_b": (_I, [_P, _D]),
    "dn_sys_generate_b": (_I, [_P, ct.c_uint64]),
    "dn_sys_ri_metric": (_I, [_P, _D, _D]),
    "dn_sys_set_symmetry": (_I, [_P, _I]),
    "dn_sys_b_symmetric": (_I, [_P, ct.POINTER(ct.c_int)]),
    "dn_jkh_mo": (_I, [_P, _D, _D, _D, _D, _D]),
    "dn_lagrange2": (_I, [_P, _D, _D, _D, _D, _D, _I, _I, _D, _D, _D]),
    "dn_lagrange2_jkh": (_I, [_P, _D, _D, _D, _D, _D, _I, _I, _D, _D, _D, _D, _D, _D]),
    "dn_sys_set_orbitals": (_I, [_P, _D]),
    "dn_sys_set_hcore": (_I, [_P, _D]),
    "dn_jkh_mo_resident": (_I, [_P]),
    "dn_sys_get_jkh": (_I, [_P, _D, _D, _D]),
    "dn_sys_sync": (_I, [_P]),
    "dn_sys_set_timing": (_I, [_P, _I]),
    "dn_sys_get_timing": (_I, [_P, ct.POINTER(ct.c_double), ct.POINTER(ct.c_long)]),
    "dn_ocupacion": (_I, [_P, ct.POINTER(dn_functional), _D, _D, _D]),
    "dn_pnof_coefficients": (_I, [_P, ct.POINTER(dn_functional), _D, _D, _D]),
    "dn_der_pnof_coefficients": (_I, [_P, ct.POINTER(dn_functional), _D, _D, _D, _D]),
    "dn_calce": (_I, [_P, ct.POINTER(dn_functional), _D, _D, _D, _D, _D, _D, _D]),
    "dn_calcocce": (_I, [_P, ct.POINTER(dn_functional), _D, _D, _D, _D, _D]),
    "dn_occoptr": (_I, [_P, ct.POINTER(dn_functional), _D, _D, _D, _D, ct.c_double, _I, _D,
                        _D, ct.POINTER(ct.c_int), ct.POINTER(ct.c_int)]),
    "dn_order_occupations_softmax": (_I, [_P, _D, _D, _D, _D]),
    "dn_calcoccg": (_I, [_P, ct.POINTER(dn_functional), _D, _D, _D, _D, _D]),
    "dn_rotate_orbital": (_I, [_P, _D, _D, _D]),
    "dn_orbopt": (_I, [_P, ct.POINTER(dn_functional), ct.POINTER(dn_orbopt_state), _D, _D, _D]),
    "dn_sys_set_gamma": (_I, [_P, _D]),
    "dn_sys_get_orbitals": (_I, [_P, _D]),
    "dn_orbopt_resident": (_I, [_P, ct.POINTER(dn_functional), ct.POINTER(dn_orbopt_state)]),
    "dn_orbopt_start": (_I, [_P, ct.POINTER(dn_functional), ct.POINTER(dn_orbopt_state)]),
    "dn_orbopt_done": (_I, [_P, ct.POINTER(ct.c_int)]),
    "dn_orbopt_finish": (_I, [_P, ct.POINTER(dn_orbopt_state)]),
    "dn_ogroup_create": (_I, [_P, ct.POINTER(_P), _I, ct.POINTER(_P)]),
    "dn_ogroup_destroy": (_I, [_P]),
    "dn_ogroup_step": (_I, [_P, ct.POINTER(ct.c_int)]),
    "dn_ogroup_run": (_I, [_P, _I, ct.POINTER(ct.c_int), ct.POINTER(ct.c_int),
                           ct.POINTER(ct.c_int)]),
    "dn_group_create": (_I, [_P, ct.POINTER(_P), _I, ct.POINTER(_P)]),
    "dn_group_destroy": (_I, [_P]),
    "dn_group_tune": (_I, [_P, ct.POINTER(ct.c_int)]),
    "dn_group_place": (_I, [_P, _I]),
    "dn_group_jkh": (_I, [_P]),
    "dn_group_plan_text": (_I, [_P, ct.c_char_p, _I]),
    "dn_group_sync": (_I, [_P]),
    "dn_group_set_timing": (_I, [_P, _I]),
    "dn_group_get_timing": (_I, [_P, ct.POINTER(ct.c_double), ct.POINTER(ct.c_long)]),
    "dn_jkh_mo_noalloc": (_I, [_I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P]),
    "dn_basis_create": (_I, [_I, ct.POINTER(ct.c_int), ct.POINTER(ct.c_int),
                             ct.POINTER(ct.c_int), _D, _D, _D, ct.POINTER(_P)]),
    "dn_basis_destroy": (_I, [_P]),
    "dn_basis_nbf": (_I, [_P, ct.POINTER(ct.c_int)]),
    "dn_int1e": (_I, [_P, _I, _D, _D, _D, _D, _D]),
    "dn_int2c2e": (_I, [_P, _D]),
    "dn_int3c2e": (_I, [_P, _P, _D]),
    "dn_int4c2e": (_I, [_P, _D]),
    "dn_int3c2e_device": (_I, [_P, _P, _D]),
    "dn_sys_ri_from_basis": (_I, [_P, _P, _P]),
    "dn_int1e_deriv": (_I, [_P, ct.POINTER(ct.c_int), _I, _D, _D, _I, _D, _D, _D]),
    "dn_int2c2e_deriv": (_I, [_P, ct.POINTER(ct.c_int), _I, _D]),
    "dn_int3c2e_deriv": (_I, [_P, ct.POINTER(ct.c_int), _P, ct.POINTER(ct.c_int), _I, _D]),
}

_lib = None


def library_symbols(header=HEADER):
    """Names of the functions declared in include/donof_hip.h."""
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dn_[a-z0-9_]+)\s*\(", text)))


def load_library(path=LIB_PATH):
    """Load libdonofhip.so and attach signatures.  Raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise DnError(DN_ERR_STATE, "libdonofhip.so not built (run __graft_entry__.build())")
    lib = ct.CDLL(path)
    for name, (res, args) in _SIGS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc != DN_OK:
        msg = _lib.dn_last_error().decode(errors="replace") if _lib else ""
        raise DnError(rc, msg)


def dptr(a):
    """Pointer to a float64 array's data (must be contiguous in the ABI's order)."""
    if a is None:
        return None
    return a.ctypes.data_as(_D)


def as_f64(a, order="F"):
    """Column-major float64 view/copy of a host array, as the ABI expects."""
    return np.require(np.asarray(a, dtype=np.float64), requirements=[order])


class Basis:
    """A dn_basis (contracted Gaussian shells) and the host integrals over it."""

    def __init__(self, l, nprim, pure, centers, exps, coefs):
        lib = load_library()
        ints = lambda v: (ct.c_int * len(v))(*[int(x) for x in v])
        self._keep = [as_f64(np.asarray(centers).reshape(-1), "C"), as_f64(exps, "C"),
                      as_f64(coefs, "C")]
        h = ct.c_void_p()
        check(lib.dn_basis_create(len(l), ints(l), ints(nprim), ints(pure), dptr(self._keep[0]),
                                  dptr(self._keep[1]), dptr(self._keep[2]), ct.byref(h)))
        self.h = h
        n = ct.c_int()
        check(lib.dn_basis_nbf(self.h, ct.byref(n)))
        self.nbf = n.value

    def int1e(self, Z, xyz):
        """(S, T, V) for charges Z at xyz (natom x 3, bohr)."""
        n = self.nbf
        S, T, V = (np.zeros((n, n), order="F") for _ in range(3))
        Z = as_f64(Z, "C")
        xyz = as_f64(np.asarray(xyz).reshape(-1), "C")
        check(_lib.dn_int1e(self.h, len(Z), dptr(Z), dptr(xyz), dptr(S), dptr(T), dptr(V)))
        return S, T, V

    def int2c2e(self):
        G = np.zeros((self.nbf, self.nbf), order="F")
        check(_lib.dn_int2c2e(self.h, dptr(G)))
        return G

    def int3c2e(self, aux):
        I3 = np.zeros((self.nbf, self.nbf, aux.nbf), order="F")
        check(_lib.dn_int3c2e(self.h, aux.h, dptr(I3)))
        return I3

    def int3c2e_device(self, aux):
        """(mn|P) computed on the GPU (orbital l <= 2, fitting l <= 4)."""
        I3 = np.zeros((self.nbf, self.nbf, aux.nbf), order="F")
        check(_lib.dn_int3c2e_device(self.h, aux.h, dptr(I3)))
        return I3

    def int4c2e(self):
        n = self.nbf
        eri = np.zeros((n, n, n, n), order="F")
        check(_lib.dn_int4c2e(self.h, dptr(eri)))
        return eri

    @staticmethod
    def _ints(v):
        return (ct.c_int * len(v))(*[int(x) for x in v])

    def int1e_deriv(self, shell_atom, Z, xyz, atom):
        """(dS, dT, dV), each nbf x nbf x 3: d/dR of S, T, V for atom `atom`'s position."""
        n = self.nbf
        dS, dT, dV = (np.zeros((n, n, 3), order="F") for _ in range(3))
        Z = as_f64(Z, "C")
        xyz = as_f64(np.asarray(xyz).reshape(-1), "C")
        check(_lib.dn_int1e_deriv(self.h, self._ints(shell_atom), len(Z), dptr(Z), dptr(xyz),
                                  int(atom), dptr(dS), dptr(dT), dptr(dV)))
        return dS, dT, dV

    def int2c2e_deriv(self, aux_atom, atom):
        dG = np.zeros((self.nbf, self.nbf, 3), order="F")
        check(_lib.dn_int2c2e_deriv(self.h, self._ints(aux_atom), int(atom), dptr(dG)))
        return dG

    def int3c2e_deriv(self, shell_atom, aux, aux_atom, atom):
        dI = np.zeros((self.nbf, self.nbf, aux.nbf, 3), order="F")
        check(_lib.dn_int3c2e_deriv(self.h, self._ints(shell_atom), aux.h, self._ints(aux_atom),
                                    int(atom), dptr(dI)))
        return dI

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.dn_basis_destroy(self.h)
            self.h = None
