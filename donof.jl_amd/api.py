"""Host-side mirror of DoNOF.jl's hot-path interface, running on libdonofhip.

Names and argument meaning follow the reference so calling code reads the same:

  computeJKH_MO(C, H, I, p)                    integrals.jl:131-140
  JKH_MO(C, H, b, nbf, nbf5, nbfaux)           integrals.jl:144-172
  compute_Lagrange2(C, n, H, b, cj12, ck12, nalpha, nbeta)   utils.jl:1-59
  compute_geom_gradients(bset, n, C, cj12, ck12, elag, p, I)  grads.jl:1-100 (RI)

`I`/`b` is a `System`: the device-resident three-index tensor of one solve, the analogue
of the CuArray the reference's GPU extension dispatches on (ext/DoNOFGPUExt.jl:52,63,150).
Host arrays are float64; they are passed column-major, as Julia stores them.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import _lib
from ._lib import DN_FP64, DN_SPLIT_F16, DnError, check, dptr, dn_dims


class Context:
    """One GPU with an HBM budget (bytes; 0 = all free memory)."""

    def __init__(self, device=0, hbm_budget=0):
        self.lib = _lib.load_library()
        h = ct.c_void_p()
        check(self.lib.dn_ctx_create(int(device), ct.c_uint64(int(hbm_budget)), ct.byref(h)))
        self.handle = h
        self.device = device

    @property
    def hbm_budget(self):
        u = ct.c_uint64()
        check(self.lib.dn_ctx_hbm_budget(self.handle, ct.byref(u)))
        return u.value

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
        if precision in ("fp64", "FP64"):
            precision = DN_FP64
        elif precision in ("split", "split_f16"):
            precision = DN_SPLIT_F16
        self.precision = precision
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

    def ri_metric(self, G, I3c):
        """b = I3c G^{-1/2} on the device (integrals.jl:68-91): the FP64 Cholesky route
        for a positive-definite G, else the reference's eigen route (eigenvalues below
        zero contribute zero)."""
        G = _colmajor(G, (self.nbfaux, self.nbfaux), "G")
        I3c = _colmajor(I3c, (self.nbf, self.nbf, self.nbfaux), "I3c")
        check(self.lib.dn_sys_ri_metric(self.handle, dptr(G), dptr(I3c)))
        return self

    def ri_from_basis(self, basis, aux):
        """The RI setup from a basis and a fitting basis (`_lib.Basis` handles): (P|Q) on
        the host, (mn|P) computed on the device into the metric solve (integrals.jl:47-53,
        68-91)."""
        check(self.lib.dn_sys_ri_from_basis(self.handle, basis.h, aux.h))
        return self

    def set_ri_factor(self, Wt):
        """The RI factor of an uploaded b = I3c W for the nuclear gradients:
        Wt = W^{-1} G^{-1} (naux x naux), so that I3c G^{-1} = b Wt; G^{-1/2} for the
        reference's eigen-route b (integrals.jl:68-91)."""
        Wt = _colmajor(Wt, (self.nbfaux, self.nbfaux), "Wt")
        check(self.lib.dn_sys_set_ri_factor(self.handle, dptr(Wt)))
        return self

    def set_gamma(self, gamma):
        g = np.ascontiguousarray(gamma, dtype=np.float64)
        check(self.lib.dn_sys_set_gamma(self.handle, dptr(g)))
        self.sync()

    def get_orbitals(self):
        out = np.empty((self.nbf, self.nbf), order="F")
        check(self.lib.dn_sys_get_orbitals(self.handle, dptr(out)))
        return out

    # -- step-wise ADAM family (dn_orbopt_start / dn_ogroup_step / dn_orbopt_finish) ----
    def orbopt_start(self, method, p):
        """Start an ADAM-family run at the resident gamma, C and H (minimization.jl:344-360)."""
        f = functional_of(p)
        st = _lib.dn_orbopt_state()
        st.method = _ORB[method]
        st.maxloop, st.alpha, st.threshgorb = int(p.maxloop), float(p.alpha), float(p.threshgorb)
        check(self.lib.dn_orbopt_start(self.handle, ct.byref(f), ct.byref(st)))
        return self

    def orbopt_done(self):
        v = ct.c_int()
        check(self.lib.dn_orbopt_done(self.handle, ct.byref(v)))
        return bool(v.value)

    def orbopt_finish(self, p=None):
        """(E, nit, success) of the finished run; p.alpha / p.maxloop updated as the
        reference does when p is given; the orbitals stay resident (get_orbitals)."""
        st = _lib.dn_orbopt_state()
        check(self.lib.dn_orbopt_finish(self.handle, ct.byref(st)))
        if p is not None:
            p.maxloop, p.alpha = int(st.maxloop), float(st.alpha)
        return st.E, int(st.nit), bool(st.success)

    def set_symmetry(self, use):
        """Use (default) or ignore the (m, n) symmetry of b in the J/K build."""
        check(self.lib.dn_sys_set_symmetry(self.handle, 1 if use else 0))
        return self

    @property
    def b_symmetric(self):
        v = ct.c_int()
        check(self.lib.dn_sys_b_symmetric(self.handle, ct.byref(v)))
        return bool(v.value)

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
        """Frees the system; refused (DnError, DN_ERR_STATE) while a group holds it."""
        if getattr(self, "handle", None):
            check(self.lib.dn_sys_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Group:
    """Resident systems of uneven size built together: grouped persistent launches
    captured into one HIP graph (dn_group_*).  Each system's result is its own
    computeJKH_MO (integrals.jl:131-172), read with System.get_jkh."""

    STAGES = ("K1", "K2", "J", "HC")
    CONFIGS = {0: "128x128", 1: "64x64", 2: "64x128", 3: "128x64", 4: "64x32", 5: "32x64",
               6: "lds64x128", 7: "lds128x128", 8: "lds128x256", 9: "lds64x192",
               10: "64x192",
               100: "s128x128", 101: "s64x128", 102: "s64x64", 103: "s32x64"}

    def __init__(self, systems, ctx=None):
        self.systems = list(systems)
        if not self.systems:
            raise ValueError("a group needs at least one system")
        self.ctx = ctx if ctx is not None else self.systems[0].ctx
        self.lib = self.ctx.lib
        arr = (ct.c_void_p * len(self.systems))(*[s.handle for s in self.systems])
        h = ct.c_void_p()
        check(self.lib.dn_group_create(self.ctx.handle, arr, len(self.systems), ct.byref(h)))
        self.handle = h
        self.configs = None

    def place(self, max_grouped_nbf5):
        """Fix the placement: systems with nbf5 <= max_grouped_nbf5 grouped, the others
        built by the single-system build on their own streams (dn_group_place)."""
        check(self.lib.dn_group_place(self.handle, int(max_grouped_nbf5)))
        return self

    def tune(self):
        """Autotune the tile configuration of every grouped GEMM stage, capture the build
        graph and (unless place() fixed it) choose the placement; returns
        {stage: tile shape} (None for a stage when no system is grouped)."""
        cfg = (ct.c_int * 4)()
        check(self.lib.dn_group_tune(self.handle, cfg))
        self.configs = {st: self.CONFIGS.get(c) for st, c in zip(self.STAGES, cfg)}
        return self.configs

    def plan(self):
        """Placement, then per stage each class's system count, tile shape and GFLOP."""
        buf = ct.create_string_buffer(16384)
        check(self.lib.dn_group_plan_text(self.handle, buf, len(buf)))
        return buf.value.decode()

    def jkh(self):
        check(self.lib.dn_group_jkh(self.handle))
        return self

    def sync(self):
        check(self.lib.dn_group_sync(self.handle))

    def set_timing(self, enable):
        check(self.lib.dn_group_set_timing(self.handle, 1 if enable else 0))

    def get_timing(self):
        ms = ct.c_double()
        n = ct.c_long()
        check(self.lib.dn_group_get_timing(self.handle, ct.byref(ms), ct.byref(n)))
        return ms.value, n.value

    def close(self):
        if getattr(self, "handle", None):
            self.lib.dn_group_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class OGroup:
    """Started systems advanced together, one iteration per step (dn_ogroup_*): the
    Lagrangian of all of them runs as grouped GEMMs in one HIP graph."""

    def __init__(self, systems, ctx=None):
        self.systems = list(systems)
        self.ctx = ctx if ctx is not None else self.systems[0].ctx
        self.lib = self.ctx.lib
        arr = (ct.c_void_p * len(self.systems))(*[s.handle for s in self.systems])
        h = ct.c_void_p()
        check(self.lib.dn_ogroup_create(self.ctx.handle, arr, len(self.systems), ct.byref(h)))
        self.handle = h

    def step(self):
        n = ct.c_int()
        check(self.lib.dn_ogroup_step(self.handle, ct.byref(n)))
        return n.value

    def run(self, max_steps, stop=None):
        """Steps in native code until a system finishes, none steps, `stop` (a ctypes
        c_int another thread may set) is non-zero, or max_steps ran (dn_ogroup_run).
        Returns (steps run, systems that stepped in the last one)."""
        steps, active = ct.c_int(), ct.c_int()
        check(self.lib.dn_ogroup_run(self.handle, int(max_steps),
                                     ct.byref(stop) if stop is not None else None,
                                     ct.byref(steps), ct.byref(active)))
        return steps.value, active.value

    def close(self):
        if getattr(self, "handle", None):
            self.lib.dn_ogroup_destroy(self.handle)
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
# integrals.jl:68-91, 131-172
# --------------------------------------------------------------------------------------


def Iaux(G, I, p, ctx=None, precision=DN_FP64):
    """integrals.jl:68-91: the RI tensor from the metric G and the 3-centre integrals I.
    Returns the device-resident System holding b (the reference returns b itself; every
    later call takes that System where the reference takes b).  b is I L^{-T} with
    G = L L^T, i.e. the reference's I G^{-1/2} rotated by an orthogonal matrix in the
    auxiliary index, which no J, K, Lagrangian or energy can see."""
    return System(p, ctx, precision).ri_metric(G, I)


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


def compute_Lagrange2_JKH(C, n, H, b, cj12, ck12, nalpha, nbeta):
    """(elag, H_mat, E, J_MO, K_MO, H_core): compute_Lagrange2 (utils.jl:1-59) and
    computeJKH_MO (integrals.jl:131-172) at the same orbitals from one transform of b,
    as the diagnostics of an outer iteration need them (energy.jl:172-190)."""
    s = _require_system(b)
    nbf, nbf5 = s.nbf, s.nbf5
    C = _colmajor(C, (nbf, nbf), "C")
    H = _colmajor(H, (nbf, nbf), "H")
    n = _vec(n, nbf5, "n")
    cj12 = _colmajor(cj12, (nbf5, nbf5), "cj12")
    ck12 = _colmajor(ck12, (nbf5, nbf5), "ck12")
    elag = np.empty((nbf, nbf), order="F")
    Hmat = np.empty((nbf, nbf5), order="F")
    J = np.empty((nbf5, nbf5), order="F")
    K = np.empty((nbf5, nbf5), order="F")
    Hc = np.empty(nbf5)
    E = ct.c_double()
    check(s.lib.dn_lagrange2_jkh(s.handle, dptr(C), dptr(n), dptr(H), dptr(cj12), dptr(ck12),
                                 int(nalpha), int(nbeta), dptr(elag), dptr(Hmat), ct.byref(E),
                                 dptr(J), dptr(K), dptr(Hc)))
    return elag, Hmat, E.value, J, K, Hc


# --------------------------------------------------------------------------------------
# grads.jl:1-100 (RI branch), utils.jl:315-336
# --------------------------------------------------------------------------------------

BOHR_PER_ANGSTROM = 1.88973      # the conversion utils.jl:305,320 use


class DeviceArray:
    """A column-major float64 array in HBM, allocated with the HIP runtime libdonofhip.so
    links (hipMalloc / hipMemcpy / hipFree of libamdhip64.so.7): inputs such as derivative
    integrals kept resident and read in place.  (A second HIP runtime in the process, as
    torch wheels bundle, does not share devices with this one.)"""
    _rt = None

    @classmethod
    def _runtime(cls):
        if cls._rt is None:
            _lib.load_library()
            rt = ct.CDLL("libamdhip64.so.7")
            rt.hipMalloc.argtypes = [ct.POINTER(ct.c_void_p), ct.c_size_t]
            rt.hipFree.argtypes = [ct.c_void_p]
            rt.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
            rt.hipMemset.argtypes = [ct.c_void_p, ct.c_int, ct.c_size_t]
            rt.hipSetDevice.argtypes = [ct.c_int]
            cls._rt = rt
        return cls._rt

    def __init__(self, shape, device=0):
        rt = self._runtime()
        self.shape = tuple(int(x) for x in shape)
        self.nbytes = 8 * int(np.prod(self.shape))
        self.device = device
        self.ptr = ct.c_void_p()
        if rt.hipSetDevice(int(device)) != 0 or \
                rt.hipMalloc(ct.byref(self.ptr), max(self.nbytes, 8)) != 0:
            self.ptr = None
            raise DnError(_lib.DN_ERR_NOMEM, "hipMalloc of %d bytes" % self.nbytes)
        if rt.hipMemset(self.ptr, 0, max(self.nbytes, 8)) != 0:
            raise DnError(_lib.DN_ERR_HIP, "hipMemset")

    @classmethod
    def from_host(cls, a, device=0):
        a = np.asfortranarray(np.asarray(a, dtype=np.float64))
        d = cls(a.shape, device)
        d.upload(a)
        return d

    def upload(self, a, offset_bytes=0):
        a = np.asfortranarray(np.asarray(a, dtype=np.float64))
        if offset_bytes + a.nbytes > self.nbytes:
            raise ValueError("upload past the end of the device array")
        dst = ct.c_void_p(self.ptr.value + offset_bytes)
        if self._runtime().hipMemcpy(dst, a.ctypes.data_as(ct.c_void_p), a.nbytes, 1) != 0:
            raise DnError(_lib.DN_ERR_HIP, "hipMemcpy host to device")

    def download(self):
        out = np.empty(self.shape, order="F")
        if self._runtime().hipMemcpy(out.ctypes.data_as(ct.c_void_p), self.ptr, self.nbytes,
                                     2) != 0:
            raise DnError(_lib.DN_ERR_HIP, "hipMemcpy device to host")
        return out

    def free(self):
        if self.ptr:
            self._runtime().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _host_or_device(a, shape, name):
    """A derivative-integral array as the ABI takes it: a DeviceArray is read in place,
    anything else becomes a column-major float64 host array."""
    if isinstance(a, DeviceArray):
        if a.shape != tuple(shape):
            raise ValueError("%s must have shape %s, got %s" % (name, shape, a.shape))
        return a, a.ptr
    h = _colmajor(a, shape, name)
    return h, h.ctypes.data_as(ct.c_void_p)


class Gradient:
    """The densities of compute_geom_gradients' RI branch (grads.jl:1-100), formed on the
    device from the resident b: RDM1, lag, the three-index density d and the metric
    density m.  `atom(...)` contracts them with one atom's derivative integrals.

    Wt: None when b was set by System.ri_metric (b = I3c L^{-T}) or its factor was given
    with System.set_ri_factor; otherwise the matrix with I3c G^{-1} = b Wt (G^{-1/2} for
    the reference's eigen-route b), kept on the system."""

    def __init__(self, b, n, C, cj12, ck12, elag, Wt=None):
        s = _require_system(b)
        self.system = s
        nbf, nbf5, naux = s.nbf, s.nbf5, s.nbfaux
        C = _colmajor(C, (nbf, nbf), "C")
        n = _vec(np.asarray(n)[:nbf5], nbf5, "n")
        cj12 = _colmajor(cj12, (nbf5, nbf5), "cj12")
        ck12 = _colmajor(ck12, (nbf5, nbf5), "ck12")
        elag = _colmajor(elag, (nbf, nbf), "elag")
        if Wt is not None:
            s.set_ri_factor(Wt)
        h = ct.c_void_p()
        check(s.lib.dn_grad_create(s.handle, dptr(C), dptr(n), dptr(cj12), dptr(ck12),
                                   dptr(elag), ct.byref(h)))
        self.handle = h

    def atom(self, dS, dT, dV, geri3c, geri2c):
        """-sum lag dS + sum RDM1 (dT + dV) + 2 sum d geri3c - sum m geri2c per Cartesian
        component (grads.jl:13-21, 91-96)."""
        s = self.system
        nbf, naux = s.nbf, s.nbfaux
        keep = [_host_or_device(dS, (nbf, nbf, 3), "dS"),
                _host_or_device(dT, (nbf, nbf, 3), "dT"),
                _host_or_device(dV, (nbf, nbf, 3), "dV"),
                _host_or_device(geri3c, (nbf, nbf, naux, 3), "geri3c"),
                _host_or_device(geri2c, (naux, naux, 3), "geri2c")]
        out = np.zeros(3)
        check(s.lib.dn_grad_atom(self.handle, *[k[1] for k in keep], dptr(out)))
        return out

    def densities(self, with_d=False):
        """(RDM1, lag, m[, d]) copied to the host."""
        s = self.system
        nbf, naux = s.nbf, s.nbfaux
        R = np.empty((nbf, nbf), order="F")
        L = np.empty((nbf, nbf), order="F")
        m = np.empty((naux, naux), order="F")
        d = np.empty((nbf, nbf, naux), order="F") if with_d else None
        check(s.lib.dn_grad_get_densities(self.handle, dptr(R), dptr(L), dptr(m), dptr(d)))
        return (R, L, m, d) if with_d else (R, L, m)

    def close(self):
        if getattr(self, "handle", None):
            check(self.system.lib.dn_grad_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def compute_grad_nuc(atoms):
    """Nuclear-repulsion gradient (utils.jl:315-336) in Hartree/Bohr, natoms x 3; each
    atom has .Z and .xyz (Angstrom)."""
    Z = np.array([float(a.Z) for a in atoms])
    R = np.array([np.asarray(a.xyz, dtype=np.float64) for a in atoms]) * BOHR_PER_ANGSTROM
    g = np.zeros((len(atoms), 3))
    for k in range(len(atoms)):
        for j in range(len(atoms)):
            if j != k:
                r = R[k] - R[j]
                g[k] -= Z[k] * Z[j] * r / np.linalg.norm(r) ** 3
    return g


def compute_geom_gradients(bset, n, C, cj12, ck12, elag, p, I, Wt=None):
    """grads.jl:1-100 with p.RI: natoms x 3 nuclear gradient.

    `I` is the System holding b (as everywhere in this module); `bset` supplies the atoms
    and the derivative integrals the reference takes from GaussianBasis: bset.atoms
    (each with .Z, .xyz in Angstrom) and, for atom i (0-based), bset.overlap_deriv(i),
    kinetic_deriv(i), nuclear_deriv(i) (nbf x nbf x 3), eri_2e3c_deriv(i)
    (nbf x nbf x naux x 3) and eri_2e2c_deriv(i) (naux x naux x 3), as NumPy arrays or
    DeviceArrays (read in place).  Unlike the reference, cj12 and ck12 are not modified
    (grads.jl:24-25 zero their diagonals in place; the device ignores them)."""
    if not p.RI:
        raise NotImplementedError("only the RI branch (p.RI = true) is implemented")
    grad = compute_grad_nuc(bset.atoms)
    g = Gradient(I, n, C, cj12, ck12, elag, Wt)
    try:
        for i in range(len(bset.atoms)):
            grad[i] += g.atom(bset.overlap_deriv(i), bset.kinetic_deriv(i),
                              bset.nuclear_deriv(i), bset.eri_2e3c_deriv(i),
                              bset.eri_2e2c_deriv(i))
    finally:
        g.close()
    return grad


# --------------------------------------------------------------------------------------
# pnof.jl: occupations, coefficients, occupation energy and gradient
# --------------------------------------------------------------------------------------

_OCC = {"Softmax": _lib.DN_OCC_SOFTMAX, "Trigonometric": _lib.DN_OCC_TRIGONOMETRIC}
_ORB = {"ADAM": _lib.DN_ADAM, "ADABelief": _lib.DN_ADABELIEF, "YOGI": _lib.DN_YOGI,
        "Demon": _lib.DN_DEMON}


def functional_of(p):
    """The dn_functional record of a Param (p.ipnof, p.ista, p.occ_method, ...)."""
    if p.occ_method not in _OCC:
        raise NotImplementedError("occ_method %r (EBI is out of scope)" % p.occ_method)
    f = _lib.dn_functional()
    f.ipnof, f.ista, f.occ_method = int(p.ipnof), int(p.ista), _OCC[p.occ_method]
    f.highspin, f.mspin, f.h_cut = int(bool(p.HighSpin)), int(p.MSpin), float(p.h_cut)
    return f


def ocupacion(gamma, p, I):
    """(n, dn_dgamma) of pnof.jl:1024-1045 on the device of system I."""
    s = _require_system(I)
    f = functional_of(p)
    g = _vec(gamma, p.nv, "gamma")
    n = np.empty(p.nbf5)
    dn = np.empty((p.nbf5, p.nv), order="F")
    check(s.lib.dn_ocupacion(s.handle, ct.byref(f), dptr(g), dptr(n), dptr(dn)))
    return n, dn


def ocupacion_softmax(gamma, p, I):
    """pnof.jl:1122-1163: ocupacion with the softmax parametrisation."""
    q = p.copy()
    q.occ_method = "Softmax"
    return ocupacion(gamma, q, I)


def ocupacion_trigonometric(gamma, p, I):
    """pnof.jl:1047-1120: ocupacion with the trigonometric parametrisation."""
    q = p.copy()
    q.occ_method = "Trigonometric"
    return ocupacion(gamma, q, I)


def PNOFi_selector(n, p, I):
    """(cj12, ck12) with zero diagonal (pnof.jl:1-53)."""
    s = _require_system(I)
    f = functional_of(p)
    n = _vec(n, p.nbf5, "n")
    cj = np.empty((p.nbf5, p.nbf5), order="F")
    ck = np.empty((p.nbf5, p.nbf5), order="F")
    check(s.lib.dn_pnof_coefficients(s.handle, ct.byref(f), dptr(n), dptr(cj), dptr(ck)))
    return cj, ck


def der_PNOFi_selector(n, dn_dgamma, p, I):
    """(Dcj12r, Dck12r), nbf5 x nbf5 x nv (pnof.jl:55-117)."""
    s = _require_system(I)
    f = functional_of(p)
    n = _vec(n, p.nbf5, "n")
    dn = _colmajor(dn_dgamma, (p.nbf5, p.nv), "dn_dgamma")
    Dcj = np.empty((p.nbf5, p.nbf5, p.nv), order="F")
    Dck = np.empty((p.nbf5, p.nbf5, p.nv), order="F")
    check(s.lib.dn_der_pnof_coefficients(s.handle, ct.byref(f), dptr(n), dptr(dn), dptr(Dcj),
                                         dptr(Dck)))
    return Dcj, Dck


def _jkh_args(p, J, K, H_core):
    if J is None and K is None and H_core is None:
        return None, None, None
    return (_colmajor(J, (p.nbf5, p.nbf5), "J_MO"), _colmajor(K, (p.nbf5, p.nbf5), "K_MO"),
            _vec(H_core, p.nbf5, "H_core"))


def calce(n, cj12, ck12, J_MO, K_MO, H_core, p, I):
    """pnof.jl:1273-1300.  J_MO, K_MO, H_core = None uses the resident ones."""
    s = _require_system(I)
    f = functional_of(p)
    n = _vec(n, p.nbf5, "n")
    cj = _colmajor(cj12, (p.nbf5, p.nbf5), "cj12")
    ck = _colmajor(ck12, (p.nbf5, p.nbf5), "ck12")
    J, K, Hc = _jkh_args(p, J_MO, K_MO, H_core)
    E = ct.c_double()
    check(s.lib.dn_calce(s.handle, ct.byref(f), dptr(n), dptr(cj), dptr(ck), dptr(J), dptr(K),
                         dptr(Hc), ct.byref(E)))
    return E.value


def calcocce(gamma, J_MO, K_MO, H_core, p, I):
    """pnof.jl:1302-1322."""
    s = _require_system(I)
    f = functional_of(p)
    g = _vec(gamma, p.nv, "gamma")
    J, K, Hc = _jkh_args(p, J_MO, K_MO, H_core)
    E = ct.c_double()
    check(s.lib.dn_calcocce(s.handle, ct.byref(f), dptr(g), dptr(J), dptr(K), dptr(Hc),
                            ct.byref(E)))
    return E.value


def occ_conjugate_gradient(gamma, J_MO, K_MO, H_core, p, I, g_abstol, maxiter=1000):
    """optimize(calcocce, calcoccg, gamma, ConjugateGradient(), g_abstol) of occoptr
    (minimization.jl:100-139) on the device: (gamma, minimum, iterations, converged).
    J_MO, K_MO, H_core = None uses the resident ones."""
    s = _require_system(I)
    f = functional_of(p)
    g = _vec(gamma, p.nv, "gamma")
    J, K, Hc = _jkh_args(p, J_MO, K_MO, H_core)
    out = np.empty(p.nv)
    E = ct.c_double()
    nit, conv = ct.c_int(), ct.c_int()
    check(s.lib.dn_occoptr(s.handle, ct.byref(f), dptr(g), dptr(J), dptr(K), dptr(Hc),
                           float(g_abstol), int(maxiter), dptr(out), ct.byref(E),
                           ct.byref(nit), ct.byref(conv)))
    return out, E.value, nit.value, bool(conv.value)


def order_occupations_softmax(old_C, old_gamma, p, I):
    """utils.jl:586-616 on the device: (C, gamma) with each pair sorted by gamma."""
    s = _require_system(I)
    C = _colmajor(old_C, (s.nbf, s.nbf), "C")
    g = _vec(old_gamma, p.nv, "gamma")
    Cout = np.empty((s.nbf, s.nbf), order="F")
    gout = np.empty(p.nv)
    check(s.lib.dn_order_occupations_softmax(s.handle, dptr(C), dptr(g), dptr(Cout), dptr(gout)))
    return Cout, gout


def calcoccg(gamma, J_MO, K_MO, H_core, p, I):
    """pnof.jl:1324-1377."""
    s = _require_system(I)
    f = functional_of(p)
    g = _vec(gamma, p.nv, "gamma")
    J, K, Hc = _jkh_args(p, J_MO, K_MO, H_core)
    grad = np.empty(p.nv)
    check(s.lib.dn_calcoccg(s.handle, ct.byref(f), dptr(g), dptr(J), dptr(K), dptr(Hc),
                            dptr(grad)))
    return grad


# --------------------------------------------------------------------------------------
# utils.jl:499-521, pnof.jl:1389-1407
# --------------------------------------------------------------------------------------


def rotate_orbital(y, C, p, I):
    """Cnew = C exp(Y(y)) in FP64 on the device (utils.jl:499-521)."""
    s = _require_system(I)
    y = _vec(y, p.nvar, "y")
    C = _colmajor(C, (p.nbf, p.nbf), "C")
    out = np.empty((p.nbf, p.nbf), order="F")
    check(s.lib.dn_rotate_orbital(s.handle, dptr(y), dptr(C), dptr(out)))
    return out


def pack_grad(elag, nbf5):
    """grads = 4 elag - 4 elag^T packed row-wise over i < nbf5, j > i
    (pnof.jl:1395-1403)."""
    nbf = elag.shape[0]
    g = 4.0 * elag - 4.0 * elag.T
    rows = np.repeat(np.arange(nbf5), [nbf - i - 1 for i in range(nbf5)])
    cols = np.concatenate([np.arange(i + 1, nbf) for i in range(nbf5)]) if nbf5 else \
        np.zeros(0, dtype=int)
    return g[rows, cols]


def calcorbg(y, n, cj12, ck12, C, H, I, pa):
    """Packed orbital gradient at rotate(y, C) (pnof.jl:1389-1407)."""
    Cnew = rotate_orbital(y, C, pa, I)
    elag, _ = compute_Lagrange2(Cnew, n, H, I, cj12, ck12, pa.nalpha, pa.nbeta)
    return pack_grad(elag, pa.nbf5)


def calcorbe(y, n, cj12, ck12, C, H, I, p):
    """pnof.jl:1379-1387: the energy at C exp(Y(y)) from J/K-MO of the rotated orbitals."""
    Cnew = rotate_orbital(y, C, p, I)
    J, K, Hc = computeJKH_MO(Cnew, H, I, p)
    return calce(n, cj12, ck12, J, K, Hc, p, I)


def calcorbeg(F, G, y, n, cj12, ck12, C, H, I, pa):
    """pnof.jl:1409-1433: one Lagrangian at C exp(Y(y)); with G (an nvar array) the packed
    orbital gradient is written into it, with F not None the energy
    computeE_elec(Hmat, n, elag) is returned."""
    Cnew = rotate_orbital(y, C, pa, I)
    elag, Hmat = compute_Lagrange2(Cnew, n, H, I, cj12, ck12, pa.nalpha, pa.nbeta)
    if G is not None:
        G[:] = pack_grad(elag, pa.nbf5)
    if F is not None:
        E = float(np.trace(elag[:pa.nbf5, :pa.nbf5]))
        return E + float(np.dot(np.asarray(n)[:pa.nbf5], np.diag(Hmat[:pa.nbf5, :pa.nbf5])))
    return None


def orbopt(method, gamma, C, H, I, p):
    """ADAM-family orbital optimiser (minimization.jl:242-630) run entirely on the
    device: returns (E, C, nit, success) and updates p.alpha / p.maxloop as the
    reference does."""
    s = _require_system(I)
    f = functional_of(p)
    st = _lib.dn_orbopt_state()
    st.method = _ORB[method]
    st.maxloop, st.alpha, st.threshgorb = int(p.maxloop), float(p.alpha), float(p.threshgorb)
    g = _vec(gamma, p.nv, "gamma")
    Cio = np.array(_colmajor(C, (p.nbf, p.nbf), "C"), order="F", copy=True)
    H = _colmajor(H, (p.nbf, p.nbf), "H")
    check(s.lib.dn_orbopt(s.handle, ct.byref(f), ct.byref(st), dptr(g), dptr(Cio), dptr(H)))
    p.maxloop, p.alpha = int(st.maxloop), float(st.alpha)
    return st.E, Cio, int(st.nit), bool(st.success)


def orbopt_adam(gamma, C, H, I, p):
    """minimization.jl:344-438."""
    return orbopt("ADAM", gamma, C, H, I, p)


def orbopt_adabelief(gamma, C, H, I, p):
    """minimization.jl:441-536."""
    return orbopt("ADABelief", gamma, C, H, I, p)


def orbopt_yogi(gamma, C, H, I, p):
    """minimization.jl:539-630."""
    return orbopt("YOGI", gamma, C, H, I, p)


def orbopt_demon(gamma, C, H, I, p):
    """minimization.jl:242-340."""
    return orbopt("Demon", gamma, C, H, I, p)


__all__ = ["Context", "System", "default_context", "footprint", "JKH_MO", "computeJKH_MO",
           "compute_Lagrange2", "DnError", "functional_of", "ocupacion", "PNOFi_selector",
           "der_PNOFi_selector", "calce", "calcocce", "calcoccg", "rotate_orbital",
           "calcorbg", "pack_grad", "orbopt", "orbopt_adam", "orbopt_adabelief",
           "orbopt_yogi", "orbopt_demon"]
