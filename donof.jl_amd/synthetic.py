"""Synthetic RI systems of the benchmark shapes, generated from a counter-based RNG.

The normal stream is Philox4x32-10 (Salmon, Moraes, Dror & Shaw, SC'11) followed by
Box-Muller.  The same definition is implemented on the device (csrc/dn_jk.hip), so a
16 GB three-index tensor can be generated in HBM and spot-checked on the host without a
transfer.  Element e of stream s under seed k uses counter
(lo32(e>>1), hi32(e>>1), s, 0x0D0F) and key (lo32(k), hi32(k)); the 4 output words give
u = ((x0>>5)*2^26 + (x1>>6) + 0.5) * 2^-53 and v likewise from (x2, x3);
z = sqrt(-2 ln u) * (cos(2 pi v) if e is even else sin(2 pi v)); uniforms are u (even e)
or v (odd e).

Shapes follow the reference's RI contraction (integrals.jl:144-172): b[m,n,Q] is
symmetric in (m, n) and column-major.  Generators (1-based Q as in Julia):
  b[:,:,Q] = (A_Q + A_Q^T) / (2 sqrt(nbf)),  A_Q ~ N(0,1) * exp(-Q/400)
  C        = Q factor of QR(N(0,1) nbf x nbf)
  H        = (A + A^T) / 2,  A ~ N(0,1)
  n        = strong orbitals at 0.99; the weak orbitals of each pair carry
             0.01 * Dirichlet(1,...,1)  (so every pair sums to 1)
"""
from __future__ import annotations

import math

import numpy as np

STREAM_B, STREAM_C, STREAM_H, STREAM_OCC, STREAM_METRIC, STREAM_SHAPE = 0, 1, 2, 3, 4, 5
B_DECAY = 400.0

_M0 = np.uint64(0xD2511F53)
_M1 = np.uint64(0xCD9E8D57)
_W0 = 0x9E3779B9
_W1 = 0xBB67AE85
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Vectorised Philox4x32-10 on uint64 arrays holding 32-bit words."""
    c0, c1, c2, c3 = (np.asarray(x, dtype=np.uint64) for x in (c0, c1, c2, c3))
    k0 = int(k0) & 0xFFFFFFFF
    k1 = int(k1) & 0xFFFFFFFF
    for r in range(10):
        p0 = _M0 * c0
        p1 = _M1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK
        c0, c1, c2, c3 = (hi1 ^ c1 ^ np.uint64(k0)), lo1, (hi0 ^ c3 ^ np.uint64(k1)), lo0
        k0 = (k0 + _W0) & 0xFFFFFFFF
        k1 = (k1 + _W1) & 0xFFFFFFFF
    return c0, c1, c2, c3


def _uniform_pair(e, stream, seed):
    e = np.asarray(e, dtype=np.uint64)
    ctr = e >> np.uint64(1)
    x0, x1, x2, x3 = philox4x32_10(ctr & _MASK, ctr >> np.uint64(32),
                                   np.full(e.shape, stream, np.uint64),
                                   np.full(e.shape, 0x0D0F, np.uint64),
                                   seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)
    s53 = 2.0 ** -53
    u = ((x0 >> np.uint64(5)).astype(np.float64) * 67108864.0
         + (x1 >> np.uint64(6)).astype(np.float64) + 0.5) * s53
    v = ((x2 >> np.uint64(5)).astype(np.float64) * 67108864.0
         + (x3 >> np.uint64(6)).astype(np.float64) + 0.5) * s53
    return u, v


def normal(e, stream, seed):
    """N(0,1) deviates for element indices e of a stream."""
    e = np.asarray(e, dtype=np.uint64)
    u, v = _uniform_pair(e, stream, seed)
    r = np.sqrt(-2.0 * np.log(u))
    ang = 2.0 * math.pi * v
    return np.where((e & np.uint64(1)) == 0, r * np.cos(ang), r * np.sin(ang))


def uniform(e, stream, seed):
    """U(0,1) deviates for element indices e (u for even e, v for odd e)."""
    e = np.asarray(e, dtype=np.uint64)
    u, v = _uniform_pair(e, stream, seed)
    return np.where((e & np.uint64(1)) == 0, u, v)


def make_b(nbf, naux, seed, q0=0, q1=None):
    """Host generation of b[:, :, q0:q1] (Fortran order).  Use for small sizes; the
    device generator produces the same tensor in HBM."""
    q1 = naux if q1 is None else q1
    nq = q1 - q0
    m = np.arange(nbf, dtype=np.uint64)
    out = np.empty((nbf, nbf, nq), order="F")
    inv = 0.5 / math.sqrt(nbf)
    for t, Q in enumerate(range(q0, q1)):
        e = m[:, None] + np.uint64(nbf) * m[None, :] + np.uint64(nbf * nbf) * np.uint64(Q)
        A = normal(e, STREAM_B, seed)
        out[:, :, t] = (A + A.T) * (math.exp(-(Q + 1) / B_DECAY) * inv)
    return out


def make_metric(nbf, naux, seed, cond=1e6):
    """RI setup inputs with a known answer: a positive-definite 2-centre metric G
    (naux x naux, eigenvalues log-spaced over [1/cond, 1], seeded orthogonal
    eigenvectors) and 3-centre integrals I3c = b G^{1/2} with b = make_b(nbf, naux, seed),
    so that I3c G^{-1/2} = b.  Returns (G, I3c, b)."""
    e = np.arange(naux * naux, dtype=np.uint64)
    A = normal(e, STREAM_METRIC, seed).reshape(naux, naux, order="F")
    V, _ = np.linalg.qr(A)
    lam = np.logspace(0.0, -math.log10(cond), naux)
    G = np.asfortranarray(0.5 * ((V * lam) @ V.T + ((V * lam) @ V.T).T))
    Gh = (V * np.sqrt(lam)) @ V.T
    b = make_b(nbf, naux, seed)
    I3c = (b.reshape(nbf * nbf, naux, order="F") @ Gh).reshape(nbf, nbf, naux, order="F")
    return G, np.asfortranarray(I3c), b


def make_C(nbf, seed):
    e = np.arange(nbf * nbf, dtype=np.uint64)
    A = normal(e, STREAM_C, seed).reshape(nbf, nbf, order="F")
    Q, R = np.linalg.qr(A)
    return np.asfortranarray(Q)


def make_H(nbf, seed):
    e = np.arange(nbf * nbf, dtype=np.uint64)
    A = normal(e, STREAM_H, seed).reshape(nbf, nbf, order="F")
    return np.asfortranarray(0.5 * (A + A.T))


def default_pairing(nbf5):
    """ndoc = largest d in {5,4,3,2,1} dividing nbf5; ncwo = nbf5/ndoc - 1."""
    for d in (5, 4, 3, 2, 1):
        if nbf5 % d == 0 and nbf5 // d >= 2:
            return d, nbf5 // d - 1
    raise ValueError("nbf5 too small for a pairing: %d" % nbf5)


def make_occupations(p, seed, strong=0.99):
    """GNOF occupations: strong orbitals at `strong`; the ncwo weak orbitals of each
    pair share 1-strong as a Dirichlet(1,...,1) draw.  Returns n (nbf5)."""
    n = np.zeros(p.nbf5)
    n[:p.no1] = 1.0
    n[p.no1 + p.ndoc:p.no1 + p.ndns] = 0.5
    for l in range(1, p.ndoc + 1):
        n[p.no1 + l - 1] = strong
        w = p.no1 + p.ndns + (p.ndoc - l) + p.ndoc * np.arange(p.ncwo)
        e = np.arange(p.ncwo, dtype=np.uint64) + np.uint64(l * 1_000_003)
        x = -np.log(uniform(e, STREAM_OCC, seed))
        n[w] = (1.0 - strong) * x / x.sum()
    return n


def gammas_from_occupations(n, p):
    """Softmax gammas reproducing n (utils.jl:539-561, with n[i] as the reference)."""
    g = np.zeros(p.nv)
    for i in range(1, p.ndoc + 1):
        w = p.no1 + p.ndns + (p.ndoc - i) + p.ndoc * np.arange(p.ncwo)
        gi = (p.ndoc - i) + p.ndoc * np.arange(p.ncwo)
        g[gi] = np.log(np.maximum(n[w], 1e-16)) - math.log(n[i - 1])
    return g


def synthetic_param(nbf, naux, nbf5, ndoc=None, ipnof=8, orb_method="ADAM",
                    occ_method="Softmax"):
    """A Param for a closed-shell synthetic system with nbf5 active orbitals:
    ndoc doubly occupied pairs (ne = 2 ndoc) and ncwo = nbf5/ndoc - 1 (set_ncwo)."""
    from .param import Param
    if ndoc is None:
        ndoc, ncwo = default_pairing(nbf5)
    else:
        if nbf5 % ndoc:
            raise ValueError("nbf5 must be a multiple of ndoc")
        ncwo = nbf5 // ndoc - 1
    p = Param(nbf=nbf, nbfaux=naux, ne=2 * ndoc, mul=1)
    p.set_ncwo(ncwo)
    if p.nbf5 != nbf5:
        raise ValueError("pairing gives nbf5=%d, wanted %d" % (p.nbf5, nbf5))
    p.ipnof = ipnof
    p.orb_method = orb_method
    p.occ_method = occ_method
    return p


def grouped_shapes(nsys=64, seed=2024, nbf_min=48, nbf_max=400):
    """Sizes of a grouped batch: nbf uniform on the integers [nbf_min, nbf_max] (seeded
    Philox), naux = 4 nbf, nbf5 = ceil(nbf / 3).  Returns a list of (nbf, naux, nbf5)."""
    u = uniform(np.arange(nsys, dtype=np.uint64), STREAM_SHAPE, seed)
    nbf = nbf_min + np.minimum((u * (nbf_max - nbf_min + 1)).astype(np.int64),
                               nbf_max - nbf_min)
    return [(int(n), 4 * int(n), -(-int(n) // 3)) for n in nbf]


def build_flops(nbf, naux, nbf5):
    """Algorithmic FLOP of one J/K-MO build in the reference's contraction order
    (integrals.jl:156-169): both transform halves, J, K and H_core."""
    return (2.0 * nbf5 * nbf * nbf * naux + 2.0 * nbf5 * nbf5 * nbf * naux
            + 4.0 * nbf5 * nbf5 * naux + 2.0 * nbf * nbf * nbf5 + 2.0 * nbf * nbf5)


def build_bytes(nbf, naux, nbf5, e=8):
    """Minimum HBM bytes of one J/K-MO build: b read once (e bytes per element: 8 for FP64,
    4 for a hi+lo pair of 16-bit planes) plus C5, H, J, K and H_core in FP64."""
    return e * nbf * nbf * naux + 8 * (nbf * nbf5 + nbf * nbf + 2 * nbf5 * nbf5 + nbf5)


class SyntheticSystem:
    """Inputs of one synthetic solve: C, H, n, gamma and the Param; b is generated
    on the device (or on the host with make_b for small sizes)."""

    def __init__(self, nbf, naux, nbf5, seed, ndoc=None, ipnof=8, orb_method="ADAM"):
        self.nbf, self.naux, self.nbf5, self.seed = nbf, naux, nbf5, seed
        self.p = synthetic_param(nbf, naux, nbf5, ndoc=ndoc, ipnof=ipnof,
                                 orb_method=orb_method)
        self.C = make_C(nbf, seed)
        self.H = make_H(nbf, seed)
        self.n = make_occupations(self.p, seed)
        self.gamma = gammas_from_occupations(self.n, self.p)

    def host_b(self):
        return make_b(self.nbf, self.naux, self.seed)
