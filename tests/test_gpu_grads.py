"""RI nuclear gradients on the device (dn_grad_*; grads.jl:1-100) against the oracle.

Small systems: the densities RDM1, lag, m, d and each atom's gradient against the
oracle's restatement (which forms IG = I3c inv(G) and contracts over (l, s) as the
reference does), to the FP64 bar; b from the device RI setup (Cholesky route, its kept
factor) and b uploaded as the reference's eigen-route b with Wt = G^{-1/2}; derivative
integrals from host arrays and from device tensors; the compute_geom_gradients mirror end
to end with the nuclear term.  At 400/1600/120 the oracle would need hours, so the full
size is checked by what the formula is: the derivative of the energy.  b(x) = b0 + x b1
(I3c(x) = I0 + x I1 with I1 = b1 L^T and G = L L^T fixed), C(x) = C0 (C0^T S(x) C0)^{-1/2}
with S(x) = 1 + x S1, H(x) = H0 + x H1: central differences of the device energy (J/K
build plus calce) against the device gradient with geri3c = I1, dS = S1, dT = H1."""
import numpy as np
import pytest

from donof_amd import api, synthetic
from donof_amd._lib import DN_ERR_STATE
from donof_amd.param import Param

pytestmark = pytest.mark.gpu

# (nbf, naux, ne, mul): a ragged Cholesky block, odd nbf (padded b), an open shell
SHAPES = [(10, 24, 6, 1), (17, 70, 6, 1), (24, 130, 7, 2), (40, 200, 10, 1)]


def _param(nbf, naux, ne, mul):
    p = Param(nbf=nbf, nbfaux=naux, ne=ne, mul=mul)
    p.ipnof = 8
    return p


def _inputs(oracle, p, seed):
    nbf, naux = p.nbf, p.nbfaux
    rng = np.random.default_rng(seed)
    G, I3c, _ = synthetic.make_metric(nbf, naux, seed, cond=1e3)
    C = synthetic.make_C(nbf, seed + 1)
    H = synthetic.make_H(nbf, seed + 2)
    gamma = oracle.guess_gamma_softmax(p.ndoc, p.ncwo) + 0.3 * rng.standard_normal(p.nv)
    n, _ = oracle.ocupacion(gamma, p)
    cj12, ck12 = oracle.pnof_selector(n, p)
    elag, _ = oracle.compute_lagrange2(C, n, H, oracle.iaux(G, I3c), cj12, ck12, p.nalpha,
                                       p.nbeta)
    return G, I3c, C, n, cj12, ck12, elag


def _derivs(nbf, naux, seed):
    rng = np.random.default_rng(seed)
    f = lambda *s: np.asfortranarray(rng.standard_normal(s))
    return f(nbf, nbf, 3), f(nbf, nbf, 3), f(nbf, nbf, 3), f(nbf, nbf, naux, 3), f(naux, naux, 3)


def _bound(dens, dv):
    """sum of |products| of the five contractions: the scale of the rounding of a dot"""
    R, L, d, m = dens
    dS, dT, dV, g3, g2 = dv
    a = lambda x, y: (np.abs(x)[..., None] * np.abs(y)).reshape(-1, 3).sum(axis=0)
    return a(L, dS) + a(R, dT) + a(R, dV) + 2 * a(d, g3) + a(m, g2)


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("nbf,naux,ne,mul", SHAPES)
def test_gradient_parity(oracle, nbf, naux, ne, mul):
    print("case grad", nbf, naux, ne, mul, flush=True)
    p = _param(nbf, naux, ne, mul)
    G, I3c, C, n, cj12, ck12, elag = _inputs(oracle, p, 11 + nbf)
    dens = oracle.geom_gradient_densities(n, C, cj12, ck12, elag, p, G, I3c)
    dev = api.Iaux(G, I3c, p)
    g = api.Gradient(dev, n, C, cj12, ck12, elag)
    R, L, m, d = g.densities(with_d=True)
    assert _rel(R, dens[0]) <= 1e-12 and _rel(L, dens[1]) <= 1e-12
    assert _rel(d, dens[2]) <= 1e-10 and _rel(m, dens[3]) <= 1e-10
    for atom in range(2):
        dv = _derivs(nbf, naux, 100 * atom + nbf)
        got = g.atom(*dv)
        ref = oracle.geom_gradient_atom(*dens, *dv)
        assert np.all(np.abs(got - ref) <= 1e-11 * _bound(dens, dv)), (atom, got, ref)
    g.close()
    dev.close()


def test_device_arrays_read_in_place(oracle):
    """Derivative integrals in HBM give exactly the host-array result."""
    p = _param(17, 70, 6, 1)
    G, I3c, C, n, cj12, ck12, elag = _inputs(oracle, p, 5)
    dev = api.Iaux(G, I3c, p)
    g = api.Gradient(dev, n, C, cj12, ck12, elag)
    dv = _derivs(p.nbf, p.nbfaux, 9)
    host = g.atom(*dv)
    dd = [api.DeviceArray.from_host(a) for a in dv]
    assert np.array_equal(dd[3].download(), dv[3])
    assert np.array_equal(g.atom(*dd), host)
    for a in dd:
        a.free()
    g.close()
    dev.close()


def test_eigen_route_b_with_its_factor(oracle):
    """b uploaded as the reference's I3c G^{-1/2} (integrals.jl:68-91) with Wt = G^{-1/2}."""
    p = _param(24, 130, 6, 1)
    G, I3c, C, n, cj12, ck12, elag = _inputs(oracle, p, 3)
    lam, V = np.linalg.eigh(G)
    Gmh = (V / np.sqrt(lam)) @ V.T
    dev = api.System(p).upload_b(oracle.iaux(G, I3c))
    with pytest.raises(api.DnError) as e:
        api.Gradient(dev, n, C, cj12, ck12, elag)       # no factor kept for an uploaded b
    assert e.value.code == DN_ERR_STATE
    g = api.Gradient(dev, n, C, cj12, ck12, elag, Wt=Gmh)
    dens = oracle.geom_gradient_densities(n, C, cj12, ck12, elag, p, G, I3c)
    _, _, m, d = g.densities(with_d=True)
    assert _rel(d, dens[2]) <= 1e-10 and _rel(m, dens[3]) <= 1e-10
    dv = _derivs(p.nbf, p.nbfaux, 4)
    assert np.all(np.abs(g.atom(*dv) - oracle.geom_gradient_atom(*dens, *dv))
                  <= 1e-11 * _bound(dens, dv))
    with pytest.raises(api.DnError) as e:
        dev.close()                                     # held by the gradient object
    assert e.value.code == DN_ERR_STATE
    g.close()
    dev.close()


class _Atom:
    def __init__(self, Z, xyz):
        self.Z, self.xyz = Z, xyz


class _BasisSet:
    """The integral source compute_geom_gradients takes (seeded synthetic derivatives)."""

    def __init__(self, nbf, naux, natoms, seed):
        rng = np.random.default_rng(seed)
        self.atoms = [_Atom(float(z), rng.standard_normal(3) * 1.2)
                      for z in rng.integers(1, 9, natoms)]
        self.d = [_derivs(nbf, naux, seed + 7 * i) for i in range(natoms)]

    def overlap_deriv(self, i):
        return self.d[i][0]

    def kinetic_deriv(self, i):
        return self.d[i][1]

    def nuclear_deriv(self, i):
        return self.d[i][2]

    def eri_2e3c_deriv(self, i):
        return self.d[i][3]

    def eri_2e2c_deriv(self, i):
        return self.d[i][4]


def test_compute_geom_gradients_mirror(oracle):
    p = _param(17, 70, 6, 1)
    G, I3c, C, n, cj12, ck12, elag = _inputs(oracle, p, 21)
    bset = _BasisSet(p.nbf, p.nbfaux, 3, 8)
    cj0, ck0 = cj12.copy(), ck12.copy()
    dev = api.Iaux(G, I3c, p)
    got = api.compute_geom_gradients(bset, n, C, cj12, ck12, elag, p, dev)
    dev.close()
    assert np.array_equal(cj12, cj0) and np.array_equal(ck12, ck0)
    Z = np.array([a.Z for a in bset.atoms])
    xyz = np.array([a.xyz for a in bset.atoms])
    ref = oracle.compute_geom_gradients(n, C, cj12, ck12, elag, p, G, I3c, bset.d, Z, xyz)
    assert got.shape == (3, 3)
    dens = oracle.geom_gradient_densities(n, C, cj12, ck12, elag, p, G, I3c)
    for i in range(3):
        assert np.all(np.abs(got[i] - ref[i]) <= 1e-11 * _bound(dens, bset.d[i]))


def test_full_size_gradient_is_the_energy_derivative(oracle):
    nbf, naux, nbf5, seed = 400, 1600, 120, 2024
    print("case grad full 400/1600/120", flush=True)
    p = synthetic.synthetic_param(nbf, naux, nbf5)
    sd = synthetic.SyntheticSystem(nbf, naux, nbf5, seed)
    rng = np.random.default_rng(1)
    # G = L L^T with a lower bidiagonal L: I3c G^{-1} = b L^{-1} (L^{-1} dense lower), and
    # I1 = b1 L^T is a two-term sum per Q
    a = 1.0 + rng.random(naux)
    c = 0.3 * rng.standard_normal(naux)
    Lf = np.diag(a) + np.diag(c[1:], -1)
    Wt = np.linalg.inv(Lf)
    dev = api.System(p)
    dev.generate_b(seed)
    b0 = dev.download_b()
    b1 = synthetic.make_b(nbf, naux, seed + 1)
    n = sd.n
    cj12, ck12 = oracle.pnof_selector(n, p)
    C0, H0 = sd.C, sd.H
    elag, _ = api.compute_Lagrange2(C0, n, H0, dev, cj12, ck12, p.nalpha, p.nbeta)
    g = api.Gradient(dev, n, C0, cj12, ck12, elag, Wt=Wt)
    rng = np.random.default_rng(2)
    # perturbations with O(1) spectra (a random symmetric matrix has norm ~ 2 sqrt(nbf)),
    # so that the difference quotients' truncation stays far below the tolerance
    S1 = rng.standard_normal((nbf, nbf)) / np.sqrt(nbf)
    S1 = 0.5 * (S1 + S1.T)
    H1 = rng.standard_normal((nbf, nbf)) / np.sqrt(nbf)
    H1 = 0.5 * (H1 + H1.T)
    z2 = np.zeros((nbf, nbf, 3), order="F")
    k0 = lambda x: np.asfortranarray(np.stack([x, 0 * x, 0 * x], axis=-1))
    g2 = np.zeros((naux, naux, 3), order="F")
    # geri3c = I1 (component 0) from the host: streamed through the staging buffer
    g3 = np.zeros((nbf, nbf, naux, 3), order="F")
    g3[:, :, :, 0] = b1 * a[None, None, :]
    g3[:, :, 1:, 0] += b1[:, :, :-1] * c[None, None, 1:]
    g_I = g.atom(z2, z2, z2, g3, g2)[0]
    del g3
    z3 = api.DeviceArray((nbf, nbf, naux, 3))           # zeros in HBM
    g_S = g.atom(k0(S1), z2, z2, z3, g2)[0]
    g_H = g.atom(z2, k0(H1), z2, z3, g2)[0]
    z3.free()
    g.close()

    def energy(b_sys, C, H):
        J, K, Hc = api.JKH_MO(C, H, b_sys, nbf, nbf5, naux)
        return oracle.calce(n, cj12, ck12, J, K, Hc, p)

    def lowdin(x):
        lam, U = np.linalg.eigh(C0.T @ (np.eye(nbf) + x * S1) @ C0)
        return C0 @ ((U / np.sqrt(lam)) @ U.T)

    # I term: E is quadratic in x, the central difference is exact up to rounding
    h = 0.1
    ex = []
    for x in (h, -h):
        bx = api.System(p).upload_b(b0 + x * b1)
        ex.append(energy(bx, C0, H0))
        bx.close()
    fd_I = (ex[0] - ex[1]) / (2 * h)
    # S term (C0^T C0 = 1 for S(0) = 1): fourth-order central difference
    hs = 1e-3
    e = {x: energy(dev, lowdin(x), H0) for x in (hs, -hs, 2 * hs, -2 * hs)}
    fd_S = (8 * (e[hs] - e[-hs]) - (e[2 * hs] - e[-2 * hs])) / (12 * hs)
    # H term: linear
    fd_H = (energy(dev, C0, H0 + H1) - energy(dev, C0, H0 - H1)) / 2.0
    dev.close()
    print("full size: I %.12e / %.12e  S %.12e / %.12e  H %.12e / %.12e"
          % (g_I, fd_I, g_S, fd_S, g_H, fd_H), flush=True)
    assert abs(g_I - fd_I) <= 1e-9 * max(1.0, abs(fd_I))
    assert abs(g_S - fd_S) <= 1e-7 * max(1.0, abs(fd_S))
    assert abs(g_H - fd_H) <= 1e-9 * max(1.0, abs(fd_H))
