"""The oracle's RI nuclear gradient (grads.jl:1-100) pinned by the derivative of the total
energy it differentiates.

A one-parameter family of integrals S(x) = S0 + x S1, H(x) = H0 + x (T1 + V1),
I3c(x) = I0 + x I1, G(x) = G0 + x G1 with orbitals kept orthonormal by the symmetric
(Loewdin) re-orthonormalisation C(x) = C0 (C0^T S(x) C0)^{-1/2} and fixed occupations:
dE/dx at x = 0, from the oracle's own energy (iaux, jkh_mo, calce), must equal the
gradient formula with dS = S1, dT = T1, dV = V1, geri3c = I1, geri2c = G1.  That holds at
any orbitals (not only at a stationary point): C0^T S1 C0 is symmetric, so only the
symmetric part of the Lagrangian enters, as in -sum lag dS.  Each term is checked alone
(one integral varied at a time), so a dropped term, a wrong factor or sign fails here.
The derivative is a fourth-order central difference (Richardson): error ~h^4."""
import numpy as np
import pytest

h = 1e-3


def _sym(a):
    return 0.5 * (a + a.T)


def _sym3(a):
    return 0.5 * (a + a.transpose(1, 0, 2))


def _family(oracle, nbf, naux, ne, mul, ipnof, seed):
    rng = np.random.default_rng(seed)
    p = oracle.make_param(nbf, naux, ne, mul=mul, ipnof=ipnof)
    S0 = np.eye(nbf) + 0.1 * _sym(rng.standard_normal((nbf, nbf)))
    lam, V = np.linalg.eigh(S0)
    Q, _ = np.linalg.qr(rng.standard_normal((nbf, nbf)))
    C0 = (V / np.sqrt(lam)) @ V.T @ Q                     # C0^T S0 C0 = 1
    A = rng.standard_normal((naux, naux))
    f = dict(p=p, C0=C0, S0=S0, H0=_sym(rng.standard_normal((nbf, nbf))),
             G0=A @ A.T + naux * np.eye(naux),
             I0=_sym3(rng.standard_normal((nbf, nbf, naux))),
             S1=_sym(rng.standard_normal((nbf, nbf))), T1=_sym(rng.standard_normal((nbf, nbf))),
             V1=_sym(rng.standard_normal((nbf, nbf))),
             I1=_sym3(rng.standard_normal((nbf, nbf, naux))),
             G1=_sym(rng.standard_normal((naux, naux))))
    gam = oracle.guess_gamma_softmax(p.ndoc, p.ncwo) + 0.3 * rng.standard_normal(p.nv)
    f["n"], _ = oracle.ocupacion(gam, p)
    f["cj12"], f["ck12"] = oracle.pnof_selector(f["n"], p)
    return f


def _energy(oracle, f, x, which):
    S = f["S0"] + (x * f["S1"] if "S" in which else 0.0)
    H = f["H0"] + (x * (f["T1"] + f["V1"]) if "H" in which else 0.0)
    I3c = f["I0"] + (x * f["I1"] if "I" in which else 0.0)
    G = f["G0"] + (x * f["G1"] if "G" in which else 0.0)
    lam, U = np.linalg.eigh(f["C0"].T @ S @ f["C0"])
    C = f["C0"] @ ((U / np.sqrt(lam)) @ U.T)
    J, K, Hc = oracle.jkh_mo(C, H, oracle.iaux(G, I3c), f["p"].nbf5)
    return oracle.calce(f["n"], f["cj12"], f["ck12"], J, K, Hc, f["p"])


def _fd(oracle, f, which):
    e = lambda x: _energy(oracle, f, x, which)
    return (8 * (e(h) - e(-h)) - (e(2 * h) - e(-2 * h))) / (12 * h)


def _formula(oracle, f, which):
    p, nbf, naux = f["p"], f["p"].nbf, f["p"].nbfaux
    b0 = oracle.iaux(f["G0"], f["I0"])
    elag, _ = oracle.compute_lagrange2(f["C0"], f["n"], f["H0"], b0, f["cj12"], f["ck12"],
                                       p.nalpha, p.nbeta)
    z = lambda *s: np.zeros(s + (3,))
    x0 = lambda a: np.stack([a, 0 * a, 0 * a], axis=-1)     # the term in Cartesian k = 0
    dS = x0(f["S1"]) if "S" in which else z(nbf, nbf)
    dT = x0(f["T1"]) if "H" in which else z(nbf, nbf)
    dV = x0(f["V1"]) if "H" in which else z(nbf, nbf)
    g3 = x0(f["I1"]) if "I" in which else z(nbf, nbf, naux)
    g2 = x0(f["G1"]) if "G" in which else z(naux, naux)
    dens = oracle.geom_gradient_densities(f["n"], f["C0"], f["cj12"], f["ck12"], elag, p,
                                          f["G0"], f["I0"])
    g = oracle.geom_gradient_atom(*dens, dS, dT, dV, g3, g2)
    assert g[1] == 0.0 and g[2] == 0.0
    return g[0]


@pytest.mark.parametrize("ne,mul,ipnof", [(6, 1, 8), (6, 1, 7), (7, 2, 8), (6, 1, 5)])
@pytest.mark.parametrize("which", ["S", "H", "I", "G", "SHIG"])
def test_gradient_is_the_energy_derivative(oracle, ne, mul, ipnof, which):
    f = _family(oracle, 10, 24, ne, mul, ipnof, seed=ne + 10 * ipnof)
    fd = _fd(oracle, f, which)
    g = _formula(oracle, f, which)
    assert abs(g - fd) <= 1e-8 * max(1.0, abs(fd)), (g, fd)


def test_singly_occupied_block_is_excluded(oracle):
    """With nalpha > nbeta the singly occupied orbitals carry no diagonal Coulomb term
    (grads.jl:76-84 use p < nbeta and p >= nalpha), as in calce (pnof.jl:1273-1300)."""
    f = _family(oracle, 10, 24, 7, 2, 8, seed=3)
    assert f["p"].nalpha > f["p"].nbeta
    assert abs(_formula(oracle, f, "I") - _fd(oracle, f, "I")) <= 1e-8


def test_nuclear_repulsion_gradient(oracle):
    rng = np.random.default_rng(1)
    Z = np.array([8.0, 1.0, 1.0, 6.0])
    xyz = rng.standard_normal((4, 3)) * 1.5
    g = oracle.grad_nuc(Z, xyz)
    eps = 1e-5
    for k in range(4):
        for c in range(3):
            xp, xm = xyz.copy(), xyz.copy()
            xp[k, c] += eps
            xm[k, c] -= eps
            fd = (oracle.e_nuc(Z, xp) - oracle.e_nuc(Z, xm)) / (2 * eps)
            fd /= oracle.BOHR_PER_ANGSTROM          # per Bohr
            assert abs(g[k, c] - fd) <= 1e-7 * max(1.0, abs(fd))
    assert np.abs(g.sum(axis=0)).max() <= 1e-12     # translation invariance


def test_coefficient_diagonals_are_ignored(oracle):
    """grads.jl:24-25 zero the diagonals of cj12 and ck12 before the RI terms: values put
    on them change nothing."""
    f = _family(oracle, 10, 24, 6, 1, 8, seed=5)
    p = f["p"]
    b0 = oracle.iaux(f["G0"], f["I0"])
    elag, _ = oracle.compute_lagrange2(f["C0"], f["n"], f["H0"], b0, f["cj12"], f["ck12"],
                                       p.nalpha, p.nbeta)
    base = oracle.geom_gradient_densities(f["n"], f["C0"], f["cj12"], f["ck12"], elag, p,
                                          f["G0"], f["I0"])
    cj, ck = f["cj12"].copy(), f["ck12"].copy()
    np.fill_diagonal(cj, 0.7)
    np.fill_diagonal(ck, -1.3)
    got = oracle.geom_gradient_densities(f["n"], f["C0"], cj, ck, elag, p, f["G0"], f["I0"])
    for a, b in zip(base, got):
        assert np.array_equal(a, b)
    assert np.all(np.diag(cj) == 0.7)            # the caller's arrays are left alone
