"""GPU parity of the occupation functional, the orbital rotation and the ADAM-family
orbital optimiser against the CPU oracle, through the C ABI."""
import numpy as np
import pytest

from donof_amd import api, synthetic, Param

pytestmark = pytest.mark.gpu

CASES = [(4, 0), (5, 0), (7, 0), (7, 1), (8, 0), (8, 1), (8, 2), (8, 3), (8, 4), (8, 5), (9, 0)]


def closed(ipnof=8, ista=0, occ="Softmax", nbf=30, naux=40, nbf5=15, ndoc=3, seed=3):
    s = synthetic.SyntheticSystem(nbf, naux, nbf5, seed, ndoc=ndoc)
    p = s.p
    p.ipnof, p.ista, p.occ_method = ipnof, ista, occ
    return s, p


def open_shell(ipnof=8, ista=0, occ="Softmax"):
    """Triplet: ndoc=3, nsoc=2, ncwo=3 -> nbf5=14 (exercises the singly occupied blocks)."""
    p = Param(nbf=24, nbfaux=30, ne=8, mul=3)
    p.set_ncwo(3)
    p.ipnof, p.ista, p.occ_method = ipnof, ista, occ
    return p


def gammas(p, occ, seed=0):
    rng = np.random.default_rng(seed)
    if occ == "Softmax":
        return rng.standard_normal(p.nv) - 2.0
    return rng.uniform(0.2, 1.3, p.nv)


def jk_of(oracle, p, seed=5):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((p.nbf5, 3 * p.nbf5)) * 0.3
    J = X @ X.T
    K = np.abs(rng.standard_normal((p.nbf5, p.nbf5))) * 0.2
    K = 0.5 * (K + K.T)
    np.fill_diagonal(K, np.diag(J))
    Hc = rng.standard_normal(p.nbf5)
    return J, K, Hc


def close(a, b, tol):
    return np.allclose(a, b, rtol=0, atol=tol, equal_nan=True)


@pytest.mark.parametrize("occ", ["Softmax", "Trigonometric"])
@pytest.mark.parametrize("shell", ["closed", "open"])
def test_ocupacion(oracle, occ, shell):
    p = closed(occ=occ)[1] if shell == "closed" else open_shell(occ=occ)
    dev = api.System(p)
    g = gammas(p, occ)
    n, dn = api.ocupacion(g, p, dev)
    nr, dnr = oracle.ocupacion(g, p)
    assert close(n, nr, 1e-14) and close(dn, dnr, 1e-14)
    dev.close()


@pytest.mark.parametrize("ipnof,ista", CASES)
@pytest.mark.parametrize("shell", ["closed", "open"])
def test_coefficients_and_derivatives(oracle, ipnof, ista, shell):
    p = closed(ipnof, ista)[1] if shell == "closed" else open_shell(ipnof, ista)
    dev = api.System(p)
    g = gammas(p, "Softmax", seed=ipnof + ista)
    n, dn = oracle.ocupacion(g, p)
    cj, ck = api.PNOFi_selector(n, p, dev)
    cjr, ckr = oracle.pnof_selector(n, p)
    assert close(cj, cjr, 1e-14) and close(ck, ckr, 1e-14)
    Dcj, Dck = api.der_PNOFi_selector(n, dn, p, dev)
    Dcjr, Dckr = oracle.der_pnof_selector(n, dn, p)
    assert close(Dcj, Dcjr, 1e-12) and close(Dck, Dckr, 1e-12)
    dev.close()


@pytest.mark.parametrize("ipnof,ista", CASES)
@pytest.mark.parametrize("occ", ["Softmax", "Trigonometric"])
@pytest.mark.parametrize("shell", ["closed", "open"])
def test_calce_calcocce_calcoccg(oracle, ipnof, ista, occ, shell):
    p = closed(ipnof, ista, occ)[1] if shell == "closed" else open_shell(ipnof, ista, occ)
    dev = api.System(p)
    J, K, Hc = jk_of(oracle, p)
    g = gammas(p, occ, seed=ipnof * 7 + ista)
    n, _ = oracle.ocupacion(g, p)
    cj, ck = oracle.pnof_selector(n, p)
    assert abs(api.calce(n, cj, ck, J, K, Hc, p, dev)
               - oracle.calce(n, cj, ck, J, K, Hc, p)) < 1e-11
    Er = oracle.calcocce(g, J, K, Hc, p)
    grr = oracle.calcoccg(g, J, K, Hc, p)
    if np.isfinite(Er):
        assert abs(api.calcocce(g, J, K, Hc, p, dev) - Er) < 1e-11
    else:  # a non-finite result is reported, not returned
        with pytest.raises(api.DnError):
            api.calcocce(g, J, K, Hc, p, dev)
    if np.all(np.isfinite(grr)):
        assert np.abs(api.calcoccg(g, J, K, Hc, p, dev) - grr).max() < 1e-10
    else:
        with pytest.raises(api.DnError):
            api.calcoccg(g, J, K, Hc, p, dev)
    dev.close()


def test_calcoccg_resident_jk(oracle):
    """J/K/H_core left on the device by the J/K build feed calcocce/calcoccg."""
    s, p = closed(8, 0, nbf=40, naux=60, nbf5=15, ndoc=5)
    b = synthetic.make_b(40, 60, s.seed)
    dev = api.System(p).upload_b(b)
    J, K, Hc = api.computeJKH_MO(s.C, s.H, dev, p)
    g = s.gamma + 0.1
    assert abs(api.calcocce(g, None, None, None, p, dev) - oracle.calcocce(g, J, K, Hc, p)) < 1e-11
    assert np.abs(api.calcoccg(g, None, None, None, p, dev)
                  - oracle.calcoccg(g, J, K, Hc, p)).max() < 1e-10
    dev.close()


@pytest.mark.parametrize("nbf,nbf5,scale", [(20, 6, 0.01), (57, 19, 0.3), (130, 40, 2.0),
                                            (200, 60, 20.0)])
def test_rotate_orbital(oracle, nbf, nbf5, scale):
    """FP64 exp on the device against scipy's expm route (the reference's GPU path does
    this in FP32)."""
    p = synthetic.synthetic_param(nbf, 2 * nbf, nbf5, ndoc=1 if nbf5 % 5 and nbf5 % 3 else None)
    dev = api.System(p)
    rng = np.random.default_rng(nbf)
    C, _ = np.linalg.qr(rng.standard_normal((nbf, nbf)))
    y = scale * rng.standard_normal(p.nvar) / np.sqrt(nbf)
    Cn = api.rotate_orbital(y, C, p, dev)
    Cr = oracle.rotate_orbital(y, C, nbf5)
    assert np.abs(Cn - Cr).max() < 1e-12 * max(1.0, scale)
    assert np.abs(Cn.T @ Cn - np.eye(nbf)).max() < 1e-12 * max(1.0, scale)
    dev.close()


def test_rotations_keep_orbitals_orthonormal(oracle):
    """exp(Y) is re-orthonormalised (a Newton-Schulz step) before C <- C exp(Y): over 300
    rotations with many squarings each, C^T C stays at the identity to rounding, below the
    drift of the same rotations through scipy's expm."""
    nbf, nbf5 = 160, 50
    p = synthetic.synthetic_param(nbf, 2 * nbf, nbf5, ndoc=5)
    dev = api.System(p)
    rng = np.random.default_rng(7)
    C, _ = np.linalg.qr(rng.standard_normal((nbf, nbf)))
    Cr = C.copy()
    for _ in range(300):
        y = 8.0 * rng.standard_normal(p.nvar) / np.sqrt(nbf)
        C = api.rotate_orbital(y, C, p, dev)
        Cr = oracle.rotate_orbital(y, Cr, nbf5)
    dev.close()
    e_dev = np.abs(C.T @ C - np.eye(nbf)).max()
    e_ref = np.abs(Cr.T @ Cr - np.eye(nbf)).max()
    print("orthonormality after 300 rotations: device %.2e, expm %.2e" % (e_dev, e_ref))
    assert e_dev <= 2e-14
    assert e_dev <= max(e_ref, 1e-14)


@pytest.mark.parametrize("method", ["ADAM", "ADABelief", "YOGI", "Demon"])
def test_orbopt_matches_oracle(oracle, method):
    s, p = closed(8, 0, nbf=36, naux=72, nbf5=15, ndoc=5, seed=9)
    b = synthetic.make_b(36, 72, 9)
    dev = api.System(p).upload_b(b)
    pr = oracle.make_param(36, 72, 10, ncwo=2, ipnof=8)
    pr.maxloop, pr.alpha = p.maxloop, p.alpha
    E, C, nit, ok = api.orbopt(method, s.gamma, s.C, s.H, dev, p)
    Er, Cr, nitr, okr = oracle.orbopt(method, s.gamma, s.C, s.H, b, pr)
    assert (nit, ok) == (nitr, okr)
    assert (p.maxloop, p.alpha) == (pr.maxloop, pr.alpha)
    assert abs(E - Er) < 1e-9
    assert np.abs(C - Cr).max() < 1e-9
    # a second call continues from the returned orbitals with the mutated p
    E2, C2, nit2, ok2 = api.orbopt(method, s.gamma, C, s.H, dev, p)
    Er2, Cr2, nitr2, okr2 = oracle.orbopt(method, s.gamma, Cr, s.H, b, pr)
    assert (nit2, ok2) == (nitr2, okr2) and abs(E2 - Er2) < 1e-9
    assert E2 <= E + 1e-12
    dev.close()


def test_calcorbg(oracle):
    s, p = closed(8, 0, nbf=30, naux=50, nbf5=15, ndoc=5)
    b = synthetic.make_b(30, 50, s.seed)
    dev = api.System(p).upload_b(b)
    cj, ck = oracle.pnof_selector(s.n, p)
    y = 0.05 * np.random.default_rng(2).standard_normal(p.nvar)
    g = api.calcorbg(y, s.n, cj, ck, s.C, s.H, dev, p)
    gr = oracle.calcorbg(y, s.n, cj, ck, s.C, s.H, b, p)
    assert np.abs(g - gr).max() < 1e-10
    dev.close()
