"""The oracle's RI setup (integrals.jl:68-91) against known answers: I3c = b G^{1/2}
must return b; a negative eigenvalue of G must be clamped (its direction removed, not
reflected); the Cholesky route must give the same sum_Q b b and the same J/K."""
import numpy as np

from donof_amd import synthetic


def test_iaux_known_answer(oracle):
    G, I3c, b = synthetic.make_metric(9, 70, 3, cond=1e6)
    out = oracle.iaux(G, I3c)
    assert np.abs(out - b).max() <= 1e-9 * np.abs(b).max()


def test_iaux_clamps_negative_eigenvalues(oracle):
    rng = np.random.default_rng(0)
    naux, nbf = 12, 4
    V, _ = np.linalg.qr(rng.standard_normal((naux, naux)))
    lam = np.linspace(1.0, 0.1, naux)
    lam[3] = -0.5
    G = (V * lam) @ V.T
    I3c = rng.standard_normal((nbf, nbf, naux))
    out = oracle.iaux(G, I3c).reshape(nbf * nbf, naux)
    # the clamped direction carries nothing; the others are scaled by 1/sqrt(lambda)
    assert np.abs(out @ V[:, 3]).max() < 1e-12
    X = I3c.reshape(nbf * nbf, naux)
    assert np.allclose(out @ V[:, 0], X @ V[:, 0] / np.sqrt(lam[0]), atol=1e-12)


def test_cholesky_route_same_contractions(oracle):
    nbf, naux, nbf5 = 10, 40, 6
    G, I3c, b = synthetic.make_metric(nbf, naux, 8, cond=1e4)
    be = oracle.iaux(G, I3c)
    bc = oracle.iaux_cholesky(G, I3c)
    Xe = be.reshape(nbf * nbf, naux, order="F")
    Xc = bc.reshape(nbf * nbf, naux, order="F")
    assert np.abs(Xe @ Xe.T - Xc @ Xc.T).max() <= 1e-10 * np.abs(Xe @ Xe.T).max()
    assert np.abs(Xe - Xc).max() > 1e-3          # the two b differ by a rotation
    C = synthetic.make_C(nbf, 1)
    H = synthetic.make_H(nbf, 2)
    Je, Ke, _ = oracle.jkh_mo(C, H, be, nbf5)
    Jc, Kc, _ = oracle.jkh_mo(C, H, bc, nbf5)
    assert np.abs(Je - Jc).max() <= 1e-10 * np.abs(Je).max()
    assert np.abs(Ke - Kc).max() <= 1e-10 * np.abs(Ke).max()
