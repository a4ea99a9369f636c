"""CPU checks of the host-side logic of the package (no GPU): the conjugate-gradient
occupation optimiser, the Softmax re-ordering and the gamma conversions, against the
oracle's functions."""
import numpy as np
import pytest

from donof_amd import minimization as mini, synthetic


def system(occ="Softmax"):
    s = synthetic.SyntheticSystem(20, 40, 15, 21, ndoc=3)
    s.p.occ_method = occ
    return s


def test_gamma_guesses_and_conversions(oracle):
    p = system().p
    assert np.array_equal(mini.guess_gamma_softmax(p.ndoc, p.ncwo),
                          oracle.guess_gamma_softmax(p.ndoc, p.ncwo))
    assert np.array_equal(mini.guess_gamma_trigonometric(p.ndoc, p.ncwo),
                          oracle.guess_gamma_trigonometric(p.ndoc, p.ncwo))
    rng = np.random.default_rng(0)
    g = rng.standard_normal(p.nv) - 2
    n, _ = oracle.ocupacion_softmax(g, p)
    assert np.abs(mini.n_to_gammas_softmax(n, p) - oracle.n_to_gammas_softmax(n, p)).max() == 0
    pt = system("Trigonometric").p
    gt = rng.uniform(0.2, 1.2, pt.nv)
    nt, _ = oracle.ocupacion_trigonometric(gt, pt)
    # round trip through the trigonometric parametrisation
    g2 = mini.n_to_gammas_trigonometric(nt, pt)
    assert np.abs(oracle.ocupacion_trigonometric(g2, pt)[0] - nt).max() < 1e-12


@pytest.mark.parametrize("ipnof", [7, 8, 9])
@pytest.mark.parametrize("occ", ["Softmax", "Trigonometric"])
def test_conjugate_gradient_reaches_stationary_point(oracle, ipnof, occ):
    """(PNOF5 on these random integrals has several stationary points and unbounded
    directions, so only the physical-like functionals are compared.)"""
    s = system(occ)
    p = s.p
    p.ipnof = ipnof
    J, K, Hc = oracle.jkh_mo(s.C, s.H, s.host_b(), p.nbf5)
    g0 = (mini.guess_gamma_softmax(p.ndoc, p.ncwo) if occ == "Softmax"
          else mini.guess_gamma_trigonometric(p.ndoc, p.ncwo))
    res = mini.conjugate_gradient(lambda g: oracle.calcocce(g, J, K, Hc, p),
                                  lambda g: oracle.calcoccg(g, J, K, Hc, p), g0, g_abstol=1e-7)
    assert res.g_converged
    assert np.max(np.abs(oracle.calcoccg(res.x, J, K, Hc, p))) <= 1e-7
    assert res.minimum <= oracle.calcocce(g0, J, K, Hc, p)
    # scipy's CG (the oracle's occoptr, gtol = threshen = 1e-4) reaches the same minimum
    Eo, _, _ = oracle.occoptr(g0, J, K, Hc, p)
    assert abs(res.minimum - Eo) < 1e-7


def test_conjugate_gradient_quadratic():
    rng = np.random.default_rng(3)
    A = rng.standard_normal((12, 12))
    A = A @ A.T + 12 * np.eye(12)
    bvec = rng.standard_normal(12)
    res = mini.conjugate_gradient(lambda x: 0.5 * x @ A @ x - bvec @ x, lambda x: A @ x - bvec,
                                  np.zeros(12), g_abstol=1e-12)
    assert res.g_converged
    assert np.abs(res.x - np.linalg.solve(A, bvec)).max() < 1e-10


def test_grouped_shapes():
    from donof_amd import synthetic
    sh = synthetic.grouped_shapes(64, seed=2024)
    assert len(sh) == 64
    assert all(48 <= n <= 400 and a == 4 * n and n5 == -(-n // 3) for n, a, n5 in sh)
    assert len({n for n, _, _ in sh}) > 40                     # spread over the range
    assert sh == synthetic.grouped_shapes(64, seed=2024)       # seeded
    assert synthetic.build_flops(400, 1600, 120) == pytest.approx(80.0e9, rel=1e-3)
