"""The occupation optimiser on the device (dn_occoptr; minimization.jl:100-139): the
nonlinear CG with every evaluation, the line search and the iterate in HBM, against the
same CG driven from the host (one device calcocce/calcoccg call per evaluation) and the
oracle's occoptr: the same stationary point (max|calcoccg| <= threshen) and energy to
1e-9 Ha.  The iterates need not agree bitwise: the host sums its dot products in another
order, which can move a line-search decision."""
import numpy as np
import pytest

from donof_amd import api, minimization as mini, synthetic

pytestmark = pytest.mark.gpu


def _setup(nbf, naux, nbf5, ipnof, occ, seed):
    s = synthetic.SyntheticSystem(nbf, naux, nbf5, seed, ndoc=5 if nbf5 % 5 == 0 else 3)
    p = s.p
    p.ipnof, p.occ_method = ipnof, occ
    dev = api.System(p).generate_b(seed)
    api.computeJKH_MO(s.C, s.H, dev, p)          # J, K, H_core resident
    g0 = (mini.guess_gamma_softmax(p.ndoc, p.ncwo) if occ == "Softmax"
          else mini.guess_gamma_trigonometric(p.ndoc, p.ncwo))
    return s, p, dev, g0


@pytest.mark.parametrize("ipnof", [7, 8])
@pytest.mark.parametrize("occ", ["Softmax", "Trigonometric"])
@pytest.mark.parametrize("nbf,naux,nbf5", [(40, 160, 15), (120, 480, 40)])
def test_device_cg_matches_host_cg(oracle, nbf, naux, nbf5, ipnof, occ):
    print("case occoptr", nbf, naux, nbf5, ipnof, occ, flush=True)
    s, p, dev, g0 = _setup(nbf, naux, nbf5, ipnof, occ, 11 + nbf)
    tol = 1e-7
    x, E, nit, conv = api.occ_conjugate_gradient(g0, None, None, None, p, dev, tol)
    host = mini.occoptr_host_cg(g0, p, dev, g_abstol=tol)
    assert conv and host.g_converged and nit >= 1
    g = api.calcoccg(x, None, None, None, p, dev)
    assert np.max(np.abs(g)) <= tol
    # the minimum is E(gamma) (the device sums the energy in other slices)
    assert abs(E - api.calcocce(x, None, None, None, p, dev)) <= 1e-12 * max(1.0, abs(E))
    assert abs(E - host.minimum) <= 1e-9
    # and the oracle's energy at the device's gamma
    J, K, Hc = oracle.jkh_mo(s.C, s.H, dev.download_b(), nbf5)
    assert abs(oracle.calcocce(x, J, K, Hc, p) - E) <= 1e-9
    dev.close()


def test_occoptr_uses_the_device_cg(oracle):
    """occoptr at 400/1600/120 (GNOF, Softmax, threshen = 1e-4): the returned gamma is a
    stationary point and cj12/ck12 are those of its occupations."""
    s, p, dev, g0 = _setup(400, 1600, 120, 8, "Softmax", 2024)
    E, nit, conv, gamma, n, cj12, ck12 = mini.occoptr(g0, s.C, s.H, dev, False, p)
    assert conv and nit >= 1
    assert np.max(np.abs(api.calcoccg(gamma, None, None, None, p, dev))) <= p.threshen
    nr, _ = oracle.ocupacion(gamma, p)
    assert np.abs(n - nr).max() <= 1e-14
    cjr, ckr = oracle.pnof_selector(nr, p)
    assert np.abs(cj12 - cjr).max() <= 1e-13 and np.abs(ck12 - ckr).max() <= 1e-13
    host = mini.occoptr_host_cg(g0, p, dev)
    assert abs(E - host.minimum) <= 1e-9
    dev.close()


@pytest.mark.parametrize("ties", [False, True])
def test_order_occupations_softmax_on_device(oracle, ties):
    """utils.jl:586-616 on the device against the oracle, bitwise: pairs re-ordered when
    some gamma > 0 (with ties: equal gammas, and gammas equal to the strong orbital's 0,
    keep the reverse of their stable order), unchanged otherwise."""
    s = synthetic.SyntheticSystem(40, 80, 15, 3, ndoc=3)
    p = s.p
    dev = api.System(p).generate_b(3)
    rng = np.random.default_rng(1)
    g = rng.standard_normal(p.nv)
    if ties:
        g[::3] = 0.0
        g[1::4] = g[1]
    C2, g2 = mini.order_occupations_softmax(s.C, g, p, dev)
    Cr, gr = oracle.order_occupations_softmax(s.C, g, p)
    assert np.array_equal(C2, Cr) and np.array_equal(g2, gr)
    n1, _ = oracle.ocupacion_softmax(g, p)
    n2, _ = oracle.ocupacion_softmax(g2, p)
    for l in range(1, p.ndoc + 1):
        cols = np.r_[p.no1 + l - 1, oracle.pair_weak(p, l)]
        assert np.allclose(np.sort(n1[cols]), np.sort(n2[cols]))
        assert n2[p.no1 + l - 1] == n2[cols].max()
    C3, g3 = mini.order_occupations_softmax(s.C, -np.abs(g) - 0.1, p, dev)
    assert np.array_equal(C3, s.C) and np.array_equal(g3, -np.abs(g) - 0.1)
    dev.close()
