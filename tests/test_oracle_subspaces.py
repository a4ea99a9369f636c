"""Pins the oracle's ENERGY1r and order_subspaces (guess.jl:58-106) by invariances: the
energy of a natural-orbital functional does not depend on how its electron pairs are
numbered, so sorting the pairs leaves E and the Lagrangian's eigen-structure unchanged and
only permutes its diagonal."""
import numpy as np
import pytest

from donof_amd import synthetic


@pytest.mark.parametrize("ipnof", [5, 7, 8, 9])
def test_order_subspaces_relabels_pairs(oracle, ipnof):
    nbf, naux, nbf5, ndoc = 18, 30, 12, 3
    s = synthetic.SyntheticSystem(nbf, naux, nbf5, 31, ndoc=ndoc, ipnof=ipnof)
    p = s.p
    b = synthetic.make_b(nbf, naux, 31)
    rng = np.random.default_rng(2)
    n, _ = oracle.ocupacion(rng.standard_normal(p.nv) - 2.0, p)
    cj, ck = oracle.pnof_selector(n, p)
    E0, elag0, _, _ = oracle.energy1r(s.C, n, s.H, b, cj, ck, p)
    C1, n1, elag1 = oracle.order_subspaces(s.C, n, elag0, s.H, b, p)
    d0 = np.diag(elag0)[:ndoc]
    d1 = np.diag(elag1)[:ndoc]
    assert np.all(np.diff(d1) >= 0)                       # sorted by increasing elag_ii
    assert np.allclose(np.sort(d0), d1, atol=1e-12)
    order = np.argsort(d0, kind="stable")
    assert not np.array_equal(order, np.arange(ndoc))     # the case really permutes
    # each pair moved as a whole: strong column and its weak columns, with occupations
    for i, i_old in enumerate(order, start=1):
        assert np.array_equal(C1[:, i - 1], s.C[:, i_old])
        assert np.array_equal(C1[:, oracle.pair_weak(p, i)], s.C[:, oracle.pair_weak(p, i_old + 1)])
        assert np.array_equal(n1[oracle.pair_weak(p, i)], n[oracle.pair_weak(p, i_old + 1)])
    assert np.array_equal(C1[:, nbf5:], s.C[:, nbf5:])
    cj1, ck1 = oracle.pnof_selector(n1, p)
    E1, elag1b, _, _ = oracle.energy1r(C1, n1, s.H, b, cj1, ck1, p)
    assert abs(E1 - E0) < 1e-11 * max(1.0, abs(E0))      # pair labels do not change E
    assert np.abs(elag1b - elag1).max() == 0.0
    # the Lagrangian of the relabelled system is the permuted Lagrangian
    perm = np.arange(nbf)
    for i, i_old in enumerate(order, start=1):
        perm[i - 1] = i_old
        perm[oracle.pair_weak(p, i)] = oracle.pair_weak(p, i_old + 1)
    assert np.abs(elag1 - elag0[np.ix_(perm, perm)]).max() < 1e-11


def test_energy1r_matches_calce(oracle):
    """ENERGY1r's E equals calce at the same orbitals (the identity that pins the
    Lagrangian), and maxdiff is the largest |elag - elag^T|."""
    nbf, naux, nbf5 = 16, 24, 12
    s = synthetic.SyntheticSystem(nbf, naux, nbf5, 3, ndoc=3)
    p = s.p
    b = synthetic.make_b(nbf, naux, 3)
    cj, ck = oracle.pnof_selector(s.n, p)
    E, elag, sumdiff, maxdiff = oracle.energy1r(s.C, s.n, s.H, b, cj, ck, p)
    J, K, Hc = oracle.jkh_mo(s.C, s.H, b, nbf5)
    assert abs(E - oracle.calce(s.n, cj, ck, J, K, Hc, p)) < 1e-11
    a = np.abs(elag - elag.T)
    assert maxdiff == a.max() and abs(sumdiff - a.sum()) < 1e-12
