"""Pins the oracle's occupation functions, PNOF coefficients, their gamma-derivatives and
the occupation energy/gradient with closed forms and finite differences."""
import numpy as np
import pytest

from donof_amd import synthetic


def system(ipnof=8, ista=0, occ="Softmax", nbf=16, naux=20, nbf5=12, ndoc=3, seed=4):
    s = synthetic.SyntheticSystem(nbf, naux, nbf5, seed, ndoc=ndoc)
    p = s.p
    p.ipnof, p.ista, p.occ_method = ipnof, ista, occ
    return s, p


def jk(oracle, s):
    return oracle.jkh_mo(s.C, s.H, synthetic.make_b(s.nbf, s.naux, s.seed), s.nbf5)


def test_softmax_pairs_and_jacobian(oracle):
    s, p = system()
    rng = np.random.default_rng(0)
    g = rng.standard_normal(p.nv)
    n, dn = oracle.ocupacion_softmax(g, p)
    for l in range(1, p.ndoc + 1):
        assert abs(n[p.no1 + l - 1] + n[oracle.pair_weak(p, l)].sum() - 1) < 1e-14
    h = 1e-6
    fd = np.stack([(oracle.ocupacion_softmax(g + h * e, p)[0]
                    - oracle.ocupacion_softmax(g - h * e, p)[0]) / (2 * h)
                   for e in np.eye(p.nv)], axis=1)
    assert np.abs(fd - dn).max() < 1e-9
    # the gamma guess and n -> gamma round trip
    n0, _ = oracle.ocupacion_softmax(oracle.guess_gamma_softmax(p.ndoc, p.ncwo), p)
    assert np.allclose(n0[:p.ndoc], 0.999)
    g2 = oracle.n_to_gammas_softmax(n, p)
    assert np.abs(oracle.ocupacion_softmax(g2, p)[0] - n).max() < 1e-13


def test_trigonometric_pairs_and_jacobian(oracle):
    s, p = system(occ="Trigonometric")
    rng = np.random.default_rng(1)
    g = rng.uniform(0.2, 1.2, p.nv)
    n, dn = oracle.ocupacion_trigonometric(g, p)
    for l in range(1, p.ndoc + 1):
        assert abs(n[p.no1 + l - 1] + n[oracle.pair_weak(p, l)].sum() - 1) < 1e-14
    h = 1e-6
    fd = np.stack([(oracle.ocupacion_trigonometric(g + h * e, p)[0]
                    - oracle.ocupacion_trigonometric(g - h * e, p)[0]) / (2 * h)
                   for e in np.eye(p.nv)], axis=1)
    assert np.abs(fd - dn).max() < 1e-9
    n0, _ = oracle.ocupacion_trigonometric(oracle.guess_gamma_trigonometric(p.ndoc, p.ncwo), p)
    assert np.allclose(n0[:p.ndoc], 0.999)


CASES = [(5, 0), (7, 0), (7, 1), (8, 0), (8, 1), (8, 2), (8, 3), (8, 4), (8, 5), (9, 0)]


@pytest.mark.parametrize("ipnof,ista", CASES)
@pytest.mark.parametrize("occ", ["Softmax", "Trigonometric"])
def test_calcoccg_finite_differences(oracle, ipnof, ista, occ):
    s, p = system(ipnof, ista, occ)
    J, K, Hc = jk(oracle, s)
    rng = np.random.default_rng(ipnof * 10 + ista)
    g = (rng.standard_normal(p.nv) - 2.0) if occ == "Softmax" else rng.uniform(0.2, 1.3, p.nv)
    grad = oracle.calcoccg(g, J, K, Hc, p)
    h = 1e-6
    fd = np.array([(oracle.calcocce(g + h * e, J, K, Hc, p)
                    - oracle.calcocce(g - h * e, J, K, Hc, p)) / (2 * h) for e in np.eye(p.nv)])
    assert np.abs(fd - grad).max() < 1e-7 * max(1.0, np.abs(grad).max())


@pytest.mark.parametrize("ipnof,ista", CASES)
def test_coefficients_symmetric_and_row_derivatives(oracle, ipnof, ista):
    """ck12/cj12 are symmetric and Dc[i,j,k] + Dc[j,i,k] is the total gamma-derivative
    (the derivative arrays carry the row index's part only)."""
    s, p = system(ipnof, ista)
    rng = np.random.default_rng(7)
    g = rng.standard_normal(p.nv) - 2.0
    n, dn = oracle.ocupacion(g, p)
    cj, ck = oracle.pnof_selector(n, p)
    assert np.allclose(cj, cj.T, atol=1e-15) and np.allclose(ck, ck.T, atol=1e-15)
    Dcj, Dck = oracle.der_pnof_selector(n, dn, p)
    h = 1e-6
    for k in range(p.nv):
        e = np.zeros(p.nv)
        e[k] = h
        cjp, ckp = oracle.pnof_selector(oracle.ocupacion(g + e, p)[0], p)
        cjm, ckm = oracle.pnof_selector(oracle.ocupacion(g - e, p)[0], p)
        assert np.abs((cjp - cjm) / (2 * h) - (Dcj[:, :, k] + Dcj[:, :, k].T)).max() < 1e-7
        assert np.abs((ckp - ckm) / (2 * h) - (Dck[:, :, k] + Dck[:, :, k].T)).max() < 1e-7


@pytest.mark.parametrize("ipnof", [5, 7, 8, 9])
def test_hartree_fock_limit(oracle, ipnof):
    """With integer occupations the functionals reduce to closed-shell HF:
    E = 2 sum_i H_ii + sum_ij (2 J_ij - K_ij) over the doubly occupied orbitals."""
    s, p = system(ipnof)
    J, K, Hc = jk(oracle, s)
    n = np.zeros(p.nbf5)
    n[:p.ndoc] = 1.0
    cj, ck = oracle.pnof_selector(n, p)
    E = oracle.calce(n, cj, ck, J, K, Hc, p)
    o = slice(0, p.ndoc)
    Ehf = 2 * Hc[o].sum() + (2 * J[o, o] - K[o, o]).sum()
    assert abs(E - Ehf) < 1e-12


def test_gnof_static_and_dynamic_terms(oracle):
    """Closed forms of single GNOF entries (ista = 0)."""
    s, p = system(8, 0)
    rng = np.random.default_rng(3)
    n, _ = oracle.ocupacion(rng.standard_normal(p.nv) - 2.0, p)
    cj, ck = oracle.pnof_selector(n, p)
    hc = p.h_cut
    w1 = oracle.pair_weak(p, 1)          # weak orbitals of strong orbital 0
    w2 = oracle.pair_weak(p, 2)
    i, a, b2 = 0, w1[0], w2[1]
    # strong 0 with its own weak orbital: intra-pair sqrt(n_i n_a)
    assert ck[i, a] == pytest.approx(np.sqrt(n[i] * n[a]), abs=1e-15)
    assert cj[i, a] == 0.0
    # weak orbitals of different pairs: n n + fi fi - nd12 nd12 - nd nd
    fi = np.sqrt(n * (1 - n))
    F1 = np.exp(-((1 - n[0]) / hc) ** 2)
    F2 = np.exp(-((1 - n[1]) / hc) ** 2)
    nda, ndb = n[a] * F1, n[b2] * F2
    expect = n[a] * n[b2] + fi[a] * fi[b2] - np.sqrt(nda * ndb) - nda * ndb
    assert ck[a, b2] == pytest.approx(expect, abs=1e-15)
    assert cj[a, b2] == pytest.approx(2 * n[a] * n[b2], abs=1e-15)
    # strong 0 with a weak orbital of pair 2: beta-cwo block
    nd0 = n[0] * F1
    expect = n[0] * n[b2] + fi[0] * fi[b2] + np.sqrt(nd0 * ndb) - nd0 * ndb
    assert ck[0, b2] == pytest.approx(expect, abs=1e-15)


# ---- pins that do not restate the coefficient formulas --------------------------------

def _two_electron_energy(n, J, K, Hc):
    """Energy of the two-electron Loewdin-Shull wave function sum_i c_i |i alpha i beta|
    in natural orbitals, c_strong = +sqrt(n), c_weak = -sqrt(n):
    E = sum_i n_i (2 h_ii + (ii|ii)) + sum_{i != j} c_i c_j (ij|ij)."""
    c = -np.sqrt(n)
    c[0] = np.sqrt(n[0])
    E = float(np.sum(n * (2.0 * Hc + np.diag(J))))
    cc = np.outer(c, c)
    np.fill_diagonal(cc, 0.0)
    return E + float(np.sum(cc * K))


@pytest.mark.parametrize("ipnof,ista", [(4, 0), (5, 0), (7, 0), (7, 1), (8, 0), (8, 1),
                                         (8, 2), (8, 3), (8, 4), (8, 5), (9, 0)])
def test_one_pair_reduces_to_loewdin_shull(oracle, ipnof, ista):
    """With one electron pair (ndoc = 1) every functional is exact: its energy is the
    Loewdin-Shull two-electron energy built from (ij|ij) = K_ij alone."""
    s, p = system(ipnof, ista, nbf=14, naux=30, nbf5=9, ndoc=1)
    J, K, Hc = jk(oracle, s)
    rng = np.random.default_rng(11 + ipnof)
    n, _ = oracle.ocupacion(rng.standard_normal(p.nv) - 1.5, p)
    assert abs(n.sum() - 1.0) < 1e-14
    cj, ck = oracle.pnof_selector(n, p)
    E = oracle.calce(n, cj, ck, J, K, Hc, p)
    assert abs(E - _two_electron_energy(n, J, K, Hc)) < 1e-12 * max(1.0, abs(E))


@pytest.mark.parametrize("ipnof,ista", [(4, 0), (5, 0), (7, 0), (8, 0), (8, 3), (9, 0)])
def test_cumulant_sum_rule(oracle, ipnof, ista):
    """The cumulant part Delta = n n^T - cj12/2 of the reconstructed two-matrix obeys the
    sum rule sum_{q != p} Delta_qp = n_p (1 - n_p) for every p (the pairing of the
    occupations and of the coefficients agree)."""
    s, p = system(ipnof, ista, ndoc=3)
    rng = np.random.default_rng(21)
    n, _ = oracle.ocupacion(rng.standard_normal(p.nv) - 2.0, p)
    cj, _ = oracle.pnof_selector(n, p)
    D = np.outer(n, n) - cj / 2.0
    np.fill_diagonal(D, 0.0)
    assert np.abs(D.sum(axis=0) - n * (1.0 - n)).max() < 1e-14


@pytest.mark.parametrize("ista", [1, 5])
def test_gnof_quenched_dynamics_is_pnof7(oracle, ista):
    """GNOF with the interpair static term on every block (ista 1, 5) and its dynamic
    factor quenched (1 - n_g >= 10 h_cut for every pair) is PNOF7 (ista 0)."""
    s, p = system(8, ista, ndoc=3)
    p.h_cut = 1e-3
    rng = np.random.default_rng(5)
    n, _ = oracle.ocupacion(rng.standard_normal(p.nv) - 1.0, p)
    assert (1.0 - n[:p.ndoc]).min() >= 10 * p.h_cut
    cj8, ck8 = oracle.pnof_selector(n, p)
    p.ipnof, p.ista = 7, 0
    cj7, ck7 = oracle.pnof_selector(n, p)
    assert np.abs(cj8 - cj7).max() == 0.0
    assert np.abs(ck8 - ck7).max() < 1e-15


@pytest.mark.parametrize("seed", [2, 9])
def test_gnof_dynamic_terms_carry_the_pairing_signs(oracle, seed):
    """The dynamic part of GNOF's exchange-type coefficients (GNOF minus PNOF7 at the same
    occupations, ista 1, where the quenched GNOF is PNOF7 exactly) acts only between a
    strong orbital and the weak orbitals of another pair and between weak orbitals of
    different pairs, and there equals -(c c^T + n_d n_d^T) with c the dynamic amplitudes
    signed as the two-electron (Loewdin-Shull) amplitudes: c = +sqrt(n_d) on strong and
    -sqrt(n_d) on weak orbitals; n_d = n exp(-((1 - n_g) / h_cut)^2) with n_g the strong
    occupation of the orbital's pair (pnof.jl:506-519).  One sign vector for both blocks
    fixes their relative sign; the energy it adds with the molecule's K >= 0 is negative
    (dynamic correlation lowers the energy), which fixes the absolute one."""
    s, p = system(8, 1, nbf=18, naux=40, nbf5=12, ndoc=3, seed=seed)
    J, K, Hc = jk(oracle, s)
    rng = np.random.default_rng(seed)
    # strong occupations within a few h_cut of 1, where the dynamic factor is of order one
    n, _ = oracle.ocupacion(0.3 * rng.standard_normal(p.nv) - 5.0, p)
    _, ck8 = oracle.pnof_selector(n, p)
    p.ipnof, p.ista = 7, 0
    cj7, ck7 = oracle.pnof_selector(n, p)
    p.ipnof, p.ista = 8, 1
    D = ck8 - ck7
    pair = np.full(p.nbf5, -1)
    strong = np.zeros(p.nbf5, dtype=bool)
    for l in range(1, p.ndoc + 1):
        g = p.no1 + l - 1
        pair[g] = l
        strong[g] = True
        pair[oracle.pair_weak(p, l)] = l
    F = np.exp(-((1.0 - n[p.no1:p.no1 + p.ndoc]) / p.h_cut) ** 2)
    assert F.min() > 0.05
    nd = n * F[pair - 1]
    c = np.where(strong, 1.0, -1.0) * np.sqrt(nd)
    inter = pair[:, None] != pair[None, :]
    dyn = inter & ~(strong[:, None] & strong[None, :])
    expect = np.where(dyn, -(np.outer(c, c) + np.outer(nd, nd)), 0.0)
    assert np.abs(D - expect).max() < 1e-15
    # the energy the dynamic terms add: negative for the molecule-like K >= 0
    assert K.min() >= 0.0
    E8 = oracle.calce(n, cj7, ck8, J, K, Hc, p)
    E7 = oracle.calce(n, cj7, ck7, J, K, Hc, p)
    assert E8 - E7 < -1e-4


@pytest.mark.parametrize("ipnof", [8, 9])
@pytest.mark.parametrize("moved", [1, 2, 3])
def test_dynamic_factor_is_local_to_its_pair(oracle, ipnof, moved):
    """The dynamic occupations of pair g depend on pair g's occupations only: moving
    weight inside one pair leaves every coefficient between orbitals of the two other
    pairs unchanged, and changes those between the moved pair and the others."""
    s, p = system(ipnof, 0, ndoc=3)
    rng = np.random.default_rng(8)
    # strong occupations near 1 - h_cut, where the dynamic factor is of order one
    g = 0.3 * rng.standard_normal(p.nv) - 5.0
    n1, _ = oracle.ocupacion(g, p)
    F = np.exp(-((1.0 - n1[:p.ndoc]) / p.h_cut) ** 2)
    assert F.min() > 0.05
    g2 = g.copy()
    g2[oracle.pair_gamma(p, moved)] += 0.7
    n2, _ = oracle.ocupacion(g2, p)
    members = {l: np.r_[p.no1 + l - 1, oracle.pair_weak(p, l)] for l in (1, 2, 3)}
    a, c = [members[l] for l in (1, 2, 3) if l != moved]
    b = members[moved]
    assert np.abs(n1[a] - n2[a]).max() == 0.0 and np.abs(n1[c] - n2[c]).max() == 0.0
    assert np.abs(n1[b] - n2[b]).max() > 1e-3
    _, ck1 = oracle.pnof_selector(n1, p)
    _, ck2 = oracle.pnof_selector(n2, p)
    assert np.abs(ck1[np.ix_(a, c)] - ck2[np.ix_(a, c)]).max() == 0.0
    assert np.abs(ck1[np.ix_(a, b)] - ck2[np.ix_(a, b)]).max() > 1e-6


def test_pnof4_derivative_against_its_factors(oracle):
    """der_CJCKD4 (pnof.jl:175-272) against finite differences.  cj12's derivative is the
    total derivative.  In ck12 the occupied-virtual block of Pi = -sqrt(R) sqrt(n_p - n_q
    + R), R = h_p n_q / S_F (pnof.jl:128-139) is differentiated with the denominator
    2 sqrt(n_p n_q + R) where the chain rule gives 2 sqrt(n_p - n_q + R) (pnof.jl:255 as
    written); every other entry is the total derivative.  The finite differences of the
    two factors sqrt(R) and n_p - n_q + R, combined with that denominator, predict the
    difference entry by entry."""
    s, p = system(4, 0, nbf=14, naux=20, nbf5=12, ndoc=3)
    rng = np.random.default_rng(9)
    g = rng.standard_normal(p.nv) - 2.0
    n, dn = oracle.ocupacion(g, p)
    Dcj, Dck = oracle.der_pnof_selector(n, dn, p)
    nd = p.ndoc
    occ, vir = np.arange(nd), np.arange(nd, p.nbf5)

    def factors(nn):
        h = 1.0 - nn
        R = np.outer(h[occ], nn[vir]) / h[occ].sum()
        return np.sqrt(R), nn[occ][:, None] - nn[vir][None, :] + R, R

    sR, arg, R = factors(n)
    prod = np.outer(n[occ], n[vir]) + R
    hstep = 1e-6
    for k in range(p.nv):
        e = np.zeros(p.nv)
        e[k] = hstep
        np_, _ = oracle.ocupacion(g + e, p)
        nm_, _ = oracle.ocupacion(g - e, p)
        cjp, ckp = oracle.pnof_selector(np_, p)
        cjm, ckm = oracle.pnof_selector(nm_, p)
        fd_cj = (cjp - cjm) / (2 * hstep)
        fd_ck = (ckp - ckm) / (2 * hstep)
        assert np.abs(fd_cj - (Dcj[:, :, k] + Dcj[:, :, k].T)).max() < 1e-7
        sRp, argp, _ = factors(np_)
        sRm, argm, _ = factors(nm_)
        dsR = (sRp - sRm) / (2 * hstep)
        darg = (argp - argm) / (2 * hstep)
        # Pi_pq's chain rule against the written denominator, on the occupied-virtual block
        dpi_true = -(dsR * np.sqrt(arg) + sR * darg / (2 * np.sqrt(arg)))
        dpi_code = -(dsR * np.sqrt(arg) + sR * darg / np.maximum(2 * np.sqrt(prod), 1e-7))
        expect = np.zeros((p.nbf5, p.nbf5))
        expect[np.ix_(occ, vir)] = dpi_true - dpi_code
        expect[np.ix_(vir, occ)] = (dpi_true - dpi_code).T
        got = (Dck[:, :, k] + Dck[:, :, k].T) - fd_ck
        assert np.abs(got - expect).max() < 1e-6
    # the denominators differ here, so the quirk is visible
    assert np.abs(np.sqrt(prod) - np.sqrt(arg)).max() > 1e-3
