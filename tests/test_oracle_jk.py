"""Pins the oracle's transform, J/K/H_core, Lagrangian, energies and rotation with checks
that do not restate its formulas: brute-force loops, invariants, finite differences and
library routines."""
import numpy as np
import pytest
import scipy.linalg

from donof_amd import synthetic


def tiny(nbf=7, naux=5, nbf5=6, seed=11, ndoc=2):
    s = synthetic.SyntheticSystem(nbf, naux, nbf5, seed, ndoc=ndoc)
    return s, synthetic.make_b(nbf, naux, seed)


def test_jkh_bruteforce(oracle):
    s, b = tiny()
    nbf, naux, nbf5 = s.nbf, s.naux, s.nbf5
    C, H = s.C, s.H
    bpq = np.zeros((nbf5, nbf5, naux))
    for p in range(nbf5):
        for q in range(nbf5):
            for l in range(naux):
                acc = 0.0
                for m in range(nbf):
                    for n in range(nbf):
                        acc += C[m, p] * b[m, n, l] * C[n, q]
                bpq[p, q, l] = acc
    J = np.zeros((nbf5, nbf5))
    K = np.zeros((nbf5, nbf5))
    for p in range(nbf5):
        for q in range(nbf5):
            J[p, q] = sum(bpq[p, p, l] * bpq[q, q, l] for l in range(naux))
            K[p, q] = sum(bpq[p, q, l] ** 2 for l in range(naux))
    Hc = np.array([sum(C[m, i] * C[n, i] * H[m, n] for m in range(nbf) for n in range(nbf))
                   for i in range(nbf5)])
    Jo, Ko, Ho = oracle.jkh_mo(C, H, b, nbf5)
    assert np.abs(Jo - J).max() < 1e-13
    assert np.abs(Ko - K).max() < 1e-13
    assert np.abs(Ho - Hc).max() < 1e-13


def test_jk_invariants(oracle):
    s, b = tiny(nbf=16, naux=24, nbf5=9, ndoc=3)
    J, K, _ = oracle.jkh_mo(s.C, s.H, b, s.nbf5)
    assert np.allclose(J, J.T, atol=1e-14)
    assert np.allclose(K, K.T, atol=1e-14)
    assert np.linalg.eigvalsh(J).min() > -1e-12          # J = X X^T is PSD
    assert K.min() >= 0.0
    assert np.allclose(np.diag(J), np.diag(K), atol=1e-14)
    # J and K are invariant under b -> b Q for orthogonal Q over the auxiliary index
    Q, _ = np.linalg.qr(np.random.default_rng(0).standard_normal((s.naux, s.naux)))
    bq = np.einsum("mnk,kl->mnl", b, Q)
    J2, K2, _ = oracle.jkh_mo(s.C, s.H, bq, s.nbf5)
    assert np.abs(J2 - J).max() < 1e-12 and np.abs(K2 - K).max() < 1e-12


def test_lagrange_bruteforce(oracle):
    s, b = tiny(nbf=6, naux=3, nbf5=6, ndoc=2)
    p = s.p
    nbf, naux, nbf5 = s.nbf, s.naux, s.nbf5
    C, H, n = s.C, s.H, s.n
    rng = np.random.default_rng(5)
    cj = rng.standard_normal((nbf5, nbf5))
    ck = rng.standard_normal((nbf5, nbf5))
    bMO = np.einsum("mp,mnl,nq->pql", C, b, C[:, :nbf5])
    Hm = C.T @ H @ C[:, :nbf5]
    el = np.zeros((nbf, nbf))
    for a in range(nbf):
        for bb in range(nbf5):
            v = n[bb] * Hm[a, bb]
            if p.nbeta > 0 and (bb < p.nbeta or bb >= p.nalpha):
                v += n[bb] * sum(bMO[a, bb, k] * bMO[bb, bb, k] for k in range(naux))
            for k in range(naux):
                v += bMO[a, bb, k] * sum(cj[bb, q] * bMO[q, q, k] for q in range(nbf5))
                for q in range(nbf5):
                    v -= ck[bb, q] * bMO[a, q, k] * bMO[bb, q, k]
            el[a, bb] = v
    elo, Hmo = oracle.compute_lagrange2(C, n, H, b, cj, ck, p.nalpha, p.nbeta)
    assert np.abs(elo - el).max() < 1e-12
    assert np.abs(Hmo - Hm).max() < 1e-13


@pytest.mark.parametrize("ipnof", [5, 7, 8, 9])
def test_energy_identity(oracle, ipnof):
    """computeE_elec(Hmat, n, elag) equals calce(n, cj12, ck12, J, K, H_core) at the same
    orbitals, and equals the (H, C) form of ENERGY1r."""
    s, b = tiny(nbf=14, naux=20, nbf5=12, ndoc=3)
    p = s.p
    p.ipnof = ipnof
    cj, ck = oracle.pnof_selector(s.n, p)
    el, Hm = oracle.compute_lagrange2(s.C, s.n, s.H, b, cj, ck, p.nalpha, p.nbeta)
    J, K, Hc = oracle.jkh_mo(s.C, s.H, b, s.nbf5)
    E1 = oracle.computeE_elec_hmat(Hm, s.n, el, s.nbf5)
    E2 = oracle.calce(s.n, cj, ck, J, K, Hc, p)
    E3 = oracle.computeE_elec_hc(s.H, s.C, s.n, el, p)
    assert abs(E1 - E2) < 1e-12 and abs(E1 - E3) < 1e-12


def test_orbital_gradient_finite_differences(oracle):
    """The packed 4(elag - elag^T) is the gradient of E(C exp(Y(y))) at y = 0."""
    s, b = tiny(nbf=9, naux=12, nbf5=6, ndoc=2)
    p = s.p
    cj, ck = oracle.pnof_selector(s.n, p)
    el, _ = oracle.compute_lagrange2(s.C, s.n, s.H, b, cj, ck, p.nalpha, p.nbeta)
    g = oracle.pack_grad(el, p.nbf5)

    def E(y):
        C2 = oracle.rotate_orbital(y, s.C, p.nbf5)
        e2, H2 = oracle.compute_lagrange2(C2, s.n, s.H, b, cj, ck, p.nalpha, p.nbeta)
        return oracle.computeE_elec_hmat(H2, s.n, e2, p.nbf5)

    h = 1e-5
    fd = np.array([(E(h * e) - E(-h * e)) / (2 * h) for e in np.eye(p.nvar)])
    assert np.abs(fd - g).max() < 1e-7 * max(1.0, np.abs(g).max())
    # calcorbg at y = 0 is the same packing
    g2 = oracle.calcorbg(np.zeros(p.nvar), s.n, cj, ck, s.C, s.H, b, p)
    assert np.abs(g2 - g).max() < 1e-14


def test_rotation(oracle):
    rng = np.random.default_rng(3)
    nbf, nbf5 = 10, 4
    nvar = nbf * nbf5 - nbf5 * (nbf5 + 1) // 2
    C, _ = np.linalg.qr(rng.standard_normal((nbf, nbf)))
    y = 0.3 * rng.standard_normal(nvar)
    Cn = oracle.rotate_orbital(y, C, nbf5)
    Y = oracle.rotation_generator(y, nbf, nbf5)
    assert np.allclose(Y, -Y.T)
    assert np.count_nonzero(Y[nbf5:, nbf5:]) == 0
    # independent route: exp of a skew matrix from the Hermitian eigendecomposition of iY
    w, V = np.linalg.eigh(1j * Y)
    U = (V * np.exp(-1j * w)) @ V.conj().T
    assert np.abs(Cn - C @ U.real).max() < 1e-13
    assert np.abs(Cn.T @ Cn - np.eye(nbf)).max() < 1e-13
    assert np.abs(oracle.rotate_orbital(np.zeros(nvar), C, nbf5) - C).max() == 0.0
    r, c = oracle.triu_pairs(nbf, nbf5)
    assert len(r) == nvar and all(r < c) and all(r < nbf5)


def test_generators():
    s = synthetic.SyntheticSystem(20, 30, 10, 9, ndoc=2)
    b = s.host_b()
    assert np.allclose(b, np.transpose(b, (1, 0, 2)))
    assert np.abs(s.C.T @ s.C - np.eye(20)).max() < 1e-13
    assert np.allclose(s.H, s.H.T)
    p = s.p
    for l in range(1, p.ndoc + 1):
        w = p.no1 + p.ndns + (p.ndoc - l) + p.ndoc * np.arange(p.ncwo)
        assert abs(s.n[p.no1 + l - 1] + s.n[w].sum() - 1.0) < 1e-14
    assert len(set(np.round(s.n[p.ndoc:], 12))) == p.nbf5 - p.ndoc   # distinct weak draws
    z = synthetic.normal(np.arange(200000, dtype=np.uint64), 0, 1)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01


def test_param_sizes(dn):
    """set_ncwo and the derived sizes (param.jl:242-266) for the benchmark shapes."""
    for nbf, naux, nbf5, ncwo in ((400, 1600, 120, 23), (800, 3200, 240, 47)):
        p = synthetic.synthetic_param(nbf, naux, nbf5)
        assert (p.ndoc, p.ncwo, p.nbf5, p.nv) == (5, ncwo, nbf5, 5 * ncwo)
        assert p.nvar == nbf * nbf5 - nbf5 * (nbf5 + 1) // 2
    # H2O / cc-pVDZ sizes of the reference's precompile workload: nbf=24, ne=10
    p = dn.Param(nbf=24, nbfaux=116, ne=10)
    assert (p.ndoc, p.ncwo, p.nbf5, p.nv, p.nvar) == (5, 3, 20, 15, 270)


def test_param_has_every_reference_field(dn):
    """Every field of the reference's Param (param.jl:3-57) exists, so code that sets or
    reads any of them runs unchanged; closed follows param.jl:120."""
    import dataclasses
    fields = {f.name for f in dataclasses.fields(dn.Param)}
    reference = (
        "title mol natoms spherical nbf nbfaux nalpha nbeta ne mul ndoc nsoc ndns nvir closed "
        "ipnof ista maxit maxitid maxloop threshl threshe threshec threshen threshgorb "
        "threshgocc thresheid scaling nzeros nzerosm nzerosr itziter diis thdiis ndiis "
        "perdiis ncwo noptorb nac nbf5 no0 no1 nv nvar HighSpin MSpin alpha occ_method "
        "orb_method RI gpu gpu_bits h_cut").split()
    assert not [f for f in reference if f not in fields]
    assert dn.Param(nbf=24, nbfaux=116, ne=10).closed
    assert not dn.Param(nbf=24, nbfaux=116, ne=9, mul=2).closed
