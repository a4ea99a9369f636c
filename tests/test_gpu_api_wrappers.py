"""The reference's orbital energy and gradient entry points calcorbe (pnof.jl:1379-1387) and
calcorbeg (pnof.jl:1409-1433), and the two occupation parametrisations by name
(pnof.jl:1047-1163), on the device against the oracle and the general calls."""
import numpy as np
import pytest

from donof_amd import api, energy, synthetic

pytestmark = pytest.mark.gpu


def system(seed=11):
    s = synthetic.SyntheticSystem(40, 90, 15, seed, ndoc=3)
    b = synthetic.make_b(s.nbf, s.naux, s.seed)
    dev = api.System(s.p).upload_b(b)
    return s, s.p, b, dev


def test_calcorbe_and_calcorbeg_against_oracle(oracle):
    print("case calcorbe/calcorbeg 40/90/15", flush=True)
    s, p, b, dev = system()
    cj, ck = api.PNOFi_selector(s.n, p, dev)
    rng = np.random.default_rng(2)
    y = 1e-2 * rng.standard_normal(p.nvar)
    E = api.calcorbe(y, s.n, cj, ck, s.C, s.H, dev, p)
    Cr = oracle.rotate_orbital(y, s.C, p.nbf5)
    J, K, Hc = oracle.jkh_mo(Cr, s.H, b, p.nbf5)
    cjr, ckr = oracle.pnof_selector(s.n, p)
    Er = oracle.calce(s.n, cjr, ckr, J, K, Hc, p)
    assert abs(E - Er) < 1e-9 * max(1.0, abs(Er))
    G = np.zeros(p.nvar)
    Eg = api.calcorbeg(True, G, y, s.n, cj, ck, s.C, s.H, dev, p)
    gr = oracle.calcorbg(y, s.n, cjr, ckr, s.C, s.H, b, p)
    assert np.abs(G - gr).max() < 1e-10 * max(1.0, np.abs(gr).max())
    assert np.array_equal(G, api.calcorbg(y, s.n, cj, ck, s.C, s.H, dev, p))
    # the energy of the Lagrangian equals the J/K energy at the same orbitals (closed shell)
    assert abs(Eg - E) < 1e-9 * max(1.0, abs(E))
    E1 = energy.ENERGY1r(api.rotate_orbital(y, s.C, p, dev), s.n, s.H, dev, cj, ck, p)[0]
    assert abs(Eg - E1) < 1e-12 * max(1.0, abs(E1))
    assert api.calcorbeg(None, G, y, s.n, cj, ck, s.C, s.H, dev, p) is None
    dev.close()


def test_occupation_parametrisations_by_name(oracle):
    s, p, b, dev = system(12)
    rng = np.random.default_rng(4)
    g = rng.standard_normal(p.nv) - 2.0
    n1, d1 = api.ocupacion_softmax(g, p, dev)
    q = p.copy()
    q.occ_method = "Softmax"
    n2, d2 = api.ocupacion(g, q, dev)
    assert np.array_equal(n1, n2) and np.array_equal(d1, d2)
    t = rng.uniform(0.2, 1.3, p.nv)
    n3, _ = api.ocupacion_trigonometric(t, p, dev)
    q.occ_method = "Trigonometric"
    nr, _ = oracle.ocupacion(t, q)
    assert np.abs(n3 - nr).max() < 1e-13
    assert p.occ_method == s.p.occ_method       # the caller's Param is left as it was
    dev.close()
