"""The outer energy() loop (energy.jl:140-220) on the device against the oracle's loop on
the same inputs.  With frozen occupations every step is a deterministic recurrence, so
the per-iteration energies must agree to roundoff.  With the occupation optimiser on,
the two CG implementations differ (the reference's Optim.jl CG is not available; the
oracle uses scipy's CG, the package its own Hager-Zhang CG): both are driven to a tight
gradient tolerance and must reach the same stationary occupations after the first
iteration; later iterations are not compared, since the optimisers' discrete decisions
(best energy, step-size cuts) can flip on roundoff (parity unpinned there)."""
import numpy as np
import pytest

from donof_amd import api, energy, synthetic

pytestmark = pytest.mark.gpu


def _h2o_sized(orb_method, seed=11):
    nbf, naux, nbf5 = 24, 60, 20       # the H2O / cc-pVDZ sizes: ndoc = 5, ncwo = 3
    s = synthetic.SyntheticSystem(nbf, naux, nbf5, seed, ndoc=5, orb_method=orb_method)
    b = synthetic.make_b(nbf, naux, seed)
    return s, b


@pytest.mark.parametrize("orb_method", ["ADAM", "ADABelief", "YOGI", "Demon"])
def test_energy_frozen_occupations_match_oracle(oracle, orb_method):
    s, b = _h2o_sized(orb_method)
    p = s.p
    p.maxit, p.threshgorb, p.threshgocc = 3, 1e-12, 1e-12      # run all maxit iterations
    dev = api.System(p).upload_b(b)
    pr = oracle.make_param(24, 60, 10, ncwo=3, ipnof=8, orb_method=orb_method)
    pr.maxit, pr.threshgorb, pr.threshgocc = 3, 1e-12, 1e-12
    Et, C, n, elag, hist = energy.energy(dev, s.H, s.C, p, n=s.n, freeze_occ=True,
                                         printmode=False)
    Er, Cr, nr, histr, elagr = oracle.energy_loop(b, s.H, s.C, pr, n=s.n, freeze_occ=True)
    assert len(hist) == len(histr) == 3
    for h, hr in zip(hist, histr):
        assert abs(h["E"] - hr[0]) < 1e-8
        assert abs(h["E1r"] - hr[3]) < 1e-8          # ENERGY1r's E and maxdiff
        assert abs(h["maxdiff"] - hr[4]) < 1e-8
    # after order_subspaces (energy.jl:222-224): the same sorted orbitals and occupations
    assert np.abs(C - Cr).max() < 1e-7
    assert np.abs(n - nr).max() < 1e-12
    assert np.abs(elag - elagr).max() < 1e-7
    dev.close()


def test_energy_first_iteration_matches_oracle(oracle):
    s, b = _h2o_sized("ADAM")
    p = s.p
    p.maxit, p.threshen, p.threshgorb, p.threshgocc = 1, 1e-10, 1e-12, 1e-12
    dev = api.System(p).upload_b(b)
    pr = oracle.make_param(24, 60, 10, ncwo=3, ipnof=8, orb_method="ADAM")
    pr.maxit, pr.threshen, pr.threshgorb, pr.threshgocc = 1, 1e-10, 1e-12, 1e-12
    Et, C, n, elag, hist = energy.energy(dev, s.H, s.C, p, printmode=False)
    Er, Cr, nr, histr, _ = oracle.energy_loop(b, s.H, s.C, pr)
    assert abs(hist[0]["E"] - histr[0][0]) < 1e-9
    assert np.abs(n - nr).max() < 1e-6
    dev.close()


def test_energy_driver_decreases_energy():
    s = synthetic.SyntheticSystem(40, 120, 15, 5, ndoc=5)
    p = s.p
    p.maxit = 4
    dev = api.System(p).generate_b(5)
    _, _, n, _, hist = energy.energy(dev, s.H, s.C, p, printmode=False)
    Es = [h["E"] for h in hist]
    assert Es[-1] < Es[0]
    assert abs(n[:p.ndoc].sum() + n[p.ndoc:].sum() - p.ndoc) < 1e-12   # pairs sum to 1
    dev.close()


def test_energy1r_and_order_subspaces_match_oracle(oracle):
    """ENERGY1r (utils.jl:117-136) and order_subspaces (guess.jl:58-106) on the device
    against the oracle at the same orbitals and occupations (FP64 bar)."""
    from donof_amd import api as dapi
    nbf, naux, nbf5 = 30, 90, 24
    s = synthetic.SyntheticSystem(nbf, naux, nbf5, 17, ndoc=4)
    p = s.p
    b = synthetic.make_b(nbf, naux, 17)
    dev = api.System(p).upload_b(b)
    rng = np.random.default_rng(4)
    n0, _ = oracle.ocupacion(rng.standard_normal(p.nv) - 2.0, p)
    cj, ck = oracle.pnof_selector(n0, p)
    E, elag, sumdiff, maxdiff = energy.ENERGY1r(s.C, n0, s.H, dev, cj, ck, p)
    Er, elagr, sumr, maxr = oracle.energy1r(s.C, n0, s.H, b, cj, ck, p)
    assert abs(E - Er) < 1e-9
    assert np.abs(elag - elagr).max() < 1e-10
    assert abs(maxdiff - maxr) < 1e-10 and abs(sumdiff - sumr) < 1e-8
    C1, n1, elag1 = energy.order_subspaces(s.C, n0, elag, s.H, dev, p)
    C1r, n1r, elag1r = oracle.order_subspaces(s.C, n0, elagr, s.H, b, p)
    assert not np.array_equal(C1, s.C)                 # the pairs were reordered
    assert np.array_equal(C1, C1r) and np.array_equal(n1, n1r)
    assert np.abs(elag1 - elag1r).max() < 1e-10
    dev.close()


def test_checkpoint_resume(oracle, tmp_path):
    """energy() with checkpoints, read back by read_C / read_n (guess.jl:1-56) and resumed,
    continues the run: the resumed iteration's energy equals the uninterrupted run's."""
    import copy
    s, b = _h2o_sized("ADAM", seed=13)
    p = s.p
    p.maxit, p.threshgorb, p.threshgocc = 3, 1e-12, 1e-12
    p.title = str(tmp_path / "h2o")
    dev = api.System(p).upload_b(b)
    p_full = copy.deepcopy(p)
    _, _, _, _, hist = energy.energy(dev, s.H, s.C, p_full, n=s.n, freeze_occ=True,
                                     printmode=False)
    p.maxit = 2
    _, C2, n2, _, _ = energy.energy(dev, s.H, s.C, p, n=s.n, freeze_occ=True, printmode=False,
                                    checkpoint=True)
    Cr, nr = energy.read_all(p.title)
    assert np.array_equal(Cr, C2) and np.array_equal(nr, n2)
    p.maxit = 1
    _, _, _, _, hist2 = energy.energy(dev, s.H, Cr, p, n=nr, freeze_occ=True, printmode=False)
    assert abs(hist2[0]["E"] - hist[2]["E"]) < 1e-8
    assert energy.read_C(str(tmp_path / "missing")) is None
    dev.close()
