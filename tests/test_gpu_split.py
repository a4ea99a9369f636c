"""Split-precision J/K-MO build and Lagrangian (FP16 hi+lo operands, FP32 accumulation,
FP64 reductions) against the FP64 oracle.  Stated error bounds (relative to max|J|, max|K|,
max|elag|): 5e-6 for J, K and the Lagrangian (and so for the orbital gradient
4 (elag - elag^T)); energies within 1e-6 Ha of FP64 on the same orbitals."""
import numpy as np
import pytest

from donof_amd import api, synthetic
from donof_amd._lib import DN_SPLIT_F16

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("use_sym", [True, False])
@pytest.mark.parametrize("nbf,naux,nbf5,ndoc", [(64, 128, 20, 5), (130, 300, 45, 5),
                                                (200, 800, 60, 5), (256, 512, 120, 5),
                                                (37, 70, 10, 5)])
def test_split_jk_against_oracle(oracle, nbf, naux, nbf5, ndoc, use_sym):
    print("case split", nbf, naux, nbf5, use_sym, flush=True)
    s = synthetic.SyntheticSystem(nbf, naux, nbf5, 31 + nbf, ndoc=ndoc)
    p = s.p
    b = synthetic.make_b(nbf, naux, s.seed)
    dev = api.System(p, precision=DN_SPLIT_F16).upload_b(b)
    assert dev.b_symmetric
    dev.set_symmetry(use_sym)
    J, K, Hc = api.computeJKH_MO(s.C, s.H, dev, p)
    Jr, Kr, Hr = oracle.jkh_mo(s.C, s.H, b, nbf5)
    ej = np.abs(J - Jr).max() / np.abs(Jr).max()
    ek = np.abs(K - Kr).max() / np.abs(Kr).max()
    print("split rel err J %.2e K %.2e" % (ej, ek))
    assert ej < 5e-6 and ek < 5e-6
    assert np.abs(Hc - Hr).max() < 1e-10          # H_core stays FP64
    cj, ck = oracle.pnof_selector(s.n, p)
    dE = abs(oracle.calce(s.n, cj, ck, J, K, Hc, p) - oracle.calce(s.n, cj, ck, Jr, Kr, Hr, p))
    print("split |dE| %.2e" % dE)
    assert dE < 1e-6
    dev.close()


def test_split_generated_matches_fp64_device():
    """Same generated tensor in both precisions: split J/K against the FP64 device path."""
    s = synthetic.SyntheticSystem(160, 640, 40, 77, ndoc=5)
    p = s.p
    d64 = api.System(p).generate_b(77)
    J64, K64, H64 = api.computeJKH_MO(s.C, s.H, d64, p)
    d64.close()
    d16 = api.System(p, precision=DN_SPLIT_F16).generate_b(77)
    J16, K16, H16 = api.computeJKH_MO(s.C, s.H, d16, p)
    assert np.abs(J16 - J64).max() / np.abs(J64).max() < 5e-6
    assert np.abs(K16 - K64).max() / np.abs(K64).max() < 5e-6
    d16.close()


@pytest.mark.parametrize("nbf,naux,nbf5", [(64, 128, 20), (130, 300, 45), (37, 70, 10),
                                           (200, 800, 60)])
def test_split_lagrangian_against_oracle(oracle, nbf, naux, nbf5):
    """The Lagrangian of a split system (utils.jl:1-59): both transforms on FP16 planes,
    the C^J/C^K terms and the contraction in FP64; b is not resident in FP64."""
    from donof_amd._lib import DN_ERR_STATE
    print("case split lagrangian", nbf, naux, nbf5, flush=True)
    s = synthetic.SyntheticSystem(nbf, naux, nbf5, 5 + nbf, ndoc=5)
    p = s.p
    b = synthetic.make_b(nbf, naux, s.seed)
    dev = api.System(p, precision=DN_SPLIT_F16).upload_b(b)
    with pytest.raises(api.DnError) as e:
        dev.download_b()
    assert e.value.code == DN_ERR_STATE
    cj, ck = oracle.pnof_selector(s.n, p)
    el, Hm, E = api.compute_Lagrange2(s.C, s.n, s.H, dev, cj, ck, p.nalpha, p.nbeta,
                                      return_energy=True)
    elr, Hmr = oracle.compute_lagrange2(s.C, s.n, s.H, b, cj, ck, p.nalpha, p.nbeta)
    Er = oracle.computeE_elec_hmat(Hmr, s.n, elr, nbf5)
    assert np.abs(Hm - Hmr).max() <= 1e-10 * max(1.0, np.abs(Hmr).max())   # FP64
    scale = np.abs(elr).max()
    e_lag = np.abs(el - elr).max() / scale
    g, gr = 4 * (el - el.T), 4 * (elr - elr.T)
    e_grad = np.abs(g - gr).max() / scale
    print("split lagrangian rel err %.2e, gradient %.2e of max|elag|, |dE| %.2e"
          % (e_lag, e_grad, abs(E - Er)), flush=True)
    assert e_lag <= 5e-6 and e_grad <= 4e-5
    assert abs(E - Er) <= 1e-6
    # the J/K build afterwards (its own scales) is unaffected
    J, K, Hc = api.computeJKH_MO(s.C, s.H, dev, p)
    Jr, Kr, _ = oracle.jkh_mo(s.C, s.H, b, nbf5)
    assert np.abs(J - Jr).max() <= 5e-6 * np.abs(Jr).max()
    assert np.abs(K - Kr).max() <= 5e-6 * np.abs(Kr).max()
    dev.close()


@pytest.mark.parametrize("nbf,naux,nbf5", [(130, 300, 45), (37, 70, 10)])
def test_split_lagrange2_jkh_against_oracle(oracle, nbf, naux, nbf5):
    """dn_lagrange2_jkh on a split system: the Lagrangian and J/K/H_core of the same
    orbitals from the one FP32 b_pq (utils.jl:1-59, integrals.jl:144-172), to the split
    bounds; H_core and Hmat stay FP64."""
    print("case split lag+jkh", nbf, naux, nbf5, flush=True)
    s = synthetic.SyntheticSystem(nbf, naux, nbf5, 9 + nbf, ndoc=5)
    p = s.p
    b = synthetic.make_b(nbf, naux, s.seed)
    dev = api.System(p, precision=DN_SPLIT_F16).upload_b(b)
    cj, ck = oracle.pnof_selector(s.n, p)
    el, Hm, E, J, K, Hc = api.compute_Lagrange2_JKH(s.C, s.n, s.H, dev, cj, ck, p.nalpha,
                                                    p.nbeta)
    elr, Hmr = oracle.compute_lagrange2(s.C, s.n, s.H, b, cj, ck, p.nalpha, p.nbeta)
    Jr, Kr, Hcr = oracle.jkh_mo(s.C, s.H, b, nbf5)
    assert np.abs(el - elr).max() <= 5e-6 * np.abs(elr).max()
    assert np.abs(Hm - Hmr).max() <= 1e-10 * max(1.0, np.abs(Hmr).max())
    assert abs(E - oracle.computeE_elec_hmat(Hmr, s.n, elr, nbf5)) <= 1e-6
    assert np.abs(J - Jr).max() <= 5e-6 * np.abs(Jr).max()
    assert np.abs(K - Kr).max() <= 5e-6 * np.abs(Kr).max()
    assert np.abs(Hc - Hcr).max() <= 1e-10 * max(1.0, np.abs(Hcr).max())
    dev.close()
