"""GPU parity of the J/K-MO build and the Lagrangian against the CPU oracle, through the
C ABI (libdonofhip.so).  FP64 bar: max|err| <= 1e-10 for J, K, H_core and elag, energy
within 1e-9 Ha."""
import numpy as np
import pytest

from donof_amd import api, synthetic
from donof_amd._lib import DnError, DN_ERR_BUDGET, DN_ERR_STATE, DN_ERR_NONFINITE

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _system(nbf, naux, nbf5, seed, ndoc=None):
    s = synthetic.SyntheticSystem(nbf, naux, nbf5, seed, ndoc=ndoc)
    b = synthetic.make_b(nbf, naux, seed)
    return s, b


# ragged shapes exercise every tile edge of the MFMA kernels; (150, 1200, 130) has more
# (tile pair, Q range) workgroups than the K/X pass allows at 8 slabs per range, so its
# ranges are doubled
SHAPES = [(8, 4, 3, 1), (24, 96, 20, 5), (37, 19, 9, 3), (65, 130, 33, 3), (130, 70, 66, 3),
          (200, 300, 60, 5), (150, 1200, 130, 5)]


@pytest.mark.parametrize("nbf,naux,nbf5,ndoc", SHAPES)
def test_jkh_mo_parity(oracle, nbf, naux, nbf5, ndoc):
    print("case jkh", nbf, naux, nbf5, flush=True)
    s, b = _system(nbf, naux, nbf5, 100 + nbf, ndoc)
    dev = api.System(s.p)
    dev.upload_b(b)
    J, K, Hc = api.computeJKH_MO(s.C, s.H, dev, s.p)
    Jr, Kr, Hr = oracle.jkh_mo(s.C, s.H, b, nbf5)
    assert np.abs(J - Jr).max() <= TOL
    assert np.abs(K - Kr).max() <= TOL
    assert np.abs(Hc - Hr).max() <= TOL
    dev.close()


@pytest.mark.parametrize("nbf,naux,nbf5,ndoc", SHAPES)
def test_jkh_mo_full_contraction_parity(oracle, nbf, naux, nbf5, ndoc):
    """The same build with the (m, n) symmetry of b switched off (integrals.jl:159-160
    as written)."""
    print("case jkh full", nbf, naux, nbf5, flush=True)
    s, b = _system(nbf, naux, nbf5, 100 + nbf, ndoc)
    dev = api.System(s.p)
    dev.upload_b(b)
    assert dev.b_symmetric
    dev.set_symmetry(False)
    J, K, Hc = api.computeJKH_MO(s.C, s.H, dev, s.p)
    Jr, Kr, Hr = oracle.jkh_mo(s.C, s.H, b, nbf5)
    assert np.abs(J - Jr).max() <= TOL
    assert np.abs(K - Kr).max() <= TOL
    dev.set_symmetry(True)
    J2, K2, _ = api.computeJKH_MO(s.C, s.H, dev, s.p)
    assert np.abs(J2 - J).max() <= TOL and np.abs(K2 - K).max() <= TOL
    dev.close()


def test_nonsymmetric_b_uses_full_contraction(oracle):
    nbf, naux, nbf5 = 37, 45, 9
    s, b = _system(nbf, naux, nbf5, 5, 3)
    b = b + 1e-3 * np.random.default_rng(3).standard_normal(b.shape)   # b[m,n] != b[n,m]
    dev = api.System(s.p)
    dev.upload_b(b)
    assert not dev.b_symmetric
    J, K, Hc = api.computeJKH_MO(s.C, s.H, dev, s.p)
    Jr, Kr, Hr = oracle.jkh_mo(s.C, s.H, b, nbf5)
    assert np.abs(J - Jr).max() <= TOL
    assert np.abs(K - Kr).max() <= TOL
    dev.close()


@pytest.mark.parametrize("nbf,naux,nbf5,ndoc", SHAPES)
def test_lagrange2_parity(oracle, nbf, naux, nbf5, ndoc):
    print("case lag", nbf, naux, nbf5, flush=True)
    s, b = _system(nbf, naux, nbf5, 200 + nbf, ndoc)
    p = s.p
    dev = api.System(p)
    dev.upload_b(b)
    cj, ck = oracle.pnof_selector(s.n, p)
    el, Hm, E = api.compute_Lagrange2(s.C, s.n, s.H, dev, cj, ck, p.nalpha, p.nbeta,
                                      return_energy=True)
    elr, Hmr = oracle.compute_lagrange2(s.C, s.n, s.H, b, cj, ck, p.nalpha, p.nbeta)
    Er = oracle.computeE_elec_hmat(Hmr, s.n, elr, nbf5)
    assert np.abs(el - elr).max() <= TOL
    assert np.abs(Hm - Hmr).max() <= TOL
    assert abs(E - Er) <= 1e-9
    dev.close()


def test_lagrange2_general_coefficients(oracle):
    """Arbitrary (non-symmetric, nonzero-diagonal) cj12/ck12 and nbeta = 0."""
    s, b = _system(40, 50, 15, 17, 3)
    p = s.p
    dev = api.System(p)
    dev.upload_b(b)
    rng = np.random.default_rng(1)
    cj = rng.standard_normal((15, 15))
    ck = rng.standard_normal((15, 15))
    for nalpha, nbeta in ((p.nalpha, p.nbeta), (0, 0), (4, 2)):
        el, Hm = api.compute_Lagrange2(s.C, s.n, s.H, dev, cj, ck, nalpha, nbeta)
        elr, Hmr = oracle.compute_lagrange2(s.C, s.n, s.H, b, cj, ck, nalpha, nbeta)
        assert np.abs(el - elr).max() <= TOL
    dev.close()


@pytest.mark.parametrize("nbf", [37, 38])
def test_b_upload_download_round_trip(nbf):
    """b goes through the resident layout (a padded leading dimension for odd nbf) and
    comes back bit for bit; a symmetric b is recognised as such."""
    naux = 11
    p = synthetic.synthetic_param(nbf, naux, 12, ndoc=3)
    rng = np.random.default_rng(nbf)
    b = np.asfortranarray(rng.standard_normal((nbf, nbf, naux)))
    dev = api.System(p)
    dev.upload_b(b)
    assert not dev.b_symmetric
    assert np.array_equal(dev.download_b(), b)
    bs = np.asfortranarray(b + b.transpose(1, 0, 2))
    dev.upload_b(bs)
    assert dev.b_symmetric
    assert np.array_equal(dev.download_b(), bs)
    dev.close()


def test_device_generator_matches_host():
    nbf, naux, seed = 33, 21, 42
    p = synthetic.synthetic_param(nbf, naux, 12, ndoc=3)
    dev = api.System(p)
    dev.generate_b(seed)
    bd = dev.download_b()
    bh = synthetic.make_b(nbf, naux, seed)
    assert np.abs(bd - bh).max() <= 1e-14 * max(1.0, np.abs(bh).max())
    dev.close()


def test_config2_full_size(oracle):
    """nbf=400, nbfaux=1600, nbf5=120: b generated in HBM, J/K/H_core and the Lagrangian
    compared with the oracle on the downloaded tensor."""
    print("case cfg2 full", flush=True)
    s = synthetic.SyntheticSystem(400, 1600, 120, 2024)
    p = s.p
    dev = api.System(p)
    dev.generate_b(2024)
    b = dev.download_b()
    J, K, Hc = api.computeJKH_MO(s.C, s.H, dev, p)
    Jr, Kr, Hr = oracle.jkh_mo(s.C, s.H, b, 120)
    assert np.abs(J - Jr).max() <= TOL and np.abs(K - Kr).max() <= TOL
    assert np.abs(Hc - Hr).max() <= TOL
    cj, ck = oracle.pnof_selector(s.n, p)
    el, Hm, E = api.compute_Lagrange2(s.C, s.n, s.H, dev, cj, ck, p.nalpha, p.nbeta,
                                      return_energy=True)
    elr, Hmr = oracle.compute_lagrange2(s.C, s.n, s.H, b, cj, ck, p.nalpha, p.nbeta)
    assert np.abs(el - elr).max() <= TOL
    assert abs(E - oracle.computeE_elec_hmat(Hmr, s.n, elr, 120)) <= 1e-9
    # the energy identity holds on the device results too
    assert abs(E - oracle.calce(s.n, cj, ck, J, K, Hc, p)) <= 1e-9
    dev.close()


def test_resident_repeat_is_bitwise_stable():
    s = synthetic.SyntheticSystem(96, 200, 30, 5, ndoc=5)
    dev = api.System(s.p)
    dev.generate_b(5)
    dev.set_orbitals(s.C)
    dev.set_hcore(s.H)
    dev.jkh_mo_resident()
    J1, K1, H1 = dev.get_jkh()
    for _ in range(3):
        dev.jkh_mo_resident()
    J2, K2, H2 = dev.get_jkh()
    assert np.array_equal(J1, J2) and np.array_equal(K1, K2) and np.array_equal(H1, H2)
    dev.close()


def test_errors_are_reported():
    ctx = api.Context(0, hbm_budget=1 << 20)
    p = synthetic.synthetic_param(64, 256, 20, ndoc=5)
    with pytest.raises(DnError) as e:
        api.System(p, ctx=ctx)
    assert e.value.code == DN_ERR_BUDGET
    ctx.close()
    s = synthetic.SyntheticSystem(32, 40, 10, 1, ndoc=5)
    dev = api.System(s.p)
    with pytest.raises(DnError) as e:
        api.computeJKH_MO(s.C, s.H, dev, s.p)
    assert e.value.code == DN_ERR_STATE
    dev.generate_b(1)
    C = s.C.copy()
    C[3, 2] = np.nan
    with pytest.raises(DnError) as e:
        api.computeJKH_MO(C, s.H, dev, s.p)
    assert e.value.code == DN_ERR_NONFINITE
    dev.close()


@pytest.mark.parametrize("nbf,naux,nbf5,ndoc", [SHAPES[1], SHAPES[2], SHAPES[5]])
def test_lagrange2_jkh_one_transform(oracle, nbf, naux, nbf5, ndoc):
    """dn_lagrange2_jkh: the Lagrangian and J/K/H_core of the same orbitals from one
    transform of b, each against the oracle (utils.jl:1-59, integrals.jl:144-172)."""
    print("case lag+jkh", nbf, naux, nbf5, flush=True)
    s, b = _system(nbf, naux, nbf5, 300 + nbf, ndoc)
    p = s.p
    dev = api.System(p)
    dev.upload_b(b)
    cj12, ck12 = oracle.pnof_selector(s.n, p)
    elag, Hmat, E, J, K, Hc = api.compute_Lagrange2_JKH(s.C, s.n, s.H, dev, cj12, ck12,
                                                        p.nalpha, p.nbeta)
    er, Hr = oracle.compute_lagrange2(s.C, s.n, s.H, b, cj12, ck12, p.nalpha, p.nbeta)
    Jr, Kr, Hcr = oracle.jkh_mo(s.C, s.H, b, nbf5)
    assert np.abs(elag - er).max() <= TOL
    assert np.abs(Hmat - Hr).max() <= TOL
    assert abs(E - oracle.computeE_elec_hmat(Hr, s.n, er, nbf5)) <= 1e-9
    assert np.abs(J - Jr).max() <= TOL
    assert np.abs(K - Kr).max() <= TOL
    assert np.abs(Hc - Hcr).max() <= TOL
    dev.close()
