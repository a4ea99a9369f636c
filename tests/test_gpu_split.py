"""Split-precision J/K-MO build (FP16 hi+lo operands, FP32 accumulation, FP64 J/K
reduction) against the FP64 oracle.  Stated error bounds (relative to max|J|, max|K|):
5e-6 for J and K; occupation energy within 1e-6 Ha of FP64 on the same orbitals."""
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
    # the Lagrangian of a split system runs in FP64
    from oracle import donof_oracle as orc
    cj, ck = orc.pnof_selector(s.n, p)
    el, _ = api.compute_Lagrange2(s.C, s.n, s.H, d16, cj, ck, p.nalpha, p.nbeta)
    assert np.all(np.isfinite(el))
    d16.close()
