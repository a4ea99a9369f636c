"""RI setup on the device (dn_sys_ri_metric; integrals.jl:68-91) against the oracle.
b itself follows the Cholesky route (an orthogonal rotation of the reference's b in the
auxiliary index), so parity is checked on what consumers see: sum_Q b b and J/K, to
the FP64 bar relative to their scale, and b against the oracle's Cholesky route."""
import numpy as np
import pytest

from donof_amd import api, synthetic
from donof_amd._lib import DnError

pytestmark = pytest.mark.gpu

# naux below, at and across the 64-wide Cholesky blocks; one with a ragged last block; odd
# nbf (b resident with a padded leading dimension)
SHAPES = [(6, 20, 4), (12, 64, 8), (20, 150, 7), (16, 300, 9), (40, 513, 14), (15, 130, 6)]


@pytest.mark.parametrize("nbf,naux,nbf5", SHAPES)
def test_ri_metric_parity(oracle, nbf, naux, nbf5):
    print("case ri", nbf, naux, nbf5, flush=True)
    G, I3c, b = synthetic.make_metric(nbf, naux, 20 + naux, cond=1e6)
    p = synthetic.synthetic_param(nbf, naux, nbf5)
    dev = api.Iaux(G, I3c, p)
    bd = dev.download_b()
    bc = oracle.iaux_cholesky(G, I3c)
    assert np.abs(bd - bc).max() <= 1e-10 * np.abs(bc).max()
    X = bd.reshape(nbf * nbf, naux, order="F")
    R = b.reshape(nbf * nbf, naux, order="F")
    assert np.abs(X @ X.T - R @ R.T).max() <= 1e-10 * np.abs(R @ R.T).max()
    C = synthetic.make_C(nbf, 3)
    H = synthetic.make_H(nbf, 4)
    J, K, Hc = api.JKH_MO(C, H, dev, nbf, nbf5, naux)
    Jr, Kr, Hr = oracle.jkh_mo(C, H, oracle.iaux(G, I3c), nbf5)
    assert np.abs(J - Jr).max() <= 1e-10 * max(1.0, np.abs(Jr).max())
    assert np.abs(K - Kr).max() <= 1e-10 * max(1.0, np.abs(Kr).max())
    assert np.abs(Hc - Hr).max() <= 1e-10
    dev.close()


def test_ri_metric_not_positive_definite():
    nbf, naux = 6, 40
    G, I3c, _ = synthetic.make_metric(nbf, naux, 1)
    w, V = np.linalg.eigh(G)
    G = G - 2.0 * w[-1] * np.outer(V[:, -1], V[:, -1])     # flip the largest eigenvalue
    dev = api.System(synthetic.synthetic_param(nbf, naux, 3))
    with pytest.raises(DnError, match="positive definite"):
        dev.ri_metric(G, I3c)
    dev.close()


def test_ri_metric_split_precision(oracle):
    nbf, naux, nbf5 = 30, 200, 10
    G, I3c, b = synthetic.make_metric(nbf, naux, 7, cond=1e4)
    p = synthetic.synthetic_param(nbf, naux, nbf5)
    dev = api.Iaux(G, I3c, p, precision="split")
    C = synthetic.make_C(nbf, 3)
    H = synthetic.make_H(nbf, 4)
    J, K, _ = api.JKH_MO(C, H, dev, nbf, nbf5, naux)
    Jr, Kr, _ = oracle.jkh_mo(C, H, b, nbf5)
    assert np.abs(J - Jr).max() <= 5e-6 * np.abs(Jr).max()
    assert np.abs(K - Kr).max() <= 5e-6 * np.abs(Kr).max()
    dev.close()
