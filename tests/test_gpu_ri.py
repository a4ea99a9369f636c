"""RI setup on the device (dn_sys_ri_metric; integrals.jl:68-91) against the oracle.
For a positive-definite metric b follows the Cholesky route (an orthogonal rotation of the
reference's b in the auxiliary index), so parity is checked on what consumers see:
sum_Q b b and J/K, to the FP64 bar relative to their scale, and b against the oracle's
Cholesky route.  Otherwise the eigen route gives the reference's b itself."""
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


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_ri_metric_not_positive_definite_takes_the_eigen_route(oracle, precision):
    """A metric with a negative eigenvalue: the reference's own route (integrals.jl:74-88,
    eigenvalues below zero contribute zero), so b is the reference's b itself."""
    from donof_amd._lib import DN_ERR_STATE
    nbf, naux, nbf5 = 12, 70, 6
    G, I3c, _ = synthetic.make_metric(nbf, naux, 1)
    w, V = np.linalg.eigh(G)
    G = G - 2.0 * w[-1] * np.outer(V[:, -1], V[:, -1])     # flip the largest eigenvalue
    G = 0.5 * (G + G.T)
    p = synthetic.synthetic_param(nbf, naux, nbf5)
    dev = api.System(p, precision=precision).ri_metric(G, I3c)
    br = oracle.iaux(G, I3c)
    C = synthetic.make_C(nbf, 3)
    H = synthetic.make_H(nbf, 4)
    J, K, _ = api.JKH_MO(C, H, dev, nbf, nbf5, naux)
    Jr, Kr, _ = oracle.jkh_mo(C, H, br, nbf5)
    tol = 1e-10 if precision == "fp64" else 5e-6
    assert np.abs(J - Jr).max() <= tol * np.abs(Jr).max()
    assert np.abs(K - Kr).max() <= tol * np.abs(Kr).max()
    if precision == "fp64":
        bd = dev.download_b()
        assert np.abs(bd - br).max() <= 1e-10 * np.abs(br).max()
        # the clamped eigenvalue makes I3c unrecoverable from b: no nuclear gradients
        n = np.full(nbf5, 0.5)
        cj = ck = np.zeros((nbf5, nbf5))
        with pytest.raises(DnError) as e:
            api.Gradient(dev, n, C, cj, ck, np.zeros((nbf, nbf)))
        assert e.value.code == DN_ERR_STATE
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
