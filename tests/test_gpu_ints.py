"""Three-centre integrals on the device (dn_int3c2e_device, dn_sys_ri_from_basis) against
the host integral engine (pinned in tests/test_integrals.py) for every angular momentum
class the device path covers (orbital s-d, fitting s-g), and the RI tensor built from them
against the one dn_sys_ri_metric builds from host integrals."""
import numpy as np
import pytest

from donof_amd import _lib, api, basis, synthetic

pytestmark = pytest.mark.gpu

MOL = "O 0 0 0.121\nH 0 0.751 -0.485\nH 0 -0.751 -0.485"


def _bases():
    bs = basis.basis_set("cc-pvdz", basis.parse_molecule(MOL)[0], True)
    return bs, bs.fitting_set()


def test_device_three_centre_matches_host():
    print("case int3c cc-pvdz/autofit", flush=True)
    bs, aux = _bases()
    h, ha = bs.handle(), aux.handle()
    ref = h.int3c2e(ha)
    got = h.int3c2e_device(ha)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("pure", [1, 0])
def test_device_three_centre_every_class(pure):
    # orbital s, p, d and fitting s .. g shells on three centres, contracted and not
    print("case int3c classes", pure, flush=True)
    ctr = np.array([[0.1, -0.2, 0.3], [1.1, 0.4, -0.6], [-0.7, 0.9, 0.2]])
    orb = _lib.Basis([0, 1, 2, 2, 1, 0], [3, 2, 1, 2, 1, 2], [pure] * 6,
                     ctr[[0, 0, 0, 1, 2, 2]],
                     np.array([5.0, 1.2, 0.3, 2.0, 0.5, 0.9, 1.5, 0.4, 0.7, 3.0, 0.6]),
                     np.array([0.2, 0.5, 0.4, 0.6, 0.5, 1.0, 0.7, 0.4, 1.0, 0.3, 0.8]))
    fit = _lib.Basis([0, 1, 2, 3, 4, 4], [1, 1, 2, 1, 1, 1], [pure] * 6,
                     ctr[[0, 1, 2, 0, 1, 2]],
                     np.array([2.0, 1.1, 0.8, 3.0, 0.9, 1.3, 0.6]),
                     np.array([1.0, 1.0, 0.6, 0.5, 1.0, 1.0, 1.0]))
    ref = orb.int3c2e(fit)
    got = orb.int3c2e_device(fit)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_ri_from_basis_matches_host_route():
    print("case ri_from_basis", flush=True)
    bs, aux = _bases()
    h, ha = bs.handle(), aux.handle()
    p = synthetic.synthetic_param(h.nbf, ha.nbf, 20)
    G = ha.int2c2e()
    G = (G + G.T) / 2
    host = api.Iaux(G, h.int3c2e(ha), p)
    dev = api.System(p).ri_from_basis(h, ha)
    bh, bd = host.download_b(), dev.download_b()
    assert np.abs(bd - bh).max() <= 1e-11 * np.abs(bh).max()
    host.close()
    dev.close()
