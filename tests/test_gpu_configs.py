"""The benchmark configurations at their full sizes on the GPU:

  800/3200/240  split precision (FP16 hi+lo products, FP32 accumulation) against the FP64
                device build on the same tensor and orbitals, both against the oracle on a
                block of J and K, and full-size properties of the FP64 build;
  64 systems    nbf ~ U[48, 400] (seeded), naux = 4 nbf, nbf5 = ceil(nbf/3): the grouped
                build against each system's own build, FP64 and split, and the smallest
                and largest systems against the oracle;
  served        requests with nbf in [50, 600] and up to 200 ADAM iterations, every size
                class of the engines: each result alone equals the one in the mix.

The FP64 build is pinned to the oracle at 400/1600/120 and ragged shapes (test_gpu_jk.py);
here the oracle computes the entries of J and K for a few orbitals from the full tensor
(b downloaded from HBM): J_pq, K_pq for p, q in P need only C5[:, P]^T b C5[:, P]."""
import numpy as np
import pytest

from donof_amd import api, serve, synthetic

pytestmark = pytest.mark.gpu


def _oracle_block(b, C5, P):
    """J[P,P], K[P,P] from the full b (integrals.jl:159-166 restricted to orbitals P)."""
    nbf, _, naux = b.shape
    T = C5[:, P].T @ b.reshape(nbf, nbf * naux, order="F")          # [p, (n, Q)]
    T = T.reshape(len(P), nbf, naux, order="F")
    bpq = np.einsum("pnQ,nq->pqQ", T, C5[:, P])
    X = np.einsum("ppQ->pQ", bpq)
    return X @ X.T, np.einsum("pqQ,pqQ->pq", bpq, bpq)


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def test_split_800_3200_240_full_size(oracle):
    nbf, naux, nbf5, seed = 800, 3200, 240, 2024
    print("case split 800/3200/240", flush=True)
    sd = synthetic.SyntheticSystem(nbf, naux, nbf5, seed)
    p = sd.p
    out, lag = {}, {}
    cj, ck = oracle.pnof_selector(sd.n, p)
    for prec in ("fp64", "split"):
        dev = api.System(p, precision=prec)
        dev.generate_b(seed)
        dev.set_orbitals(sd.C)
        dev.set_hcore(sd.H)
        dev.jkh_mo_resident()
        out[prec] = dev.get_jkh()
        if prec == "fp64":
            b = dev.download_b()
        # the Lagrangian (general transform, utils.jl:1-59) at the same orbitals
        lag[prec] = api.compute_Lagrange2(sd.C, sd.n, sd.H, dev, cj, ck, p.nalpha, p.nbeta,
                                          return_energy=True)
        dev.close()
    E_lag = lag["fp64"][2]
    J64, K64, H64 = out["fp64"]
    J16, K16, H16 = out["split"]
    # full-size properties of the FP64 build
    assert np.abs(J64 - J64.T).max() <= 1e-12 * np.abs(J64).max()
    assert np.linalg.eigvalsh(J64).min() >= -1e-10 * np.abs(J64).max()
    assert K64.min() >= 0.0
    assert np.abs(np.diag(J64) - np.diag(K64)).max() <= 1e-12 * np.abs(J64).max()
    # calce of the J/K build equals computeE_elec of the Lagrangian (two code paths)
    cj, ck = oracle.pnof_selector(sd.n, p)
    E64 = oracle.calce(sd.n, cj, ck, J64, K64, H64, p)
    assert abs(E64 - E_lag) <= 1e-9 * max(1.0, abs(E64))
    # a block of J and K from the oracle on the full tensor
    P = np.array([0, 1, 117, 239])
    Jb, Kb = _oracle_block(b, sd.C[:, :nbf5], P)
    del b
    assert np.abs(J64[np.ix_(P, P)] - Jb).max() <= 1e-10
    assert np.abs(K64[np.ix_(P, P)] - Kb).max() <= 1e-10
    # split precision against FP64 (and the oracle block): J/K to 5e-6 of their scale,
    # the GNOF energy to 1e-6 Ha
    assert _rel(J16, J64) <= 5e-6 and _rel(K16, K64) <= 5e-6
    assert _rel(J16[np.ix_(P, P)], Jb) <= 5e-6 and _rel(K16[np.ix_(P, P)], Kb) <= 5e-6
    # H_core stays FP64 (the split build runs H C5 unsplit on its second stream, so the
    # summation order, not the precision, differs)
    assert np.abs(H16 - H64).max() <= 1e-12 * max(1.0, np.abs(H64).max())
    E16 = oracle.calce(sd.n, cj, ck, J16, K16, H16, p)
    print("split 800: |dJ|/|J| %.2e |dK|/|K| %.2e |dE| %.2e Ha"
          % (_rel(J16, J64), _rel(K16, K64), abs(E16 - E64)), flush=True)
    assert abs(E16 - E64) <= 1e-6
    # the split Lagrangian against the FP64 one: elag to 5e-6 of its scale, the orbital
    # gradient 4 (elag - elag^T) stated against the same scale, E_elec to 1e-6 Ha
    el64, Hm64, El64 = lag["fp64"]
    el16, Hm16, El16 = lag["split"]
    scale = np.abs(el64).max()
    e_grad = np.abs(4 * (el16 - el16.T) - 4 * (el64 - el64.T)).max() / scale
    print("split 800 lagrangian: |d elag|/|elag| %.2e, |d grad|/|elag| %.2e, |dE| %.2e Ha"
          % (_rel(el16, el64), e_grad, abs(El16 - El64)), flush=True)
    assert _rel(el16, el64) <= 5e-6 and e_grad <= 4e-5
    assert abs(El16 - El64) <= 1e-6
    assert np.abs(Hm16 - Hm64).max() <= 1e-12 * max(1.0, np.abs(Hm64).max())


def test_grouped_64_systems_full_size(oracle):
    seed = 2024
    shapes = synthetic.grouped_shapes(64, seed)
    assert len(shapes) == 64 and min(s[0] for s in shapes) >= 48
    ref64 = None
    for prec in ("fp64", "split"):
        print("case grouped 64", prec, flush=True)
        systems = []
        for i, (nbf, naux, nbf5) in enumerate(shapes):
            s = api.System(synthetic.synthetic_param(nbf, naux, nbf5), precision=prec)
            s.generate_b(seed + i)
            s.set_orbitals(synthetic.make_C(nbf, seed + 1000 + i))
            s.set_hcore(synthetic.make_H(nbf, seed + 2000 + i))
            systems.append(s)
        single = []
        for s in systems:
            s.jkh_mo_resident()
            single.append(s.get_jkh())
        g = api.Group(systems)
        g.tune()
        # FP64: the same arithmetic as the single-system build; split: the grouped build
        # keeps the FP32 intermediate T with its exact per-Q scale, the single-system build
        # splits T into FP16 planes under a Cauchy-Schwarz scale, so they agree to the
        # split rounding
        tol = 1e-12 if prec == "fp64" else 1e-6
        for rep in range(2):
            g.jkh()
            g.sync()
            got = [s.get_jkh() for s in systems]
            for i, ((J, K, Hc), (Js, Ks, Hs)) in enumerate(zip(got, single)):
                assert _rel(J, Js) <= tol and _rel(K, Ks) <= tol, (prec, rep, i)
                assert np.abs(Hc - Hs).max() <= 1e-12 * max(1.0, np.abs(Hs).max())
        g.close()
        if prec == "fp64":
            ref64 = got
            order = np.argsort([sh[0] for sh in shapes])
            for i in (int(order[0]), int(order[-1])):
                nbf, naux, nbf5 = shapes[i]
                b = systems[i].download_b()
                Jr, Kr, Hr = oracle.jkh_mo(synthetic.make_C(nbf, seed + 1000 + i),
                                           synthetic.make_H(nbf, seed + 2000 + i), b, nbf5)
                J, K, Hc = got[i]
                assert np.abs(J - Jr).max() <= 1e-10 and np.abs(K - Kr).max() <= 1e-10
                assert np.abs(Hc - Hr).max() <= 1e-10
                del b
        else:
            for i, ((J, K, _), (J64, K64, _)) in enumerate(zip(got, ref64)):
                assert _rel(J, J64) <= 5e-6 and _rel(K, K64) <= 5e-6, i
        for s in systems:
            s.close()


@pytest.mark.parametrize("mode", ["grouped", "streams"])
def test_served_full_range_alone_equals_mix(mode):
    reqs = serve.synthetic_requests(6, 8.0, seed=5, nbf_min=50, nbf_max=600, max_iter=200,
                                    gtol=1e-4)
    sizes = sorted(r.nbf for r in reqs)
    assert sizes[0] < 160 and any(160 <= x < 320 for x in sizes) and sizes[-1] >= 320
    print("case served", mode, sizes, flush=True)
    alone = {}
    for r in reqs:
        srv = serve.Server(mode=mode, max_live=1)
        (a,) = srv.run([r])
        srv.close()
        assert a.error == ""
        alone[r.rid] = (a.E, a.nit, a.success)
    srv = serve.Server(mode=mode, max_live=6)
    res = srv.run(reqs)
    srv.close()
    assert len(res) == len(reqs)
    for r in res:
        assert r.error == "", (r.rid, r.error)
        assert (r.E, r.nit, r.success) == alone[r.rid]        # bitwise the same
