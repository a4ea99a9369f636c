"""Grouped J/K/H_core builds (dn_group_*): every system of a group of uneven sizes gets
its own computeJKH_MO (integrals.jl:131-172) to the FP64 bar, on every graph replay,
with new orbitals set between replays."""
import numpy as np
import pytest

from donof_amd import api, synthetic
from donof_amd._lib import DnError

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _group(oracle_needed_shapes, seed=7):
    systems, inputs = [], []
    for i, (nbf, naux, nbf5) in enumerate(oracle_needed_shapes):
        p = synthetic.synthetic_param(nbf, naux, nbf5)
        b = synthetic.make_b(nbf, naux, seed + i)
        s = api.System(p).upload_b(b)
        C = synthetic.make_C(nbf, 100 + i)
        H = synthetic.make_H(nbf, 200 + i)
        s.set_orbitals(C)
        s.set_hcore(H)
        systems.append(s)
        inputs.append((b, C, H, nbf5))
    return systems, inputs


SHAPES = [(20, 80, 7), (48, 192, 16), (33, 132, 11), (70, 280, 24), (9, 36, 3), (130, 520, 44)]


def test_group_matches_oracle_on_every_replay(oracle):
    systems, inputs = _group(SHAPES)
    g = api.Group(systems)
    cfg = g.tune()
    print("configs", cfg, flush=True)
    assert set(cfg) == {"K1", "K2", "J", "HC"}
    for rep in range(3):
        if rep == 2:   # new orbitals: the replayed graph must read them
            for i, s in enumerate(systems):
                C = synthetic.make_C(s.nbf, 300 + i)
                s.set_orbitals(C)
                inputs[i] = (inputs[i][0], C, inputs[i][2], inputs[i][3])
        g.jkh()
        g.sync()
        for s, (b, C, H, nbf5) in zip(systems, inputs):
            J, K, Hc = s.get_jkh()
            Jr, Kr, Hr = oracle.jkh_mo(C, H, b, nbf5)
            assert np.abs(J - Jr).max() <= TOL, (rep, s.nbf)
            assert np.abs(K - Kr).max() <= TOL, (rep, s.nbf)
            assert np.abs(Hc - Hr).max() <= TOL, (rep, s.nbf)
        # eager single-system work between replays
        systems[0].jkh_mo_resident()
        systems[0].sync()
    g.close()
    for s in systems:
        s.close()


def test_group_untuned_replay_and_timing(oracle):
    systems, inputs = _group(SHAPES[:3], seed=30)
    g = api.Group(systems)
    g.set_timing(True)
    g.jkh()
    g.jkh()
    ms, n = g.get_timing()
    assert n == 2 and ms > 0
    for s, (b, C, H, nbf5) in zip(systems, inputs):
        J, K, Hc = s.get_jkh()
        Jr, Kr, Hr = oracle.jkh_mo(C, H, b, nbf5)
        assert np.abs(J - Jr).max() <= TOL and np.abs(K - Kr).max() <= TOL
    g.close()
    for s in systems:
        s.close()


SPLIT_SHAPES = [(64, 256, 22), (37, 148, 13), (130, 520, 44), (20, 80, 7), (96, 384, 32),
                (200, 800, 67)]


def test_split_group_matches_oracle_on_every_replay(oracle):
    """Split-precision grouped builds (grouped FP16 hi+lo transforms, FP64 J and H_core):
    every system within the split bound of the FP64 oracle (5e-6 of max|J|, max|K|),
    tuned, on every replay, with new orbitals between replays."""
    systems, inputs = [], []
    for i, (nbf, naux, nbf5) in enumerate(SPLIT_SHAPES):
        p = synthetic.synthetic_param(nbf, naux, nbf5)
        b = synthetic.make_b(nbf, naux, 50 + i)
        s = api.System(p, precision="split").upload_b(b)
        assert s.b_symmetric
        C = synthetic.make_C(nbf, 150 + i)
        H = synthetic.make_H(nbf, 250 + i)
        s.set_orbitals(C)
        s.set_hcore(H)
        systems.append(s)
        inputs.append((b, C, H, nbf5))
    g = api.Group(systems)
    cfg = g.tune()
    print("split configs", cfg, g.plan(), flush=True)
    assert cfg["K1"].startswith("s") and cfg["K2"].startswith("s")
    for rep in range(3):
        if rep == 2:
            for i, s in enumerate(systems):
                C = synthetic.make_C(s.nbf, 350 + i)
                s.set_orbitals(C)
                inputs[i] = (inputs[i][0], C, inputs[i][2], inputs[i][3])
        g.jkh()
        g.sync()
        for s, (b, C, H, nbf5) in zip(systems, inputs):
            J, K, Hc = s.get_jkh()
            Jr, Kr, Hr = oracle.jkh_mo(C, H, b, nbf5)
            ej = np.abs(J - Jr).max() / np.abs(Jr).max()
            ek = np.abs(K - Kr).max() / np.abs(Kr).max()
            assert ej < 5e-6 and ek < 5e-6, (rep, s.nbf, ej, ek)
            assert np.abs(Hc - Hr).max() <= TOL, (rep, s.nbf)
    g.close()
    for s in systems:
        s.close()


def test_split_group_needs_symmetric_build():
    p = synthetic.synthetic_param(24, 96, 8)
    s = api.System(p, precision="split").upload_b(synthetic.make_b(24, 96, 3))
    s.set_symmetry(False)
    with pytest.raises(DnError, match="symmetric-b"):
        api.Group([s])
    s.close()


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_chained_group_matches_per_system_builds(precision):
    """Groups with enough work run as chains of leaf groups (parallel graph branches):
    every system's J, K, H_core equal its own single-system build (to 1e-12 relative in
    FP64, 1e-6 in split precision, where the two paths use different tile shapes)."""
    shapes = [(300 + 6 * i, 4 * (300 + 6 * i), 100 + 2 * i) for i in range(6)]
    systems, ref = [], []
    for i, (nbf, naux, nbf5) in enumerate(shapes):
        s = api.System(synthetic.synthetic_param(nbf, naux, nbf5), precision=precision)
        s.generate_b(900 + i)
        s.set_orbitals(synthetic.make_C(nbf, 910 + i))
        s.set_hcore(synthetic.make_H(nbf, 920 + i))
        s.jkh_mo_resident()
        s.sync()
        ref.append(s.get_jkh())
        systems.append(s)
    g = api.Group(systems).place(10 ** 6)      # every system grouped (no placement search)
    g.tune()
    plan = g.plan()
    print(plan[:300], flush=True)
    assert plan.startswith("chain 0")
    tol = 1e-12 if precision == "fp64" else 1e-6
    for rep in range(2):
        g.jkh()
        g.sync()
        for s, (J, K, Hc) in zip(systems, ref):
            Jg, Kg, Hg = s.get_jkh()
            assert np.abs(Jg - J).max() <= tol * np.abs(J).max(), (rep, s.nbf)
            assert np.abs(Kg - K).max() <= tol * np.abs(K).max(), (rep, s.nbf)
            assert np.abs(Hg - Hc).max() <= 1e-12 * max(1.0, np.abs(Hc).max())
    g.close()
    for s in systems:
        s.close()


def test_group_requires_ready_systems():
    p = synthetic.synthetic_param(20, 80, 7)
    s = api.System(p)                       # b never uploaded
    with pytest.raises(DnError, match="not been uploaded"):
        api.Group([s])
    s.close()


@pytest.mark.parametrize("limit", [15, -1])
def test_group_placement_matches_oracle(oracle, limit):
    # the systems with nbf5 <= limit grouped, the others built by the single-system build
    # on their own streams beside them (dn_group_place); -1: none grouped
    print("case placement", limit, flush=True)
    systems, inputs = _group(SHAPES, seed=50)
    g = api.Group(systems).place(limit)
    cfg = g.tune()
    plan = g.plan()
    assert plan.startswith("placement: %d grouped" % sum(sh[2] <= limit for sh in SHAPES))
    if limit < 0:
        assert all(v is None for v in cfg.values())
    g.set_timing(True)
    for rep in range(3):
        if rep == 1:
            for i, s in enumerate(systems):
                C = synthetic.make_C(s.nbf, 400 + i)
                s.set_orbitals(C)
                inputs[i] = (inputs[i][0], C, inputs[i][2], inputs[i][3])
        g.jkh()
        g.sync()
        for s, (b, C, H, nbf5) in zip(systems, inputs):
            J, K, Hc = s.get_jkh()
            Jr, Kr, Hr = oracle.jkh_mo(C, H, b, nbf5)
            assert np.abs(J - Jr).max() <= TOL, (rep, s.nbf)
            assert np.abs(K - Kr).max() <= TOL, (rep, s.nbf)
            assert np.abs(Hc - Hr).max() <= TOL, (rep, s.nbf)
    ms, n = g.get_timing()
    assert n == 3 and ms > 0
    # an independent system's result is its single-system build's, bit for bit
    s = systems[-1]
    J, K, _ = s.get_jkh()
    s.jkh_mo_resident()
    s.sync()
    J2, K2, _ = s.get_jkh()
    assert np.array_equal(J, J2) and np.array_equal(K, K2)
    g.close()
    for s in systems:
        s.close()
