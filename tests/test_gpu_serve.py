"""Served solves: continuous batching on one GPU returns, for every request, exactly the
result the same request gets when it runs alone; admission honours the HBM budget.
The lock-step grouped engine (dn_ogroup_step) must also reproduce the single-system
device loop (dn_orbopt_resident) bit for bit."""
import ctypes as ct

import numpy as np
import pytest

from donof_amd import api, serve, synthetic

pytestmark = pytest.mark.gpu


def requests(n=6, rate=200.0):
    return serve.synthetic_requests(n, rate, seed=4, nbf_min=40, nbf_max=90, max_iter=15,
                                    gtol=1e-4)


def _serve(reqs, mode, **kw):
    srv = serve.Server(mode=mode, **kw)
    res = srv.run(reqs)
    srv.close()
    return res


def test_grouped_engine_matches_single_system_loop():
    reqs = requests(4)
    ctx = api.Context(0)
    resident = {r.rid: serve.solve_one(r, ctx) for r in reqs}
    ctx.close()
    for r in reqs:
        (g,) = _serve([r], "grouped", max_live=1)
        assert g.error == ""
        assert (g.E, g.nit, g.success) == resident[r.rid]     # bitwise the same


def test_native_run_equals_single_steps():
    """dn_ogroup_run (the engines' inner loop) advances exactly as repeated dn_ogroup_step
    and honours its stop flag."""
    reqs = requests(3)
    ctx = api.Context(0)
    a = [serve.start_solve(r, ctx) for r in reqs]
    b = [serve.start_solve(r, ctx) for r in reqs]
    ga = api.OGroup([d for d, _ in a], ctx)
    gb = api.OGroup([d for d, _ in b], ctx)
    stop = ct.c_int(1)
    assert ga.run(5, stop) == (0, 0)          # stop set: no iteration
    stop.value = 0
    runs = 0
    while not all(d.orbopt_done() for d, _ in a):
        steps, _ = ga.run(4, stop)
        assert 1 <= steps <= 4
        runs += 1
    while not all(d.orbopt_done() for d, _ in b):
        gb.step()
    ga.close()
    gb.close()
    for (da, pa), (db, pb) in zip(a, b):
        assert da.orbopt_finish(pa) == db.orbopt_finish(pb)          # bitwise the same
        assert np.array_equal(da.get_orbitals(), db.get_orbitals())
        da.close()
        db.close()
    ctx.close()
    assert runs >= 1


@pytest.mark.parametrize("mode", ["grouped", "streams"])
def test_alone_equals_in_the_mix(mode):
    reqs = requests()
    alone = {}
    for r in reqs:
        (a,) = _serve([r], mode, max_live=1)
        alone[r.rid] = (a.E, a.nit, a.success)
    res = _serve(reqs, mode, max_live=4)
    assert len(res) == len(reqs)
    for r in res:
        assert r.error == "", (r.rid, r.error)
        assert (r.E, r.nit, r.success) == alone[r.rid]       # bitwise the same
    st = serve.latency_stats(res, max(r.done for r in res))
    assert st["solves"] == len(reqs) and st["p99_s"] >= st["p50_s"] > 0


@pytest.mark.parametrize("mode", ["grouped", "streams"])
def test_split_served_alone_equals_in_the_mix(mode):
    """precision "split": the single-system solves (here nbf >= 60; in mode "grouped" the
    class from 60 up is capped at one) hold FP16 planes and run split Lagrangians; each
    result is bitwise the same alone and in the mix, and close to the FP64 solve's."""
    reqs = requests()
    kw = dict(precision="split", split_min_nbf=60, size_classes=(60,), group_caps=(8, 1))
    alone = {}
    for r in reqs:
        (a,) = _serve([r], mode, max_live=1, **kw)
        assert a.error == "", a.error
        alone[r.rid] = (a.E, a.nit, a.success)
    res = _serve(reqs, mode, max_live=4, **kw)
    for r in res:
        assert r.error == "", (r.rid, r.error)
        assert (r.E, r.nit, r.success) == alone[r.rid]       # bitwise the same
    fp64 = {r.rid: r.E for r in _serve(reqs, mode, max_live=4)}
    n_split = 0
    for r in res:
        d = abs(r.E - fp64[r.rid])
        print("split served rid %d nbf %d E %.10f |dE| vs fp64 %.2e" % (r.rid, r.nbf, r.E, d))
        if r.nbf >= 60:
            n_split += 1
            assert d <= 1e-5 * max(1.0, abs(fp64[r.rid]))
        else:
            assert d == 0.0                                  # FP64 either way
    assert n_split >= 1


@pytest.mark.parametrize("mode", ["grouped", "streams"])
def test_admission_by_hbm_budget(mode):
    reqs = requests(5, rate=0.0)             # all arrive at t = 0
    fp = {r.rid: api.footprint(synthetic.synthetic_param(r.nbf, r.naux, r.nbf5)) for r in reqs}
    budget = int(1.2 * max(fp.values()))
    res = _serve(reqs, mode, hbm_budget=budget, max_live=8)
    assert all(r.error == "" for r in res), [(r.rid, r.error) for r in res if r.error]
    # at every admission, the footprints of the solves then resident fit in the budget
    for r in res:
        live = sum(fp[o.rid] for o in res if o.admitted <= r.admitted < o.done)
        assert live <= budget


def test_request_larger_than_budget_is_rejected():
    res = _serve(requests(2, rate=0.0), "grouped", hbm_budget=1 << 20, max_live=2)
    assert all("budget" in r.error for r in res)


def test_group_membership_misuse_is_refused():
    """A system belongs to at most one lock-step group, is listed once, and is not
    destroyed while a group (lock-step or grouped build) holds it."""
    from donof_amd._lib import DN_ERR_ARG, DN_ERR_STATE
    reqs = requests(2)
    ctx = api.Context(0)
    (a, pa), (b, pb) = [serve.start_solve(r, ctx) for r in reqs]
    g = api.OGroup([a, b], ctx)
    with pytest.raises(api.DnError) as e:
        api.OGroup([a], ctx)                      # already in a lock-step group
    assert e.value.code == DN_ERR_STATE
    with pytest.raises(api.DnError) as e:
        a.close()                                 # destroy while grouped
    assert e.value.code == DN_ERR_STATE and a.handle
    g.step()                                      # the group still works
    g.close()
    with pytest.raises(api.DnError) as e:
        api.OGroup([a, a], ctx)                   # listed twice
    assert e.value.code == DN_ERR_ARG
    g2 = api.OGroup([a], ctx)                     # membership released by close
    g2.close()
    bg = api.Group([b], ctx)
    with pytest.raises(api.DnError) as e:
        b.close()                                 # held by a grouped build
    assert e.value.code == DN_ERR_STATE
    bg.close()
    a.orbopt_finish(pa)
    b.orbopt_finish(pb)
    a.close()
    b.close()
    ctx.close()
