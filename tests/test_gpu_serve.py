"""Served solves: continuous batching on one GPU returns, for every request, exactly the
result the same request gets when it runs alone; admission honours the HBM budget."""
import numpy as np
import pytest

from donof_amd import api, serve, synthetic

pytestmark = pytest.mark.gpu


def requests(n=6, rate=200.0):
    return serve.synthetic_requests(n, rate, seed=4, nbf_min=40, nbf_max=90, max_iter=15,
                                    gtol=1e-4)


def test_alone_equals_in_the_mix():
    reqs = requests()
    ctx = api.Context(0)
    alone = {r.rid: serve.solve_one(r, ctx) for r in reqs}
    ctx.close()
    srv = serve.Server(max_live=4)
    res = srv.run(reqs)
    srv.close()
    assert len(res) == len(reqs)
    for r in res:
        assert r.error == ""
        E, nit, ok = alone[r.rid]
        assert (r.E, r.nit, r.success) == (E, nit, ok)   # bitwise the same
    assert max(r.admitted for r in res) > 0 or len(res) <= 4
    st = serve.latency_stats(res, max(r.done for r in res))
    assert st["solves"] == len(reqs) and st["p99_s"] >= st["p50_s"] > 0


def test_admission_by_hbm_budget():
    reqs = requests(5, rate=0.0)             # all arrive at t = 0
    fp = {r.rid: api.footprint(synthetic.synthetic_param(r.nbf, r.naux, r.nbf5)) for r in reqs}
    budget = int(1.2 * max(fp.values()))
    srv = serve.Server(hbm_budget=budget, max_live=8)
    res = srv.run(reqs)
    srv.close()
    assert all(r.error == "" for r in res)
    # at every admission, the footprints of the solves then resident fit in the budget
    for r in res:
        live = sum(fp[o.rid] for o in res if o.admitted <= r.admitted < o.done)
        assert live <= budget


def test_request_larger_than_budget_is_rejected():
    reqs = requests(2, rate=0.0)
    srv = serve.Server(hbm_budget=1 << 20, max_live=2)
    res = srv.run(reqs)
    srv.close()
    assert all("budget" in r.error for r in res)
