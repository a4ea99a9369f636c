"""Host logic of the grouped serving engines (CPU): admission, size-class routing, join
and leave at step boundaries, completion of every request, and error isolation (a step
that fails reports its live solves and the engine goes on with later ones).  The device
calls are replaced by fakes with the same interface."""
import threading

import pytest

from donof_amd import serve
from donof_amd._lib import DnError


class FakeDev:
    def __init__(self, req):
        self.req = req
        self.steps = 0
        self.closed = False

    def orbopt_done(self):
        return self.steps >= self.req.max_iter

    def orbopt_finish(self, p):
        return -1.0 * self.req.nbf, self.steps, False

    def close(self):
        self.closed = True


class FakeGroup:
    fail_rids = set()
    lock = threading.Lock()
    sizes = []

    def __init__(self, systems, ctx):
        self.systems = list(systems)
        with FakeGroup.lock:
            FakeGroup.sizes.append(sorted(s.req.nbf for s in self.systems))

    def step(self):
        for s in self.systems:
            if s.req.rid in FakeGroup.fail_rids:
                raise DnError(6, "non-finite energy or orbital gradient")
            s.steps += 1
        return len(self.systems)

    def run(self, max_steps, stop=None):
        # dn_ogroup_run: until a system finishes, the stop flag is set or max_steps ran
        n = 0
        while n < max_steps and not (stop is not None and stop.value):
            before = sum(s.orbopt_done() for s in self.systems)
            self.step()
            n += 1
            if sum(s.orbopt_done() for s in self.systems) != before:
                break
        return n, len(self.systems)

    def close(self):
        pass


class FakeCtx:
    hbm_budget = 10 ** 12


@pytest.fixture
def fake_device(monkeypatch):
    monkeypatch.setattr(serve, "start_solve", lambda req, ctx: (FakeDev(req), object()))
    monkeypatch.setattr(serve, "solve_one", _fake_solo)
    monkeypatch.setattr(serve.api, "OGroup", FakeGroup)
    monkeypatch.setattr(serve.api, "footprint", lambda p, precision=0: 10 ** 9)
    FakeGroup.fail_rids = set()
    FakeGroup.sizes = []


SOLOS = []
SOLO_PRECISION = {}


def _fake_solo(req, ctx, precision=serve.DN_FP64):
    """A solve of a class with cap one (the single-system loop): max_iter iterations."""
    SOLOS.append(req.nbf)
    SOLO_PRECISION[req.rid] = precision
    if req.rid in FakeGroup.fail_rids:
        raise DnError(6, "non-finite energy or orbital gradient")
    return -1.0 * req.nbf, req.max_iter, False


def _server(max_live=4, classes=(160, 320), caps=(8, 2, 2)):
    srv = serve.Server.__new__(serve.Server)
    srv.ctx = FakeCtx()
    srv.max_live = max_live
    srv.mode = "grouped"
    srv.precision = "fp64"
    srv.split_min_nbf = 160
    srv.size_classes = tuple(classes)
    srv.group_caps = tuple(caps)
    srv.results = []
    srv.phases = {}
    srv._lock = threading.Lock()
    srv._live = 0
    srv._reserved = 0
    return srv


def _requests(n, iters=5):
    sizes = [60, 500, 200, 80, 350, 150, 600, 100, 250, 90]
    return [serve.SolveRequest(rid=i, nbf=sizes[i % len(sizes)], seed=i, arrival=0.001 * i,
                               max_iter=iters) for i in range(n)]


def test_every_request_completes_in_size_classes(fake_device):
    srv = _server()
    res = srv.run(_requests(10))
    assert [r.rid for r in res] == list(range(10))
    assert all(not r.error and r.nit == 5 for r in res)
    # no group mixes size classes (< 160, [160, 320), >= 320)
    cls = lambda n: 0 if n < 160 else (1 if n < 320 else 2)
    assert all(len({cls(n) for n in g}) == 1 for g in FakeGroup.sizes)
    assert set(srv.phases) == {"admit", "regroup", "step", "finish"}


def test_failed_step_is_reported_and_engine_goes_on(fake_device):
    FakeGroup.fail_rids = {1}          # nbf=500: its engine's step fails while it is live
    srv = _server(max_live=2)
    res = srv.run(_requests(10))
    assert len(res) == 10
    assert "non-finite" in res[1].error
    # the large class serves later requests: 4 (350) and 6 (600)
    assert not res[4].error and not res[6].error
    assert sum(1 for r in res if not r.error) >= 8


def test_admission_respects_slots(fake_device):
    srv = _server(max_live=1)
    res = srv.run(_requests(4, iters=3))
    assert all(not r.error for r in res)
    # one slot: every group holds exactly one solve
    assert all(len(g) == 1 for g in FakeGroup.sizes)


def test_group_caps(fake_device):
    """No lock-step group holds more live solves than its class's cap; with a cap of one a
    solve runs in a group of its own."""
    srv = _server(max_live=8, caps=(3, 2, 1))
    sizes = [60, 70, 80, 90, 100, 200, 210, 220, 400, 500]
    reqs = [serve.SolveRequest(rid=i, nbf=n, seed=i, arrival=0.0, max_iter=4)
            for i, n in enumerate(sizes)]
    res = srv.run(reqs)
    assert all(not r.error for r in res)
    cap = lambda n: 3 if n < 160 else (2 if n < 320 else 1)
    assert FakeGroup.sizes and all(len(g) <= cap(g[0]) for g in FakeGroup.sizes)
    assert any(len(g) == 3 for g in FakeGroup.sizes)   # small solves do share groups


def test_cap_one_runs_the_single_system_loop(fake_device):
    """With a cap of one a class's solves run the single-system device loop (no group)."""
    SOLOS.clear()
    srv = _server(max_live=4, caps=(8, 1, 1))
    res = srv.run(_requests(10))
    assert all(not r.error and r.nit == 5 for r in res)
    assert sorted(SOLOS) == sorted(n for n in (r.nbf for r in res) if n >= 160)
    assert all(max(g) < 160 for g in FakeGroup.sizes)


def test_split_precision_runs_the_single_system_solves(fake_device):
    """precision "split": solves of nbf >= split_min_nbf that run the single-system loop
    (classes capped at one) get split precision; lock-step groups and smaller solves stay
    FP64."""
    SOLOS.clear()
    SOLO_PRECISION.clear()
    srv = _server(max_live=4, caps=(8, 1, 1))
    srv.precision, srv.split_min_nbf = "split", 300
    res = srv.run(_requests(10))
    assert all(not r.error and r.nit == 5 for r in res)
    by_rid = {r.rid: r.nbf for r in res}
    assert SOLO_PRECISION and all(
        prec == (serve.DN_SPLIT_F16 if by_rid[rid] >= 300 else serve.DN_FP64)
        for rid, prec in SOLO_PRECISION.items())
    assert any(prec == serve.DN_SPLIT_F16 for prec in SOLO_PRECISION.values())
    # a class with a lock-step cap above one stays FP64 whatever its size
    srv2 = _server(caps=(8, 8, 1))
    srv2.precision, srv2.split_min_nbf = "split", 100
    assert srv2.precision_of(250) == serve.DN_FP64          # class [160, 320), cap 8
    assert srv2.precision_of(400) == serve.DN_SPLIT_F16     # class >= 320, cap 1
    assert srv2.precision_of(90) == serve.DN_FP64           # below split_min_nbf
    srv2.mode = "streams"
    assert srv2.precision_of(250) == serve.DN_SPLIT_F16     # every solve is single-system


def test_server_rejects_unknown_precision():
    with pytest.raises(ValueError):
        serve.Server.__init__(serve.Server.__new__(serve.Server), precision="bf16")
