"""Served orbital-optimisation solves with continuous batching on one GPU.

Requests (synthetic RI systems of uneven size) arrive over time.  The scheduler admits
them first-come-first-served whenever a slot is free and the system's HBM footprint
(dn_sys_footprint) fits in what the budget has left; otherwise they wait.  An admitted
request gets its own resident system (b generated in HBM, C, H and the occupations set
once) and runs the ADAM loop of minimization.jl:344-438 (maxloop = max_iter, stop when
2 max|g| < gtol).

mode "grouped" (default): lock-step engines (per nbf size class, each on its own thread
and stream, each holding at most the class's cap of live solves) advance their live
solves together (dn_ogroup_run, a native loop of dn_ogroup_step): per step every solve takes its ADAM step and rotation as grouped
launches, and the orbital Lagrangians of all of them run as grouped GEMMs in one captured
HIP graph.  Admitted requests are set up by preparer threads (on their own streams)
while the engines keep stepping; when a solve finishes it leaves, and prepared requests
join at the next step (the group is rebuilt).  Size classes keep a small solve from
advancing at the pace of a large one.  mode "streams": one host thread per live solve,
each running the whole loop on its own HIP stream (dn_orbopt_resident).

precision "split": solves of nbf >= split_min_nbf that run the single-system loop (every
solve in mode "streams"; in mode "grouped" those of the classes capped at one) hold b as
FP16 hi/lo planes only and run both transforms of every Lagrangian on FP16 MFMA (the
split-precision Lagrangian), with rotation, ADAM step and energy sum in FP64; lock-step
groups stay FP64 (their steps are launch-bound, not transform-bound).

Every kernel of a solve is deterministic, reads only its own system's buffers, and each
GEMM of a system uses one tile family whose summation order does not depend on the tile
shape (split-K boundaries depend only on the system's sizes), so a result does not depend
on what else is running: tests check that each result is bitwise the same alone and in
the mix.
"""
from __future__ import annotations

import ctypes as ct
import math
import queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import api, synthetic
from ._lib import DN_FP64, DN_SPLIT_F16, DnError


@dataclass
class SolveRequest:
    rid: int
    nbf: int
    seed: int
    arrival: float = 0.0          # seconds after the start of the run
    max_iter: int = 200
    gtol: float = 1e-4            # stop when 2 max|grad| < gtol (minimization.jl:400)
    alpha: float = 0.01

    @property
    def naux(self):
        return 4 * self.nbf

    @property
    def nbf5(self):
        return int(math.ceil(self.nbf / 3))


@dataclass
class SolveResult:
    rid: int
    nbf: int
    E: float
    nit: int
    success: bool
    arrival: float
    admitted: float
    done: float
    error: str = ""

    @property
    def latency(self):
        return self.done - self.arrival


def synthetic_requests(n, rate, seed=0, nbf_min=50, nbf_max=600, max_iter=200, gtol=1e-4):
    """n requests with Poisson arrivals at `rate` per second and nbf log-uniform in
    [nbf_min, nbf_max] (seeded)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n)) if rate > 0 else np.zeros(n)
    nbf = np.exp(rng.uniform(math.log(nbf_min), math.log(nbf_max), n)).astype(int)
    return [SolveRequest(rid=i, nbf=int(nbf[i]), seed=1000 + i, arrival=float(t[i]),
                         max_iter=max_iter, gtol=gtol) for i in range(n)]


# Structural device operations of the grouped engines -- creating and destroying systems
# (allocation, synchronous copies, frees) and lock-step groups (graph capture) -- are kept
# apart from each other across threads; steps run concurrently.
_STRUCT_LOCK = threading.Lock()


def solve_one(req, ctx, precision=DN_FP64):
    """Run one request on its own resident system (FP64 or split precision); returns
    (E, nit, success)."""
    sysd = synthetic.SyntheticSystem(req.nbf, req.naux, req.nbf5, req.seed)
    p = sysd.p
    p.maxloop, p.alpha, p.threshgorb = req.max_iter, req.alpha, req.gtol
    dev = api.System(p, ctx=ctx, precision=precision)
    try:
        dev.generate_b(req.seed)
        E, C, nit, ok = api.orbopt("ADAM", sysd.gamma, sysd.C, sysd.H, dev, p)
    finally:
        dev.close()
    return E, nit, ok


def start_solve(req, ctx):
    """A resident system for the request with its ADAM run started (gamma, C, H set)."""
    sysd = synthetic.SyntheticSystem(req.nbf, req.naux, req.nbf5, req.seed)
    p = sysd.p
    p.maxloop, p.alpha, p.threshgorb = req.max_iter, req.alpha, req.gtol
    with _STRUCT_LOCK:
        dev = api.System(p, ctx=ctx)
    try:
        dev.generate_b(req.seed)
        dev.set_orbitals(sysd.C)
        dev.set_hcore(sysd.H)
        dev.set_gamma(sysd.gamma)
        dev.orbopt_start("ADAM", p)
    except Exception:
        with _STRUCT_LOCK:
            dev.close()
        raise
    return dev, p


class _LockStepEngine(threading.Thread):
    """Advances its live solves in lock step (dn_ogroup_step) on its own thread: solves
    arriving in the inbox join at the next step boundary (the group is rebuilt), finished
    ones leave and are reported through `finished(result, need)`."""

    # iterations per native run between checks of the inbox and of finished solves
    RUN_STEPS = 64

    def __init__(self, ctx, t0, phases, lock, finished, log):
        super().__init__(daemon=True)
        self.ctx, self.t0, self.phases, self.lock = ctx, t0, phases, lock
        self.finished, self.log = finished, log
        self.inbox = queue.Queue()
        self.load = 0        # solves submitted and not finished (under `lock`)
        # set when something waits in the inbox: the native loop ends at the next
        # iteration boundary (dn_ogroup_run), the newcomers join there
        self.stop = ct.c_int(0)

    def submit(self, item):
        self.inbox.put(item)
        self.stop.value = 1

    def _flush(self, ph):
        with self.lock:
            for k, v in ph.items():
                self.phases[k] += v

    def run(self):
        live = {}            # rid -> (req, dev, p, need, admitted)
        group = None
        stop = False
        ph = {"regroup": 0.0, "step": 0.0, "finish": 0.0}
        try:
            while not stop or live:
                changed = False
                self.stop.value = 0   # before draining: a later submit sets it again
                while True:   # newcomers join at this step boundary
                    try:
                        item = self.inbox.get(timeout=0.05) if not live and not stop \
                            else self.inbox.get_nowait()
                    except queue.Empty:
                        break
                    if item is None:
                        stop = True
                        continue
                    live[item[0].rid] = item
                    changed = True
                if not live:
                    continue
                try:
                    group = self._advance(live, group, changed, ph)
                except Exception as e:  # noqa: BLE001 - reported per solve, engine goes on
                    # the live solves fail with the error; later arrivals are served anew
                    with _STRUCT_LOCK:
                        if group is not None:
                            try:
                                group.close()
                            except Exception:  # noqa: BLE001
                                pass
                            group = None
                    t = time.perf_counter() - self.t0
                    with self.lock:
                        self.load -= len(live)
                    for rid, (req, dev, p, need, admitted) in list(live.items()):
                        try:
                            with _STRUCT_LOCK:
                                dev.close()
                        except Exception:  # noqa: BLE001
                            pass
                        self.finished(SolveResult(rid, req.nbf, float("nan"), 0, False,
                                                  req.arrival, admitted, t,
                                                  "%s: %s" % (type(e).__name__, e)), need)
                    live.clear()
        finally:
            with _STRUCT_LOCK:
                if group is not None:
                    group.close()
                for v in live.values():
                    v[1].close()
            self._flush(ph)

    def _advance(self, live, group, changed, ph):
        """One step of the live solves (the group rebuilt first when membership changed);
        finished solves leave and are reported.  Returns the group to keep."""
        if changed and group is not None:   # the systems leave the group's stream first
            with _STRUCT_LOCK:
                group.close()
            group = None
        if group is None:
            tg = time.perf_counter()
            with _STRUCT_LOCK:
                group = api.OGroup([v[1] for v in live.values()], self.ctx)
            ph["regroup"] += time.perf_counter() - tg
        ts = time.perf_counter()
        group.run(self.RUN_STEPS, self.stop)
        ph["step"] += time.perf_counter() - ts
        done = [rid for rid, v in live.items() if v[1].orbopt_done()]
        if not done:
            return group
        tf = time.perf_counter()
        with _STRUCT_LOCK:
            group.close()
        t = time.perf_counter() - self.t0
        for rid in done:
            req, dev, p, need, admitted = live[rid]
            E, nit, ok = dev.orbopt_finish(p)
            with _STRUCT_LOCK:
                dev.close()
            del live[rid]
            with self.lock:
                self.load -= 1
            if self.log:
                self.log("done %d nbf=%d E=%.8f nit=%d latency %.3fs"
                         % (rid, req.nbf, E, nit, t - req.arrival))
            self.finished(SolveResult(rid, req.nbf, E, nit, ok, req.arrival, admitted, t), need)
        ph["finish"] += time.perf_counter() - tf
        return None


class Server:
    """Admission by HBM budget and slots; live solves advanced in lock step by grouped
    kernels (mode "grouped") or by one thread and stream each (mode "streams")."""

    def __init__(self, hbm_budget=0, max_live=8, device=0, mode="grouped",
                 size_classes=(80, 120, 160, 320), group_caps=(8, 8, 8, 1, 1), preparers=4,
                 precision="fp64", split_min_nbf=160):
        if mode not in ("grouped", "streams"):
            raise ValueError("mode must be 'grouped' or 'streams'")
        if precision not in ("fp64", "split"):
            raise ValueError("precision must be 'fp64' or 'split'")
        self.precision = precision
        self.split_min_nbf = int(split_min_nbf)
        # nbf edges of the engines: a lock-step group advances at the pace of its largest
        # member, so the classes below 160 are narrow (100 requests at 2/s, nbf in
        # [50, 600]: p50 0.63 s with one class below 160, 0.55 s with two, 0.53 s with
        # three; one stream per solve 0.55 s)
        self.size_classes = tuple(sorted(size_classes))
        # live solves one lock-step group of each class holds: small systems share
        # grouped kernels (their steps are launch-bound); a larger one fills the GPU on
        # its own and would only wait for its group's slowest member: with a cap of one
        # it runs the single-system device loop on a stream of its own
        if len(group_caps) != len(self.size_classes) + 1 or min(group_caps) < 1:
            raise ValueError("group_caps needs one cap >= 1 per size class")
        self.group_caps = tuple(int(c) for c in group_caps)
        # threads setting admitted solves up (host generators, system creation, b in HBM,
        # the first Lagrangian) while the engines step: one per request at a time would
        # make admission the queue at a few requests per second
        self.preparers = max(1, int(preparers))
        self.ctx = api.Context(device, hbm_budget)
        self.max_live = max_live
        self.mode = mode
        self.results = []
        self.phases = {}      # seconds spent per phase of the last run (grouped mode)
        self._lock = threading.Lock()
        self._live = 0
        self._reserved = 0

    def precision_of(self, nbf):
        """The precision a solve of this size runs at (see the module docstring)."""
        if self.precision != "split" or nbf < self.split_min_nbf:
            return DN_FP64
        if self.mode == "grouped":
            k = sum(1 for edge in self.size_classes if nbf >= edge)
            if self.group_caps[k] != 1:
                return DN_FP64          # a lock-step group (FP64)
        return DN_SPLIT_F16

    def footprint(self, req):
        """HBM bytes the request's system reserves at its precision."""
        return api.footprint(synthetic.synthetic_param(req.nbf, req.naux, req.nbf5),
                             self.precision_of(req.nbf))

    def _fits(self, bytes_needed):
        # every live solve holds its full footprint until it finishes
        return self._reserved + bytes_needed <= self.ctx.hbm_budget

    def run(self, requests, log=None):
        """Serve `requests` (arrival times relative to now).  Blocks until all are done;
        returns the results ordered by request id."""
        if self.mode == "grouped":
            return self._run_grouped(requests, log)
        return self._run_streams(requests, log)

    def _run_grouped(self, requests, log=None):
        t0 = time.perf_counter()
        ph = {"admit": 0.0, "regroup": 0.0, "step": 0.0, "finish": 0.0}
        self.phases = ph
        lock = threading.Lock()
        results = []
        state = {"reserved": 0, "slots": 0}   # admitted and not finished
        done_ev = threading.Event()

        def finished(res, need):
            with lock:
                results.append(res)
                state["reserved"] -= need
                state["slots"] -= 1
            done_ev.set()

        # lock-step engines per size class, each on its own thread and stream, holding at
        # most group_caps[class] live solves: solves of similar cost share grouped kernels,
        # small ones do not wait for large ones; an engine is started when every engine of
        # the class is full
        engines = [[] for _ in range(len(self.size_classes) + 1)]

        def class_of(nbf):
            k = 0
            while k < len(self.size_classes) and nbf >= self.size_classes[k]:
                k += 1
            return k

        def engine_for(nbf):
            k = class_of(nbf)
            with lock:
                for e in engines[k]:
                    if e.load < self.group_caps[k]:
                        e.load += 1
                        return e
            e = _LockStepEngine(self.ctx, t0, ph, lock, finished, log)
            e.load = 1
            e.start()
            with lock:
                engines[k].append(e)
            return e

        def all_engines():
            with lock:
                return [e for cls in engines for e in cls]

        # admitted requests are set up (system created, b generated, run started) by a
        # few preparer threads on their own streams while the engines keep stepping
        prep_q = queue.Queue()

        solos = []

        def solo(req, need, admitted):
            # a class whose groups hold one solve: the whole device loop on the solve's own
            # stream (dn_orbopt_resident), as in mode "streams"
            err = ""
            E, nit, ok = float("nan"), 0, False
            try:
                E, nit, ok = solve_one(req, self.ctx, self.precision_of(req.nbf))
            except Exception as e:  # noqa: BLE001 - reported as the request's error
                err = "%s: %s" % (type(e).__name__, e)
            t = time.perf_counter() - t0
            if log and not err:
                log("done %d nbf=%d E=%.8f nit=%d latency %.3fs"
                    % (req.rid, req.nbf, E, nit, t - req.arrival))
            finished(SolveResult(req.rid, req.nbf, E, nit, ok, req.arrival, admitted, t, err),
                     need)

        def preparer():
            while True:
                item = prep_q.get()
                if item is None:
                    return
                req, need, admitted = item
                if self.group_caps[class_of(req.nbf)] == 1:
                    th = threading.Thread(target=solo, args=(req, need, admitted), daemon=True)
                    th.start()
                    with lock:
                        solos.append(th)
                    continue
                ta = time.perf_counter()
                try:
                    dev, p = start_solve(req, self.ctx)
                except Exception as e:  # noqa: BLE001 - reported as the request's error
                    finished(SolveResult(req.rid, req.nbf, float("nan"), 0, False, req.arrival,
                                         admitted, time.perf_counter() - t0,
                                         "%s: %s" % (type(e).__name__, e)), need)
                    continue
                finally:
                    with lock:
                        ph["admit"] += time.perf_counter() - ta
                engine_for(req.nbf).submit((req, dev, p, need, admitted))

        workers = [threading.Thread(target=preparer, daemon=True)
                   for _ in range(getattr(self, "preparers", 1))]
        for w in workers:
            w.start()
        pending = sorted(requests, key=lambda r: r.arrival)
        waiting = []
        try:
            while True:
                now = time.perf_counter() - t0
                while pending and pending[0].arrival <= now:
                    waiting.append(pending.pop(0))
                while waiting:   # FIFO admission while slots and memory allow
                    req = waiting[0]
                    need = self.footprint(req)
                    if need > self.ctx.hbm_budget:
                        waiting.pop(0)
                        with lock:
                            results.append(SolveResult(req.rid, req.nbf, float("nan"), 0,
                                                       False, req.arrival, now, now,
                                                       "larger than the HBM budget"))
                        continue
                    with lock:
                        ok = (state["slots"] < self.max_live
                              and state["reserved"] + need <= self.ctx.hbm_budget)
                        if ok:
                            state["slots"] += 1
                            state["reserved"] += need
                    if not ok:
                        break
                    waiting.pop(0)
                    prep_q.put((req, need, now))
                    if log:
                        log("admit %d nbf=%d at %.3fs" % (req.rid, req.nbf, now))
                with lock:
                    if len(results) == len(requests):
                        break
                if not all(w.is_alive() for w in workers) or \
                        not all(e.is_alive() for e in all_engines()):
                    raise RuntimeError("a serving thread stopped with requests outstanding")
                done_ev.wait(min(0.002, max(0.0, pending[0].arrival - now)) if pending
                             else 0.002)
                done_ev.clear()
        finally:
            for _ in workers:
                prep_q.put(None)
            for w in workers:
                w.join()
            for th in list(solos):
                th.join()
            for e in all_engines():
                e.submit(None)
            for e in all_engines():
                e.join()
        self.results = sorted(results, key=lambda r: r.rid)
        return self.results

    def _run_streams(self, requests, log=None):
        t0 = time.perf_counter()
        pending = sorted(requests, key=lambda r: r.arrival)
        waiting = []
        threads = []
        done_q = queue.Queue()

        def worker(req, admitted, need):
            err = ""
            E, nit, ok = float("nan"), 0, False
            try:
                E, nit, ok = solve_one(req, self.ctx, self.precision_of(req.nbf))
            except DnError as e:
                err = str(e)
            t = time.perf_counter() - t0
            with self._lock:
                self._live -= 1
                self._reserved -= need
            done_q.put(SolveResult(req.rid, req.nbf, E, nit, ok, req.arrival, admitted, t, err))

        results = []
        while pending or waiting or len(results) < len(requests):
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.pop(0))
            # FIFO admission while slots and memory allow
            while waiting:
                req = waiting[0]
                need = self.footprint(req)
                with self._lock:
                    ok = self._live < self.max_live and self._fits(need)
                    if ok:
                        self._live += 1
                        self._reserved += need
                if not ok:
                    if need > self.ctx.hbm_budget:
                        waiting.pop(0)
                        results.append(SolveResult(req.rid, req.nbf, float("nan"), 0, False,
                                                   req.arrival, now, now,
                                                   "larger than the HBM budget"))
                        continue
                    break
                waiting.pop(0)
                th = threading.Thread(target=worker, args=(req, now, need), daemon=True)
                th.start()
                threads.append(th)
                if log:
                    log("admit %d nbf=%d at %.3fs" % (req.rid, req.nbf, now))
            try:
                r = done_q.get(timeout=0.002)
                results.append(r)
                if log:
                    log("done %d nbf=%d E=%.8f nit=%d latency %.3fs"
                        % (r.rid, r.nbf, r.E, r.nit, r.latency))
            except queue.Empty:
                pass
        for th in threads:
            th.join()
        self.results = sorted(results, key=lambda r: r.rid)
        return self.results

    def close(self):
        self.ctx.close()


def latency_stats(results, wall):
    """Latency = time from arrival to exit (converged, or stopped at the iteration cap);
    the converged solves' own percentiles beside it."""
    ok = [r for r in results if not r.error]
    lat = np.array([r.latency for r in ok])
    conv = np.array([r.latency for r in ok if r.success])
    pct = lambda a, q: float(np.percentile(a, q)) if len(a) else None
    return {"solves": len(lat), "solves_per_s": len(lat) / wall if wall > 0 else 0.0,
            "latency": "time to exit (converged or at the iteration cap)",
            "p50_s": pct(lat, 50), "p99_s": pct(lat, 99),
            "mean_iterations": float(np.mean([r.nit for r in ok])) if ok else None,
            "converged": int(sum(r.success for r in ok)),
            "converged_p50_s": pct(conv, 50), "converged_p99_s": pct(conv, 99)}
