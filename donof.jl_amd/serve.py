"""Served orbital-optimisation solves with continuous batching on one GPU.

Requests (synthetic RI systems of uneven size) arrive over time.  The scheduler admits
them first-come-first-served whenever a slot is free and the system's HBM footprint
(dn_sys_footprint) fits in what the budget has left; otherwise they wait.  An admitted
request gets its own resident system (b generated in HBM, C, H and the occupations set
once) and runs the ADAM loop of minimization.jl:344-438 (maxloop = max_iter, stop when
2 max|g| < gtol).

mode "grouped" (default): one engine advances all live solves in lock step
(dn_ogroup_step): per step every solve takes its ADAM step and rotation, and the orbital
Lagrangians of all of them run as grouped persistent GEMMs in one captured HIP graph.
When a solve finishes it leaves, and waiting requests join at the next step (the group is
rebuilt).  mode "streams": one host thread per live solve, each running the whole loop
on its own HIP stream (dn_orbopt_resident).

Every kernel of a solve is deterministic, reads only its own system's buffers, and the
tile shapes of the grouped GEMMs depend only on that system's sizes, so a result does
not depend on what else is running: tests check that each result is bitwise the same
alone and in the mix.
"""
from __future__ import annotations

import math
import queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import api, synthetic
from ._lib import DnError


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


def solve_one(req, ctx):
    """Run one request on its own resident system; returns (E, nit, success)."""
    sysd = synthetic.SyntheticSystem(req.nbf, req.naux, req.nbf5, req.seed)
    p = sysd.p
    p.maxloop, p.alpha, p.threshgorb = req.max_iter, req.alpha, req.gtol
    dev = api.System(p, ctx=ctx)
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
    dev = api.System(p, ctx=ctx)
    try:
        dev.generate_b(req.seed)
        dev.set_orbitals(sysd.C)
        dev.set_hcore(sysd.H)
        dev.set_gamma(sysd.gamma)
        dev.orbopt_start("ADAM", p)
    except Exception:
        dev.close()
        raise
    return dev, p


class Server:
    """Admission by HBM budget and slots; live solves advanced in lock step by grouped
    kernels (mode "grouped") or by one thread and stream each (mode "streams")."""

    def __init__(self, hbm_budget=0, max_live=8, device=0, mode="grouped"):
        if mode not in ("grouped", "streams"):
            raise ValueError("mode must be 'grouped' or 'streams'")
        self.ctx = api.Context(device, hbm_budget)
        self.max_live = max_live
        self.mode = mode
        self.results = []
        self._lock = threading.Lock()
        self._live = 0
        self._reserved = 0

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
        pending = sorted(requests, key=lambda r: r.arrival)
        waiting, results = [], []
        live = {}            # rid -> (req, dev, p, need, admitted)
        reserved = 0
        group = None
        changed = False
        while pending or waiting or live:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.pop(0))
            while waiting:   # FIFO admission while slots and memory allow
                req = waiting[0]
                need = api.footprint(synthetic.synthetic_param(req.nbf, req.naux, req.nbf5))
                if need > self.ctx.hbm_budget:
                    waiting.pop(0)
                    results.append(SolveResult(req.rid, req.nbf, float("nan"), 0, False,
                                               req.arrival, now, now,
                                               "larger than the HBM budget"))
                    continue
                if len(live) >= self.max_live or reserved + need > self.ctx.hbm_budget:
                    break
                waiting.pop(0)
                if group is not None:       # the systems leave the group's stream first
                    group.close()
                    group = None
                try:
                    dev, p = start_solve(req, self.ctx)
                except DnError as e:
                    results.append(SolveResult(req.rid, req.nbf, float("nan"), 0, False,
                                               req.arrival, now, time.perf_counter() - t0,
                                               str(e)))
                    continue
                live[req.rid] = (req, dev, p, need, now)
                reserved += need
                changed = True
                if log:
                    log("admit %d nbf=%d at %.3fs" % (req.rid, req.nbf, now))
            if not live:
                if pending and not waiting:
                    time.sleep(min(0.002, max(0.0, pending[0].arrival - now)))
                continue
            if changed or group is None:
                if group is not None:
                    group.close()
                group = api.OGroup([v[1] for v in live.values()], self.ctx)
                changed = False
            group.step()
            finished = [rid for rid, v in live.items() if v[1].orbopt_done()]
            if finished:
                group.close()
                group = None
                t = time.perf_counter() - t0
                for rid in finished:
                    req, dev, p, need, admitted = live.pop(rid)
                    E, nit, ok = dev.orbopt_finish(p)
                    dev.close()
                    reserved -= need
                    results.append(SolveResult(rid, req.nbf, E, nit, ok, req.arrival, admitted,
                                               t))
                    if log:
                        log("done %d nbf=%d E=%.8f nit=%d latency %.3fs"
                            % (rid, req.nbf, E, nit, t - req.arrival))
        if group is not None:
            group.close()
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
                E, nit, ok = solve_one(req, self.ctx)
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
                need = api.footprint(synthetic.synthetic_param(req.nbf, req.naux, req.nbf5))
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
    lat = np.array([r.latency for r in results if not r.error])
    return {"solves": len(lat), "solves_per_s": len(lat) / wall if wall > 0 else 0.0,
            "p50_s": float(np.percentile(lat, 50)) if len(lat) else None,
            "p99_s": float(np.percentile(lat, 99)) if len(lat) else None,
            "mean_iterations": float(np.mean([r.nit for r in results if not r.error]))
            if len(lat) else None,
            "converged": int(sum(r.success for r in results))}
