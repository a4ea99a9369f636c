"""Benchmark: RI J/K-MO builds per second on one MI355X.

Headline workload (the FP64 J/K-MO kernel on synthetic tensors): nbf=400, nbfaux=1600,
nbf5=120, b generated in HBM by the Philox generator, C = Q of a seeded QR, H symmetric
N(0,1), GNOF pairing ndoc=5, ncwo=23.  One step = one J/K-MO build (the RI transform
b_pqQ = C5^T b_mnQ C5, J, K and H_core) with b, C and H already resident in HBM.

The dominant kernel is the first transform.  b is symmetric in (m, n), so it runs on the
lower half of b: T = C5^T Lh with b_Q = Lh_Q + Lh_Q^T (and C5^T b_Q C5 = W + W^T,
W = T C5).  It is timed with HIP events on the library's stream around every launch.
roofline.achieved counts the reference's 2 nbf5 nbf^2 nbfaux FLOP per launch; the
executed_* fields count the MFMA work the kernel really does.  roofline.traffic is the
HBM bytes per launch that rocprofv3 counted (profiles/jk_build_<precision>_nbf<nbf>_
traffic.md), or null.  The CPU baseline times the NumPy oracle (BLAS-backed) on the same
tensor on the host cores for about --cpu-seconds.

Extra legs (--legs, default all; --no-extra for none):
  800     the same build at nbf=800, nbfaux=3200, nbf5=240 in FP64 and in split precision
          (FP16 hi+lo on MFMA, FP32 accumulation): their ratio and the split path's J/K and
          GNOF-energy error against the FP64 device path on the same tensor and orbitals;
  grouped 64 systems, nbf ~ U[48, 400], built together by grouped persistent kernels in
          one captured HIP graph, against per-system builds on their own streams;
  served  Poisson arrivals of orbital-optimisation solves, p50/p99 latency and solves/s;
  grad    RI nuclear gradients: the densities at 400/1600/120 and 800/3200/240 and one
          atom's contractions with derivative integrals in HBM (and from host arrays);
  occ     occoptr's conjugate gradient on the device against the host-driven iteration;
  h2o     the full energy() of H2O / cc-pVDZ (GNOF, ADAM, Softmax, default thresholds):
          integrals from the package's engine, RI metric and the whole outer loop on the
          device (the CPU baseline times the oracle's loop on the same integrals).

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--precision fp64|split]
                    [--legs 800,grouped,served,grad,occ,h2o] [--no-cpu-baseline] [--no-extra]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "RI J/K-MO builds/s + %peak at ΔE≤1e-6 Ha; p50/p99 solve latency"
FP64_PEAK_TFLOPS = 78.6
F16_PEAK_TFLOPS = 2500.0
HBM_PEAK_GBS = 8000.0


def log(msg):
    print("bench: " + msg, file=sys.stderr, flush=True)


def cpu_baseline(C, H, b, nbf5, budget_s=20.0):
    """Time the oracle's J/K-MO build on the host (bounded sample)."""
    import __graft_entry__
    orc = __graft_entry__.load_oracle()
    threads = None
    try:
        import threadpoolctl
        info = threadpoolctl.threadpool_info()
        threads = max((i.get("num_threads") or 0) for i in info) if info else None
    except Exception:
        pass
    orc.jkh_mo(C, H, b[:, :, :8], nbf5)  # warm BLAS
    t0 = time.perf_counter()
    n = 0
    while True:
        orc.jkh_mo(C, H, b, nbf5)
        n += 1
        if time.perf_counter() - t0 > budget_s or n >= 10000:
            break
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "builds/s", "cores": threads or os.cpu_count(),
            "kind": "port",
            "sample": "%d full J/K-MO builds of the same nbf=%d/nbfaux=%d/nbf5=%d tensor with the "
                      "NumPy FP64 oracle (OpenBLAS threads=%s), %.1f s"
                      % (n, b.shape[0], b.shape[2], nbf5, threads, dt)}


def executed_flops_sym_k1(nbf, naux, nbf5, precision="fp64"):
    """MFMA flops the symmetric-b first transform really issues (the count of its MFMA
    instructions times their flops; SQ_INSTS_MFMA agrees exactly at 400/1600/120 in FP64
    and 800/3200/240 split).
    fp64: T[n,p,Q] = sum_{m >= n} Lh[m,n,Q] C5[m,p]; a 16-row fragment row starting at n0
    runs the 16-wide k tiles from n0 to the end (earlier ones are masked to zero and
    skipped), rows past nbf and 16-column fragments past nbf5 are skipped.
    split: T[p,n,Q] with 128-column blocks; a 16-column fragment starting at c0 runs the
    32-wide k tiles from the one holding c0 to the end; 16-row fragments past nbf5 are
    skipped; every fragment pair issues three products (hi*hi, hi*lo, lo*hi)."""
    if precision == "fp64":
        nkt = -(-nbf // 16)
        frag_cols = -(-nbf5 // 16)
        ktiles = sum(nkt - r0 // 16 for r0 in range(0, nbf, 16))
        return 2.0 * 16 * 16 * 16 * frag_cols * ktiles * naux
    mpad = -(-nbf5 // 16) * 16        # fragment rows past nbf5 are skipped too
    nkt = -(-nbf // 32)
    ktiles = sum(max(0, nkt - c0 // 32) for c0 in range(0, -(-nbf // 128) * 128, 16))
    return 3 * 2.0 * mpad * 16 * 32 * ktiles * naux


def tri_flops(nbf, naux, nbf5, precision="fp64"):
    """Algorithmic flops of the symmetric-b first transform T = Lh^T C5: the lower triangle
    of each b slab (nbf(nbf+1)/2 entries, the diagonal halved) against nbf5 columns;
    three FP16 products per pair in split precision."""
    f = 2.0 * nbf5 * (nbf * (nbf + 1) // 2) * naux
    return f if precision == "fp64" else 3.0 * f


def profile_value(name, key):
    """A figure recorded in profiles/<name> as a line '<key>: <value>' (None if absent)."""
    try:
        with open(os.path.join(ROOT, "profiles", name)) as f:
            for line in f:
                if line.strip().startswith(key + ":"):
                    return float(line.split(":", 1)[1].split()[0])
    except (OSError, ValueError):
        pass
    return None


def time_builds(dev, steps, warmup):
    """(ms per build, dominant-kernel ms per launch, launches) with HIP-event timing."""
    for _ in range(warmup):
        dev.jkh_mo_resident()
    dev.sync()
    dev.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        dev.jkh_mo_resident()
    dev.sync()
    t1 = time.perf_counter()
    k_ms, k_n = dev.get_timing()
    dev.set_timing(False)
    return (t1 - t0) * 1e3 / steps, k_ms / max(k_n, 1), k_n


def time_lagrangian(api, sysd, dev, reps):
    """(elag, H_mat, E) of compute_Lagrange2 at the system's orbitals and its ms per call
    (host arrays in and out)."""
    p = sysd.p
    cj, ck = api.PNOFi_selector(sysd.n, p, dev)
    args = (sysd.C, sysd.n, sysd.H, dev, cj, ck, p.nalpha, p.nbeta)
    res = api.compute_Lagrange2(*args, return_energy=True)
    t0 = time.perf_counter()
    for _ in range(reps):
        api.compute_Lagrange2(*args, return_energy=True)
    return res, (time.perf_counter() - t0) * 1e3 / reps


def make_system(api, synthetic, nbf, naux, nbf5, seed, precision):
    sysd = synthetic.SyntheticSystem(nbf, naux, nbf5, seed)
    dev = api.System(sysd.p, precision=precision)
    dev.generate_b(seed)
    dev.set_orbitals(sysd.C)
    dev.set_hcore(sysd.H)
    return sysd, dev


def extra_legs(api, synthetic, steps, warmup, seed):
    """FP64 and split-precision builds at nbf=800, nbfaux=3200, nbf5=240."""
    nbf, naux, nbf5 = 800, 3200, 240
    out = {"config": {"nbf": nbf, "nbfaux": naux, "nbf5": nbf5}}
    F = synthetic.build_flops(nbf, naux, nbf5)
    sysd, dev = make_system(api, synthetic, nbf, naux, nbf5, seed, "fp64")
    ms, kms, _ = time_builds(dev, steps, warmup)
    J64, K64, H64 = dev.get_jkh()
    lag64, lag_ms64 = time_lagrangian(api, sysd, dev, max(3, steps // 2))
    dev.close()
    out["fp64"] = {"builds_per_s": 1e3 / ms, "ms_per_build": ms, "tflops": F / (ms * 1e-3) / 1e12,
                   "transform_kernel_ms": kms,
                   "transform_tflops": 2.0 * nbf5 * nbf * nbf * naux / (kms * 1e-3) / 1e12,
                   "transform_executed_frac_of_f64_peak":
                       executed_flops_sym_k1(nbf, naux, nbf5) / (kms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS}
    log("fp64 800/3200/240: %.3f ms/build" % ms)
    sysd, dev = make_system(api, synthetic, nbf, naux, nbf5, seed, "split")
    ms16, kms16, _ = time_builds(dev, steps, warmup)
    J16, K16, H16 = dev.get_jkh()
    lag16, lag_ms16 = time_lagrangian(api, sysd, dev, max(3, steps // 2))
    dev.close()
    el64, _, El64 = lag64
    el16, _, El16 = lag16
    scale = float(np.abs(el64).max())
    k1_useful = 2.0 * nbf5 * nbf * nbf * naux
    x16 = executed_flops_sym_k1(nbf, naux, nbf5, "split")
    # GNOF energy (calce, pnof.jl:1273-1300) of both builds, on the device
    p = sysd.p
    aux = api.System(synthetic.synthetic_param(nbf, 4, nbf5))
    cj, ck = api.PNOFi_selector(sysd.n, p, aux)
    E64 = api.calce(sysd.n, cj, ck, J64, K64, H64, p, aux)
    E16 = api.calce(sysd.n, cj, ck, J16, K16, H16, p, aux)
    aux.close()
    out["split_f16"] = {
        "builds_per_s": 1e3 / ms16, "ms_per_build": ms16, "tflops": F / (ms16 * 1e-3) / 1e12,
        "transform_kernel_ms": kms16,
        "transform_useful_tflops": k1_useful / (kms16 * 1e-3) / 1e12,
        # FP16 MFMA flops the kernel issues (three products over the lower half of b)
        "transform_mfma_tflops": x16 / (kms16 * 1e-3) / 1e12,
        "transform_mfma_frac_of_f16_peak": x16 / (kms16 * 1e-3) / 1e12 / F16_PEAK_TFLOPS,
        "J_max_rel_err": float(np.abs(J16 - J64).max() / np.abs(J64).max()),
        "K_max_rel_err": float(np.abs(K16 - K64).max() / np.abs(K64).max()),
        "gnof_energy_abs_err_Ha": float(abs(E16 - E64)),
    }
    # the Lagrangian (utils.jl:1-59) at the same orbitals: FP64, and split precision (both
    # transforms and the contraction Z = sum bpn bmat on FP16 planes; the C^J/C^K terms,
    # elag += C^T Z and the energy FP64), host boundary included (C, H, n, cj12, ck12 up;
    # elag, H_mat down)
    out["lagrangian"] = {
        "fp64_ms": lag_ms64, "split_ms": lag_ms16, "split_speedup": lag_ms64 / lag_ms16,
        "elag_max_rel_err": float(np.abs(el16 - el64).max()) / scale,
        "orbital_gradient_max_abs_err": float(np.abs(4 * (el16 - el16.T)
                                                     - 4 * (el64 - el64.T)).max()),
        "orbital_gradient_err_rel_to_max_elag":
            float(np.abs(4 * (el16 - el16.T) - 4 * (el64 - el64.T)).max()) / scale,
        "energy_abs_err_Ha": float(abs(El16 - El64)),
    }
    log("lagrangian 800/3200/240: fp64 %.2f ms, split %.2f ms, |d grad| %.2e"
        % (lag_ms64, lag_ms16, out["lagrangian"]["orbital_gradient_max_abs_err"]))
    out["split_speedup_vs_fp64"] = ms / ms16
    log("split 800/3200/240: %.3f ms/build (x%.2f), |dE|=%.2e"
        % (ms16, ms / ms16, abs(E16 - E64)))
    return out


def grad_flops(nbf, naux, nbf5):
    """Algorithmic FLOP of the gradient densities as built here, and of the reference's
    order of operations (grads.jl:61-89: IG = I3c inv(G), I2, d, m = d IG over (l, s))."""
    t = 2.0 * (nbf * nbf * nbf5 + nbf * nbf5 * nbf5) * naux   # one two-sided transform
    ours = 2 * t + 2 * (2.0 * nbf5 * nbf5 * naux * naux)        # b_pq, d; I2 and m GEMMs
    ref = 2 * t + 2 * (2.0 * nbf * nbf * naux * naux)           # I2, d; IG and m
    return ours, ref


def grad_leg(api, synthetic, seed, shapes=((400, 1600, 120), (800, 3200, 240)), reps=3):
    """RI nuclear gradients (grads.jl:1-100): the densities (dn_grad_create) and one
    atom's contractions (dn_grad_atom) with derivative integrals resident in HBM, and
    from host arrays at the first shape (PCIe-inclusive)."""
    out = {}
    for nbf, naux, nbf5 in shapes:
        sysd, dev = make_system(api, synthetic, nbf, naux, nbf5, seed, "fp64")
        p = sysd.p
        cj, ck = api.PNOFi_selector(sysd.n, p, dev)
        elag, _ = api.compute_Lagrange2(sysd.C, sysd.n, sysd.H, dev, cj, ck, p.nalpha,
                                        p.nbeta)
        Wt = np.eye(naux) + np.tril(np.full((naux, naux), 1e-3), -1)
        api.Gradient(dev, sysd.n, sysd.C, cj, ck, elag, Wt=Wt).close()      # warm-up
        t0 = time.perf_counter()
        for _ in range(reps):
            api.Gradient(dev, sysd.n, sysd.C, cj, ck, elag, Wt=Wt).close()
        ms_dens = (time.perf_counter() - t0) * 1e3 / reps
        g = api.Gradient(dev, sysd.n, sysd.C, cj, ck, elag, Wt=Wt)
        dS = api.DeviceArray((nbf, nbf, 3))
        g3 = api.DeviceArray((nbf, nbf, naux, 3))
        g2 = api.DeviceArray((naux, naux, 3))
        g.atom(dS, dS, dS, g3, g2)
        t0 = time.perf_counter()
        for _ in range(reps):
            g.atom(dS, dS, dS, g3, g2)
        ms_atom = (time.perf_counter() - t0) * 1e3 / reps
        # bytes one atom reads: d once and geri3c (3 components), m and geri2c, the
        # nbf^2 densities against dS, dT, dV
        atom_bytes = 8.0 * (4 * nbf * nbf * naux + 4 * naux * naux + 3 * 4 * nbf * nbf)
        ours, ref = grad_flops(nbf, naux, nbf5)
        rec = {"densities_ms": ms_dens,
               "densities_tflops": ours / (ms_dens * 1e-3) / 1e12,
               "densities_flops": ours, "reference_order_flops": ref,
               "reference_order_tflops_equiv": ref / (ms_dens * 1e-3) / 1e12,
               "atom_ms_hbm_inputs": ms_atom,
               "atom_gbs": atom_bytes / (ms_atom * 1e-3) / 1e9}
        if nbf == shapes[0][0]:
            h3 = np.zeros((nbf, nbf, naux, 3), order="F")
            h2 = np.zeros((naux, naux, 3), order="F")
            hS = np.zeros((nbf, nbf, 3), order="F")
            t0 = time.perf_counter()
            g.atom(hS, hS, hS, h3, h2)
            rec["atom_ms_host_inputs"] = (time.perf_counter() - t0) * 1e3
            del h3
        for a in (dS, g3, g2):
            a.free()
        g.close()
        dev.close()
        out["%d_%d_%d" % (nbf, naux, nbf5)] = rec
        log("gradient %d/%d/%d: densities %.2f ms (%.1f TFLOP/s), atom %.3f ms (%.0f GB/s)"
            % (nbf, naux, nbf5, ms_dens, rec["densities_tflops"], ms_atom, rec["atom_gbs"]))
    return out


def occ_leg(api, synthetic, seed, nbf=400, naux=1600, nbf5=120, reps=3):
    """occoptr's CG (minimization.jl:100-139) at 400/1600/120, GNOF, Softmax: on the
    device (dn_occoptr, one host read per 64 evaluations) against the same iteration
    driven from the host (one device call and read per energy or gradient)."""
    from donof_amd import minimization as mini
    sysd, dev = make_system(api, synthetic, nbf, naux, nbf5, seed, "fp64")
    p = sysd.p
    dev.jkh_mo_resident()
    dev.sync()
    g0 = mini.guess_gamma_softmax(p.ndoc, p.ncwo)
    tol = p.threshen
    api.occ_conjugate_gradient(g0, None, None, None, p, dev, tol)
    t0 = time.perf_counter()
    for _ in range(reps):
        x, E, nit, conv = api.occ_conjugate_gradient(g0, None, None, None, p, dev, tol)
    dev_ms = (time.perf_counter() - t0) * 1e3 / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        h = mini.occoptr_host_cg(g0, p, dev)
    host_ms = (time.perf_counter() - t0) * 1e3 / reps
    dev.close()
    out = {"config": {"nbf": nbf, "nbfaux": naux, "nbf5": nbf5, "nv": p.nv,
                      "threshen": tol},
           "device_ms": dev_ms, "host_driven_ms": host_ms, "speedup": host_ms / dev_ms,
           "iterations": nit, "host_iterations": h.iterations, "converged": bool(conv),
           "energy_diff_Ha": abs(E - h.minimum)}
    log("occoptr CG 400/1600/120: device %.2f ms, host-driven %.2f ms (%d / %d iterations)"
        % (dev_ms, host_ms, nit, h.iterations))
    return out


def grouped_leg(api, synthetic, steps, warmup, seed, nsys=64, nbf_range=(48, 400)):
    """64 systems, nbf ~ U[48, 400], naux = 4 nbf, nbf5 = ceil(nbf/3): one grouped build
    of all of them per step (graph replay), against per-system builds on their own
    streams; in FP64 and in split precision (the same tensors and orbitals, so the split
    results are compared with the FP64 ones)."""
    shapes = synthetic.grouped_shapes(nsys, seed, *nbf_range)
    F = sum(synthetic.build_flops(*sh) for sh in shapes)
    out = {"config": {"nsys": nsys, "nbf_range": list(nbf_range), "seed": seed,
                      "nbf": [sh[0] for sh in shapes],
                      "tflop_per_build_of_all": F / 1e12,
                      "b_gb": sum(8.0 * n * n * a for n, a, _ in shapes) / 1e9}}
    ref64 = None
    for prec in ("fp64", "split"):
        systems = []
        for i, (nbf, naux, nbf5) in enumerate(shapes):
            s = api.System(synthetic.synthetic_param(nbf, naux, nbf5), precision=prec)
            s.generate_b(seed + i)
            s.set_orbitals(synthetic.make_C(nbf, seed + 1000 + i))
            s.set_hcore(synthetic.make_H(nbf, seed + 2000 + i))
            systems.append(s)
        log("grouped %s: %d systems (nbf %d-%d) resident, %.3f TFLOP per build of all"
            % (prec, nsys, nbf_range[0], nbf_range[1], F / 1e12))
        for _ in range(warmup):
            for s in systems:
                s.jkh_mo_resident()
        for s in systems:
            s.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            for s in systems:
                s.jkh_mo_resident()
        for s in systems:
            s.sync()
        seq_ms = (time.perf_counter() - t0) * 1e3 / steps
        ref = [s.get_jkh() for s in systems[:4]]
        g = api.Group(systems)
        t0 = time.perf_counter()
        cfg = g.tune()
        tune_s = time.perf_counter() - t0
        plan = g.plan()
        for _ in range(warmup):
            g.jkh()
        g.sync()
        g.set_timing(True)
        t0 = time.perf_counter()
        for _ in range(steps):
            g.jkh()
        g.sync()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        ev_ms, n = g.get_timing()
        got = [s.get_jkh() for s in systems[:4]]

        def rel(a, b):
            return max(float(np.abs(x - y).max() / np.abs(y).max())
                       for (xj, xk, _), (yj, yk, _) in zip(a, b)
                       for x, y in ((xj, yj), (xk, yk)))
        g.close()
        for s in systems:
            s.close()
        tf = F / (ms * 1e-3) / 1e12
        log("grouped %s: %.2f ms per build of all %d (%.1f TFLOP/s), per-system streams %.2f ms"
            % (prec, ms, nsys, tf, seq_ms))
        leg = {"builds_per_s": nsys * 1e3 / ms, "ms_per_build_of_all": ms,
               "graph_event_ms": ev_ms / max(n, 1), "tflops": tf, "tiles": cfg, "plan": plan,
               "autotune_s": tune_s, "per_system_streams_ms": seq_ms,
               "speedup_vs_per_system": seq_ms / ms,
               "max_rel_diff_vs_per_system": rel(got, ref)}
        if prec == "fp64":
            ref64 = got
            leg["frac_fp64_peak"] = tf / FP64_PEAK_TFLOPS
        else:
            leg["JK_max_rel_err_vs_fp64_group"] = rel(got, ref64)
        out[prec] = leg
    return out


def serving_leg(api, serve, n=100, rate=2.0, seed=7, max_live=8,
                modes=("grouped", "streams", "grouped-split")):
    """Served solves: Poisson arrivals at `rate`/s of GNOF orbital optimisations (ADAM, at
    most 200 iterations, stop when 2 max|g| < 1e-4), nbf log-uniform in [50, 600],
    naux = 4 nbf, nbf5 = ceil(nbf/3); admission by the device's HBM budget.  The lock-step
    grouped engines are the served path; one stream per solve is timed beside it.
    Latency is arrival to exit (converged or at the 200-iteration cap); the converged
    count is reported (the random synthetic Hamiltonians rarely reach 1e-4 in 200).
    Mode "grouped-split" is the grouped server with the single-system solves (nbf >= 160)
    at split precision; its energies are compared with the FP64 grouped run's."""
    out = {"requests": n, "arrival_rate_per_s": rate, "nbf_range": [50, 600],
           "max_live": max_live, "max_iter": 200, "gtol": 1e-4}
    energies = {}
    for mode in modes:
        reqs = serve.synthetic_requests(n, rate, seed=seed, nbf_min=50, nbf_max=600,
                                        max_iter=200, gtol=1e-4)
        base, _, prec = mode.partition("-")
        srv = serve.Server(hbm_budget=0, max_live=max_live, mode=base,
                           precision=prec or "fp64")
        out["hbm_budget_bytes"] = srv.ctx.hbm_budget
        t0 = time.perf_counter()
        res = srv.run(reqs, log=lambda m: log("serving (%s): %s" % (mode, m)))
        wall = time.perf_counter() - t0
        srv.close()
        st = serve.latency_stats(res, wall)
        log("serving (%s): %d solves in %.1f s, p50 %.3f s, p99 %.3f s"
            % (mode, st["solves"], wall, st["p50_s"] or 0.0, st["p99_s"] or 0.0))
        errs = [r.error for r in res if r.error]
        out[mode] = {"wall_s": wall, "errors": len(errs), **st}
        energies[mode] = {r.rid: (r.nbf, r.E) for r in res if not r.error}
        if prec == "split":
            out[mode]["split_min_nbf"] = srv.split_min_nbf
            out[mode]["split_solves"] = sum(srv.precision_of(r.nbf) == serve.DN_SPLIT_F16
                                            for r in res)
        if errs:
            out[mode]["first_error"] = errs[0]
            log("serving (%s): %d errors, first: %s" % (mode, len(errs), errs[0]))
        if mode == "grouped":
            out[mode]["size_classes"] = list(srv.size_classes)
            out[mode]["group_caps"] = list(srv.group_caps)
            out[mode]["phase_s"] = dict(srv.phases)
            log("serving (grouped) phases: %s" % ", ".join(
                "%s %.2f s" % kv for kv in sorted(srv.phases.items())))
    if "grouped" in energies and "grouped-split" in energies:
        # the same requests at FP64 and split precision: 200 ADAM steps each, so the
        # difference holds the split Lagrangian's error carried along two trajectories
        a, b = energies["grouped"], energies["grouped-split"]
        d = [abs(a[k][1] - b[k][1]) for k in a if k in b and a[k][0] >= 160]
        if d:
            out["grouped-split"]["max_abs_dE_vs_fp64_Ha"] = max(d)
            out["grouped-split"]["median_abs_dE_vs_fp64_Ha"] = float(np.median(d))
            out["grouped-split"]["dE_note"] = (
                "end energies of 200 ADAM steps from the same start: one Lagrangian "
                "evaluation agrees to ~1e-7 Ha, but ADAM's normalised step lets that "
                "difference change the path of an unconverged solve")
    return out


def served_in_child(timeout_s):
    """The served-solves leg in a child process (own HIP context) under a time limit, so
    that a stalled serving run cannot hold up the bench line."""
    import subprocess
    cmd = [sys.executable, "-u", os.path.abspath(__file__), "--served-only"]
    try:
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=None, timeout=timeout_s,
                           text=True)
    except subprocess.TimeoutExpired:
        log("serving: no result within %.0f s" % timeout_s)
        return {"error": "served-solves leg did not finish within %.0f s" % timeout_s}
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("{")]
    if r.returncode != 0 or not lines:
        return {"error": "served-solves leg exited with status %d" % r.returncode}
    return json.loads(lines[-1])


H2O_MOL = "0 1\n O  0.0000   0.000   0.121\n H  0.0000   0.751  -0.485\n H  0.0000  -0.751  -0.485\n"


def h2o_leg():
    """energy(bset, p) of H2O / cc-pVDZ on the device with the reference's default
    thresholds: wall time of the whole call (integrals, RI metric, outer loop, FCHK)."""
    import tempfile
    from donof_amd import basis, energy
    out = {"molecule": "H2O O(0,0,0.121) H(0,+-0.751,-0.485) A", "basis": "cc-pvdz",
           "functional": "GNOF", "orb_method": "ADAM", "occ_method": "Softmax"}
    with tempfile.TemporaryDirectory() as tmp:
        for rep in range(2):          # the first call also loads the integral code paths
            bset, p = basis.molecule(H2O_MOL, "cc-pvdz", spherical=True)
            p.ipnof, p.orb_method = 8, "ADAM"
            p.title = os.path.join(tmp, "h2o")
            t0 = time.perf_counter()
            E_t, C, n, _, hist = energy.energy(bset, p, printmode=False, return_history=True)
            dt = time.perf_counter() - t0
        out.update({"nbf": p.nbf, "nbfaux": p.nbfaux, "nbf5": p.nbf5, "E_total": E_t,
                    "outer_iterations": len(hist), "seconds": dt,
                    "final_grad_orb": hist[-1]["grad_orb"],
                    "final_grad_occ": hist[-1]["grad_occ"]})
        # the same run to tight thresholds (orbital gradient 1e-5, occupation gradient
        # 1e-6, energy change 1e-8)
        bset, p = basis.molecule(H2O_MOL, "cc-pvdz", spherical=True)
        p.ipnof, p.orb_method = 8, "ADAM"
        p.threshgorb, p.threshgocc, p.threshen = 1e-5, 1e-6, 1e-8
        p.title = os.path.join(tmp, "h2o_tight")
        t0 = time.perf_counter()
        E_t, C, n, _, hist = energy.energy(bset, p, printmode=False, return_history=True)
        out["tight"] = {"E_total": E_t, "outer_iterations": len(hist),
                        "seconds": time.perf_counter() - t0,
                        "final_grad_orb": hist[-1]["grad_orb"],
                        "final_grad_occ": hist[-1]["grad_occ"]}
    out["note"] = ("an ADAM trajectory's end point depends on roundoff (the step "
                   "-alpha mhat/sqrt(vhat + 1e-16) turns 1e-9 gradients into alpha/10 steps), "
                   "so the device and the oracle stop at different nearby stationary points "
                   "of the same functional; the GPU tests check the device's tight solution "
                   "with the oracle (its ENERGY1r, gradients and a restarted outer loop)")
    return out


def h2o_cpu_baseline():
    """The oracle's outer loop for the same H2O / cc-pVDZ energy on the host."""
    import __graft_entry__
    orc = __graft_entry__.load_oracle()
    from donof_amd import basis, integrals
    bset, p = basis.molecule(H2O_MOL, "cc-pvdz", spherical=True)
    h = bset.handle()
    Z = np.array([a.Z for a in bset.atoms], dtype=np.float64)
    S, T, V = h.int1e(Z, bset.centers_bohr())
    S, H = (S + S.T) / 2, (T + T.T + V + V.T) / 2
    G, I3c = integrals.integrals_for_gradients(bset)
    t0 = time.perf_counter()
    b = orc.iaux(G, I3c)
    _, C = integrals.core_guess(H, S)
    C = integrals.check_ortho(C, S, p, log=None)
    pr = orc.make_param(p.nbf, G.shape[0], p.ne, ipnof=8, orb_method="ADAM")
    E, _, _, hist, _ = orc.energy_loop(b, H, C, pr)
    dt = time.perf_counter() - t0
    return {"seconds": dt, "E_total": E + integrals.compute_E_nuc(bset, p),
            "outer_iterations": len(hist),
            "sample": "the oracle's whole outer loop (NumPy, scipy CG) for H2O / cc-pVDZ "
                      "from the same integrals and core guess, default thresholds"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nbf", type=int, default=400)
    ap.add_argument("--nbfaux", type=int, default=1600)
    ap.add_argument("--nbf5", type=int, default=120)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--precision", choices=("fp64", "split"), default="fp64",
                    help="arithmetic of the headline build (fp64 is the metric's)")
    ap.add_argument("--legs", default="800,grouped,served,grad,occ,h2o",
                    help="extra legs to run: any of 800, grouped, served, grad, occ, h2o "
                         "(or none)")
    ap.add_argument("--served-only", action="store_true",
                    help="run only the served-solves leg and print its JSON (the main run "
                         "starts this in a child process under a time limit)")
    ap.add_argument("--served-timeout", type=float, default=300.0)
    ap.add_argument("--served-modes", default="grouped,streams,grouped-split",
                    help="modes of the served-solves leg (--served-only)")
    ap.add_argument("--served-requests", type=int, default=100)
    ap.add_argument("--served-rate", type=float, default=2.0,
                    help="Poisson arrival rate of the served-solves leg (requests per second)")
    args = ap.parse_args()
    if args.served_only:
        import __graft_entry__
        __graft_entry__.load_package()
        from donof_amd import api, serve
        print(json.dumps(serving_leg(api, serve, n=args.served_requests, rate=args.served_rate,
                                     modes=tuple(args.served_modes.split(",")))))
        return 0
    if args.gpus != 1:
        print("bench.py measures one GPU: the hot path runs on a single MI355X and this "
              "build has no multi-GPU path (--gpus %d requested)" % args.gpus, file=sys.stderr)
        return 2

    import __graft_entry__
    __graft_entry__.load_package()
    from donof_amd import api, synthetic

    nbf, naux, nbf5 = args.nbf, args.nbfaux, args.nbf5
    sysd, dev = make_system(api, synthetic, nbf, naux, nbf5, args.seed, args.precision)
    log("system resident, warmup %d" % args.warmup)
    ms_step, k_avg_ms, k_n = time_builds(dev, args.steps, args.warmup)
    F = synthetic.build_flops(nbf, naux, nbf5)
    # the reference's count for the first transform (no symmetry credit; can pass the peak)
    ref_flops = 2.0 * nbf5 * nbf * nbf * naux
    ref_tflops = ref_flops / (k_avg_ms * 1e-3) / 1e12
    # algorithmic work of the kernel as designed: the lower triangle of each b slab
    # (m >= n, nbf(nbf+1)/2 entries) against nbf5 columns; split precision issues three
    # products (hi*hi, hi*lo, lo*hi) per pair
    k_flops = tri_flops(nbf, naux, nbf5, args.precision)
    k_tflops = k_flops / (k_avg_ms * 1e-3) / 1e12
    x_flops = executed_flops_sym_k1(nbf, naux, nbf5, args.precision)
    x_tflops = x_flops / (k_avg_ms * 1e-3) / 1e12
    peak = FP64_PEAK_TFLOPS if args.precision == "fp64" else F16_PEAK_TFLOPS
    tprof = "jk_build_%s_nbf%d_traffic.md" % (args.precision, nbf)
    traffic = profile_value(tprof, "traffic_bytes_per_launch")
    build_traffic = profile_value(tprof, "build_traffic_bytes_per_build")
    b_build = synthetic.build_bytes(nbf, naux, nbf5, 8 if args.precision == "fp64" else 4)
    out = {
        "metric": METRIC,
        "value": 1e3 / ms_step,
        "unit": "builds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f16x2 (hi+lo) products, f32 accumulate",
        "data": "synthetic (Philox-generated b in HBM, seeded QR orbitals)",
        "config": {"workload": "%s J/K-MO build nbf=%d nbfaux=%d nbf5=%d"
                               % (args.precision, nbf, naux, nbf5),
                   "nbf": nbf, "nbfaux": naux, "nbf5": nbf5, "precision": args.precision},
        "build_tflops": F / (ms_step * 1e-3) / 1e12,
        "pct_fp64_peak": 100.0 * F / (ms_step * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
        "roofline": {"bound": "mfma",
                     "kernel": ("dgemm_lds" if args.precision == "fp64" else "split_lds_kernel")
                               + " symmetric-b first transform (lower half of b)",
                     "achieved": k_tflops, "peak": peak, "unit": "TFLOP/s",
                     "frac": k_tflops / peak, "traffic": traffic,
                     "kernel_ms": k_avg_ms, "launches": k_n,
                     "algorithmic_flops_per_launch": k_flops,
                     "executed_flops_per_launch": x_flops,
                     "executed_tflops": x_tflops,
                     "executed_frac": x_tflops / peak,
                     "reference_order_flops_per_launch": ref_flops,
                     "achieved_reference_order": ref_tflops,
                     "build_traffic_bytes": build_traffic,
                     "build_min_bytes": b_build,
                     "build_traffic_ratio": (build_traffic / b_build) if build_traffic else None,
                     "note": "achieved/frac count the kernel's algorithmic work, the lower "
                             "triangle of b: 2*nbf5*nbf(nbf+1)/2*nbfaux flops per launch (x3 "
                             "products in split precision); executed_* count the MFMA "
                             "instructions issued (16-wide fragments along the diagonal); "
                             "achieved_reference_order counts the reference's "
                             "2*nbf5*nbf^2*nbfaux, which the half contraction can exceed the "
                             "peak on; traffic and build_traffic_bytes are rocprofv3 "
                             "FETCH_SIZE+WRITE_SIZE per launch and per build"},
    }
    log("%.3f ms/build, transform kernel %.3f ms (%.1f TFLOP/s)" % (ms_step, k_avg_ms, k_tflops))
    # the same build through the host-array boundary (dn_jkh_mo: C and H uploaded, J, K,
    # H_core downloaded on every call), for the PCIe-inclusive rate
    api.computeJKH_MO(sysd.C, sysd.H, dev, sysd.p)
    t0 = time.perf_counter()
    for _ in range(5):
        api.computeJKH_MO(sysd.C, sysd.H, dev, sysd.p)
    out["host_boundary_ms_per_build"] = (time.perf_counter() - t0) * 1e3 / 5
    b = None
    if not args.no_cpu_baseline:
        log("cpu baseline")
        # a split system keeps b only as FP16 planes: the host generator makes the same b
        b = (dev.download_b() if args.precision == "fp64"
             else synthetic.make_b(nbf, naux, args.seed))
    dev.close()
    if b is not None:
        out["cpu_baseline"] = cpu_baseline(sysd.C, sysd.H, b, nbf5, args.cpu_seconds)
        del b
    else:
        out["cpu_baseline"] = None
    legs = set() if args.no_extra else set(x for x in args.legs.split(",") if x)
    if "800" in legs:
        out["extra_800_3200_240"] = extra_legs(api, synthetic, max(5, args.steps // 2),
                                               args.warmup, args.seed)
    if "grouped" in legs:
        out["extra_grouped_64"] = grouped_leg(api, synthetic, max(3, args.steps // 4),
                                              args.warmup, args.seed)
        # beside the configuration: many small systems, where per-system launches are
        # latency-bound and grouping pays most
        out["extra_grouped_64_small"] = grouped_leg(api, synthetic, max(5, args.steps // 2),
                                                    args.warmup, args.seed,
                                                    nbf_range=(16, 64))
    if "grad" in legs:
        out["extra_gradients"] = grad_leg(api, synthetic, args.seed)
    if "occ" in legs:
        out["extra_occoptr"] = occ_leg(api, synthetic, args.seed)
    if "h2o" in legs:
        out["extra_h2o_energy"] = h2o_leg()
        if out["cpu_baseline"] is not None:
            out["cpu_baseline"]["h2o_energy"] = h2o_cpu_baseline()
    if "served" in legs:
        out["extra_served"] = served_in_child(args.served_timeout)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
