"""The outer macro-iteration driver of DoNOF.energy (energy.jl:3-320) over a
device-resident system, and ENERGY1r (utils.jl:117-136).

The integral setup (GaussianBasis, energy.jl:29) is outside this package: the caller
supplies the RI tensor (a `System` with b uploaded or generated), H_core, starting
orbitals and the nuclear repulsion.  Each outer iteration runs, as the reference does:
the orbital optimiser (p.orb_method) on the device, the occupation optimiser, the
Softmax re-ordering, and the diagnostics (ENERGY1r, calcorbg, J/K and calcoccg), then
the convergence test max|grad_orb| < threshgorb and ||grad_occ|| < threshgocc.
"""
from __future__ import annotations

import time

import numpy as np

from . import api
from . import minimization as mini


def computeE_elec_hc(Hmat, n, elag, p):
    """E of utils.jl:191-214 from the device's Hmat = C^T H C5 (its diagonal is
    (C^T H C)_ii for i < nbf5)."""
    hd = np.diag(np.asarray(Hmat)[:p.nbf5, :p.nbf5])
    E = float(np.trace(elag[:p.nbf5, :p.nbf5]))
    E += float(np.dot(n[:p.nbeta], hd[:p.nbeta]))
    if not p.HighSpin:
        E += float(np.dot(n[p.nbeta:p.nalpha], hd[p.nbeta:p.nalpha]))
    else:
        E += 0.5 * float(np.sum(hd[p.nbeta:p.nalpha]))
    E += float(np.dot(n[p.nalpha:p.nbf5], hd[p.nalpha:p.nbf5]))
    return E


def ENERGY1r(C, n, H, I, cj12, ck12, p):
    """(E, elag, sumdiff, maxdiff) of utils.jl:117-136 (no1 = 0 branch)."""
    if p.no1 != 0:
        raise NotImplementedError("the no1 > 0 branch (computeJKj/computeF_RC) is out of scope")
    elag, Hmat = api.compute_Lagrange2(C, n, H, I, cj12, ck12, p.nalpha, p.nbeta)
    E = computeE_elec_hc(Hmat, n, elag, p)
    d = np.abs(elag - elag.T)
    return E, elag, float(d.sum()), float(d.max())


def energy(I, H, C, p, E_nuc=0.0, n=None, freeze_occ=False, printmode=True, checkpoint=False,
           log=print):
    """Outer loop of energy.jl:140-220.  Returns (E_nuc + E, C, n, elag, history)."""
    if p.orb_method not in mini.ORBOPT:
        raise NotImplementedError("orb_method %r (ID and Rotations are broken in the "
                                  "reference)" % p.orb_method)
    C = np.array(C, dtype=np.float64, order="F", copy=True)
    if n is None:
        gamma = (mini.guess_gamma_softmax(p.ndoc, p.ncwo) if p.occ_method == "Softmax"
                 else mini.guess_gamma_trigonometric(p.ndoc, p.ncwo))
    else:
        gamma = (mini.n_to_gammas_softmax(n, p) if p.occ_method == "Softmax"
                 else mini.n_to_gammas_trigonometric(n, p))
    orbopt = mini.ORBOPT[p.orb_method]
    E = E_old = 9999.0
    elag = None
    history = []
    if printmode:
        log("  Nitext Nit_orb    Time Nit_occ    Time        Eelec          Etot"
            "          Ediff    maxdiff  MaxGrad_orb  Grad_occ")
    for i_ext in range(1, p.maxit + 1):
        t1 = time.time()
        E_orb, C, nit_orb, _ = orbopt(gamma, C, H, I, p)
        t2 = time.time()
        _, nit_occ, _, gamma, n, cj12, ck12 = mini.occoptr(gamma, C, H, I, freeze_occ, p)
        if p.occ_method == "Softmax":
            C, gamma = mini.order_occupations_softmax(C, gamma, p)
        t3 = time.time()
        E = E_orb
        E_diff = E - E_old
        E_old = E
        Etmp, elag, _, maxdiff = ENERGY1r(C, n, H, I, cj12, ck12, p)
        n, _ = api.ocupacion(gamma, p, I)
        cj12, ck12 = api.PNOFi_selector(n, p, I)
        grad_orb = api.calcorbg(np.zeros(p.nvar), n, cj12, ck12, C, H, I, p)
        J, K, Hc = api.computeJKH_MO(C, H, I, p)
        grad_occ = api.calcoccg(gamma, J, K, Hc, p, I)
        gmax, gocc = float(np.max(np.abs(grad_orb))), float(np.linalg.norm(grad_occ))
        history.append({"i_ext": i_ext, "E": E, "Etot": E + E_nuc, "nit_orb": nit_orb,
                        "nit_occ": nit_occ, "maxdiff": maxdiff, "grad_orb": gmax,
                        "grad_occ": gocc})
        if printmode:
            log("%6i %7i %10.1e %4i %10.1e %14.8f %14.8f %14.8f %10.6f   %4.1e   %4.1e"
                % (i_ext, nit_orb, t2 - t1, nit_occ, t3 - t2, E, E + E_nuc, E_diff, maxdiff,
                   gmax, gocc))
        if checkpoint:  # energy.jl:209-213 writes (E, C, n) every iteration
            np.savez(p.title + ".npz", E=Etmp, C=C, n=n)
        if gmax < p.threshgorb and gocc < p.threshgocc:
            break
    return E_nuc + E, C, n, elag, history
