"""The outer macro-iteration driver of DoNOF.energy (energy.jl:3-320) over a
device-resident system, and ENERGY1r (utils.jl:117-136).

energy(bset, p) sets up the integrals as energy.jl:29-80 does (integrals.py: the host
integral engine, the RI metric on the device, the core guess and check_ortho) and writes
the FCHK file at the end; energy(I, H, C, p) takes a device-resident system the caller set
up (b uploaded or generated).  Each outer iteration runs, as the reference does:
the orbital optimiser (p.orb_method) on the device, the occupation optimiser, the
Softmax re-ordering, and the diagnostics (ENERGY1r, calcorbg, J/K and calcoccg), then
the convergence test max|grad_orb| < threshgorb and ||grad_occ|| < threshgocc.  After
the loop the pairs are sorted (order_subspaces, guess.jl:58-106) for ipnof > 4.

Checkpoints: the reference saves (E, C, n) to `<title>.jld2` after every outer iteration
and (C, n) at the end (energy.jl:209-213, 226-232), and resumes through read_C / read_n
(guess.jl:1-56) whose results the caller passes back to energy(C=..., n=...).  JLD2 is an
HDF5 file and no HDF5 library is available to this package, so the same records are
written as NumPy `.npz` files (`<title>.npz`, arrays E, C, n; column-major C as given)
and read back by read_C / read_n / read_all.
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


def computeE_elec(*args):
    """The two methods of utils.jl:191-227: computeE_elec(H, C, n, elag, p) (one-electron
    energies (C^T H C)_ii, half of them for the singly occupied orbitals in high spin) and
    computeE_elec(Hmat, n, elag, p) (sum n_i Hmat_ii over i < nbf5), each plus tr elag."""
    if len(args) == 5:
        H, C, n, elag, p = args
        C5 = np.asarray(C)[:, :p.nbf5]
        hd = np.sum(C5 * (np.asarray(H) @ C5), axis=0)
        return computeE_elec_hc(np.diag(hd), n, elag, p)
    Hmat, n, elag, p = args
    E = float(np.trace(np.asarray(elag)[:p.nbf5, :p.nbf5]))
    return E + float(np.dot(np.asarray(n)[:p.nbf5], np.diag(np.asarray(Hmat)[:p.nbf5, :p.nbf5])))


def computeLagrangeConvergency(elag):
    """utils.jl:229-236: (sum and max of |elag - elag^T|)."""
    d = np.abs(np.asarray(elag) - np.asarray(elag).T)
    return float(d.sum()), float(d.max())


def ENERGY1r(C, n, H, I, cj12, ck12, p, want_jkh=False):
    """(E, elag, sumdiff, maxdiff) of utils.jl:117-136 (no1 = 0 branch).  With want_jkh
    also (J_MO, K_MO, H_core) at the same orbitals, taken from the same transform of b."""
    if p.no1 != 0:
        raise NotImplementedError("the no1 > 0 branch (computeJKj/computeF_RC) is out of scope")
    if want_jkh:
        elag, Hmat, _, J, K, Hc = api.compute_Lagrange2_JKH(C, n, H, I, cj12, ck12, p.nalpha,
                                                           p.nbeta)
    else:
        elag, Hmat = api.compute_Lagrange2(C, n, H, I, cj12, ck12, p.nalpha, p.nbeta)
    E = computeE_elec_hc(Hmat, n, elag, p)
    d = np.abs(elag - elag.T)
    if want_jkh:
        return E, elag, float(d.sum()), float(d.max()), (J, K, Hc)
    return E, elag, float(d.sum()), float(d.max())


def order_subspaces(old_C, old_n, elag, H, I, p):
    """guess.jl:58-106: the strong orbitals sorted by increasing elag_ii (stable, as
    Julia's sortperm), each pair's weak orbitals (columns ll:ndoc:ul) and occupations
    moving with their strong orbital; the singly occupied block sorted the same way;
    frozen-core columns and columns past nbf5 kept.  Returns (C, n, elag) with elag from
    ENERGY1r at the sorted orbitals."""
    C = np.zeros_like(np.asarray(old_C))
    n = np.zeros_like(np.asarray(old_n))
    d = np.diag(elag)
    C[:, :p.no1] = old_C[:, :p.no1]
    n[:p.no1] = old_n[:p.no1]
    for i, i_old in enumerate(np.argsort(d[p.no1:p.nbeta], kind="stable"), start=1):
        i_old = int(i_old) + 1
        C[:, p.no1 + i - 1] = old_C[:, p.no1 + i_old - 1]
        n[p.no1 + i - 1] = old_n[p.no1 + i_old - 1]
        C[:, mini.pair_weak(p, i)] = old_C[:, mini.pair_weak(p, i_old)]
        n[mini.pair_weak(p, i)] = old_n[mini.pair_weak(p, i_old)]
    s0 = p.no1 + p.ndoc
    for i, i_old in enumerate(np.argsort(d[s0:p.no1 + p.ndns], kind="stable")):
        C[:, s0 + i] = old_C[:, s0 + int(i_old)]
        n[s0 + i] = old_n[s0 + int(i_old)]
    C[:, p.nbf5:] = old_C[:, p.nbf5:]
    C = np.asfortranarray(C)
    cj12, ck12 = api.PNOFi_selector(n, p, I)
    _, elag, _, _ = ENERGY1r(C, n, H, I, cj12, ck12, p)
    return C, n, elag


def _checkpoint_path(title):
    return title if title.endswith(".npz") else title + ".npz"


def save_checkpoint(title, C, n, E=None):
    """(E, C, n) as energy.jl:209-213 saves them, or (C, n) as energy.jl:226-232 does."""
    arrays = {"C": np.asarray(C), "n": np.asarray(n)}
    if E is not None:
        arrays["E"] = np.asarray(E)
    np.savez(_checkpoint_path(title), **arrays)


def _read(title, key):
    path = _checkpoint_path(title)
    try:
        with np.load(path, allow_pickle=False) as f:
            return np.array(f[key], order="F")
    except (OSError, KeyError):
        print("%s: no %s found" % (path, key))
        return None


def read_C(title="donof"):
    """guess.jl:1-15: the orbitals of a checkpoint, or None."""
    return _read(title, "C")


def read_n(title="donof"):
    """guess.jl:17-31: the occupations of a checkpoint, or None."""
    return _read(title, "n")


def read_all(title="donof"):
    """guess.jl:49-56 without fmiug0 (the ID orbital method is not supported): (C, n)."""
    return read_C(title), read_n(title)


def energy(*args, **kwargs):
    """DoNOF.energy.  energy(bset, p, ...) is the reference's entry point (energy.jl:3-320):
    integrals, guess and the outer loop for a `basis.BasisSet` (see energy_bset).
    energy(I, H, C, p, ...) runs the outer loop on a device-resident system the caller
    set up (see energy_resident)."""
    from .basis import BasisSet
    if args and isinstance(args[0], BasisSet):
        return energy_bset(*args, **kwargs)
    return energy_resident(*args, **kwargs)


def energy_bset(bset, p, C=None, n=None, aux=None, freeze_occ=False, printmode=True,
                do_fchk=True, checkpoint=True, ctx=None, precision="fp64", log=print,
                return_history=False, do_gradients=False, fmiug0=None, do_hfidr=False,
                do_nofmp2=False, nofmp2strategy="numerical", tolnofmp2=1e-10, do_ekt=False,
                do_mulliken_pop=False, do_lowdin_pop=True, do_m_diagnostic=True,
                do_mbpt=False, do_translate_to_donofsw=False, do_erpa=False):
    """energy.jl:3-320 for the RI path: compute_integrals (the integral engine and the
    device RI metric), E_nuc, the guess C = eigen(H, S) unless C is given, check_ortho,
    the outer loop (energy_resident), and the FCHK file `<p.title>.fchk`
    (energy.jl:274).  Returns (E_nuc + E, C, n, None) as the reference returns
    (E_t, C, n, fmiug0) (fmiug0 belongs to the ID orbital method, not supported), plus
    the per-iteration history with return_history.  do_gradients returns
    (grad, E_t, C, n) as energy.jl:309-314 does: the RI nuclear gradient (natoms x 3,
    Hartree/bohr) of the final state on the device (api.compute_geom_gradients, with the
    derivative integrals of the package's integral engine).

    The reference's other keywords (energy.jl:3-23) are accepted with its defaults: the
    post-loop analyses on by default there (Lowdin populations, M diagnostic) are skipped
    with a note, and asking for any other (an HF/ID guess or fmiug0, NOF-MP2, EKT, MBPT,
    ERPA, Mulliken populations, the DoNOFsw export) raises NotImplementedError: they are
    outside the orbital-optimisation path this package covers."""
    from . import integrals as ints
    from . import output
    unsupported = [k for k, v in (("fmiug0", fmiug0 is not None), ("do_hfidr", do_hfidr),
                                  ("do_nofmp2", do_nofmp2), ("do_ekt", do_ekt),
                                  ("do_mulliken_pop", do_mulliken_pop), ("do_mbpt", do_mbpt),
                                  ("do_translate_to_donofsw", do_translate_to_donofsw),
                                  ("do_erpa", do_erpa)) if v]
    if unsupported:
        raise NotImplementedError("energy(): %s not supported (post-SCF methods, the ID "
                                  "orbital method and DoNOFsw export are out of scope)"
                                  % ", ".join(unsupported))
    if printmode and (do_lowdin_pop or do_m_diagnostic):
        log("(Lowdin populations and the M diagnostic are not computed here)")
    S, T, V, H, I = ints.compute_integrals(bset, p, aux=aux, ctx=ctx, precision=precision,
                                           log=log if printmode else None)
    try:
        E_nuc = ints.compute_E_nuc(bset, p)
        if C is None:
            _, C = ints.core_guess(H, S)
        C = ints.check_ortho(C, S, p, log=log if printmode else None)
        E_t, C, n, elag, hist = energy_resident(I, H, C, p, E_nuc=E_nuc, n=n,
                                                freeze_occ=freeze_occ, printmode=printmode,
                                                checkpoint=checkpoint, log=log)
        if printmode:
            log("Final NOF Total Energy = %15.7f" % E_t)
        if do_fchk:
            output.fchk(p.title, p, bset, "Energy", E_t, elag, n, C)
        grad = None
        if do_gradients:
            cj12, ck12 = api.PNOFi_selector(n, p, I)
            grad = api.compute_geom_gradients(bset, n, C, cj12, ck12, elag, p, I)
    finally:
        I.close()
    if do_gradients:
        return (grad, E_t, C, n, hist) if return_history else (grad, E_t, C, n)
    if return_history:
        return E_t, C, n, None, hist
    return E_t, C, n, None


def energy_resident(I, H, C, p, E_nuc=0.0, n=None, freeze_occ=False, printmode=True,
                    checkpoint=False, log=print):
    """Outer loop of energy.jl:140-224.  Returns (E_nuc + E, C, n, elag, history); C, n
    and elag are sorted by order_subspaces for ipnof > 4.  checkpoint writes
    `<p.title>.npz` every iteration and at the end; a later call resumes with
    energy(..., C=read_C(title), n=read_n(title))."""
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
            C, gamma = mini.order_occupations_softmax(C, gamma, p, I)
        t3 = time.time()
        E = E_orb
        E_diff = E - E_old
        E_old = E
        # diagnostics (energy.jl:172-190) from one transform of b: ENERGY1r's Lagrangian
        # also gives J/K/H_core at these orbitals, and when the re-ordering left the
        # occupations as they were, calcorbg(0) (pnof.jl:1389-1407) is the packed
        # 4 (elag - elag^T) of that same Lagrangian
        Etmp, elag, sumdiff, maxdiff, (J, K, Hc) = ENERGY1r(C, n, H, I, cj12, ck12, p,
                                                            want_jkh=True)
        n_new, _ = api.ocupacion(gamma, p, I)
        if np.array_equal(n_new, n):
            grad_orb = api.pack_grad(elag, p.nbf5)
        else:
            n = n_new
            cj12, ck12 = api.PNOFi_selector(n, p, I)
            grad_orb = api.calcorbg(np.zeros(p.nvar), n, cj12, ck12, C, H, I, p)
        grad_occ = api.calcoccg(gamma, J, K, Hc, p, I)
        gmax, gocc = float(np.max(np.abs(grad_orb))), float(np.linalg.norm(grad_occ))
        history.append({"i_ext": i_ext, "E": E, "Etot": E + E_nuc, "nit_orb": nit_orb,
                        "nit_occ": nit_occ, "E1r": Etmp, "sumdiff": sumdiff,
                        "maxdiff": maxdiff, "grad_orb": gmax, "grad_occ": gocc})
        if printmode:
            log("%6i %7i %10.1e %4i %10.1e %14.8f %14.8f %14.8f %10.6f   %4.1e   %4.1e"
                % (i_ext, nit_orb, t2 - t1, nit_occ, t3 - t2, E, E + E_nuc, E_diff, maxdiff,
                   gmax, gocc))
        if checkpoint:  # energy.jl:209-213 writes (E, C, n) every iteration
            save_checkpoint(p.title, C, n, Etmp)
        if gmax < p.threshgorb and gocc < p.threshgocc:
            break
    if p.ipnof > 4 and elag is not None:  # energy.jl:222-224
        C, n, elag = order_subspaces(C, n, elag, H, I, p)
    if checkpoint:  # energy.jl:226-232 rewrites the file with (C, n)
        save_checkpoint(p.title, C, n)
    return E_nuc + E, C, n, elag, history
