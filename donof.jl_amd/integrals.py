"""The integral set-up of DoNOF.energy: compute_integrals (integrals.jl:1-65),
compute_E_nuc (utils.jl:299-312), and the orbital guess helpers check_ortho /
orthonormalize (utils.jl:431-497).

The one- and two-electron integrals come from libdonofhip's host integral engine over the
shells of a `basis.BasisSet`; the RI tensor b = I3c G^{-1/2} is formed on the device by
`api.Iaux` (dn_sys_ri_metric), which returns the device-resident System every later call
takes where the reference takes b.
"""
from __future__ import annotations

import numpy as np

from . import api
from . import basis as basis_mod

BOHR_PER_ANGSTROM_E_NUC = 1.88973  # utils.jl:306


def compute_integrals(bset, p, aux=None, ctx=None, precision="fp64", log=print):
    """integrals.jl:1-65, RI branch: S, T, V (each symmetrised as (X + X^T)/2), H = T + V;
    the fitting set (`aux`, default `bset.fitting_set()`), G = (P|Q) symmetrised and
    I3c = (mn|P) (on the device for orbital shells up to d and fitting shells up to g,
    else on the host); b = Iaux(G, I3c) on the device.  Sets p.nbfaux and returns
    (S, T, V, H, I) with I the device-resident System."""
    if not p.RI:
        raise NotImplementedError("the four-index (p.RI = false) path is out of scope; "
                                  "dn_int4c2e gives its integrals")
    h = bset.handle()
    Z = np.array([a.Z for a in bset.atoms], dtype=np.float64)
    S, T, V = h.int1e(Z, bset.centers_bohr())
    S = (S + S.T) / 2
    T = (T + T.T) / 2
    V = (V + V.T) / 2
    H = T + V
    if aux is None:
        aux = bset.fitting_set()
    if log:
        log("Basis Set                                            = %s" % bset.name)
        log("Auxiliary Basis Set                                  = %s" % aux.name)
        log("Gaussian Type                                        = %s"
            % ("Spherical" if p.spherical else "Cartesian"))
    ha = aux.handle()
    p.nbfaux = ha.nbf
    if max(s.l for s in bset.basis) <= 2 and max(s.l for s in aux.basis) <= 4:
        # (mn|P) on the device straight into the metric solve; (P|Q) on the host
        I = api.System(p, ctx, precision).ri_from_basis(h, ha)
    else:
        G = ha.int2c2e()
        G = (G + G.T) / 2
        I = api.Iaux(G, h.int3c2e(ha), p, ctx=ctx, precision=precision)
    return S, T, V, H, I


def integrals_for_gradients(bset, aux=None):
    """(G, I3c) of compute_integrals, for the RI nuclear gradients (grads.jl:1-100)."""
    if aux is None:
        aux = bset.fitting_set()
    ha = aux.handle()
    G = ha.int2c2e()
    return (G + G.T) / 2, bset.handle().int3c2e(ha)


def compute_E_nuc(bset, p):
    """utils.jl:299-312 with the reference's Angstrom-to-bohr factor 1.88973."""
    E = 0.0
    for i in range(bset.natoms):
        for j in range(i + 1, bset.natoms):
            ai, aj = bset.atoms[i], bset.atoms[j]
            E += ai.Z * aj.Z / (np.linalg.norm(ai.xyz - aj.xyz) * BOHR_PER_ANGSTROM_E_NUC)
    return E


def orthonormalize(C, S, p):
    """utils.jl:466-497: C in the S-orthonormal basis S^{-1/2} (eigenvalues below zero
    dropped), Gram-Schmidt there, and back."""
    evals, evecs = np.linalg.eigh(S)
    inv = np.where(evals < 0.0, 0.0, 1.0 / np.sqrt(np.where(evals > 0.0, evals, 1.0)))
    S_12 = evecs * inv
    Cnew2 = S_12.T @ (S @ C)
    for i in range(Cnew2.shape[1]):
        Cnew2[:p.nbf, i] /= np.linalg.norm(Cnew2[:p.nbf, i])
        for j in range(i + 1, Cnew2.shape[0]):
            val = -np.dot(Cnew2[:p.nbf, i], Cnew2[:p.nbf, j])
            Cnew2[:p.nbf, j] += val * Cnew2[:p.nbf, i]
    return np.asfortranarray(S_12 @ Cnew2)


def check_ortho(C, S, p, log=print):
    """utils.jl:431-464: orthonormalise C if max|C^T S C - 1| > 1e-6, then give every
    column the sign of its largest-magnitude coefficient (the first such, as the
    reference's strict comparison keeps)."""
    C = np.array(C, dtype=np.float64, order="F", copy=True)
    dev = np.abs(C.T @ S @ C - np.eye(C.shape[1]))
    if dev.max() > 1e-6:
        if log:
            log("Orthonormality violations %i, Maximum Violation %f"
                % (int(np.sum(dev > 1e-6)), dev.max()))
            log("Trying to orthonormalize")
        C = check_ortho(orthonormalize(C, S, p), S, p, log=log)
    elif log:
        log("No violations of the orthonormality")
    for j in range(p.nbf):
        idx = int(np.argmax(np.abs(C[:p.nbf, j])))
        C[:p.nbf, j] *= np.sign(C[idx, j])
    return C


def core_guess(H, S):
    """eigen(H, S) of energy.jl:77: the generalised eigenvectors, ascending eigenvalues,
    each S-normalised."""
    evals, evecs = np.linalg.eigh(S)
    X = evecs / np.sqrt(evals)
    e, U = np.linalg.eigh(X.T @ H @ X)
    return e, np.asfortranarray(X @ U)
