"""CPU oracle: FP64 NumPy restatement of DoNOF.jl's orbital-optimization hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product imports this module; only the
tests, the smoke check and the CPU-baseline leg of the benchmark use it, and only
as the checker.  Every function is restated from the mathematics of the cited
reference lines (DoNOF.jl v1.1.2, GPL-3.0; no source is copied).  Indices in the
reference are 1-based and column-major; here arrays keep the same axis order
(b[m, n, Q], C[m, p], Dcj12r[i, j, k]) and indices are 0-based.

How each part is pinned (tests/test_oracle_*.py):
  iaux              G^{-1/2} G G^{-1/2} = I; the Cholesky route gives the same
                    b b^T (auxiliary index contracted) and the same J/K.
  jkh_mo            brute-force loops at tiny sizes, symmetry/PSD invariants,
                    invariance under b -> b.Q (orthogonal Q), diag J == diag K.
  compute_lagrange2 brute-force loops; energy identity computeE_elec == calce;
                    central finite differences of E(rotate(y)) vs calcorbg.
  rotate_orbital    scipy.linalg.expm (Al-Mohy & Higham 2009) vs an independent
                    eigen-route, U^T U = I, y = 0 -> C.
  pnof_coefficients HF limit (integer n); one electron pair (ndoc = 1) reduces every
                    functional to the Loewdin-Shull two-electron energy; the
                    cumulant sum rule sum_q Delta_qp = n_p (1 - n_p); GNOF with the
                    dynamic factor quenched equals PNOF7; the dynamic factor of a pair
                    depends on that pair only; GNOF's dynamic terms are -(c c^T +
                    n_d n_d^T) on the strong-weak and weak-weak interpair blocks with c
                    the Loewdin-Shull-signed dynamic amplitudes (one sign vector for
                    both blocks) and lower the energy; symmetry.  der_* checked by
                    finite differences of cj12/ck12 in gamma through the occupations;
                    PNOF4's by finite differences of the factors of Pi combined with
                    the denominator pnof.jl:255 writes.  Parity unpinned: the block
                    pattern of GNOF's ista = 0 static term (pnof.jl:540-578) is pinned
                    only by this transcription and by finite-difference consistency.
  ocupacion_*       pair sums == 1 (softmax), finite-difference Jacobians.
  calcoccg          finite differences of calcocce.
  orbopt_*          deterministic recurrence; replayed step by step.
  geom_gradient_*   the RI nuclear gradient equals the derivative of the energy
                    (iaux, jkh_mo, calce) along S, H, I3c and G varied linearly with
                    Loewdin-orthonormalised orbitals, each term alone and all together,
                    closed and open shells, PNOF5/7 and GNOF; the coefficient diagonals
                    are ignored (grads.jl:24-25).  grad_nuc: finite differences of e_nuc
                    and translation invariance.
  occoptr           parity unpinned for the iterates (the reference uses Optim.jl
                    ConjugateGradient/HagerZhang, not available here); only the
                    stationary point |calcoccg| < threshen is pinned.
"""
from __future__ import annotations

import math
from types import SimpleNamespace

import numpy as np
import scipy.linalg
import scipy.optimize

# --------------------------------------------------------------------------------------
# Sizes (param.jl:59-222 constructor, param.jl:242-266 set_ncwo)
# --------------------------------------------------------------------------------------


def make_param(nbf, nbfaux, ne, mul=1, ncwo=-1, no1=0, ipnof=8, ista=0,
               occ_method="Softmax", orb_method="ADAM"):
    """Size record with the reference's defaults (param.jl:59-165) and set_ncwo
    (param.jl:242-266).  Returns a SimpleNamespace carrying the reference field names."""
    nalpha = (ne + mul - 1) // 2
    nbeta = (ne - mul + 1) // 2
    p = SimpleNamespace()
    p.nbf, p.nbfaux, p.ne, p.mul = nbf, nbfaux, ne, mul
    p.nalpha, p.nbeta, p.no1 = nalpha, nbeta, no1
    p.ndoc = nbeta - no1
    p.nsoc = nalpha - nbeta
    p.ndns = p.ndoc + p.nsoc
    p.nvir = nbf - nalpha
    p.maxit, p.maxloop, p.ipnof, p.ista = 1000, 10, ipnof, ista
    p.threshen, p.threshgorb, p.threshgocc = 1e-4, 4e-3, 1e-3
    p.alpha = 0.01
    p.RI, p.HighSpin, p.MSpin = True, False, 0
    p.occ_method, p.orb_method = occ_method, orb_method
    p.h_cut = 0.02 * math.sqrt(2.0)
    set_ncwo(p, ncwo)
    return p


def set_ncwo(p, ncwo):
    """param.jl:242-266."""
    if p.ndns != 0:
        if p.ndoc > 0:
            if ncwo != 1:
                if ncwo == -1 or ncwo > p.nvir / p.ndoc:
                    ncwo = int(p.nvir / p.ndoc)
        else:
            ncwo = 0
    p.ncwo = ncwo
    p.nac = p.ndoc * (1 + ncwo)
    p.nbf5 = p.no1 + p.nac + p.nsoc
    p.no0 = p.nbf - p.nbf5
    p.nv = p.ncwo * p.ndoc
    p.nvar = round(p.nbf * (p.nbf - 1) / 2 - p.no0 * (p.no0 - 1) / 2)
    p.noptorb = p.nbf


def pair_weak(p, l):
    """0-based indices of the weakly occupied orbitals coupled to strong orbital l
    (l is 1-based as in the reference): ll:ndoc:ul with ll = no1+ndns+(ndoc-l)+1
    (pnof.jl:509-510)."""
    ll0 = p.no1 + p.ndns + (p.ndoc - l)
    return ll0 + p.ndoc * np.arange(p.ncwo)


def pair_gamma(p, l):
    """0-based gamma indices of pair l for Softmax: llg = (ndoc-l)+1, stride ndoc
    (pnof.jl:1142-1143)."""
    return (p.ndoc - l) + p.ndoc * np.arange(p.ncwo)


# --------------------------------------------------------------------------------------
# RI metric (integrals.jl:68-91)
# --------------------------------------------------------------------------------------


def iaux(G, I3c):
    """b[m,n,:] = I3c[m,n,:] G^{-1/2} by the eigen route of integrals.jl:68-91: eigen of
    the symmetric G, 1/sqrt(lambda) for lambda >= 0 and 0 for lambda < 0 (74-83)."""
    lam, V = np.linalg.eigh(0.5 * (G + G.T))
    s = np.where(lam < 0.0, 0.0, 1.0 / np.sqrt(np.where(lam < 0.0, 1.0, lam)))
    Gms = (V * s) @ V.T
    nbf, _, naux = I3c.shape
    return (I3c.reshape(nbf * nbf, naux, order="F") @ Gms).reshape(nbf, nbf, naux, order="F")


def iaux_cholesky(G, I3c):
    """b = I3c L^{-T} with G = L L^T: the reference's b rotated by the orthogonal
    L^T G^{-1/2} in the auxiliary index (same sum_Q b b)."""
    L = np.linalg.cholesky(0.5 * (G + G.T))
    nbf, _, naux = I3c.shape
    X = I3c.reshape(nbf * nbf, naux, order="F")
    Y = scipy.linalg.solve_triangular(L, X.T, lower=True).T
    return Y.reshape(nbf, nbf, naux, order="F")


# --------------------------------------------------------------------------------------
# RI transform, J, K, H_core (integrals.jl:144-172)
# --------------------------------------------------------------------------------------


def mo_transform(C5, b):
    """b_pnl[p,n,l] = sum_m C5[m,p] b[m,n,l]  (integrals.jl:159)
    b_pql[p,q,l] = sum_n C5[n,q] b_pnl[p,n,l] (integrals.jl:160)."""
    nbf, _, naux = b.shape
    nbf5 = C5.shape[1]
    b_pn = (C5.T @ b.reshape(nbf, nbf * naux, order="F")).reshape(nbf5, nbf, naux, order="F")
    b_pq = np.einsum("pnl,nq->pql", b_pn, C5, optimize=True)
    return b_pn, b_pq


def jkh_mo(C, H, b, nbf5):
    """J_MO, K_MO, H_core for the RI case (integrals.jl:144-172, computeJKH_MO 131-140).

    J[p,q] = sum_l b_ppl b_qql (163); K[p,q] = sum_l b_pql^2 (166);
    H_core[i] = sum_mn C5[m,i] C5[n,i] H[m,n] (156, 169)."""
    C5 = C[:, :nbf5]
    _, b_pq = mo_transform(C5, b)
    X = np.einsum("ppl->pl", b_pq)
    J = X @ X.T
    K = np.einsum("pql,pql->pq", b_pq, b_pq)
    H_core = np.einsum("mi,mn,ni->i", C5, H, C5)
    return J, K, H_core


# --------------------------------------------------------------------------------------
# Lagrangian and energies (utils.jl:1-59, 117-136, 191-236)
# --------------------------------------------------------------------------------------


def compute_lagrange2(C, n, H, b, cj12, ck12, nalpha, nbeta):
    """Orbital Lagrangian elag (nbf x nbf, columns >= nbf5 zero) and H_mat = C^T H C5
    (utils.jl:1-59)."""
    nbf = C.shape[0]
    nbf5 = n.shape[0]
    naux = b.shape[2]
    C5 = C[:, :nbf5]
    H_mat = C.T @ (H @ C5)                                            # 17-18
    b_pn, _ = mo_transform(C5, b)   # b_pn[q,m,l] = tmp[m,q,l] since b is symmetric in (m,n)
    # b_MO[p,q,l] = sum_m C[m,p] tmp[m,q,l]                            # 19-20
    tmp = np.transpose(b_pn, (1, 0, 2))                                 # [m,q,l]
    b_MO = (C.T @ tmp.reshape(nbf, nbf5 * naux, order="F")).reshape(nbf, nbf5, naux, order="F")
    elag = np.zeros((nbf, nbf))
    elag[:, :nbf5] += n[None, :] * H_mat[:, :nbf5]                     # 39
    if nbeta > 0:                                                       # 42-47
        diag = np.einsum("bbk->bk", b_MO[:nbf5])                        # b_MO[b,b,k]
        for cols in (np.arange(0, nbeta), np.arange(nalpha, nbf5)):
            if cols.size:
                elag[:, cols] += n[cols][None, :] * np.einsum(
                    "abk,bk->ab", b_MO[:, cols, :], diag[cols])
    X = np.einsum("qqk->qk", b_MO[:nbf5])                               # 50
    tmp2 = cj12 @ X
    elag[:, :nbf5] += np.einsum("abk,bk->ab", b_MO, tmp2)               # 51
    # -sum_{q,k} ck12[b,q] b_MO[a,q,k] b_MO[b,q,k]                      # 54-55
    W = ck12[:, :, None] * b_MO[:nbf5]                                  # [b,q,k]
    elag[:, :nbf5] -= b_MO.reshape(nbf, -1, order="F") @ W.reshape(nbf5, -1, order="F").T
    return elag, H_mat


def computeE_elec_hmat(Hmat, n, elag, nbf5):
    """E = sum_{i<nbf5} elag_ii + sum_i n_i Hmat_ii (utils.jl:216-227)."""
    return float(np.trace(elag[:nbf5, :nbf5]) + np.dot(n[:nbf5], np.diag(Hmat)[:nbf5]))


def computeE_elec_hc(H, C, n, elag, p):
    """The (H, C) form used by ENERGY1r (utils.jl:191-214)."""
    E = float(np.trace(elag[:p.nbf5, :p.nbf5]))
    hdiag = np.einsum("mi,mn,ni->i", C[:, :p.nbf5], H, C[:, :p.nbf5])
    E += float(np.dot(n[:p.nbeta], hdiag[:p.nbeta]))
    if not p.HighSpin:
        E += float(np.dot(n[p.nbeta:p.nalpha], hdiag[p.nbeta:p.nalpha]))
    else:
        E += float(0.5 * np.sum(hdiag[p.nbeta:p.nalpha]))
    E += float(np.dot(n[p.nalpha:p.nbf5], hdiag[p.nalpha:p.nbf5]))
    return E


def lagrange_convergency(elag):
    """sumdiff, maxdiff of |elag - elag^T| (utils.jl:229-236)."""
    d = np.abs(elag - elag.T)
    return float(d.sum()), float(d.max())


def pack_grad(elag, nbf5):
    """grads[n] = 4 elag[i,j] - 4 elag[j,i] for i < nbf5, j > i, row-wise
    (pnof.jl:1395-1403; minimization.jl:365-373)."""
    nbf = elag.shape[0]
    g = 4.0 * elag - 4.0 * elag.T
    iu = triu_pairs(nbf, nbf5)
    return g[iu]


def triu_pairs(nbf, nbf5):
    """(rows, cols) of the packed rotation variables in reference order."""
    rows, cols = [], []
    for i in range(nbf5):
        rows.extend([i] * (nbf - i - 1))
        cols.extend(range(i + 1, nbf))
    return np.asarray(rows, dtype=np.int64), np.asarray(cols, dtype=np.int64)


def rotation_generator(y, nbf, nbf5):
    """Antisymmetric Y with Y[i,j] = y_n, Y[j,i] = -y_n (utils.jl:501-510)."""
    Y = np.zeros((nbf, nbf))
    r, c = triu_pairs(nbf, nbf5)
    Y[r, c] = y
    Y[c, r] = -y
    return Y


def rotate_orbital(y, C, nbf5):
    """Cnew = C exp(Y) (utils.jl:499-521)."""
    nbf = C.shape[0]
    U = scipy.linalg.expm(rotation_generator(y, nbf, nbf5))
    return C @ U


def calcorbg(y, n, cj12, ck12, C, H, b, p):
    """Packed orbital gradient at Cnew = rotate(y, C) (pnof.jl:1389-1407)."""
    Cnew = rotate_orbital(y, C, p.nbf5)
    elag, _ = compute_lagrange2(Cnew, n, H, b, cj12, ck12, p.nalpha, p.nbeta)
    return pack_grad(elag, p.nbf5)


# --------------------------------------------------------------------------------------
# Occupations (pnof.jl:1024-1163) and gamma helpers (guess.jl:184-213, utils.jl:539-584)
# --------------------------------------------------------------------------------------


def ocupacion(gamma, p):
    """pnof.jl:1024-1045 (EBI is out of scope)."""
    if p.occ_method == "Softmax":
        return ocupacion_softmax(gamma, p)
    if p.occ_method == "Trigonometric":
        return ocupacion_trigonometric(gamma, p)
    raise ValueError("occ_method %r not supported" % p.occ_method)


def ocupacion_softmax(gamma, p):
    """Softmax occupations and Jacobian (pnof.jl:1122-1163).  Keeps the reference's
    indexing of line 1152 (n[i] rather than n[no1+i])."""
    n = np.zeros(p.nbf5)
    dn = np.zeros((p.nbf5, p.nv))
    n[:p.no1] = 1.0
    n[p.no1 + p.ndoc:p.no1 + p.ndns] = 1.0 if p.HighSpin else 0.5
    eg = np.exp(gamma)
    for i in range(1, p.ndoc + 1):
        w = pair_weak(p, i)
        g = pair_gamma(p, i)
        den = 1.0 + eg[g].sum()
        n[p.no1 + i - 1] = 1.0 / den
        n[w] = eg[g] / den
        dn[p.no1 + i - 1, g] = -n[i - 1] * n[w]
        blk = -np.outer(n[w], n[w]) + np.diag(n[w])
        dn[np.ix_(w, g)] = blk
    return n, dn


def ocupacion_trigonometric(gamma, p):
    """Trigonometric occupations and Jacobian (pnof.jl:1047-1120)."""
    no1, ndoc, ncwo = p.no1, p.ndoc, p.ncwo
    n = np.zeros(p.nbf5)
    dni = np.zeros(p.nbf5)
    n[:no1] = 1.0
    n[no1:no1 + ndoc] = 0.5 * (1.0 + np.cos(gamma[:ndoc]) ** 2)
    dni[no1:no1 + ndoc] = -0.5 * np.sin(2.0 * gamma[:ndoc])
    n[no1 + ndoc:no1 + p.ndns] = 1.0 if p.HighSpin else 0.5
    dn = np.zeros((p.nbf5, p.nv))
    h = 1.0 - n
    for i in range(1, ndoc + 1):
        w = pair_weak(p, i)
        gi = ndoc + (ndoc - i) + ndoc * np.arange(ncwo - 1)   # ll_gamma:ndoc:ul_gamma
        gp = gamma[gi]
        npi = np.full(ncwo, h[no1 + i - 1])
        for kw in range(ncwo - 1):
            npi[kw] *= math.sin(gp[kw]) ** 2
            npi[kw + 1:] *= math.cos(gp[kw]) ** 2
        n[w] = npi
        ig = no1 + i - 1
        dn[ig, i - 1] = dni[ig]
        col = np.full(ncwo, -dni[ig])
        for kw in range(ncwo - 1):
            col[kw] *= math.sin(gp[kw]) ** 2
            col[kw + 1:] *= math.cos(gp[kw]) ** 2
        dn[w, i - 1] = col
        blk = np.zeros((ncwo, ncwo - 1))
        for jw in range(ncwo - 1):
            blk[jw + 1:, jw] = n[ig] - 1.0
            for kw in range(jw):
                blk[jw + 1:, jw] *= math.cos(gp[kw]) ** 2
            blk[jw + 1:, jw] *= math.sin(2.0 * gp[jw])
            for kw in range(jw + 1, ncwo - 1):
                blk[kw, jw] *= math.sin(gp[kw]) ** 2
                blk[kw + 1:, jw] *= math.cos(gp[kw]) ** 2
        for jw in range(ncwo - 1):
            blk[jw, jw] = 1.0 - n[ig]
        for kw in range(ncwo - 1):
            blk[kw, kw] *= math.sin(2.0 * gp[kw])
            for lw in range(kw + 1, ncwo - 1):
                blk[lw, lw] *= math.cos(gp[kw]) ** 2
        dn[np.ix_(w, gi)] = blk
    return n, dn


def guess_gamma_softmax(ndoc, ncwo):
    """guess.jl:200-213."""
    g = np.zeros(ndoc * ncwo)
    g[:] = math.log(max(0.001 / ncwo, 1e-16)) - math.log(0.999)
    return g


def guess_gamma_trigonometric(ndoc, ncwo):
    """guess.jl:184-198."""
    g = np.zeros(ndoc * ncwo)
    for i in range(1, ndoc + 1):
        g[i - 1] = math.acos(math.sqrt(2.0 * 0.999 - 1.0))
        ll0 = ndoc + (ndoc - i)
        for j in range(1, ncwo):
            g[ll0 + ndoc * (j - 1)] = math.asin(math.sqrt(1.0 / (ncwo - j + 1)))
    return g


def n_to_gammas_softmax(n, p):
    """utils.jl:539-561 (uses n[i], not n[no1+i], as the reference does)."""
    g = np.zeros(p.nv)
    for i in range(1, p.ndoc + 1):
        w = pair_weak(p, i)
        gi = pair_gamma(p, i)
        g[gi] = np.log(np.maximum(n[w], 1e-16)) - math.log(n[i - 1])
    return g


# --------------------------------------------------------------------------------------
# PNOF coefficients (pnof.jl:1-117, 119-1022)
# --------------------------------------------------------------------------------------


def _intrapair(p, n, cj12, ck12):
    """Intra-pair overrides shared by PNOF5/7/8/9 (e.g. pnof.jl:580-597)."""
    for l in range(1, p.ndoc + 1):
        ldx = p.no1 + l - 1
        w = pair_weak(p, l)
        cj12[ldx, w] = 0.0
        cj12[w, ldx] = 0.0
        cj12[np.ix_(w, w)] = 0.0
        ck12[ldx, w] = np.sqrt(n[ldx] * n[w])
        ck12[w, ldx] = np.sqrt(n[ldx] * n[w])
        ck12[np.ix_(w, w)] = -np.sqrt(np.outer(n[w], n[w]))


def _nd(p, n, c=1.0):
    """Dynamic occupations n_d = n * c*exp(-((1-n_g)/h_cut)^2) per pair (pnof.jl:506-519)."""
    nd = np.zeros(n.shape[0])
    for i in range(1, p.ndoc + 1):
        idx = p.no1 + i - 1
        w = pair_weak(p, i)
        coc = (1.0 - n[idx]) / p.h_cut
        F = c * math.exp(-coc * coc)
        nd[idx] = n[idx] * F
        nd[w] = n[w] * F
    return nd


def cjck_pnof4(n, p):
    """pnof.jl:119-173."""
    nbf5, ndoc = n.shape[0], p.ndoc
    h = 1.0 - n
    S_F = h[:ndoc].sum()
    R = np.outer(h[:ndoc], n[ndoc:]) / S_F
    D = np.zeros((nbf5, nbf5))
    D[:ndoc, :ndoc] = np.outer(h[:ndoc], h[:ndoc])
    D[:ndoc, ndoc:] = (1 - S_F) / S_F * np.outer(h[:ndoc], n[ndoc:])
    D[ndoc:, :ndoc] = (1 - S_F) / S_F * np.outer(n[ndoc:], h[:ndoc])
    D[ndoc:, ndoc:] = np.outer(n[ndoc:], n[ndoc:])
    Pi = np.zeros((nbf5, nbf5))
    Pi[:ndoc, :ndoc] = -np.sqrt(np.outer(h[:ndoc], h[:ndoc]))
    Pi[:ndoc, ndoc:] = -np.sqrt(np.outer(h[:ndoc], n[ndoc:]) / S_F) * np.sqrt(
        (n[:ndoc, None] - n[None, ndoc:]) + R)
    Pi[ndoc:, :ndoc] = -np.sqrt(np.outer(n[ndoc:], h[:ndoc]) / S_F) * np.sqrt(
        -(n[ndoc:, None] - n[None, :ndoc]) + R.T)
    Pi[ndoc:, ndoc:] = np.sqrt(np.outer(n[ndoc:], n[ndoc:]))
    nn = np.outer(n, n)
    return 2.0 * (nn - D), nn - D - Pi


def cjck_pnof5(n, p):
    """pnof.jl:274-310."""
    cj12 = 2.0 * np.outer(n, n)
    ck12 = np.outer(n, n)
    if p.MSpin == 0 and p.nsoc > 1:
        s = slice(p.nbeta, p.nalpha)
        ck12[s, s] = 2.0 * np.outer(n[s], n[s])
    _intrapair(p, n, cj12, ck12)
    return cj12, ck12


def _fi_pnof7(n, ista):
    fi = n * (1.0 - n)
    fi[fi <= 0] = 0.0
    return np.sqrt(fi) if ista == 0 else 2.0 * n * (1.0 - n)


def cjck_pnof7(n, p):
    """pnof.jl:356-411."""
    fi = _fi_pnof7(n, p.ista)
    cj12 = 2.0 * np.outer(n, n)
    ck12 = np.outer(n, n) + np.outer(fi, fi)
    if p.nsoc > 1:
        s = slice(p.nbeta, p.nalpha)
        ck12[s, s] = 2.0 * np.outer(n[s], n[s])
    _intrapair(p, n, cj12, ck12)
    return cj12, ck12


def _fi_gnof(n, ista):
    """fi variants of pnof.jl:522-538."""
    fi = n * (1.0 - n)
    fi[fi <= 0] = 0.0
    if ista in (0, 1, 5):
        return np.sqrt(fi)
    if ista == 2:
        return 0.9 * np.sqrt(fi)
    if ista == 3:
        alpha = 0.55
        aa = (2.0 ** (2.0 * alpha - 1)) ** (1 / alpha)
        return (aa * fi) ** alpha
    if ista == 4:
        return np.sqrt(fi) * np.exp(-(n - 0.5) ** 2)
    raise ValueError("ista")


def _blocks(p):
    beta = slice(p.no1, p.nbeta)
    soc = slice(p.nbeta, p.nalpha)
    cwo = slice(p.nalpha, p.nbf5)
    return beta, soc, cwo


def cjck_gnof(n, p):
    """GNOF, ipnof = 8 (pnof.jl:488-601)."""
    c = 0.8 if p.ista == 5 else 1.0
    nd = _nd(p, n, c)
    nd12 = np.sqrt(nd)
    fi = _fi_gnof(n, p.ista)
    beta, soc, cwo = _blocks(p)
    cj12 = 2.0 * np.outer(n, n)
    if p.ista == 0:
        ck12 = np.outer(n, n)
        ck12[beta, cwo] += np.outer(fi[beta], fi[cwo])
        ck12[cwo, beta] += np.outer(fi[cwo], fi[beta])
        ck12[cwo, cwo] += np.outer(fi[cwo], fi[cwo])
        if p.nsoc > 0:
            ck12[beta, soc] += (0.25 * fi[beta])[:, None]
            ck12[soc, beta] += (0.25 * fi[beta])[None, :]
            ck12[soc, cwo] += (0.5 * fi[cwo])[None, :]
            ck12[cwo, soc] += (0.5 * fi[cwo])[:, None]
        if p.nsoc > 1:
            ck12[soc, soc] = 0.5
    else:
        ck12 = np.outer(n, n) + np.outer(fi, fi)
    ck12[beta, cwo] += np.outer(nd12[beta], nd12[cwo]) - np.outer(nd[beta], nd[cwo])
    ck12[cwo, beta] += np.outer(nd12[cwo], nd12[beta]) - np.outer(nd[cwo], nd[beta])
    ck12[cwo, cwo] += -np.outer(nd12[cwo], nd12[cwo]) - np.outer(nd[cwo], nd[cwo])
    _intrapair(p, n, cj12, ck12)
    return cj12, ck12


def cjck_pnof9(n, p):
    """pnof.jl:795-879."""
    nd = _nd(p, n, 1.0)
    nd12 = np.sqrt(nd)
    fi = _fi_gnof(n, 0)
    beta, soc, cwo = _blocks(p)
    cj12 = 2.0 * np.outer(n, n)
    ck12 = np.outer(n, n) + np.outer(fi, fi)
    if p.nsoc > 1:
        ck12[soc, soc] = 0.5
    ck12[beta, cwo] += np.outer(nd12[beta], nd12[cwo]) - np.outer(nd[beta], nd[cwo])
    ck12[cwo, beta] += np.outer(nd12[cwo], nd12[beta]) - np.outer(nd[cwo], nd[beta])
    ck12[cwo, cwo] += -np.outer(nd12[cwo], nd12[cwo]) - np.outer(nd[cwo], nd[cwo])
    _intrapair(p, n, cj12, ck12)
    return cj12, ck12


def pnof_selector(n, p):
    """PNOFi_selector (pnof.jl:1-53): coefficients, diagonals zeroed (49-50)."""
    f = {4: cjck_pnof4, 5: cjck_pnof5, 7: cjck_pnof7, 8: cjck_gnof, 9: cjck_pnof9}[p.ipnof]
    cj12, ck12 = f(n, p)
    np.fill_diagonal(cj12, 0.0)
    np.fill_diagonal(ck12, 0.0)
    return cj12, ck12


# ---- derivatives ---------------------------------------------------------------------


def _der_intrapair(p, n, dn, Dcj, Dck):
    """Intra-pair derivative overrides shared by PNOF5/7/8/9 (e.g. pnof.jl:759-789)."""
    for l in range(1, p.ndoc + 1):
        ldx = p.no1 + l - 1
        w = pair_weak(p, l)
        Dcj[ldx, w, :] = 0.0
        Dcj[w, ldx, :] = 0.0
        Dcj[np.ix_(w, w, np.arange(p.nv))] = 0.0
        a = max(n[ldx], 1e-15)
        bw = np.maximum(n[w], 1e-15)
        Dck[ldx, w, :] = 0.5 / math.sqrt(a) * dn[ldx, None, :] * np.sqrt(n[w])[:, None]
        Dck[w, ldx, :] = 0.5 / np.sqrt(bw)[:, None] * dn[w, :] * math.sqrt(n[ldx])
        Dck[np.ix_(w, w, np.arange(p.nv))] = (
            -0.5 / np.sqrt(bw)[:, None, None] * dn[w][:, None, :] * np.sqrt(n[w])[None, :, None])


def der_pnof4(n, dn, p):
    """pnof.jl:175-272."""
    nbf5, ndoc, nv = n.shape[0], p.ndoc, p.nv
    h = 1.0 - n
    dh = -dn
    S_F = max(h[:ndoc].sum(), 1e-7)
    D_S_F = dh[:ndoc].sum(axis=0)
    T = (1 - S_F) / S_F
    D_T = -D_S_F / S_F ** 2
    R = np.outer(h[:ndoc], n[ndoc:]) / S_F
    D_R = np.zeros((ndoc, nbf5 - ndoc, nv))
    for k in range(nv):
        D_R[:, :, k] = ((np.outer(dh[:ndoc, k], n[ndoc:]) + np.outer(h[:ndoc], dn[ndoc:, k])) * S_F
                        - np.outer(h[:ndoc], n[ndoc:]) * D_S_F[k]) / S_F ** 2
    D_D = np.zeros((nbf5, nbf5, nv))
    D_Pi = np.zeros((nbf5, nbf5, nv))
    for k in range(nv):
        D_D[:ndoc, :ndoc, k] = np.outer(dh[:ndoc, k], h[:ndoc])
        D_D[:ndoc, ndoc:, k] = D_T[k] * np.outer(h[:ndoc], n[ndoc:]) + T * np.outer(dh[:ndoc, k], n[ndoc:])
        D_D[ndoc:, :ndoc, k] = T * np.outer(dn[ndoc:, k], h[:ndoc])
        D_D[ndoc:, ndoc:, k] = np.outer(dn[ndoc:, k], n[ndoc:])
        D_Pi[:ndoc, :ndoc, k] = np.outer(
            -(1.0 / np.maximum(2 * np.sqrt(h[:ndoc]), 1e-7) * dh[:ndoc, k]), np.sqrt(h[:ndoc]))
        D_Pi[:ndoc, ndoc:, k] = -(
            1.0 / np.maximum(2 * np.sqrt(R), 1e-7) * D_R[:, :, k]
            * np.sqrt((n[:ndoc, None] - n[None, ndoc:]) + R)
            + np.sqrt(R) * ((dn[:ndoc, k][:, None] - dn[ndoc:, k][None, :]) + D_R[:, :, k])
            / np.maximum(2 * np.sqrt(np.outer(n[:ndoc], n[ndoc:]) + R), 1e-7))
        D_Pi[ndoc:, ndoc:, k] = np.outer(
            1.0 / np.maximum(2 * np.sqrt(n[ndoc:]), 1e-7) * dn[ndoc:, k], np.sqrt(n[ndoc:]))
    Dcj = np.zeros((nbf5, nbf5, nv))
    Dck = np.zeros((nbf5, nbf5, nv))
    for k in range(nv):
        Dcj[:, :, k] = 2.0 * (np.outer(dn[:, k], n) - D_D[:, :, k])
        Dck[:, :, k] = np.outer(dn[:, k], n) - D_D[:, :, k] - D_Pi[:, :, k]
    return Dcj, Dck


def der_pnof5(n, dn, p):
    """pnof.jl:312-354."""
    Dcj = 2.0 * dn[:, None, :] * n[None, :, None]
    Dck = dn[:, None, :] * n[None, :, None]
    _der_intrapair(p, n, dn, Dcj, Dck)
    return Dcj, Dck


def der_pnof7(n, dn, p):
    """pnof.jl:413-486."""
    fi = _fi_pnof7(n, p.ista)
    dfi = np.zeros_like(dn)
    for i in range(p.no1, n.shape[0]):
        a = max(fi[i], 1e-15)
        if p.ista == 0:
            dfi[i] = 1.0 / (2 * a) * (1 - 2 * n[i]) * dn[i]
        else:
            dfi[i] = 2.0 * (1 - 2 * n[i]) * dn[i]
    Dcj = 2.0 * dn[:, None, :] * n[None, :, None]
    Dck = dn[:, None, :] * n[None, :, None] + dfi[:, None, :] * fi[None, :, None]
    _der_intrapair(p, n, dn, Dcj, Dck)
    return Dcj, Dck


def _der_nd(p, n, dn, c):
    """n_d and its gamma-derivatives (pnof.jl:631-654)."""
    nbf5 = n.shape[0]
    nd = np.zeros(nbf5)
    dnd = np.zeros((nbf5, p.nv))
    for i in range(1, p.ndoc + 1):
        idx = p.no1 + i - 1
        w = pair_weak(p, i)
        coc = (1.0 - n[idx]) / p.h_cut
        F = c * math.exp(-coc * coc)
        nd[idx] = n[idx] * F
        nd[w] = n[w] * F
        dnd[idx] = F * dn[idx] * (1 - n[idx] * (-2 * coc / p.h_cut))
        dnd[w] = F * dn[w] + F * (2 * coc / p.h_cut) * n[w][:, None] * dn[idx][None, :]
    nd12 = np.sqrt(nd)
    dnd12 = 0.5 * dnd / np.maximum(nd12, 1e-15)[:, None]
    return nd, nd12, dnd, dnd12


def _dfi_gnof(n, dn, p, ista):
    """fi and dfi/dgamma variants (pnof.jl:656-702)."""
    nbf5 = n.shape[0]
    fi = n * (1.0 - n)
    fi[fi <= 0] = 0.0
    dfi = np.zeros((nbf5, p.nv))
    if ista in (0, 1, 5):
        fi = np.sqrt(fi)
        for i in range(p.no1, nbf5):
            a = max(fi[i], 1e-15)
            dfi[i] = 1.0 / (2 * a) * (1 - 2 * n[i]) * dn[i]
    elif ista == 2:
        alpha = 0.9
        fi = alpha * np.sqrt(fi)
        for i in range(p.no1, nbf5):
            a = max(fi[i], 1e-15)
            dfi[i] = alpha ** 2 / (2 * a) * (1 - 2 * n[i]) * dn[i]
    elif ista == 3:
        alpha = 0.55
        aa = (2.0 ** (2.0 * alpha - 1)) ** (1 / alpha)
        fi = (aa * fi) ** alpha
        for i in range(p.no1, nbf5):
            a = max(aa * n[i] * (1 - n[i]), 1e-15)
            dfi[i] = alpha * a ** (alpha - 1) * aa * (1 - 2 * n[i]) * dn[i]
    elif ista == 4:
        fi = np.sqrt(fi)
        for i in range(p.no1, nbf5):
            a = max(fi[i], 1e-15)
            e = math.exp(-(n[i] - 0.5) ** 2)
            dfi[i] = 1.0 / (2 * a) * (1 - 2 * n[i]) * dn[i] * e
            dfi[i] += -fi[i] * 2 * (n[i] - 0.5) * dn[i] * e
        fi = fi * np.exp(-(n - 0.5) ** 2)
    return fi, dfi


def der_gnof(n, dn, p):
    """GNOF derivatives, ipnof = 8 (pnof.jl:603-793)."""
    c = 0.8 if p.ista == 5 else 1.0
    nd, nd12, dnd, dnd12 = _der_nd(p, n, dn, c)
    fi, dfi = _dfi_gnof(n, dn, p, p.ista)
    beta, soc, cwo = _blocks(p)
    Dcj = 2.0 * dn[:, None, :] * n[None, :, None]
    Dck = dn[:, None, :] * n[None, :, None]
    if p.ista == 0:
        Dck[beta, cwo] += dfi[beta][:, None, :] * fi[cwo][None, :, None]
        Dck[cwo, beta] += fi[beta][None, :, None] * dfi[cwo][:, None, :]
        Dck[cwo, cwo] += fi[cwo][None, :, None] * dfi[cwo][:, None, :]
        if p.nsoc > 0:
            Dck[beta, soc] += 0.25 * dfi[beta][:, None, :]
            Dck[cwo, soc] += 0.5 * dfi[cwo][:, None, :]
        if p.nsoc > 1:
            Dck[soc, soc] = 0.0
    else:
        Dck = dn[:, None, :] * n[None, :, None] + dfi[:, None, :] * fi[None, :, None]
        if p.nsoc > 1:
            Dck[soc, soc] = 0.0
    Dck[beta, cwo] += dnd12[beta][:, None, :] * nd12[cwo][None, :, None] - dnd[beta][:, None, :] * nd[cwo][None, :, None]
    Dck[cwo, beta] += nd12[beta][None, :, None] * dnd12[cwo][:, None, :] - nd[beta][None, :, None] * dnd[cwo][:, None, :]
    Dck[cwo, cwo] += -nd12[cwo][None, :, None] * dnd12[cwo][:, None, :] - nd[cwo][None, :, None] * dnd[cwo][:, None, :]
    _der_intrapair(p, n, dn, Dcj, Dck)
    return Dcj, Dck


def der_pnof9(n, dn, p):
    """pnof.jl:881-1022."""
    nd, nd12, dnd, dnd12 = _der_nd(p, n, dn, 1.0)
    fi, dfi = _dfi_gnof(n, dn, p, 0)
    beta, soc, cwo = _blocks(p)
    Dcj = 2.0 * dn[:, None, :] * n[None, :, None]
    Dck = dn[:, None, :] * n[None, :, None] + dfi[:, None, :] * fi[None, :, None]
    if p.nsoc > 1:
        Dck[soc, soc] = 0.0
    Dck[beta, cwo] += dnd12[beta][:, None, :] * nd12[cwo][None, :, None] - dnd[beta][:, None, :] * nd[cwo][None, :, None]
    Dck[cwo, beta] += nd12[beta][None, :, None] * dnd12[cwo][:, None, :] - nd[beta][None, :, None] * dnd[cwo][:, None, :]
    Dck[cwo, cwo] += -nd12[cwo][None, :, None] * dnd12[cwo][:, None, :] - nd[cwo][None, :, None] * dnd[cwo][:, None, :]
    _der_intrapair(p, n, dn, Dcj, Dck)
    return Dcj, Dck


def der_pnof_selector(n, dn, p):
    """der_PNOFi_selector (pnof.jl:55-117): diagonals zeroed (109-114)."""
    f = {4: der_pnof4, 5: der_pnof5, 7: der_pnof7, 8: der_gnof, 9: der_pnof9}[p.ipnof]
    Dcj, Dck = f(n, dn, p)
    d = np.arange(n.shape[0])
    Dcj[d, d, :] = 0.0
    Dck[d, d, :] = 0.0
    return Dcj, Dck


# --------------------------------------------------------------------------------------
# Occupation energy and gradient (pnof.jl:1273-1377)
# --------------------------------------------------------------------------------------


def calce(n, cj12, ck12, J, K, H_core, p):
    """pnof.jl:1273-1300."""
    b, a = p.nbeta, p.nalpha
    E = float(np.sum(n[:b] * (2 * H_core[:b] + np.diag(J)[:b])))
    if p.nsoc > 0:
        E += float(np.sum(n[b:a] * 2 * H_core[b:a]))
    E += float(np.sum(n[a:p.nbf5] * (2 * H_core[a:p.nbf5] + np.diag(J)[a:p.nbf5])))
    E += float(np.sum(cj12 * J.T))
    E -= float(np.sum(ck12 * K.T))
    return E


def calcocce(gamma, J, K, H_core, p):
    """pnof.jl:1302-1322."""
    n, _ = ocupacion(gamma, p)
    cj12, ck12 = pnof_selector(n, p)
    return calce(n, cj12, ck12, J, K, H_core, p)


def calcoccg(gamma, J, K, H_core, p):
    """pnof.jl:1324-1377."""
    n, dn = ocupacion(gamma, p)
    Dcj, Dck = der_pnof_selector(n, dn, p)
    rows = np.r_[np.arange(p.no1, p.nbeta), np.arange(p.nalpha, p.nbf5)]
    Jd = np.diag(J)
    grad = dn[rows].T @ (2 * H_core[rows] + Jd[rows])
    grad += 2 * np.einsum("ijk,ji->k", Dcj[rows], J[:, rows])
    grad -= 2 * np.einsum("ijk,ji->k", Dck[rows], K[:, rows])
    return grad


def occoptr(gamma, J, K, H_core, p):
    """Occupation optimisation with J/K/H_core frozen (minimization.jl:100-139).

    The reference minimises with Optim.jl ConjugateGradient (HagerZhang line search,
    g_abstol = p.threshen).  That library is absent here, so this restatement uses
    scipy's nonlinear CG to the same gradient tolerance: parity unpinned for the
    iterates; only the stationary point is pinned."""
    res = scipy.optimize.minimize(
        lambda g: calcocce(g, J, K, H_core, p), gamma,
        jac=lambda g: calcoccg(g, J, K, H_core, p), method="CG",
        options={"gtol": p.threshen, "maxiter": 10000})
    return float(res.fun), int(res.nit), res.x


# --------------------------------------------------------------------------------------
# ADAM-family orbital optimisers (minimization.jl:242-630)
# --------------------------------------------------------------------------------------


def orbopt(method, gamma, C, H, b, p, trace=None):
    """ADAM / ADABelief / YOGI / Demon orbital rotation loop (minimization.jl:242-630).
    Mutates p.alpha and p.maxloop as the reference does (426-431)."""
    n, _ = ocupacion(gamma, p)
    cj12, ck12 = pnof_selector(n, p)
    elag, Hmat = compute_lagrange2(C, n, H, b, cj12, ck12, p.nalpha, p.nbeta)
    E = computeE_elec_hmat(Hmat, n, elag, p.nbf5)
    alpha = p.alpha
    if method == "Demon":
        beta1, beta2 = 0.6, 0.999
    else:
        beta1, beta2 = 0.5, 0.9
    binit = beta1
    m = np.zeros(p.nvar)
    v = np.zeros(p.nvar)
    vmax = np.zeros(p.nvar)
    improved = success = False
    best_E, best_C = E, C
    nit = 0
    for i in range(1, p.maxloop + 1):
        nit += 1
        g = pack_grad(elag, p.nbf5)
        if 2 * np.max(np.abs(g)) < p.threshgorb and improved:
            success = True
            break
        if method == "Demon":
            rate = (i - 1.0) / (p.maxloop - 1.0)
            beta1 = binit * (1.0 - rate) / ((1.0 - binit) + binit * (1.0 - rate))
        m = beta1 * m + (1.0 - beta1) * g
        if method == "ADABelief":
            v = beta2 * v + (1.0 - beta2) * (g - m) ** 2
        elif method == "YOGI":
            v = v - (1.0 - beta2) * np.sign(v - g ** 2) * g ** 2
        else:
            v = beta2 * v + (1.0 - beta2) * g ** 2
        mhat = m / (1.0 - beta1 ** i)
        vhat = v / (1.0 - beta2 ** i)
        vmax = np.maximum(vmax, vhat)
        y = -alpha * mhat / np.sqrt(vmax + 1e-16)
        C = rotate_orbital(y, C, p.nbf5)
        elag, Hmat = compute_lagrange2(C, n, H, b, cj12, ck12, p.nalpha, p.nbeta)
        E = computeE_elec_hmat(Hmat, n, elag, p.nbf5)
        if trace is not None:
            trace.append((E, y.copy()))
        if E < best_E:
            best_C, best_E = C, E
            improved = True
    if not improved:
        p.alpha = 0.3 * p.alpha
        p.maxloop = p.maxloop + 10
    if p.maxloop < 100:
        return best_E, best_C, nit, success
    return E, C, nit, success


# --------------------------------------------------------------------------------------
# Outer driver (energy.jl:140-220) with the helpers it calls
# --------------------------------------------------------------------------------------


def order_occupations_softmax(old_C, old_gamma, p):
    """utils.jl:586-616: each pair's columns and gammas sorted by decreasing gamma
    (stable, as Julia's sortperm), gammas shifted so the leader has gamma 0."""
    C = np.array(old_C, copy=True)
    gamma = np.array(old_gamma, copy=True)
    if np.any(gamma > 0):
        for i in range(1, p.ndoc + 1):
            cols = np.r_[p.no1 + i - 1, pair_weak(p, i)]
            gi = pair_gamma(p, i)
            gt = np.r_[0.0, gamma[gi]]
            idx = np.argsort(gt, kind="stable")[::-1]
            gt = gt[idx] - gt.max()
            gamma[gi] = gt[1:]
            C[:, cols] = old_C[:, cols][:, idx]
    return C, gamma


def energy1r(C, n, H, b, cj12, ck12, p):
    """ENERGY1r (utils.jl:117-136), no1 = 0 branch: the Lagrangian, E by the (H, C) form
    of computeE_elec (utils.jl:191-214) and sumdiff/maxdiff of |elag - elag^T|."""
    elag, _ = compute_lagrange2(C, n, H, b, cj12, ck12, p.nalpha, p.nbeta)
    E = computeE_elec_hc(H, C, n, elag, p)
    sumdiff, maxdiff = lagrange_convergency(elag)
    return E, elag, sumdiff, maxdiff


def order_subspaces(old_C, old_n, elag, H, b, p):
    """guess.jl:58-106: the strong orbitals sorted by increasing elag_ii (stable, as
    Julia's sortperm), each carrying its weak orbitals (pair columns ll:ndoc:ul) and
    occupations along; the singly occupied block sorted the same way; frozen-core columns
    kept (the reference sorts them and then copies them back); columns past nbf5 kept.
    Then ENERGY1r at the sorted (C, n) gives the returned elag."""
    C = np.zeros_like(old_C)
    n = np.zeros_like(old_n)
    d = np.diag(elag)
    C[:, :p.no1] = old_C[:, :p.no1]
    n[:p.no1] = old_n[:p.no1]
    order = np.argsort(d[p.no1:p.nbeta], kind="stable")
    for i, i_old in enumerate(order, start=1):
        i_old = int(i_old) + 1
        C[:, p.no1 + i - 1] = old_C[:, p.no1 + i_old - 1]
        n[p.no1 + i - 1] = old_n[p.no1 + i_old - 1]
        C[:, pair_weak(p, i)] = old_C[:, pair_weak(p, i_old)]
        n[pair_weak(p, i)] = old_n[pair_weak(p, i_old)]
    s0 = p.no1 + p.ndoc
    order = np.argsort(d[s0:p.no1 + p.ndns], kind="stable")
    for i, i_old in enumerate(order):
        C[:, s0 + i] = old_C[:, s0 + int(i_old)]
        n[s0 + i] = old_n[s0 + int(i_old)]
    C[:, p.nbf5:] = old_C[:, p.nbf5:]
    cj12, ck12 = pnof_selector(n, p)
    _, elag_new, _, _ = energy1r(C, n, H, b, cj12, ck12, p)
    return C, n, elag_new


def energy_loop(b, H, C, p, n=None, freeze_occ=False):
    """energy.jl:140-224 for given integrals (no1 = 0).  Returns (E, C, n, hist, elag);
    hist rows are (E, max|grad_orb|, ||grad_occ||, ENERGY1r E, maxdiff)."""
    if n is None:
        gamma = (guess_gamma_softmax(p.ndoc, p.ncwo) if p.occ_method == "Softmax"
                 else guess_gamma_trigonometric(p.ndoc, p.ncwo))
    else:
        gamma = n_to_gammas_softmax(n, p)
    E = 9999.0
    hist = []
    elag = None
    for i_ext in range(1, p.maxit + 1):
        E_orb, C, nit_orb, _ = orbopt(p.orb_method, gamma, C, H, b, p)
        if p.ndoc > 0 and not freeze_occ:
            J, K, Hc = jkh_mo(C, H, b, p.nbf5)
            _, nit_occ, gamma = occoptr(gamma, J, K, Hc, p)
            if nit_occ < 1:
                p.maxloop += 10
        n, _ = ocupacion(gamma, p)
        cj12, ck12 = pnof_selector(n, p)
        if p.occ_method == "Softmax":
            C, gamma = order_occupations_softmax(C, gamma, p)
        E = E_orb
        # ENERGY1r with the occupations of occoptr and the re-ordered orbitals
        # (energy.jl:157-172)
        Etmp, elag, _, maxdiff = energy1r(C, n, H, b, cj12, ck12, p)
        n, _ = ocupacion(gamma, p)
        cj12, ck12 = pnof_selector(n, p)
        g_orb = calcorbg(np.zeros(p.nvar), n, cj12, ck12, C, H, b, p)
        J, K, Hc = jkh_mo(C, H, b, p.nbf5)
        g_occ = calcoccg(gamma, J, K, Hc, p)
        hist.append((E, float(np.max(np.abs(g_orb))), float(np.linalg.norm(g_occ)), Etmp,
                     maxdiff))
        if np.max(np.abs(g_orb)) < p.threshgorb and np.linalg.norm(g_occ) < p.threshgocc:
            break
    if p.ipnof > 4:                                             # energy.jl:222-224
        C, n, elag = order_subspaces(C, n, elag, H, b, p)
    return E, C, n, hist, elag


# --------------------------------------------------------------------------------------
# Nuclear gradients, RI branch (grads.jl:1-100; utils.jl:299-336)
# --------------------------------------------------------------------------------------

BOHR_PER_ANGSTROM = 1.88973      # the factor utils.jl:305,320 uses


def e_nuc(Z, xyz):
    """compute_E_nuc (utils.jl:299-313): sum_{i<j} Z_i Z_j / |r_i - r_j|, xyz in Angstrom."""
    E = 0.0
    for i in range(len(Z)):
        for j in range(i + 1, len(Z)):
            E += Z[i] * Z[j] / (np.linalg.norm(xyz[i] - xyz[j]) * BOHR_PER_ANGSTROM)
    return E


def grad_nuc(Z, xyz):
    """compute_grad_nuc (utils.jl:315-336): dE_nuc/dR_k in Hartree/Bohr, natoms x 3."""
    na = len(Z)
    g = np.zeros((na, 3))
    for k in range(na):
        for j in range(na):
            if j != k:
                r = (xyz[k] - xyz[j]) * BOHR_PER_ANGSTROM
                g[k] -= Z[k] * Z[j] * r / np.linalg.norm(r) ** 3
    return g


def geom_gradient_densities(n, C, cj12, ck12, elag, p, G, I3c):
    """The densities of compute_geom_gradients' RI branch (grads.jl:1-100):
    RDM1[m,l] = 2 sum_p n_p C5[m,p] C5[l,p] (5); lag = 2 C elag C^T (6);
    cj12, ck12 with zero diagonals (24-25; the reference zeroes the caller's arrays in
    place, this copy does not); G symmetrised (55) and IG = I3c inv(G) (61);
    I2 = C5^T IG C5 per Q (66); d[l,s,Q] = sum_pq C5[l,p] C5[s,q] D[p,q,Q] with
    D = -ck12 I2 (68-70) plus, on the diagonal, bj = cj12^T diag(I2) (72-74) and
    n_p diag(I2) for p < nbeta (76-79) and for nalpha <= p < nbf5 (81-84);
    m[Q,R] = sum_{l,s} d[l,s,Q] IG[s,l,R] (89).  Returns (RDM1, lag, d, m)."""
    nbf5 = p.nbf5
    C5 = C[:, :nbf5]
    RDM1 = 2.0 * (C5 * n[None, :nbf5]) @ C5.T
    lag = 2.0 * C @ elag @ C.T
    cj = np.array(cj12, dtype=float)
    ck = np.array(ck12, dtype=float)
    np.fill_diagonal(cj, 0.0)
    np.fill_diagonal(ck, 0.0)
    Gs = 0.5 * (G + G.T)
    nbf, _, naux = I3c.shape
    IG = (I3c.reshape(nbf * nbf, naux, order="F") @ np.linalg.inv(Gs)).reshape(
        nbf, nbf, naux, order="F")
    _, I2 = mo_transform(C5, IG)                                         # [p,q,Q]
    D = -ck[:, :, None] * I2
    X = np.einsum("ppQ->pQ", I2)
    diag = cj.T @ X                                                      # bj[q,Q]
    diag[:p.nbeta] += n[:p.nbeta, None] * X[:p.nbeta]
    diag[p.nalpha:nbf5] += n[p.nalpha:nbf5, None] * X[p.nalpha:nbf5]
    idx = np.arange(nbf5)
    D[idx, idx, :] += diag
    # d_Q = C5 D_Q C5^T
    T = np.einsum("lp,pqQ->lqQ", C5, D, optimize=True)
    d = np.einsum("lqQ,sq->lsQ", T, C5, optimize=True)
    m = np.einsum("lsQ,slR->QR", d, IG, optimize=True)
    return RDM1, lag, d, m


def geom_gradient_atom(RDM1, lag, d, m, dS, dT, dV, geri3c, geri2c):
    """One atom's electronic terms (grads.jl:13-21, 94-95):
    -sum lag dS + sum RDM1 (dT + dV) + 2 sum d geri3c - sum m geri2c, per Cartesian k."""
    g = -np.einsum("mn,mnk->k", lag, dS) + np.einsum("mn,mnk->k", RDM1, dT + dV)
    g += 2.0 * np.einsum("slQ,slQk->k", d, geri3c, optimize=True)
    g -= np.einsum("QR,QRk->k", m, geri2c)
    return g


def compute_geom_gradients(n, C, cj12, ck12, elag, p, G, I3c, atoms, Z, xyz):
    """compute_geom_gradients (grads.jl:1-100) with RI: the nuclear repulsion gradient
    (utils.jl:315-336) plus each atom's electronic terms; `atoms` yields per atom the
    derivative integrals (dS, dT, dV: nbf x nbf x 3; geri3c: nbf x nbf x naux x 3;
    geri2c: naux x naux x 3).  Returns natoms x 3."""
    RDM1, lag, d, m = geom_gradient_densities(n, C, cj12, ck12, elag, p, G, I3c)
    grad = grad_nuc(Z, xyz)
    for i, (dS, dT, dV, g3, g2) in enumerate(atoms):
        grad[i] += geom_gradient_atom(RDM1, lag, d, m, dS, dT, dV, g3, g2)
    return grad
