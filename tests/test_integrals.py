"""The integral source (libdonofhip's host integral engine; compute_integrals,
integrals.jl:1-65) pinned without the reference: textbook Hartree-Fock energies of H2 and
H2O in STO-3G from four-index integrals, closed forms, quadrature, invariances, the
point-charge limit that ties the three-centre code to the nuclear-attraction code, and
the fitting error of the RI set against exact four-index integrals.  CPU only (the
integral engine is host code)."""
import json
import math
import os

import numpy as np
import pytest

from donof_amd import _lib, basis, integrals, output

HERE = os.path.dirname(os.path.abspath(__file__))
BOHR = 0.529177210903


def _mol_ints(mol, name, spherical=True):
    atoms, charge, _ = basis.parse_molecule(mol)
    bs = basis.basis_set(name, atoms, spherical)
    h = bs.handle()
    Z = np.array([a.Z for a in atoms], dtype=np.float64)
    xyz = bs.centers_bohr()
    S, T, V = h.int1e(Z, xyz)
    eri = h.int4c2e()
    E_nuc = sum(Z[i] * Z[j] / np.linalg.norm(xyz[i] - xyz[j])
                for i in range(len(Z)) for j in range(i))
    return bs, S, T, V, eri, E_nuc, int(Z.sum()) - charge


def rhf(S, H, eri, nocc, E_nuc):
    """Closed-shell Roothaan iterations to 1e-13 Ha (test helper)."""
    s, U = np.linalg.eigh(S)
    X = U / np.sqrt(s)
    F = H.copy()
    E_old = 0.0
    for it in range(500):
        _, Cp = np.linalg.eigh(X.T @ F @ X)
        Co = (X @ Cp)[:, :nocc]
        D = Co @ Co.T
        J = np.einsum("mnls,ls->mn", eri, D)
        K = np.einsum("mlns,ls->mn", eri, D)
        F = H + 2 * J - K
        E = float(np.sum(D * (H + F))) + E_nuc
        if abs(E - E_old) < 1e-13 and it > 5:
            return E
        E_old = E
    raise AssertionError("RHF did not converge")


def test_h2_sto3g_textbook():
    # Szabo & Ostlund, Modern Quantum Chemistry, sec. 3.5.2: R = 1.4 bohr, S12 = 0.6593,
    # E0 + 1/R = -1.1167 Ha (-1.116714 to the digits the STO-3G exponents carry)
    bs, S, T, V, eri, E_nuc, _ = _mol_ints("H 0 0 0\nH 0 0 %.15f" % (1.4 * BOHR), "sto-3g")
    assert abs(S[0, 1] - 0.6593) < 5e-5
    assert abs(rhf(S, T + V, eri, 1, E_nuc) - (-1.116714)) < 1e-6


def test_h2o_sto3g_known_answer():
    # H2O/STO-3G at the geometry of T. D. Crawford's programming project 3 (bohr):
    # E_nuc = 8.002367061810450, E_SCF = -74.942079928192
    g = np.array([[0.0, -0.143225816552, 0.0], [1.638036840407, 1.136548822547, 0.0],
                  [-1.638036840407, 1.136548822547, 0.0]]) * BOHR
    mol = "\n".join("%s %.15f %.15f %.15f" % (s, *r) for s, r in zip("OHH", g))
    bs, S, T, V, eri, E_nuc, ne = _mol_ints(mol, "sto-3g")
    assert bs.nbas == 7 and ne == 10
    assert abs(E_nuc - 8.002367061810450) < 1e-10
    assert abs(rhf(S, T + V, eri, 5, E_nuc) - (-74.942079928192)) < 1e-9


def test_nuclear_attraction_closed_form():
    # a normalised s primitive at A and a unit charge at C: <s|1/r_C|s> = erf(sqrt(2a) R)/R
    for a, R in ((0.7, 1.3), (3.0, 0.05), (0.2, 9.0)):
        h = _lib.Basis([0], [1], [0], np.zeros(3), np.array([a]), np.array([1.0]))
        _, T, V = h.int1e(np.array([1.0]), np.array([0.0, 0.0, R]))
        assert abs(-V[0, 0] - math.erf(math.sqrt(2 * a) * R) / R) < 1e-14
        assert abs(T[0, 0] - 1.5 * a) < 1e-14      # <s| -nabla^2/2 |s> = 3a/2


def test_single_centre_shells_orthonormal():
    for l in range(7):
        hp = _lib.Basis([l], [2], [1], np.zeros(3), np.array([1.3, 0.4]), np.array([0.6, 0.5]))
        S, _, _ = hp.int1e(np.zeros(0), np.zeros(0))
        assert S.shape == (2 * l + 1, 2 * l + 1)
        assert np.abs(S - np.eye(2 * l + 1)).max() < 1e-14
        hc = _lib.Basis([l], [2], [0], np.zeros(3), np.array([1.3, 0.4]), np.array([0.6, 0.5]))
        Sc, _, _ = hc.int1e(np.zeros(0), np.zeros(0))
        assert np.abs(np.diag(Sc) - 1).max() < 1e-14


def test_solid_harmonic_convention():
    # which function is which: overlaps of a pure shell with the Cartesian components of the
    # same shell at the same centre (x^l, x^(l-1)y, x^(l-1)z, ...) show each real solid
    # harmonic's polynomial (m = -l..l; p as x, y, z): d_-2 ~ xy, d_-1 ~ yz, d_0 ~ 2zz-xx-yy,
    # d_1 ~ xz, d_2 ~ xx-yy; f_-3 ~ y(3xx-yy), f_3 ~ x(xx-3yy), f_0 ~ z(2zz-3xx-3yy)
    def ov(l):
        h = _lib.Basis([l, l], [1, 1], [1, 0], np.zeros(6), np.array([0.8, 0.8]),
                       np.array([1.0, 1.0]))
        S, _, _ = h.int1e(np.zeros(0), np.zeros(0))
        return S[:2 * l + 1, 2 * l + 1:]

    def comp(l, lx, ly):
        return [(a, b) for a in range(l, -1, -1) for b in range(l - a, -1, -1)].index((lx, ly))

    Sp = ov(1)
    assert np.allclose(Sp, np.eye(3), atol=1e-14)
    Sd = ov(2)
    xx, xy, xz, yy, yz, zz = (comp(2, a, b) for a, b in ((2, 0), (1, 1), (1, 0), (0, 2), (0, 1), (0, 0)))
    # the pure functions are orthogonal to the spherical part of the Cartesian d set, so
    # each row of Sd holds that function's polynomial in the (non-orthogonal) components
    assert Sd[0, xy] > 0.5 and Sd[1, yz] > 0.5 and Sd[3, xz] > 0.5
    assert Sd[4, xx] > 0 and Sd[4, yy] < 0 and abs(Sd[4, zz]) < 1e-14
    assert Sd[2, zz] > 0 and Sd[2, xx] < 0 and Sd[2, yy] < 0
    for row, keep in ((0, xy), (1, yz), (3, xz)):
        assert np.abs(np.delete(Sd[row], keep)).max() < 1e-14
    Sf = ov(3)
    c = lambda a, b: comp(3, a, b)
    assert Sf[0, c(2, 1)] > 0 and Sf[0, c(0, 3)] < 0          # y(3xx - yy)
    assert Sf[6, c(3, 0)] > 0 and Sf[6, c(1, 2)] < 0          # x(xx - 3yy)
    assert Sf[3, c(0, 0)] > 0 and Sf[3, c(2, 0)] < 0          # z(2zz - 3xx - 3yy)


def test_overlap_kinetic_quadrature():
    # Cartesian d on A against Cartesian f (two primitives) on B: every integrand is a
    # product of one-dimensional factors per primitive pair, integrated here on a fine 1D
    # grid (no Hermite recursion): S, and T = 1/2 <grad a|grad b>
    A, B = np.array([0.1, -0.2, 0.3]), np.array([-0.4, 0.5, -0.1])
    ea, ca = np.array([0.9]), np.array([1.0])
    eb, cb = np.array([1.1, 0.35]), np.array([0.4, 0.7])
    h = _lib.Basis([2, 3], [1, 2], [0, 0], np.concatenate([A, B]), np.concatenate([ea, eb]),
                   np.concatenate([ca, cb]))
    S, T, _ = h.int1e(np.zeros(0), np.zeros(0))
    # basis-file coefficients multiply primitives normalised as x^l exp(-a r^2)
    prim = lambda a, l: (2 * a / math.pi) ** 0.75 * (4 * a) ** (l / 2) / math.sqrt(
        float(np.prod(np.arange(2 * l - 1, 0, -2))))
    ca = ca * np.array([prim(a, 2) for a in ea])
    cb = cb * np.array([prim(a, 3) for a in eb])
    x = np.linspace(-16.0, 16.0, 64001)
    dxg = x[1] - x[0]

    def comps(l):
        return [(lx, ly, l - lx - ly) for lx in range(l, -1, -1) for ly in range(l - lx, -1, -1)]

    def f1(k, e, c0):       # (x - c0)^k exp(-e (x - c0)^2) and its derivative
        d = x - c0
        g = np.exp(-e * d * d)
        return d ** k * g, ((k * d ** (k - 1)) if k > 0 else 0.0) * g - 2 * e * d ** (k + 1) * g

    def pair(ka, kb, e1, e2, dim):
        fa, ga = f1(ka, e1, A[dim])
        fb, gb = f1(kb, e2, B[dim])
        return np.sum(fa * fb) * dxg, np.sum(ga * gb) * dxg

    def norm(k, e, c, ctr):
        tot = 0.0
        for e1, c1 in zip(e, c):
            for e2, c2 in zip(e, c):
                t = c1 * c2
                for dim in range(3):
                    d = x - ctr[dim]
                    t *= np.sum(d ** (2 * k[dim]) * np.exp(-(e1 + e2) * d * d)) * dxg
                tot += t
        return math.sqrt(tot)

    for i, ka in enumerate(comps(2)):
        na = norm(ka, ea, ca, A)
        for j, kb in enumerate(comps(3)):
            nb = norm(kb, eb, cb, B)
            s_q = t_q = 0.0
            for e1, c1 in zip(ea, ca):
                for e2, c2 in zip(eb, cb):
                    sd, td = zip(*(pair(ka[d], kb[d], e1, e2, d) for d in range(3)))
                    s_q += c1 * c2 * sd[0] * sd[1] * sd[2]
                    t_q += 0.5 * c1 * c2 * (td[0] * sd[1] * sd[2] + sd[0] * td[1] * sd[2]
                                            + sd[0] * sd[1] * td[2])
            assert abs(S[i, 6 + j] - s_q / (na * nb)) < 1e-12
            assert abs(T[i, 6 + j] - t_q / (na * nb)) < 1e-12


def test_three_centre_point_charge_limit():
    # (ab|s_C) with s_C a unit-charge Gaussian of exponent 1e9 -> <a|1/r_C|b>: the
    # three-centre code against the nuclear-attraction code
    bs, S, T, V, eri, _, _ = _mol_ints("O 0 0 0.121\nH 0 0.751 -0.485\nH 0 -0.751 -0.485",
                                       "cc-pvdz")
    C = np.array([0.3, -0.2, 0.5])
    alpha = 1e9
    aux = _lib.Basis([0], [1], [0], C, np.array([alpha]), np.array([1.0]))
    I3 = bs.handle().int3c2e(aux)[:, :, 0]
    # unit charge: (alpha/pi)^{3/2} e^{-alpha r^2} = f * (normalised primitive)
    f = (alpha / math.pi) ** 1.5 / (2 * alpha / math.pi) ** 0.75
    _, _, Vc = bs.handle().int1e(np.array([1.0]), C)
    assert np.abs(f * I3 + Vc).max() < 1e-6 * np.abs(Vc).max()


def test_rotation_translation_invariance():
    # d functions in cc-pVDZ: the RHF energy of a rigidly moved molecule is unchanged
    xyz0 = np.array([[0, 0, 0.121], [0, 0.751, -0.485], [0, -0.751, -0.485]])
    Q, _ = np.linalg.qr(np.random.default_rng(3).normal(size=(3, 3)))
    E = []
    for g in (xyz0, xyz0 @ Q.T + np.array([0.3, -0.2, 0.5])):
        mol = "\n".join("%s %.15f %.15f %.15f" % (s, *r) for s, r in zip("OHH", g))
        bs, S, T, V, eri, E_nuc, _ = _mol_ints(mol, "cc-pvdz")
        E.append(rhf(S, T + V, eri, 5, E_nuc))
    assert abs(E[0] - E[1]) < 1e-10


def test_two_and_three_centre_symmetry_and_ri_error():
    mol = "O 0 0 0.121\nH 0 0.751 -0.485\nH 0 -0.751 -0.485"
    bs, S, T, V, eri, E_nuc, _ = _mol_ints(mol, "cc-pvdz")
    assert bs.nbas == 24
    aux = basis.aux_basis(bs)
    ha = aux.handle()
    G = ha.int2c2e()
    I3 = bs.handle().int3c2e(ha)
    assert np.abs(G - G.T).max() < 1e-13 * np.abs(G).max()
    assert np.linalg.eigvalsh(G).min() > 0
    assert np.abs(I3 - I3.transpose(1, 0, 2)).max() < 1e-13 * np.abs(I3).max()
    # exact four-index integrals are symmetric under the 8 permutations
    assert np.abs(eri - eri.transpose(1, 0, 2, 3)).max() < 1e-14
    assert np.abs(eri - eri.transpose(2, 3, 0, 1)).max() < 1e-14
    # the fitting set's RI error at the RHF level (cc-pVDZ-JKFIT's is of this order too)
    w, U = np.linalg.eigh(G)
    b = np.einsum("mnP,PQ->mnQ", I3, (U / np.sqrt(w)) @ U.T)
    eri_ri = np.einsum("mnQ,lsQ->mnls", b, b)
    E4 = rhf(S, T + V, eri, 5, E_nuc)
    Eri = rhf(S, T + V, eri_ri, 5, E_nuc)
    assert abs(E4 - Eri) < 1e-4


def test_molecule_and_param_sizes():
    # the sizes of the H2O / cc-pVDZ workload: nbf 24, nbf5 20, nv 15, nvar 270, no1 0
    mol = "0 1\n O  0.0000   0.000   0.121\n H  0.0000   0.751  -0.485\n H  0.0000  -0.751  -0.485\n"
    bset, p = basis.molecule(mol, "cc-pvdz", spherical=True)
    assert (p.nbf, p.nbf5, p.nv, p.nvar, p.no1, p.ndoc, p.ncwo) == (24, 20, 15, 270, 0, 5, 3)
    assert p.natoms == 3 and p.spherical and p.mol.startswith(" O")
    assert bset.nshells == 12 and bset.shells_per_atom == [6, 3, 3]


def test_integrals_match_golden_checksums():
    with open(os.path.join(HERE, "golden", "config1_h2o_ccpvdz.json")) as f:
        gold = json.load(f)
    bset, p = basis.molecule(gold["molecule"], gold["basis"], spherical=True)
    h = bset.handle()
    Z = np.array([a.Z for a in bset.atoms], dtype=np.float64)
    S, T, V = h.int1e(Z, bset.centers_bohr())
    ha = basis.aux_basis(bset).handle()
    G, I3c = ha.int2c2e(), h.int3c2e(ha)
    assert ha.nbf == gold["nbfaux"]
    for k, v in (("S", S), ("T", T), ("V", V), ("G", G), ("I3c", I3c)):
        assert abs(np.abs(v).sum() - gold["integral_checksums"][k]) < 1e-9 * gold["integral_checksums"][k]
    assert abs(integrals.compute_E_nuc(bset, p) - gold["E_nuc"]) < 1e-12


def test_check_ortho_and_core_guess():
    mol = "O 0 0 0.121\nH 0 0.751 -0.485\nH 0 -0.751 -0.485"
    bset, p = basis.molecule(mol, "cc-pvdz", spherical=True)
    h = bset.handle()
    S, T, V = h.int1e(np.array([8.0, 1, 1]), bset.centers_bohr())
    e, C = integrals.core_guess(T + V, S)
    assert np.all(np.diff(e) >= 0)
    assert np.abs(C.T @ S @ C - np.eye(24)).max() < 1e-10
    C2 = integrals.check_ortho(C * 1.01, S, p, log=None)       # needs orthonormalising
    assert np.abs(C2.T @ S @ C2 - np.eye(24)).max() < 1e-10
    for j in range(24):                                         # sign of the largest entry
        assert C2[np.argmax(np.abs(C2[:, j])), j] > 0


def test_fchk_layout(tmp_path):
    mol = "0 1\n O  0.0000   0.000   0.121\n H  0.0000   0.751  -0.485\n H  0.0000  -0.751  -0.485\n"
    bset, p = basis.molecule(mol, "cc-pvdz", spherical=True)
    rng = np.random.default_rng(0)
    C = np.linalg.qr(rng.normal(size=(24, 24)))[0]
    n = np.sort(rng.random(p.nbf5))[::-1]
    elag = rng.normal(size=(24, 24))
    fn = str(tmp_path / "h2o")
    output.fchk(fn, p, bset, "Energy", -76.25, elag, n, C)
    text = open(fn + ".fchk").read().split("\n")
    assert text[0] == fn and text[1] == "Energy PNOF7 cc-pvdz"

    def section(name):
        k = next(i for i, ln in enumerate(text) if ln.startswith(name))
        cnt = int(text[k].split("N=")[1])
        vals = []
        i = k + 1
        while len(vals) < cnt:
            vals += text[i].split()
            i += 1
        assert len(vals) == cnt
        return np.array([float(v) for v in vals])

    assert int(text[2].split()[-1]) == 3 and int(text[3].split()[-1]) == 0
    assert list(section("Shell types")) == [0, 0, 0, 1, 1, -2, 0, 0, 1, 0, 0, 1]
    assert list(section("Shell to atom map")) == [1] * 6 + [2] * 3 + [3] * 3
    nprim = sum(len(s.exp) for s in bset.basis)
    assert len(section("Primitive exponents")) == nprim == 34
    # the two-column coefficient lines of the O, H, H blocks, in file order
    coef = section("Contraction coefficients")
    assert coef[0] == 0.00071 and coef[8] == -0.002585 and coef[17] == 0.572759
    assert list(coef[24:29]) == [0.019685, 0.137977, 0.478148, 1.0, 1.0]
    Cf = section("Alpha MO coefficients").reshape(24, 24).T
    assert np.abs(Cf - C).max() < 1e-8 * np.abs(C).max() * 10
    DM = section("Total SCF Density")
    D = 2 * (C[:, :p.nbf5] * n) @ C[:, :p.nbf5].T
    assert np.abs(DM - D[np.tril_indices(24)]).max() < 1e-7
    assert np.abs(section("Natural orbital occupancies") - n).max() < 1e-8
    assert abs(float(text[next(i for i, ln in enumerate(text) if ln.startswith("Total Energy"))]
                     .split()[-1]) + 76.25) < 1e-12


def test_derivative_integrals_finite_differences():
    # d/dR of S, T, V, (P|Q) and (mn|P) for one atom against central differences of the
    # integrals at displaced geometries (h = 1e-4 bohr)
    xyz = np.array([[0.0, 0.0, 0.121], [0.0, 0.751, -0.485], [0.0, -0.751, -0.485]])

    def ints(g):
        mol = "\n".join("%s %.15f %.15f %.15f" % (s, *r) for s, r in zip("OHH", g))
        bs = basis.basis_set("cc-pvdz", basis.parse_molecule(mol)[0], True)
        h = bs.handle()
        S, T, V = h.int1e(np.array([8.0, 1.0, 1.0]), bs.centers_bohr())
        aux = bs.fitting_set()
        return bs, S, T, V, aux.handle().int2c2e(), h.int3c2e(aux.handle())

    bs, *_ = ints(xyz)
    aux = bs.fitting_set()
    step = 1e-4 / basis.ANGSTROM_TO_BOHR
    for atom, k in ((1, 1), (0, 2), (2, 0)):
        dS, dT, dV = bs.overlap_deriv(atom), bs.kinetic_deriv(atom), bs.nuclear_deriv(atom)
        dG, dI = bs.eri_2e2c_deriv(atom), bs.eri_2e3c_deriv(atom)
        gp, gm = xyz.copy(), xyz.copy()
        gp[atom, k] += step
        gm[atom, k] -= step
        plus, minus = ints(gp)[1:], ints(gm)[1:]
        for ana, a, b_ in zip((dS, dT, dV, dG, dI), plus, minus):
            fd = (a - b_) / 2e-4
            assert np.abs(ana[..., k] - fd).max() < 1e-6 * max(1.0, np.abs(fd).max()), (atom, k)
        # translational invariance: the derivatives of all atoms sum to zero
    tot = [sum(f(a) for a in range(3)) for f in (bs.overlap_deriv, bs.kinetic_deriv,
                                                   bs.nuclear_deriv, bs.eri_2e2c_deriv,
                                                   bs.eri_2e3c_deriv)]
    for t in tot:
        assert np.abs(t).max() < 1e-10
