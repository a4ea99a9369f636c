"""The formatted checkpoint file DoNOF writes after every energy() (output.jl:1-283,
called at energy.jl:274): basis description, natural orbitals, occupations and the
total density in Gaussian's FCHK text layout.

The file is written as the reference writes it, field by field and with its number
formats, including what it does unusually: atom coordinates are converted to bohr with
1.88973 while shell coordinates are printed as stored (Angstrom), the charge is printed
as -(sum Z) + nalpha + nbeta, the contraction coefficients are the two-column lines of
each atom's block in the basis-set file (so SP shells, written on three columns, add
none), and spherical orbitals are written in the basis' own function order.
"""
from __future__ import annotations

import math

import numpy as np

from . import basis as basis_mod


def _cart_reorder(C, bset):
    """output.jl:7-53: Cartesian d, f, g shells from x^l, x^(l-1)y, ... order to
    Gaussian's order, with the reference's scale factors."""
    Cnew = np.zeros_like(C)
    i = 0
    s3, s5, s15, s7, s35 = (math.sqrt(x) for x in (3, 5, 15, 7, 35))
    maps = {
        2: [(0, 1.0), (3, 1.0), (5, 1.0), (1, 1 / s3), (2, 1 / s3), (4, 1 / s3)],
        3: [(0, 1.0), (6, 1.0), (9, 1.0), (3, 1 / s5), (1, 1 / s5), (2, 1 / s5), (5, 1 / s5),
            (8, 1 / s5), (7, 1 / s5), (4, 1 / s15)],
        4: [(0, 1.0), (10, 1.0), (14, 1.0), (1, 1 / s7), (2, 1 / s7), (6, 1 / s7),
            (11, 1 / s7), (9, 1 / s7), (13, 1 / s7), (3, s3 / s35), (5, s3 / s35),
            (12, s3 / s35), (4, 1 / s35), (7, 1 / s35), (8, 1 / s35)],
    }
    for sh in bset.basis:
        ori = (sh.l + 1) * (sh.l + 2) // 2
        if sh.l in maps:
            for k, (src, f) in enumerate(maps[sh.l]):
                Cnew[i + k, :] = C[i + src, :] * f
        else:
            if sh.l > 4:
                print("Warning: l>4, incorrect order in fchk")
            Cnew[i:i + ori, :] = C[i:i + ori, :]
        i += ori
    return Cnew


def _file_coefficients(p, bset):
    """output.jl:57-79: for each atom line of p.mol, the second column of every
    two-column line of that element's block in the basis-set file."""
    text = basis_mod.LIBRARY[bset.name]
    lines = text.strip("\n").split("\n")
    coef = []
    for atom in p.mol[:-1].split("\n"):
        if not atom.split():
            continue
        symbol = atom.split()[0]
        k = 0
        while symbol + "     0" not in lines[k]:
            k += 1
        k += 1
        while True:
            line = lines[k]
            k += 1
            if "****" in line:
                break
            tok = line.split()
            if len(tok) == 2:
                coef.append(float(tok[1].replace("D", "E")))
    return coef


def _block(f, values, fmt, per_line):
    n = len(values)
    for i, v in enumerate(values, start=1):
        f.write(fmt % v)
        if i % per_line == 0 or i == n:
            f.write("\n")


def fchk(filename, p, bset, jobtype, E_t, elag, n, C):
    """output.jl:1-283: write `<filename>.fchk`."""
    C = np.asarray(C, dtype=np.float64)
    Cnew = C if p.spherical else _cart_reorder(C, bset)
    coef = _file_coefficients(p, bset)
    charge = -sum(a.Z for a in bset.atoms) + p.nalpha + p.nbeta
    max_l = max(s.l for s in bset.basis)
    max_np = max(len(s.exp) for s in bset.basis)
    nprim = sum(len(s.exp) for s in bset.basis)
    shells2atom = []
    for ia, ns in enumerate(bset.shells_per_atom):
        shells2atom += [ia + 1] * ns
    with open(filename + ".fchk", "w") as f:
        f.write("%s\n" % filename)
        if p.ista == 0:
            f.write("%s PNOF%d %s\n" % (jobtype, p.ipnof, bset.name))
        else:
            f.write("%s PNOF%ds %s\n" % (jobtype, p.ipnof, bset.name))
        for label, v in (("Number of atoms", p.natoms), ("Charge", charge),
                         ("Multiplicity", p.mul), ("Number of electrons", p.ne),
                         ("Number of alpha electrons", p.nalpha),
                         ("Number of beta electrons", p.nbeta),
                         ("Number of basis functions", p.nbf),
                         ("Number of independant functions", p.nbf5),
                         ("Number of contracted shells", bset.nshells),
                         ("Highest angular momentum", max_l),
                         ("Largest degree of contraction", max_np),
                         ("Number of primitive shells", nprim)):
            f.write("%-43sI           %6d\n" % (label, v))
        f.write("Atomic numbers                             I   N=      %6d\n" % bset.natoms)
        _block(f, [a.Z for a in bset.atoms], " %11d", 6)
        f.write("Nuclear Charges                            R   N=      %6d\n" % bset.natoms)
        _block(f, [float(a.Z) for a in bset.atoms], " %.8e", 6)
        f.write("Current cartesian coordinates              R   N=      %6d\n"
                % (bset.natoms * 3))
        _block(f, [x * 1.88973 for a in bset.atoms for x in a.xyz], " % .8e", 5)
        f.write("Shell types                                I   N=      %6d\n" % bset.nshells)
        _block(f, [(-s.l if (p.spherical and s.l > 1) else s.l) for s in bset.basis],
               " %11d", 6)
        f.write("Number of primitives per shell             I   N=      %6d\n" % bset.nshells)
        _block(f, [len(s.exp) for s in bset.basis], " %11d", 6)
        f.write("Shell to atom map                          I   N=      %6d\n" % bset.nshells)
        _block(f, shells2atom, " %11d", 6)
        f.write("Coordinates of each shell                  R   N=      %6d\n"
                % (bset.nshells * 3))
        _block(f, [x for s in bset.basis for x in bset.atoms[s.atom].xyz], " % 8e", 5)
        f.write("Total Energy                               R     %.15e\n" % E_t)
        f.write("Primitive exponents                        R   N=      %6d\n" % nprim)
        _block(f, [e for s in bset.basis for e in s.exp], " % .8e", 5)
        f.write("Contraction coefficients                   R   N=      %6d\n" % nprim)
        _block(f, coef[:nprim], " % .8e", 5)
        e_val = np.diag(np.asarray(elag))
        f.write("Alpha Orbital Energies                     R   N=      %6d\n" % p.nbf)
        _block(f, list(e_val[:p.nbf]), " % .8e", 5)
        f.write("Alpha MO coefficients                      R   N=      %6d\n" % (p.nbf * p.nbf))
        _block(f, list(Cnew[:p.nbf, :p.nbf].T.reshape(-1)), " % .8e", 5)
        f.write("Total SCF Density                          R   N=       %6d\n"
                % (p.nbf * (p.nbf + 1) // 2))
        C5 = Cnew[:, :p.nbf5]
        DM = 2.0 * (C5 * np.asarray(n)[:p.nbf5]) @ C5.T
        _block(f, [DM[mu, nu] for mu in range(p.nbf) for nu in range(mu + 1)], " % .8e", 5)
        f.write("Natural orbital occupancies                R   N=       %6d\n" % p.nbf5)
        _block(f, list(np.asarray(n)[:p.nbf5]), " % .8e", 5)
