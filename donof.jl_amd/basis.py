"""Molecules and Gaussian basis sets for the integral source of an RI system.

Mirrors the part of GaussianBasis that DoNOF.jl uses (`BasisSet(name, mol)`,
`bset.atoms`, `bset.basis`, `bset.nbas`, `bset.natoms`, `bset.nshells`,
`bset.shells_per_atom`) and DoNOF's own `molecule(mol, basis; spherical)` (io.jl:25-57)
with `Param(bset, mul, charge)` (param.jl:59-222).

Basis data: GaussianBasis reads its sets from Gaussian94-format files that are not in
this image.  The library below holds the published sets this package's tests and the
H2O workload need, typed in Gaussian94 format from their publications: STO-3G (Hehre,
Stewart, Pople, J. Chem. Phys. 51, 2657 (1969); H, He, C, N, O) and cc-pVDZ (Dunning,
J. Chem. Phys. 90, 1007 (1989); H, O).  The reference's RI fitting set for cc-pVDZ is
cc-pVDZ-JKFIT (Weigend, Phys. Chem. Chem. Phys. 4, 4285 (2002)), with
def2-universal-jkfit as its fallback (integrals.jl:18-23); neither is available here, so
`aux_basis` builds an even-tempered fitting set from the orbital basis by the rule in its
docstring, and RI energies differ from the reference's by the fitting error of that set
(the tests bound it against exact four-index integrals).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib

# Angstrom to bohr for the integrals (CODATA 2018 Bohr radius 0.529177210903 A).  The
# reference's nuclear repulsion uses its own rounded factor 1.88973 (utils.jl:306).
ANGSTROM_TO_BOHR = 1.0 / 0.529177210903

SYMBOLS = ["X", "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al", "Si",
           "P", "S", "Cl", "Ar"]
L_OF = {"S": 0, "P": 1, "D": 2, "F": 3, "G": 4, "H": 5, "I": 6}

_STO3G = """
H     0
S   3   1.00
      3.42525091             0.15432897
      0.62391373             0.53532814
      0.16885540             0.44463454
****
He     0
S   3   1.00
      6.36242139             0.15432897
      1.15892300             0.53532814
      0.31364979             0.44463454
****
C     0
S   3   1.00
     71.6168370              0.15432897
     13.0450960              0.53532814
      3.5305122              0.44463454
SP   3   1.00
      2.9412494             -0.09996723             0.15591627
      0.6834831              0.39951283             0.60768372
      0.2222899              0.70011547             0.39195739
****
N     0
S   3   1.00
     99.1061690              0.15432897
     18.0523120              0.53532814
      4.8856602              0.44463454
SP   3   1.00
      3.7804559             -0.09996723             0.15591627
      0.8784966              0.39951283             0.60768372
      0.2857144              0.70011547             0.39195739
****
O     0
S   3   1.00
    130.7093200              0.15432897
     23.8088610              0.53532814
      6.4436083              0.44463454
SP   3   1.00
      5.0331513             -0.09996723             0.15591627
      1.1695961              0.39951283             0.60768372
      0.3803890              0.70011547             0.39195739
****
"""

_CCPVDZ = """
H     0
S   3   1.00
     13.0100000              0.0196850
      1.9620000              0.1379770
      0.4446000              0.4781480
S   1   1.00
      0.1220000              1.0000000
P   1   1.00
      0.7270000              1.0000000
****
O     0
S   9   1.00
  11720.0000000              0.0007100
   1759.0000000              0.0054700
    400.8000000              0.0278370
    113.7000000              0.1048000
     37.0300000              0.2830620
     13.2700000              0.4487190
      5.0250000              0.2709520
      1.0130000              0.0154580
      0.3023000             -0.0025850
S   9   1.00
  11720.0000000             -0.0001600
   1759.0000000             -0.0012630
    400.8000000             -0.0062670
    113.7000000             -0.0257160
     37.0300000             -0.0709240
     13.2700000             -0.1654110
      5.0250000             -0.1169550
      1.0130000              0.5573680
      0.3023000              0.5727590
S   1   1.00
      0.3023000              1.0000000
P   3   1.00
     17.7000000              0.0430180
      3.8540000              0.2289130
      1.0460000              0.5087280
P   1   1.00
      0.2753000              1.0000000
D   1   1.00
      1.1850000              1.0000000
****
"""

LIBRARY = {"sto-3g": _STO3G, "cc-pvdz": _CCPVDZ}


def parse_g94(text):
    """{symbol: [(l, exps, coefs), ...]} from Gaussian94-format text; an SP shell becomes
    an S and a P shell on the same exponents (the S first)."""
    out = {}
    lines = [ln.strip() for ln in text.strip().splitlines()]
    i = 0
    while i < len(lines):
        if not lines[i] or lines[i] == "****":
            i += 1
            continue
        sym = lines[i].split()[0]
        shells = out.setdefault(sym, [])
        i += 1
        while i < len(lines) and lines[i] != "****":
            kind, npr = lines[i].split()[:2]
            npr = int(npr)
            rows = [[float(x.replace("D", "E")) for x in lines[i + 1 + k].split()]
                    for k in range(npr)]
            exps = [r[0] for r in rows]
            if kind.upper() == "SP":
                shells.append((0, exps, [r[1] for r in rows]))
                shells.append((1, exps, [r[2] for r in rows]))
            else:
                shells.append((L_OF[kind.upper()], exps, [r[1] for r in rows]))
            i += 1 + npr
        i += 1
    return out


@dataclass
class Atom:
    Z: int
    xyz: np.ndarray  # Angstrom, as GaussianBasis keeps atom coordinates

    @property
    def symbol(self):
        return SYMBOLS[self.Z]


@dataclass
class Shell:
    """A contracted shell: angular momentum l, exponents and coefficients (for normalised
    primitives, as the basis file lists them) on atom `atom` (0-based)."""
    l: int
    exp: np.ndarray
    coef: np.ndarray
    atom: int
    pure: bool


@dataclass
class BasisSet:
    name: str
    atoms: list
    basis: list
    spherical: bool = True
    _handle: object = field(default=None, repr=False)

    @property
    def natoms(self):
        return len(self.atoms)

    @property
    def nshells(self):
        return len(self.basis)

    @property
    def shells_per_atom(self):
        return [sum(1 for s in self.basis if s.atom == a) for a in range(self.natoms)]

    @property
    def nbas(self):
        return sum((2 * s.l + 1) if s.pure else (s.l + 1) * (s.l + 2) // 2 for s in self.basis)

    def centers_bohr(self):
        return np.array([a.xyz for a in self.atoms], dtype=np.float64) * ANGSTROM_TO_BOHR

    # derivative integrals for atom i (0-based), as api.compute_geom_gradients reads them
    # from the basis set (GaussianBasis' ∇overlap, ∇kinetic, ∇nuclear, ∇ERI_2e3c,
    # ∇ERI_2e2c in grads.jl:13-19,88-89); bohr, the Cartesian component last

    def _one_electron_derivs(self, i):
        key = ("1e", i)
        cache = self.__dict__.setdefault("_deriv_cache", {})
        if key not in cache:
            cache.clear()
            Z = np.array([a.Z for a in self.atoms], dtype=np.float64)
            cache[key] = self.handle().int1e_deriv([s.atom for s in self.basis], Z,
                                                   self.centers_bohr(), i)
        return cache[key]

    def overlap_deriv(self, i):
        return self._one_electron_derivs(i)[0]

    def kinetic_deriv(self, i):
        return self._one_electron_derivs(i)[1]

    def nuclear_deriv(self, i):
        return self._one_electron_derivs(i)[2]

    def fitting_set(self):
        """The RI fitting set the integrals of this basis use (`aux_basis`, made once)."""
        if self.__dict__.get("_aux") is None:
            self.__dict__["_aux"] = aux_basis(self)
        return self.__dict__["_aux"]

    def eri_2e3c_deriv(self, i):
        aux = self.fitting_set()
        return self.handle().int3c2e_deriv([s.atom for s in self.basis], aux.handle(),
                                           [s.atom for s in aux.basis], i)

    def eri_2e2c_deriv(self, i):
        aux = self.fitting_set()
        return aux.handle().int2c2e_deriv([s.atom for s in aux.basis], i)

    def handle(self):
        """The library's dn_basis for these shells (created once)."""
        if self._handle is None:
            self._handle = _lib.Basis(
                [s.l for s in self.basis], [len(s.exp) for s in self.basis],
                [1 if s.pure else 0 for s in self.basis],
                np.array([self.atoms[s.atom].xyz for s in self.basis], dtype=np.float64)
                * ANGSTROM_TO_BOHR,
                np.concatenate([s.exp for s in self.basis]),
                np.concatenate([s.coef for s in self.basis]))
        return self._handle


def parse_molecule(mol):
    """(atoms, charge, mul) from DoNOF's molecule string: an optional "charge mul" first
    line, then "symbol x y z" lines in Angstrom (io.jl:25-45)."""
    lines = [ln for ln in mol.strip("\n").split("\n")]
    first = lines[0].split()
    if len(first) == 2:
        charge, mul = int(first[0]), int(first[1])
        lines = lines[1:]
    else:
        charge, mul = 0, 1
    atoms = []
    for ln in lines:
        tok = ln.split()
        if not tok:
            continue
        sym = tok[0].capitalize()
        atoms.append(Atom(SYMBOLS.index(sym), np.array([float(x) for x in tok[1:4]])))
    return atoms, charge, mul


def basis_set(name, atoms, spherical=True):
    """BasisSet(name, atoms) of GaussianBasis: the library set `name` on every atom, shells
    in atom order and file order."""
    key = name.lower()
    if key not in LIBRARY:
        raise KeyError("basis set %r is not in this package's library (%s)"
                       % (name, ", ".join(sorted(LIBRARY))))
    data = parse_g94(LIBRARY[key])
    shells = []
    for ia, at in enumerate(atoms):
        if at.symbol not in data:
            raise KeyError("basis set %r has no entry for %s" % (name, at.symbol))
        for l, e, c in data[at.symbol]:
            shells.append(Shell(l, np.array(e), np.array(c), ia, bool(spherical)))
    return BasisSet(key, atoms, shells, spherical)


# highest angular momentum occupied in the ground-state atom
def _l_occ(Z):
    return 0 if Z <= 4 else (1 if Z <= 20 else 2)


def aux_basis(bset, beta=2.0):
    """Even-tempered RI fitting set built from the orbital basis.

    Per atom and per fitting angular momentum L up to max(2 l_occ, l_max + 1) (l_occ the
    highest angular momentum occupied in the ground-state atom, l_max the orbital basis'
    highest on that atom): the primitive pairs (i, j) of the atom's shells whose product
    holds L (|l_i - l_j| <= L <= l_i + l_j, l_i + l_j - L even) give the range
    [min(a_i + a_j), max(a_i + a_j)]; the set places exponents from the top of the range
    down by the ratio beta until one lies at or below its bottom.  One uncontracted
    primitive per shell, pure (spherical) as the orbital basis."""
    shells = []
    for ia, at in enumerate(bset.atoms):
        own = [s for s in bset.basis if s.atom == ia]
        lmax = max(s.l for s in own)
        Lmax = max(2 * _l_occ(at.Z), lmax + 1)
        for L in range(Lmax + 1):
            lo, hi = math.inf, 0.0
            for s in own:
                for t in own:
                    if abs(s.l - t.l) <= L <= s.l + t.l and (s.l + t.l - L) % 2 == 0:
                        sums = np.add.outer(s.exp, t.exp)
                        lo, hi = min(lo, sums.min()), max(hi, sums.max())
            if hi == 0.0:  # no orbital product reaches L: one function at the valence scale
                lo = hi = min(min(s.exp) for s in own) * 2.0
            a = hi
            while True:
                shells.append(Shell(L, np.array([a]), np.array([1.0]), ia, bset.spherical))
                if a <= lo * (1.0 + 1e-12):
                    break
                a /= beta
    return BasisSet(bset.name + "-autofit", bset.atoms, shells, bset.spherical)


def read_xyz_from_file(filename):
    """io.jl:3-22: the atom lines of an .xyz file (count on the first line, a comment line,
    then one line per atom) joined by newlines, for molecule()."""
    with open(filename) as f:
        lines = f.read().split("\n")
    natoms = int(lines[0].strip())
    return "\n".join(lines[2:2 + natoms])


def Z_to_symbol(Z):
    """utils.jl:338-: the element symbol of atomic number Z."""
    return SYMBOLS[int(Z)]


def molecule(mol, basis, spherical=False):
    """io.jl:25-57: (bset, p) for a molecule string and a basis name; p is the Param of
    param.jl:59-222 (no1 = 0; set_no1 sets the frozen core)."""
    from .param import Param
    atoms, charge, mul = parse_molecule(mol)
    bset = basis_set(basis, atoms, spherical=spherical)
    ne = sum(a.Z for a in atoms) - charge
    p = Param(nbf=bset.nbas, nbfaux=0, ne=ne, mul=mul)
    lines = mol.split("\n")
    p.mol = "\n".join(lines[1:]) if len(lines[0].split()) == 2 else mol
    p.natoms = bset.natoms
    p.spherical = spherical
    return bset, p
