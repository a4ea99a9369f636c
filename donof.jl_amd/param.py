"""Size and hyper-parameter record of a solve, mirroring DoNOF.jl's mutable Param
(param.jl:3-57) with the constructor defaults of param.jl:59-165 and set_ncwo of
param.jl:242-266.  Field names are the reference's, so code written against
p.ipnof, p.orb_method, p.maxloop, p.alpha ... reads the same."""
from __future__ import annotations

import math
from dataclasses import dataclass, field


@dataclass
class Param:
    nbf: int
    nbfaux: int
    ne: int
    mul: int = 1
    no1: int = 0
    title: str = "donof"
    maxit: int = 1000
    maxloop: int = 10
    ipnof: int = 7
    ista: int = 0
    threshen: float = 1e-4
    threshgorb: float = 4e-3
    threshgocc: float = 1e-3
    alpha: float = 0.01
    RI: bool = True
    HighSpin: bool = False
    MSpin: int = 0
    occ_method: str = "Softmax"
    orb_method: str = "ADABelief"
    h_cut: float = field(default_factory=lambda: 0.02 * math.sqrt(2.0))
    # the molecule (param.jl:4-5, 54): atom lines, atom count, spherical functions
    mol: str = ""
    natoms: int = 0
    spherical: bool = False
    # derived sizes (param.jl:71-125)
    nalpha: int = 0
    nbeta: int = 0
    ndoc: int = 0
    nsoc: int = 0
    ndns: int = 0
    nvir: int = 0
    ncwo: int = -1
    nac: int = 0
    nbf5: int = 0
    no0: int = 0
    nv: int = 0
    nvar: int = 0
    noptorb: int = 0
    closed: bool = True
    # fields of the reference's Param that this path accepts and does not read, with the
    # reference's defaults (param.jl:129-164): the iterative-diagonalisation ("ID") orbital
    # method and its DIIS, the F-matrix thresholds of the no1 > 0 branch, and the CUDA
    # device switches (here the device is always the GPU, and FP64 or split precision is
    # chosen per System)
    thresheid: float = 1e-6
    maxitid: int = 30
    threshl: float = 1e-4
    threshe: float = 1e-5
    threshec: float = 1e-8
    scaling: bool = True
    nzeros: int = 0
    nzerosm: int = 5
    nzerosr: int = 2
    itziter: int = 10
    diis: bool = True
    thdiis: float = 1e-3
    ndiis: int = 5
    perdiis: bool = True
    gpu: bool = False
    gpu_bits: int = 64

    def __post_init__(self):
        self.nalpha = (self.ne + self.mul - 1) // 2
        self.nbeta = (self.ne - self.mul + 1) // 2
        # param.jl:120: both spins paired as in a closed shell
        self.closed = (self.nbeta == (self.ne + self.mul - 1) // 2
                       and self.nalpha == (self.ne - self.mul + 1) // 2)
        self.ndoc = self.nbeta - self.no1
        self.nsoc = self.nalpha - self.nbeta
        self.ndns = self.ndoc + self.nsoc
        self.nvir = self.nbf - self.nalpha
        self.noptorb = self.nbf
        self.set_ncwo(self.ncwo)

    def set_ncwo(self, ncwo):
        """param.jl:242-266."""
        if self.ndns != 0:
            if self.ndoc > 0:
                if ncwo != 1 and (ncwo == -1 or ncwo > self.nvir / self.ndoc):
                    ncwo = int(self.nvir / self.ndoc)
            else:
                ncwo = 0
        self.ncwo = ncwo
        self.nac = self.ndoc * (1 + ncwo)
        self.nbf5 = self.no1 + self.nac + self.nsoc
        self.no0 = self.nbf - self.nbf5
        self.nv = self.ncwo * self.ndoc
        self.nvar = round(self.nbf * (self.nbf - 1) / 2 - self.no0 * (self.no0 - 1) / 2)

    def set_no1(self, bset=None, no1=-1):
        """param.jl:268-311: the frozen (fully occupied) orbitals, by default one per
        first-row atom, five per second-row atom ... from the atoms of bset; resizes the
        subspaces.  The energy of a solve with no1 > 0 needs the reference's F-matrix branch
        of ENERGY1r, which this path does not have (energy() refuses it)."""
        if no1 == -1:
            no1 = 0
            for a in (bset.atoms if bset is not None else []):
                Z = a.Z
                for lo, hi, k in ((1, 2, 0), (3, 10, 1), (11, 18, 5), (19, 36, 9),
                                  (37, 49, 18), (50, 54, 23), (55, 71, 27), (72, 81, 30),
                                  (82, 86, 39), (87, 109, 43)):
                    if lo <= Z <= hi:
                        no1 += k
                        break
        self.no1 = no1
        self.ndoc = self.nbeta - self.no1
        self.nsoc = self.nalpha - self.nbeta
        self.ndns = self.ndoc + self.nsoc
        self.nvir = self.nbf - self.nalpha
        # ncwo is kept (param.jl:300-310)
        self.nac = self.ndoc * (1 + self.ncwo)
        self.nbf5 = self.no1 + self.nac + self.nsoc
        self.no0 = self.nbf - self.nbf5
        self.nv = self.ncwo * self.ndoc
        self.nvar = round(self.nbf * (self.nbf - 1) / 2 - self.no0 * (self.no0 - 1) / 2)

    def copy(self):
        import copy
        return copy.copy(self)


def set_ncwo(p, ncwo):
    p.set_ncwo(ncwo)
