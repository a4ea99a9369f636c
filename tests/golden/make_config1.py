"""Writes tests/golden/config1_h2o_ccpvdz.json: the oracle's GNOF/ADAM/Softmax energy of
H2O (O 0 0 0.121, H 0 +-0.751 -0.485 A, cc-pVDZ spherical, the package's even-tempered
fitting set) from the core guess, run to tight thresholds (threshgorb 1e-5, threshgocc
1e-6, threshen 1e-8), and checksums of the integrals it was run on.

    python tests/golden/make_config1.py

The integrals come from the package's integral engine (they are pinned separately by the
textbook energies in tests/test_integrals.py); the oracle supplies b (its eigen-route
Iaux) and the whole outer loop.  The stationary point an ADAM run reaches depends on
roundoff (tests/test_gpu_config1.py says why), so this energy is a recorded point of the
functional that the device's solution must not lie above, and the integral checksums pin
the inputs.  About half a minute on 8 cores."""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import __graft_entry__ as ge  # noqa: E402

MOL = "0 1\n O  0.0000   0.000   0.121\n H  0.0000   0.751  -0.485\n H  0.0000  -0.751  -0.485\n"
THRESH = {"threshgorb": 1e-5, "threshgocc": 1e-6, "threshen": 1e-8}


def integrals():
    ge.load_package()
    from donof_amd import basis, integrals as ints
    bset, p = basis.molecule(MOL, "cc-pvdz", spherical=True)
    h = bset.handle()
    Z = np.array([a.Z for a in bset.atoms], dtype=np.float64)
    S, T, V = h.int1e(Z, bset.centers_bohr())
    S, T, V = (S + S.T) / 2, (T + T.T) / 2, (V + V.T) / 2
    aux = basis.aux_basis(bset)
    ha = aux.handle()
    G = ha.int2c2e()
    G = (G + G.T) / 2
    I3c = h.int3c2e(ha)
    return bset, p, S, T, V, G, I3c


def checksums(S, T, V, G, I3c):
    return {"S": float(np.abs(S).sum()), "T": float(np.abs(T).sum()), "V": float(np.abs(V).sum()),
            "G": float(np.abs(G).sum()), "I3c": float(np.abs(I3c).sum())}


def main():
    orc = ge.load_oracle()
    ge.load_package()
    from donof_amd import integrals as ints
    bset, p, S, T, V, G, I3c = integrals()
    H = T + V
    b = orc.iaux(G, I3c)
    _, C = ints.core_guess(H, S)
    C = ints.check_ortho(C, S, p, log=None)
    pr = orc.make_param(p.nbf, G.shape[0], p.ne, ipnof=8, orb_method="ADAM")
    for k, v in THRESH.items():
        setattr(pr, k, v)
    E, C2, n2, hist, elag = orc.energy_loop(b, H, C, pr)
    E_nuc = ints.compute_E_nuc(bset, p)
    out = {"molecule": MOL, "basis": "cc-pvdz", "aux": bset.name + "-autofit",
           "nbf": p.nbf, "nbfaux": int(G.shape[0]), "nbf5": p.nbf5, "ipnof": 8,
           "orb_method": "ADAM", "occ_method": "Softmax", "thresholds": THRESH,
           "E_nuc": E_nuc, "E_elec": E, "E_total": E + E_nuc, "outer_iterations": len(hist),
           "final_grad_orb": hist[-1][1], "final_grad_occ": hist[-1][2],
           "n": [float(x) for x in n2], "integral_checksums": checksums(S, T, V, G, I3c)}
    with open(os.path.join(HERE, "config1_h2o_ccpvdz.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("E_total %.12f after %d outer iterations" % (out["E_total"], len(hist)))


if __name__ == "__main__":
    main()
