"""Generate tests/golden/wigner_golden.json: exact Wigner-3j values from sympy.

sympy.physics.wigner.wigner_3j evaluates the Racah formula in exact rational
arithmetic; values are stored as 25-significant-digit decimal strings.
Entries: f(l3) = (l3 l1 l2; -m1-m2 m1 m2), the family the reference requests
from WignerFamilies.jl (src/modecoupling.jl:70).

Run:  python tests/golden/make_wigner_golden.py
"""
import json
import os

from sympy import N
from sympy.physics.wigner import wigner_3j

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "wigner_golden.json")


def family(l1, l2, m1, m2, step=1):
    mj = -m1 - m2
    nmin = max(abs(l1 - l2), abs(mj))
    vals = {}
    for l3 in range(nmin, l1 + l2 + 1, step):
        vals[str(l3)] = str(N(wigner_3j(l3, l1, l2, mj, m1, m2), 25))
    return {"l1": l1, "l2": l2, "m1": m1, "m2": m2, "values": vals}


def main():
    fams = []
    for l1 in range(0, 13):
        for l2 in range(l1, 13):
            fams.append(family(l1, l2, 0, 0))
            if l1 >= 2:
                fams.append(family(l1, l2, -2, 2))
    for l1 in range(0, 9):
        for l2 in range(0, 9):
            for m1, m2 in ((-1, 1), (1, 0), (2, -1), (0, -1)):
                if abs(m1) <= l1 and abs(m2) <= l2:
                    fams.append(family(l1, l2, m1, m2))
    for (l1, l2, step) in ((40, 57, 5), (150, 203, 23), (3, 300, 1), (2, 251, 1), (400, 401, 41),
                           (700, 800, 157)):
        fams.append(family(l1, l2, 0, 0, step))
        fams.append(family(l1, l2, -2, 2, step))
    # QuickPol families WignerF(l, l'', -s, -nu) (src/beam.jl:87-88): spins
    # s in {0, +-2}, nu up to +-4
    qp = ((-2, 0), (2, -1), (-2, -3), (0, 2), (2, 4), (-2, -4), (0, -3))
    for l1 in range(0, 9):
        for l2 in range(0, 9):
            for m1, m2 in qp:
                if abs(m1) <= l1 and abs(m2) <= l2:
                    fams.append(family(l1, l2, m1, m2))
    for (l1, l2, m1, m2, step) in ((40, 57, -2, -3, 5), (150, 203, 2, 4, 23), (5, 90, -2, -4, 1),
                                   (300, 300, 0, -3, 37)):
        fams.append(family(l1, l2, m1, m2, step))
    with open(OUT, "w") as f:
        json.dump({"generator": "sympy.physics.wigner.wigner_3j (exact), 25 digits",
                   "families": fams}, f, separators=(",", ":"))
    print(OUT, len(fams))


if __name__ == "__main__":
    main()
