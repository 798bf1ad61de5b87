"""Writes (or, with --check, verifies) tests/golden/interface_table.json.

The fixture holds the quasi-disjoint interface coefficients pbco/pbci of the
reference's numerics_base::deriv_setup (src/numerics_base.hpp:182-297): rows 8
and 9 of R P^-1 Q for the 16-point two-block system, with R = P decoupled
between points 7 and 8.  They are computed here with numpy.linalg.inv from the
P and Q the reference assembles -- as shipped (the right-end closure row of Q
with ab00/ab02 in the order of src/numerics_base.hpp:201-203) and with that row
mirrored (pbci_fixed_5/7 record the two entries the mirroring changes most).
The oracle's own Gauss-Jordan restatement is checked against this file.

usage: python tests/golden/make_interface_table.py [--check]
"""
import json
import os
import sys

import numpy as np

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "interface_table.json")
ALPHA, AA, BB = 1.0 / 3.0, 7.0 / 9.0, 1.0 / 36.0
AB00, AB01, AB02, AB10 = -2.5, 2.0, 0.5, 0.75


def table(mirrored_end_row):
    n, lmd = 16, 8
    P = np.zeros((n, n))
    Q = np.zeros((n, n))
    P[0, :2] = [1, 2]
    P[1, :3] = [0.25, 1, 0.25]
    P[n - 2, n - 3:] = [0.25, 1, 0.25]
    P[n - 1, n - 2:] = [2, 1]
    for i in range(2, n - 2):
        P[i, i - 1:i + 2] = [ALPHA, 1, ALPHA]
    Q[0, :3] = [AB00, AB01, AB02]
    Q[1, :3] = [-AB10, 0, AB10]
    Q[n - 2, n - 3:] = [-AB10, 0, AB10]
    Q[n - 1, n - 3:] = [-AB02, -AB01, -AB00] if mirrored_end_row else [-AB00, -AB01, -AB02]
    for i in range(2, n - 2):
        Q[i, i - 2:i + 3] = [-BB, -AA, 0, AA, BB]
    R = P.copy()
    R[7, 8] = 0.0
    R[8, 7] = 0.0
    M = R @ np.linalg.inv(P) @ Q
    pbco = np.concatenate([M[8, :lmd][::-1], M[9, :lmd][::-1]])   # neighbour side, nearest first
    pbci = np.concatenate([M[8, lmd:], M[9, lmd:]])               # own side
    return pbco, pbci


def build():
    pbco, pbci = table(False)
    _, pbci_f = table(True)
    return {
        "source": "numpy.linalg.inv restatement of numerics_base::deriv_setup (src/numerics_base.hpp:182-297); "
                  "written by tests/golden/make_interface_table.py",
        "pbco_shipped": pbco.tolist(),
        "pbci_shipped": pbci.tolist(),
        "pbci_fixed_5": float(pbci_f[5]),
        "pbci_fixed_7": float(pbci_f[7]),
    }


def main():
    new = build()
    if "--check" in sys.argv:
        old = json.load(open(OUT))
        err = max(np.max(np.abs(np.array(old[k]) - np.array(new[k]))) for k in new if k != "source")
        print(f"max difference to the committed fixture: {err:.3e}")
        sys.exit(0 if err < 1e-15 else 1)
    with open(OUT, "w") as f:
        json.dump(new, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
