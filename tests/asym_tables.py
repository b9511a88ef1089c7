"""Asymmetric variants of the shipped potential files (test infrastructure).

Both shipped files give every type pair the same parameters in both orders, so the
reference's tag-parity choice of whose row an unordered pair uses never shows in a result.
These variants break that symmetry: the B-C and N-C rows of BNCH.ILP differ from C-B and C-N
(z0, alpha, d, s_R, C6), and the B N N entry of BNC.tersoff differs from N B B in its
repulsion (lambda1, A).
"""
from __future__ import annotations

import os

from conftest import POT

ILP = os.path.join(POT, "BNCH.ILP")
BNC = os.path.join(POT, "BNC.tersoff")


def write(tmp_path):
    """Writes asym.ILP and asym.tersoff under tmp_path; returns (tersoff path, ILP path)."""
    out = []
    for line in open(ILP).read().splitlines():
        w = line.split()
        if len(w) == 13 and (w[0], w[1]) in (("B", "C"), ("N", "C")):
            v = [float(t) for t in w[2:]]
            v[0] *= 1.03   # beta (z0)
            v[1] *= 0.97   # alpha
            v[8] *= 1.1    # C6
            v[5] *= 1.02   # d
            v[6] *= 0.98   # sR
            line = f"{w[0]} {w[1]} " + " ".join(repr(t) for t in v)
        out.append(line)
    asym = os.path.join(str(tmp_path), "asym.ILP")
    with open(asym, "w") as fh:
        fh.write("\n".join(out) + "\n")
    tout = []
    for line in open(BNC).read().splitlines():
        w = line.split()
        if len(w) == 17 and w[:3] == ["B", "N", "N"]:
            v = [float(t) for t in w[3:]]
            v[12] *= 1.01   # lambda1
            v[13] *= 0.98   # A
            line = "B N N " + " ".join(repr(t) for t in v)
        tout.append(line)
    tasym = os.path.join(str(tmp_path), "asym.tersoff")
    with open(tasym, "w") as fh:
        fh.write("\n".join(tout) + "\n")
    return tasym, asym
