"""Deck-runner commands that need no GPU: equal-style variables (with LAMMPS' operator
precedence and "%.15g" substitution) and replicate (ids, molecule ids, unwrapped and remapped
positions in LAMMPS' copy order)."""
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import REPO
from lmff import structures as S

RUNNER = os.path.join(REPO, "lammps-lmff_amd", "lmff_run")


def run(tmp_path, text):
    if not os.path.exists(RUNNER):
        pytest.skip("lmff_run not built")
    (tmp_path / "in.deck").write_text(text)
    r = subprocess.run([RUNNER, "-in", "in.deck", "-var", "SX", "2"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=60)
    return r


def test_equal_style_variables(tmp_path):
    r = run(tmp_path, "\n".join([
        "variable SX index 1",                      # -var SX 2 wins
        "variable DX equal ${SX}*1",
        "variable BL equal 1.43876 # C-C bond length",
        "variable R0 equal ${BL}*0.5",
        "variable Xsc equal ${R0}*sin((30-0.0)/180*PI)*(-1)",
        "variable P equal -2^2+2^3^2-7%3+v_DX/4",
        "variable E equal 2.5e-3*1E2+sqrt(16)+ln(exp(1.5))+atan2(1,1)",
        'print "DX=${DX} Xsc=${Xsc} P=${P} E=${E}"',
        ""]))
    assert r.returncode == 0, r.stdout + r.stderr
    line = [l for l in r.stdout.splitlines() if l.startswith("DX=")][0]
    got = dict(kv.split("=") for kv in line.split())
    xsc = 1.43876 * 0.5 * math.sin((30 - 0.0) / 180 * math.pi) * (-1)
    # unary minus binds tighter than ^ (so -2^2 = 4) and ^ groups left to right (2^3^2 = 64)
    p = 4.0 + 64.0 - 1.0 + 2.0 / 4
    e = 2.5e-3 * 1e2 + 4.0 + 1.5 + math.atan2(1, 1)
    assert got == {"DX": "2", "Xsc": "%.15g" % xsc, "P": "%.15g" % p, "E": "%.15g" % e}


def test_equal_style_errors(tmp_path):
    r = run(tmp_path, "variable A equal 1/0\nprint ${A}\n")
    assert r.returncode != 0 and "Divide by 0" in r.stderr
    r = run(tmp_path, "variable A equal step+1\nprint ${A}\n")
    assert r.returncode != 0 and "Invalid thermo keyword" in r.stderr


def test_replicate_numbering_and_positions(tmp_path):
    lo, hi = np.array([0.0, 0.0, -5.0]), np.array([10.0, 12.0, 15.0])
    rows = [  # id mol type q x y z ix iy iz
        (3, 1, 1, 0.0, 9.5, 1.0, 0.0, 0, 0, 0),
        (7, 1, 2, 0.0, 0.5, 11.9, 0.2, -1, 1, 0),
        (5, 2, 1, 0.0, 4.0, 6.0, 3.3, 2, 0, 0),
        (1, 0, 2, 0.0, 8.0, 2.0, 3.4, 0, -1, 0),
    ]
    vel = {3: (0.1, 0.2, 0.3), 7: (-1.0, 0.0, 1.0), 5: (0.0, 0.5, 0.0), 1: (2.0, 2.0, 2.0)}
    with open(tmp_path / "in.data", "w") as fh:
        fh.write("test\n\n4 atoms\n2 atom types\n\n")
        for k, a in enumerate("xyz"):
            fh.write(f"{lo[k]} {hi[k]} {a}lo {a}hi\n")
        fh.write("\nMasses\n\n1 12.0107\n2 10.811\n\nAtoms # full\n\n")
        for r_ in rows:
            fh.write(" ".join(str(v) for v in r_) + "\n")
        fh.write("\nVelocities\n\n")
        for i, v in vel.items():
            fh.write(f"{i} {v[0]} {v[1]} {v[2]}\n")
    r = run(tmp_path, "\n".join([
        "units metal", "atom_style full", "boundary p p f", "read_data in.data",
        "variable DY equal 3", "replicate ${SX} ${DY} 1", "write_data out.data", ""]))
    assert r.returncode == 0, r.stdout + r.stderr
    out = S.read_lammps_data(str(tmp_path / "out.data"))
    nx, ny, nz = 2, 3, 1
    prd = hi - lo
    maxtag, maxmol = 7, 2
    exp = {}
    for ix in range(nx):
        for iy in range(ny):
            for iz in range(nz):
                rep = iz * ny * nx + iy * nx + ix
                for (i, m, t, q, x, y, z, jx, jy, jz) in rows:
                    p = np.array([x, y, z]) + np.array([jx, jy, jz]) * prd + np.array([ix, iy, iz]) * prd
                    for k in range(2):  # periodic x, y: remap into the replicated box
                        L = prd[k] * (nx, ny)[k]
                        while p[k] < lo[k]:
                            p[k] += L
                        while p[k] >= lo[k] + L:
                            p[k] -= L
                    exp[i + rep * maxtag] = (m + rep * maxmol if m > 0 else 0, t, p, vel[i])
    assert out.n == 24 and sorted(exp) == list(out.tag)
    np.testing.assert_array_equal(out.boxhi, lo + prd * np.array([nx, ny, nz]))
    for k, tag in enumerate(out.tag):
        m, t, p, v = exp[tag]
        assert out.mol[k] == m and out.type[k] == t
        np.testing.assert_allclose(out.x[k], p, rtol=0, atol=1e-12)
        np.testing.assert_array_equal(out.v[k], v)
