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


def _small_data(path, images):
    """Four atoms in a 10 x 12 x 20 box with velocities and the given image flags."""
    rows = [(3, 1, 1, 9.5, 1.0, 0.0), (7, 1, 2, 0.5, 11.9, 0.2), (5, 2, 1, 4.0, 6.0, 3.3), (1, 0, 2, 8.0, 2.0, 3.4)]
    with open(path, "w") as fh:
        fh.write("test\n\n4 atoms\n2 atom types\n\n0 10 xlo xhi\n0 12 ylo yhi\n-5 15 zlo zhi\n")
        fh.write("\nMasses\n\n1 12.0107\n2 10.811\n\nAtoms # full\n\n")
        for (i, m, t, x, y, z), im in zip(rows, images):
            fh.write(f"{i} {m} {t} 0.0 {x!r} {y!r} {z!r} {im[0]} {im[1]} {im[2]}\n")
        fh.write("\nVelocities\n\n")
        for k, (i, *_rest) in enumerate(rows):
            fh.write(f"{i} {0.1 * k + 0.013!r} {-0.2 * k!r} {0.7 / (k + 1)!r}\n")


def test_write_restart_read_restart_round_trip(tmp_path):
    """write_restart -> read_restart keeps the timestep, dt, box, masses, ids, molecules,
    types, image flags, positions and velocities exactly (read back through write_data, whose
    Atoms rows now carry the image flags read_data was given)."""
    images = [(0, 0, 0), (-1, 2, 0), (3, 0, -1), (0, -511, 511)]
    _small_data(tmp_path / "in.data", images)
    r = run(tmp_path, "\n".join([
        "units metal", "atom_style full", "boundary p p f", "read_data in.data", "mass 2 10.5",
        "reset_timestep 1234", "timestep 0.002", "write_data direct.data", "write_restart r.bin", ""]))
    assert r.returncode == 0, r.stdout + r.stderr
    (tmp_path / "in2.deck").write_text("read_restart r.bin\nwrite_data back.data\n")
    r2 = subprocess.run([RUNNER, "-in", "in2.deck"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r2.returncode == 0, r2.stdout + r2.stderr
    a = S.read_lammps_data(str(tmp_path / "in.data"))
    for name in ("direct.data", "back.data"):
        b = S.read_lammps_data(str(tmp_path / name))
        assert (tmp_path / name).read_text().splitlines()[0].endswith("timestep = 1234")
        np.testing.assert_array_equal(b.tag, a.tag)
        np.testing.assert_array_equal(b.mol, a.mol)
        np.testing.assert_array_equal(b.type, a.type)
        np.testing.assert_array_equal(b.x, a.x)
        np.testing.assert_array_equal(b.v, a.v)
        np.testing.assert_array_equal(b.meta["image"], a.meta["image"])
        np.testing.assert_array_equal(b.boxhi, a.boxhi)
        assert b.meta["masses"] == {1: 12.0107, 2: 10.5}


def test_restart_errors(tmp_path):
    _small_data(tmp_path / "in.data", [(0, 0, 0)] * 4)
    (tmp_path / "junk.bin").write_bytes(b"not a restart file at all")
    r = run(tmp_path, "read_restart junk.bin\n")
    assert r.returncode != 0 and "Invalid restart file" in r.stderr
    r = run(tmp_path, "read_data in.data\nread_restart junk.bin\n")
    assert r.returncode != 0 and "Cannot read_restart after simulation box is defined" in r.stderr
    r = run(tmp_path, "read_data in.data\nwrite_restart r.bin\n")
    assert r.returncode == 0, r.stderr
    data = (tmp_path / "r.bin").read_bytes()
    (tmp_path / "cut.bin").write_bytes(data[: len(data) - 20])
    r = run(tmp_path, "read_restart cut.bin\n")
    assert r.returncode != 0 and "Unexpected end of restart file" in r.stderr
    r = run(tmp_path, "read_data in.data\nrestart 10\n")
    assert r.returncode != 0 and "Illegal restart command" in r.stderr


def _temp_mom_angmom(s, v):
    m = s.masses()
    BOLTZ, MVV2E = 8.617343e-5, 1.0364269e-4
    t = (m[:, None] * v * v).sum() * MVV2E / ((3 * s.n - 3) * BOLTZ)
    p = (m[:, None] * v).sum(0)
    xcm = (m[:, None] * s.x).sum(0) / m.sum()
    L = (m[:, None] * np.cross(s.x - xcm, v)).sum(0)
    return t, p, L


def test_velocity_create_keywords_and_scale(tmp_path):
    """velocity all create T seed with dist, mom and rot (Velocity::create: draws / sqrt(m),
    zero_momentum, zero_rotation about the centre of mass, then T over 3N-3 degrees of freedom)
    and velocity all scale T."""
    s = S.layered_stack(15, 9, ["G", "BN"], seed=3, temperature=None)
    S.write_lammps_data(s, str(tmp_path / "in.data"))
    base = ["units metal", "atom_style full", "boundary p p f", "read_data in.data"]

    def velocities(lines):
        r = run(tmp_path, "\n".join(base + lines + ["write_data out.data", ""]))
        assert r.returncode == 0, r.stdout + r.stderr
        o = S.read_lammps_data(str(tmp_path / "out.data"))
        return o.v[np.argsort(np.argsort(s.tag))] if o.v is not None else None

    m = np.sqrt(s.masses())[:, None]
    v = velocities(["velocity all create 300.0 4928459 rot yes dist gaussian"])
    t, p, L = _temp_mom_angmom(s, v)
    assert abs(t - 300.0) < 1e-9 and np.abs(p).max() < 1e-10 and np.abs(L).max() < 1e-8
    w = (v * m).ravel()
    assert np.abs(w).max() / w.std() > 3.0          # Gaussian draws
    v = velocities(["velocity all create 300.0 4928459"])
    t, p, L = _temp_mom_angmom(s, v)
    assert abs(t - 300.0) < 1e-9 and np.abs(p).max() < 1e-10 and np.abs(L).max() > 1e-3
    w = (v * m).ravel()
    assert np.abs(w).max() / w.std() < 1.8          # uniform draws (the default)
    v = velocities(["velocity all create 300.0 7 mom no", "velocity all scale 450.0"])
    t, p, L = _temp_mom_angmom(s, v)
    assert abs(t - 450.0) < 1e-9 and np.abs(p).max() > 1e-6
    r = run(tmp_path, "\n".join(base + ["velocity all create 300.0 7 rot maybe", ""]))
    assert r.returncode != 0 and "Illegal velocity command" in r.stderr
