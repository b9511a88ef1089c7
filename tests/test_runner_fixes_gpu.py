"""The deck runner's fix bookkeeping, later parameter changes and neighbor statistics on the GPU.

  * fix nve / fix langevin are tracked by ID: unfix removes them, groups other than all, other
    fix styles and langevin keywords that change the dynamics end the run with an error instead
    of being skipped;
  * mass, pair_coeff, neighbor and lmff_precision given after the first run reach the library
    (a run after a mass change equals a run with those masses from the start, bit for bit);
  * Finish's neighbor lines: list builds and dangerous builds (neigh_modify check yes)."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import POT, REPO
from lmff import structures as S

pytestmark = pytest.mark.gpu

RUNNER = os.path.join(REPO, "lammps-lmff_amd", "lmff_run")


def header(data, neigh="neighbor 2.0 bin\nneigh_modify every 10 delay 0 check no"):
    return "\n".join([
        "units metal", "atom_style full", "boundary p p f", f"read_data {data}",
        "pair_style hybrid/overlay tersoff ilp/graphene/hbn 16.0 1",
        f"pair_coeff * * tersoff {os.path.join(POT, 'BNC.tersoff')} C B N",
        f"pair_coeff * * ilp/graphene/hbn {os.path.join(POT, 'BNCH.ILP')} C B N",
        neigh, "timestep 0.001"])


def run(tmp_path, name, text, ok=True):
    path = tmp_path / name
    path.write_text(text + "\n")
    print("running", name, flush=True)
    r = subprocess.run([RUNNER, "-in", str(path)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    if ok:
        assert r.returncode == 0, r.stdout + r.stderr
    return r


@pytest.fixture
def data(tmp_path):
    s = S.layered_stack(15, 9, ["G", "BN"], seed=71)
    p = str(tmp_path / "start.data")
    S.write_lammps_data(s, p)
    return p


def thermo_rows(out):
    return [l.split() for l in out.splitlines() if re.match(r"^\s+\d+\s+[-\d.eE+]+\s+[-\d.eE+]+", l)]


@pytest.mark.parametrize("fix,msg", [
    ("group l2 type 2\nfix 1 l2 nve", "group all only"),
    ("fix 1 all nve\nfix 2 all langevin 300 300 0.1 5 zero yes", "zero yes"),
    ("fix 1 all nve\nfix 2 all langevin 300 400 0.1 5", "Tstart = Tstop"),
    ("fix 1 all nve\nfix 2 all setforce 0 0 0", "not supported"),
    ("fix 1 all nve\nunfix 1\nrun 10", "needs fix ID all nve"),
    ("fix 1 all nve\nunfix 7", "Could not find fix ID 7"),
])
def test_unsupported_fixes_end_the_run(tmp_path, data, fix, msg):
    r = run(tmp_path, "bad.in", header(data) + "\n" + fix + "\nrun 10", ok=False)
    assert r.returncode != 0
    assert msg in r.stderr


def test_unfix_langevin_leaves_nve(tmp_path, data):
    out = run(tmp_path, "u.in", header(data) + "\n" + "\n".join([
        "fix 1 all nve", "fix 2 all langevin 300 300 0.05 5", "thermo 10", "run 40", "unfix 2", "run 40"])).stdout
    rows = thermo_rows(out)
    # two tables of five rows (steps 0..40 and 40..80): etotal drifts under the thermostat and is
    # conserved to the integrator's fluctuation once it is gone
    et = [float(r[3]) for r in rows]
    first, second = et[:5], et[5:]
    assert len(second) == 5
    assert max(second) - min(second) < 0.1 * (max(first) - min(first))


def test_later_masses_reach_the_library(tmp_path, data):
    heavy = "mass 1 24.0214\nmass 2 21.622\nmass 3 28.0134"
    a = run(tmp_path, "a.in", header(data) + "\n" + "\n".join([
        "fix 1 all nve", "run 0", heavy, "run 20", "write_data a.data"]))
    b = run(tmp_path, "b.in", header(data) + "\n" + "\n".join([
        heavy, "fix 1 all nve", "run 0", "run 20", "write_data b.data"]))
    sa = S.read_lammps_data(str(tmp_path / "a.data"))
    sb = S.read_lammps_data(str(tmp_path / "b.data"))
    assert np.array_equal(sa.x, sb.x) and np.array_equal(sa.v, sb.v)
    # and the masses matter: the data file's own masses give another state
    c = run(tmp_path, "c.in", header(data) + "\n" + "\n".join(["fix 1 all nve", "run 0", "run 20",
                                                                  "write_data c.data"]))
    sc = S.read_lammps_data(str(tmp_path / "c.data"))
    assert not np.array_equal(sa.v, sc.v)
    assert a.returncode == b.returncode == c.returncode == 0


def _neighbor_lines(out):
    builds = int(re.search(r"Neighbor list builds = (\d+)", out).group(1))
    m = re.search(r"Dangerous builds = (\d+)", out)
    return builds, (int(m.group(1)) if m else None), "Dangerous builds not checked" in out


def test_dangerous_builds_are_counted(tmp_path, data):
    hot = "velocity all create 3000 4928459 dist gaussian"
    # a thin skin at 3000 K, checked only every 5 steps: some atom has already moved more than
    # half the skin when the check first may trigger
    out = run(tmp_path, "d.in", header(data, "neighbor 0.3 bin\nneigh_modify every 5 delay 5 check yes") + "\n" +
              "\n".join([hot, "fix 1 all nve", "thermo 50", "run 100"])).stdout
    builds, danger, _ = _neighbor_lines(out)
    assert builds >= 10 and danger is not None and danger > 0
    # the default skin at 300 K: no dangerous build
    out = run(tmp_path, "e.in", header(data, "neighbor 2.0 bin\nneigh_modify every 1 delay 10 check yes") + "\n" +
              "\n".join(["fix 1 all nve", "thermo 50", "run 100"])).stdout
    builds, danger, _ = _neighbor_lines(out)
    assert danger == 0 and builds >= 1
    # check no: not checked
    out = run(tmp_path, "f.in", header(data) + "\n" + "\n".join(["fix 1 all nve", "thermo 50", "run 100"])).stdout
    builds, danger, unchecked = _neighbor_lines(out)
    assert unchecked and danger is None and builds == 11   # the set-up build + one every 10 steps
