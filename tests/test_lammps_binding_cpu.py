"""The LAMMPS-side binding (lammps-lmff_amd/lammps/): the style names an unchanged deck reaches
with the "hip" suffix, and a syntax check of the binding against the LAMMPS headers of the
reference tree where that tree is present (the build container; skipped elsewhere).  Nothing of
the reference is copied or shipped: its headers are only read by the compiler.

With lmp -sf hip, PairHybrid::settings creates every sub-style through
Force::new_pair(style, trysuffix=1), which looks up "<style>/hip" first
(LMFF/pair_hybrid.cpp:334, LMFF/force.cpp:255-272), so the two deck forms
  hybrid/overlay tersoff ilp/graphene/hbn   and   hybrid/overlay lmff ilp/graphene/hbn
need tersoff/hip, lmff/hip and ilp/graphene/hbn/hip."""
import glob
import os
import re
import shutil
import subprocess

import pytest

from conftest import REPO

BIND = os.path.join(REPO, "lammps-lmff_amd", "lammps")
LAMMPS_SRC = "/root/reference/LMFF"


def styles():
    out = {}
    for h in glob.glob(os.path.join(BIND, "*.h")):
        for name, cls in re.findall(r"PairStyle\(([^,]+),(\w+)\)", open(h).read()):
            out[name] = cls
    return out


def test_suffix_reaches_every_sub_style_of_both_deck_forms():
    st = styles()
    for deck in (["tersoff", "ilp/graphene/hbn"], ["lmff", "ilp/graphene/hbn"]):
        for sub in deck:
            assert f"{sub}/hip" in st, sub


def test_binding_calls_only_declared_entry_points():
    hdr = open(os.path.join(REPO, "include", "lmff_hip.h")).read()
    declared = set(re.findall(r"\b(lmff_[a-z_0-9]+)\s*\(", re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)))
    for src in glob.glob(os.path.join(BIND, "*.cpp")):
        used = set(re.findall(r"\b(lmff_[a-z_0-9]+)\s*\(", open(src).read()))
        assert used <= declared, (src, used - declared)


@pytest.mark.skipif(not os.path.exists(os.path.join(LAMMPS_SRC, "pair.h")) or not shutil.which("g++"),
                    reason="LAMMPS headers of the reference tree not present")
@pytest.mark.parametrize("src", ["pair_tersoff_hip.cpp", "pair_lmff_hip.cpp", "pair_ilp_graphene_hbn_hip.cpp"])
def test_binding_compiles_against_lammps_headers(src):
    cmd = ["g++", "-std=c++11", "-fsyntax-only", "-Wall", "-Werror", f"-I{LAMMPS_SRC}",
           f"-I{os.path.join(LAMMPS_SRC, 'STUBS')}", f"-I{os.path.join(REPO, 'include')}", f"-I{BIND}",
           os.path.join(BIND, src)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
