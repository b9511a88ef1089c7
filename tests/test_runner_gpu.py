"""The deck runner (lmff_run) end to end on the GPU: dump custom with compute pe/atom and
stress/atom, write_data, and a restart from that file with reset_timestep.  The dumped
per-atom columns are checked against Pair::compute (lmff_pair_compute) on the dumped
positions; the restarted run must follow the uninterrupted one."""
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, POT, REPO
from lmff import capi, structures as S

pytestmark = pytest.mark.gpu

RUNNER = os.path.join(REPO, "lammps-lmff_amd", "lmff_run")
NKTV2P, MVV2E = 1.6021765e6, 1.0364269e-4


def deck(data, extra_top, body):
    return "\n".join([
        "units metal", "atom_style full", "boundary p p f", extra_top,
        f"read_data {data}",
        "pair_style hybrid/overlay tersoff ilp/graphene/hbn 16.0 1",
        f"pair_coeff * * tersoff {os.path.join(POT, 'BNC.tersoff')} C B N",
        f"pair_coeff * * ilp/graphene/hbn {os.path.join(POT, 'BNCH.ILP')} C B N",
        "neighbor 2.0 bin", "neigh_modify every 10 delay 0 check no", "timestep 0.001",
        "fix 1 all nve", body, ""])


def run_deck(tmp_path, name, text):
    path = tmp_path / name
    path.write_text(text)
    print("running", name, flush=True)
    r = subprocess.run([RUNNER, "-in", str(path)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_dump_computes_write_data_and_restart(tmp_path):
    if not os.path.exists(RUNNER):
        pytest.fail("lmff_run is not built (python lammps-lmff_amd/build.py)")
    s = S.layered_stack(15, 9, ["G", "BN"], seed=61)
    data = str(tmp_path / "start.data")
    S.write_lammps_data(s, data)
    out = run_deck(tmp_path, "a.in", deck(data, "", "\n".join([
        "compute pe all pe/atom", "compute st all stress/atom NULL",
        "dump d1 all custom 10 a.dump id type x y z vx vy vz fx fy fz c_pe c_st[*]",
        "dump_modify d1 format float %.17g", "thermo 10", "run 20", "write_data mid.data", "run 20"])))
    frames = S.read_dump_custom(str(tmp_path / "a.dump"))
    assert [f[0] for f in frames] == [0, 10, 20, 30, 40]
    cols = frames[0][1]
    assert cols[:11] == ["id", "type", "x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz"]
    assert cols[11:] == ["c_pe"] + [f"c_st[{k}]" for k in range(1, 7)]
    # E_pair of step 0 from the thermo table equals the sum of c_pe
    e0 = float([l for l in out.splitlines() if l.strip().startswith("0 ")][0].split()[2])
    assert abs(frames[0][2][:, 11].sum() - e0) < 1e-9 * abs(e0)

    # per-atom columns of step 20 against one Pair::compute on the dumped positions
    step, _, a = frames[2]
    t = s.copy()
    order = np.argsort(s.tag)
    t.x = a[:, 2:5]
    t.tag, t.type, t.mol = s.tag[order], s.type[order], s.mol[order]
    with capi.Context(0) as ctx:
        capi.setup_for(ctx, t)
        r = ctx.pair_compute(t, 3, 6)
    m = t.mass_per_type[t.type - 1]
    fscale = np.abs(r["f"]).max()
    assert np.abs(a[:, 8:11] - r["f"]).max() < 1e-9 * fscale
    assert np.abs(a[:, 11] - r["eatom"]).max() < 1e-10 * np.abs(r["eatom"]).max()
    v = a[:, 5:8]
    ia, ib = [0, 1, 2, 0, 0, 1], [0, 1, 2, 1, 2, 2]
    stress = -NKTV2P * (MVV2E * m[:, None] * v[:, ia] * v[:, ib] + r["vatom"])
    assert np.abs(a[:, 12:18] - stress).max() < 1e-9 * np.abs(stress).max()

    # write_data at step 20 holds exactly the dumped state
    mid = S.read_lammps_data(str(tmp_path / "mid.data"))
    assert open(tmp_path / "mid.data").readline().strip().endswith("timestep = 20")
    assert np.array_equal(mid.x, a[:, 2:5]) and np.array_equal(mid.v, a[:, 5:8])

    # restart: read the mid-run file, reset the step, run the remaining 20 steps
    run_deck(tmp_path, "b.in", deck(str(tmp_path / "mid.data"), "", "\n".join([
        "reset_timestep 20", "dump d2 all custom 20 b.dump id x y z", "dump_modify d2 format float %.17g",
        "thermo 10", "run 20"])))
    fb = S.read_dump_custom(str(tmp_path / "b.dump"))
    assert [f[0] for f in fb] == [20, 40]
    assert np.array_equal(fb[0][2][:, 1:4], a[:, 2:5])
    assert np.abs(fb[1][2][:, 1:4] - frames[4][2][:, 2:5]).max() < 1e-8


def test_minimize_from_a_deck(tmp_path):
    """min_style fire + minimize etol ftol maxiter maxeval in a deck: the log's Minimization
    stats (LAMMPS layout) report the FIRE run of the library."""
    if not os.path.exists(RUNNER):
        pytest.fail("lmff_run is not built (python lammps-lmff_amd/build.py)")
    s = S.layered_stack(15, 9, ["G", "BN"], seed=71, jitter=0.08)
    data = str(tmp_path / "start.data")
    S.write_lammps_data(s, data)
    out = run_deck(tmp_path, "m.in", deck(data, "", "\n".join([
        "min_style fire", "minimize 0.0 0.1 400 100000", "write_data relaxed.data"])))
    lines = out.splitlines()
    k = lines.index("Minimization stats:")
    assert lines[k + 1].strip() == "Stopping criterion = force tolerance"
    e = [float(t) for t in lines[k + 3].split()]
    it = [int(t) for t in lines[k + 6].split("=")[1].split()]
    with capi.Context(0) as ctx:
        capi.setup_for(ctx, s)
        sid = ctx.add_system(s)
        ctx.integrator(sid, capi.LMFF_NVE, dt=0.001)
        st = ctx.minimize(0.0, 0.1, 400)[0]
    assert it == [st["iterations"], st["evaluations"]]
    assert abs(e[2] - st["e_final"]) < 1e-12 * abs(e[2]) + 1e-9
    r = S.read_lammps_data(str(tmp_path / "relaxed.data"))
    assert open(tmp_path / "relaxed.data").readline().strip().endswith(f"timestep = {st['iterations']}")
    assert r.n == s.n


def test_default_cg_minimize_from_a_deck(tmp_path):
    """minimize 1.0e-5 1.0e-5 1000 10000 with the default min_style (cg), as in the reference's
    Cases decks: the log reports the library's CG run."""
    if not os.path.exists(RUNNER):
        pytest.fail("lmff_run is not built (python lammps-lmff_amd/build.py)")
    s = S.layered_stack(15, 9, ["G", "BN"], seed=71, jitter=0.08)
    data = str(tmp_path / "start.data")
    S.write_lammps_data(s, data)
    out = run_deck(tmp_path, "c.in", deck(data, "", "minimize 1.0e-5 1.0e-5 1000 10000"))
    lines = out.splitlines()
    k = lines.index("Minimization stats:")
    with capi.Context(0) as ctx:
        capi.setup_for(ctx, s)
        ctx.add_system(s)
        st = ctx.minimize_cg(1e-5, 1e-5, 1000, 10000)
    assert lines[k + 1].strip() == "Stopping criterion = " + st["stop_text"]
    assert [int(t) for t in lines[k + 6].split("=")[1].split()] == [st["iterations"], st["evaluations"]]


def six_layer_deck(pair_lines, extra=""):
    return "\n".join([
        "units metal", "atom_style full", "boundary p p p",
        f"read_data {os.path.join(GOLDEN, 'inf0.data')}", extra,
        *pair_lines, "neighbor 2.0 bin", "neigh_modify every 10 delay 0 check no", "timestep 0.001",
        "fix 1 all nve", "thermo 10", "run 0", ""])


def step0_epair(out):
    return float([l for l in out.splitlines() if l.strip().startswith("0 ")][0].split()[2])


def printed_close(printed, value):
    """The thermo table prints E_pair with 10 significant digits (%16.10g)."""
    ulp10 = 10.0 ** (math.floor(math.log10(abs(value))) - 9)
    return abs(printed - value) <= 0.5 * ulp10 + 1e-10 * abs(value)


def test_lmff_deck_form_and_replicate(tmp_path):
    """pair_style hybrid/overlay lmff ilp/graphene/hbn, as in Cases/case_5016/in.replicate,
    against the oracle with the maps that form implies; the stock tersoff form with the same
    pair_coeff lines; and replicate: two copies of the periodic cell hold exactly twice the
    energy."""
    if not os.path.exists(RUNNER):
        pytest.fail("lmff_run is not built (python lammps-lmff_amd/build.py)")
    import oracle

    ct, ilp = os.path.join(POT, "C.opt.tersoff"), os.path.join(POT, "BNCH.ILP")
    null_t = "NULL " + " ".join(["C(O)"] * 5)
    lmff = run_deck(tmp_path, "l.in", six_layer_deck([
        "pair_style hybrid/overlay lmff ilp/graphene/hbn 16.0 0",
        f"pair_coeff * * lmff {ct} {null_t}", f"pair_coeff * * ilp/graphene/hbn {ilp} C C C C C C"]))
    stock = run_deck(tmp_path, "s.in", six_layer_deck([
        "pair_style hybrid/overlay tersoff ilp/graphene/hbn 16.0 0",
        f"pair_coeff * * tersoff {ct} {null_t}", f"pair_coeff * * ilp/graphene/hbn {ilp} C C C C C C"]))
    s = S.read_lammps_data(os.path.join(GOLDEN, "inf0.data"))
    for out, el_i in ((lmff, ["NULL"] + ["C"] * 5), (stock, ["C"] * 6)):
        o = oracle.Oracle(6, ct, ["NULL"] + ["C(O)"] * 5, ilp, el_i, tap_flag=0, threads=8)
        o.set_system(s)
        o.build_list()
        e = o.compute(1, 0)["energy"]
        assert printed_close(step0_epair(out), e)
    # without NULL the two forms are one force field: identical logs
    full = [f"pair_coeff * * ilp/graphene/hbn {ilp} C C C C C C"]
    a = run_deck(tmp_path, "a.in", six_layer_deck(["pair_style hybrid/overlay lmff ilp/graphene/hbn 16.0 0",
                                                   f"pair_coeff * * lmff {ct} " + " ".join(["C(O)"] * 6)] + full))
    b = run_deck(tmp_path, "b.in", six_layer_deck(["pair_style hybrid/overlay tersoff ilp/graphene/hbn 16.0 0",
                                                   f"pair_coeff * * tersoff {ct} " + " ".join(["C(O)"] * 6)] + full))
    assert step0_epair(a) == step0_epair(b)
    # replicate 2 1 1 through an equal-style variable.  LAMMPS offsets molecule ids per copy
    # (replicate.cpp:560-571), so the two halves of each sheet become different molecules and
    # ILP acts across the seam: the energy is not twice the cell's, it is the oracle's on the
    # replicated atoms.
    r = run_deck(tmp_path, "r.in", six_layer_deck(
        ["pair_style hybrid/overlay lmff ilp/graphene/hbn 16.0 0",
         f"pair_coeff * * lmff {ct} " + " ".join(["C(O)"] * 6)] + full,
        extra="variable SX index 2\nvariable DX equal ${SX}*1\nreplicate ${DX} 1 1\nwrite_data rep.data"))
    assert "10032 atoms after replicate" in r
    rep = S.read_lammps_data(str(tmp_path / "rep.data"))
    assert rep.n == 10032 and sorted(set(rep.mol.tolist())) == list(range(1, 13))
    o = oracle.Oracle(6, ct, ["C(O)"] * 6, ilp, ["C"] * 6, tap_flag=0, threads=8)
    o.set_system(rep)
    o.build_list()
    assert printed_close(step0_epair(r), o.compute(1, 0)["energy"])


def test_dump_atom_style(tmp_path):
    """dump ID all atom N file (the form in the reference's Cases decks, with dump_modify sort 1
    first no): LAMMPS' header and 'id type xs ys zs' columns, scaled by the box; with
    dump_modify scale no the unscaled x y z."""
    if not os.path.exists(RUNNER):
        pytest.fail("lmff_run is not built (python lammps-lmff_amd/build.py)")
    s = S.layered_stack(15, 9, ["G", "BN"], seed=91)
    data = str(tmp_path / "start.data")
    S.write_lammps_data(s, data)
    run_deck(tmp_path, "d.in", deck(data, "", "\n".join([
        "dump 1 all atom 10 fric.lammpstrj", "dump_modify 1 sort 1 first no",
        "dump 2 all atom 10 plain.lammpstrj", "dump_modify 2 scale no",
        "dump 3 all custom 10 c.dump id type x y z", "dump_modify 3 format float %.17g",
        "thermo 10", "run 20"])))
    text = open(tmp_path / "fric.lammpstrj").read().splitlines()
    assert text[:4] == ["ITEM: TIMESTEP", "0", "ITEM: NUMBER OF ATOMS", str(s.n)]
    assert text[4] == "ITEM: BOX BOUNDS pp pp ff" and text[8] == "ITEM: ATOMS id type xs ys zs"
    scaled = S.read_dump_custom(str(tmp_path / "fric.lammpstrj"))
    plain = S.read_dump_custom(str(tmp_path / "plain.lammpstrj"))
    custom = S.read_dump_custom(str(tmp_path / "c.dump"))
    assert [f[0] for f in scaled] == [f[0] for f in custom] == [0, 10, 20]
    lo, L = s.boxlo, s.boxhi - s.boxlo
    for a, b, c in zip(scaled, plain, custom):
        assert a[1] == ["id", "type", "xs", "ys", "zs"] and b[1] == ["id", "type", "x", "y", "z"]
        assert np.array_equal(a[2][:, :2], c[2][:, :2])
        np.testing.assert_allclose(lo + a[2][:, 2:5] * L, c[2][:, 2:5], rtol=0, atol=1e-5 * L.max())
        np.testing.assert_allclose(b[2][:, 2:5], c[2][:, 2:5], rtol=1e-5, atol=1e-5)


def test_image_flags_and_binary_restart(tmp_path):
    """Atoms drifting across the periodic x and y boundaries: the dumped image flags satisfy
    xu = x + ix * Lx and the unwrapped positions follow the drift (a missed or doubled crossing
    would be off by a box length); write_restart at step 100 + read_restart continues to the same
    state and image flags as the uninterrupted 200-step run, whose restart 100 file of step 100
    is byte for byte the write_restart file."""
    if not os.path.exists(RUNNER):
        pytest.fail("lmff_run is not built (python lammps-lmff_amd/build.py)")
    s = S.layered_stack(15, 9, ["G", "BN"], seed=81)
    drift = np.array([25.0, -15.0, 0.0])  # A/ps: 5 A and -3 A over 0.2 ps
    s.v = s.v + drift
    data = str(tmp_path / "start.data")
    S.write_lammps_data(s, data)
    run_deck(tmp_path, "a.in", deck(data, "", "\n".join([
        "dump d all custom 200 a.dump id x y z xu yu zu ix iy iz", "dump_modify d format float %.17g",
        "restart 100 ra.*.bin", "thermo 100", "run 200", "write_data a.data"])))
    run_deck(tmp_path, "b.in", deck(data, "", "\n".join(["thermo 100", "run 100", "write_restart r.bin"])))
    run_deck(tmp_path, "c.in", "\n".join([
        "units metal", "atom_style full", "read_restart r.bin",
        "pair_style hybrid/overlay tersoff ilp/graphene/hbn 16.0 1",
        f"pair_coeff * * tersoff {os.path.join(POT, 'BNC.tersoff')} C B N",
        f"pair_coeff * * ilp/graphene/hbn {os.path.join(POT, 'BNCH.ILP')} C B N",
        "neighbor 2.0 bin", "neigh_modify every 10 delay 0 check no",
        "fix 1 all nve", "thermo 100", "run 100", "write_data c.data", ""]))
    # restart N file: written at steps 100 and 200 of the uninterrupted run; the one of step 100
    # holds exactly the state write_restart saved after the 100-step run
    assert (tmp_path / "ra.100.bin").read_bytes() == (tmp_path / "r.bin").read_bytes()
    assert (tmp_path / "ra.200.bin").exists() and not (tmp_path / "ra.0.bin").exists()
    frames = S.read_dump_custom(str(tmp_path / "a.dump"))
    assert [f[0] for f in frames] == [0, 200]
    a = frames[1][2]
    L = s.boxhi - s.boxlo
    img = a[:, 7:10].astype(np.int64)
    np.testing.assert_allclose(a[:, 4:7], a[:, 1:4] + img * L, rtol=0, atol=1e-9)
    order = np.argsort(s.tag)
    expect = s.x[order] + drift * 0.2
    assert np.abs(a[:, 4:6] - expect[:, :2]).max() < 1.0
    assert (img[:, 0] != 0).sum() >= 20 and (img[:, 1] != 0).sum() >= 20 and not img[:, 2].any()
    wa = S.read_lammps_data(str(tmp_path / "a.data"))
    np.testing.assert_array_equal(wa.meta["image"], img)
    wc = S.read_lammps_data(str(tmp_path / "c.data"))
    assert open(tmp_path / "c.data").readline().strip().endswith("timestep = 200")
    np.testing.assert_array_equal(wc.meta["image"], wa.meta["image"])
    assert np.abs(wc.x - wa.x).max() < 1e-8 and np.abs(wc.v - wa.v).max() < 1e-8


def test_velocity_scale_of_a_resident_system(tmp_path):
    """velocity all scale between runs acts on the device-resident state: the next run's first
    thermo line reports the new temperature, and the rotation-free create keeps the run stable."""
    if not os.path.exists(RUNNER):
        pytest.fail("lmff_run is not built (python lammps-lmff_amd/build.py)")
    s = S.layered_stack(15, 9, ["G", "BN"], seed=91, temperature=None)
    data = str(tmp_path / "start.data")
    S.write_lammps_data(s, data)
    out = run_deck(tmp_path, "v.in", deck(data, "", "\n".join([
        "velocity all create 300.0 4928459 rot yes dist gaussian", "thermo 10", "run 20",
        "velocity all scale 600.0", "run 10"])))
    rows = [l.split() for l in out.splitlines() if l.strip() and l.split()[0].isdigit() and len(l.split()) == 5]
    temps = {int(r[0]): float(r[1]) for r in rows}
    first = [float(r[1]) for r in rows if int(r[0]) == 0][0]
    assert abs(first - 300.0) < 1e-3
    # the second run starts at step 20 with the rescaled velocities
    starts = [float(r[1]) for r in rows if int(r[0]) == 20]
    assert len(starts) == 2 and abs(starts[1] - 600.0) < 1e-3 and abs(starts[0] - 600.0) > 1.0
