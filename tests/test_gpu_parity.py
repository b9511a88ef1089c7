"""GPU parity of the HIP path against the oracle, through the C ABI.

Tolerances (stated here, used below):
  FP64:  max_i |F_gpu - F_ref| / RMS(|F_ref|) <= 1e-10; |dE|/|E| <= 1e-10 for the total and
         each of E_vdW, E_rep, E_Tersoff; max |d virial| / max |virial| <= 1e-10;
         per-atom energy and virial: max |d| <= 1e-10 * max |ref|.
  mixed: RMS_i |F_mixed - F_ref| <= 1e-5 eV/A, |dE|/|E| <= 1e-6.
  neighbor sets: bit-exact (sorted (j, image shift) rows).
"""
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN, POT
from lmff import capi, structures as S

pytestmark = pytest.mark.gpu

CT = os.path.join(POT, "C.opt.tersoff")
BNC = os.path.join(POT, "BNC.tersoff")
ILP = os.path.join(POT, "BNCH.ILP")


def oracle_for(s, tap=1, telems=None, ielems=None, threads=8):
    import oracle

    if telems is None:
        if s.elements == ["C"]:
            telems, ielems, tf = ["C(O)"], ["C"], CT
        else:
            telems, ielems, tf = list(s.elements), list(s.elements), BNC
    else:
        tf = CT if "C(O)" in telems else BNC
    o = oracle.Oracle(len(telems), tf, telems, ILP, ielems, tap_flag=tap, threads=threads)
    o.set_system(s)
    o.build_list()
    return o


def ctx_for(s, precision=capi.LMFF_FP64, tap=1, telems=None, ielems=None, masses=None):
    ctx = capi.Context(0)
    if telems is None:
        capi.setup_for(ctx, s, precision=precision, tap=tap)
    else:
        capi.setup_graphene(ctx, len(telems), tuple(telems), tuple(ielems),
                            "C.opt.tersoff" if "C(O)" in telems else "BNC.tersoff",
                            tuple(masses), tap=tap, precision=precision)
    return ctx


def ribbons():
    s = S.layered_stack(12, 9, ["G", "G"], seed=12, jitter=0.06)
    s.boxhi[0] += 30.0
    s.x[:, 0] += 5.0
    return s


def reference_stack():
    s = S.read_lammps_data(os.path.join(GOLDEN, "inf0.data"))
    s.elements = ["C"] * 6
    return s


CASES = {
    "c1_bilayer_graphene_4800": lambda: S.config_c1(),
    "graphene_hbn_bilayer_4800": lambda: S.layered_stack(30, 20, ["G", "BN"], seed=5),
    "trilayer_g_bn_g_3240": lambda: S.layered_stack(15, 9, ["G", "BN", "G"], seed=7),
    "ribbons_open_edges": ribbons,
    "twisted_bilayer_5_3": lambda: S.twisted_bilayer(5, 3, seed=3),
}


def _check_fp64(g, r):
    rms = math.sqrt((r["f"] ** 2).sum(1).mean())
    assert np.abs(g["f"] - r["f"]).max() / rms <= 1e-10
    for k in ("energy", "e_vdw", "e_rep", "e_tersoff"):
        assert abs(g[k] - r[k]) <= 1e-10 * abs(r[k]), k
    assert np.abs(g["virial"] - r["virial"]).max() <= 1e-10 * np.abs(r["virial"]).max()


@pytest.mark.parametrize("name", sorted(CASES))
def test_fp64_forces_energy_virial(name):
    s = CASES[name]()
    r = oracle_for(s).compute(3, 6)
    with ctx_for(s) as ctx:
        g = ctx.pair_compute(s, 3, 6)
    _check_fp64(g, r)
    assert np.abs(g["eatom"] - r["eatom"]).max() <= 1e-10 * np.abs(r["eatom"]).max()
    assert np.abs(g["vatom"] - r["vatom"]).max() <= 1e-10 * np.abs(r["vatom"]).max()


SIX_LAYER_MAPS = {
    # every layer in both sub-styles
    "all": (["C(O)"] * 6, ["C"] * 6),
    # Cases/case_5016/in.replicate.ilp_tersoff: the bottom layer (type 1) has no Tersoff
    # but full ILP
    "stock_null": (["NULL"] + ["C(O)"] * 5, ["C"] * 6),
    # Cases/case_5016/in.replicate (pair_style lmff): the fused style's skip list drops
    # type 1 from ILP too, so the effective ILP map is NULL for it
    "lmff_null": (["NULL"] + ["C(O)"] * 5, ["NULL"] + ["C"] * 5),
}


@pytest.mark.parametrize("tap,maps", [(0, "all"), (1, "all"), (0, "stock_null"), (0, "lmff_null")])
def test_fp64_reference_six_layer_deck(tap, maps):
    """The reference's own 5,016-atom six-layer graphene data file (periodic in x, y and z),
    with the element maps of its decks."""
    s = reference_stack()
    el_t, el_i = SIX_LAYER_MAPS[maps]
    r = oracle_for(s, tap=tap, telems=el_t, ielems=el_i).compute(3, 6)
    with ctx_for(s, tap=tap, telems=el_t, ielems=el_i, masses=[12.0107] * 6) as ctx:
        g = ctx.pair_compute(s, 3, 6)
    _check_fp64(g, r)
    assert np.abs(g["vatom"] - r["vatom"]).max() <= 1e-10 * np.abs(r["vatom"]).max()


def test_fp64_asymmetric_pair_parameters(tmp_path):
    """Both shipped ILP files give (ti,tj) and (tj,ti) the same parameters, for which the pair
    kernel has a specialised variant.  Here the B-C and N-C rows differ from C-B and C-N (z0,
    lambda, C6, d, s_R), so the general variant runs: the vdW parameters follow the tag-parity
    half selection and exp(-lambda (r - z0)) is taken separately for (i,j) and (j,i).  Likewise
    the Tersoff B N N entry's repulsion (A, lambda1) differs from N B B's, so the bonds kernel
    picks each B-N pair's repulsive parameters by tag parity as the reference does."""
    import asym_tables

    tasym, asym = asym_tables.write(tmp_path)
    import oracle

    s = S.layered_stack(30, 20, ["G", "BN"], seed=17)
    o = oracle.Oracle(3, tasym, s.elements, asym, s.elements, tap_flag=1, threads=8)
    o.set_system(s)
    o.build_list()
    r = o.compute(3, 6)
    with capi.Context(0) as ctx:
        ctx.settings(3, 16.0, 1)
        ctx.coeff_tersoff(tasym, list(s.elements))
        ctx.coeff_ilp(asym, list(s.elements))
        ctx.neighbor(2.0, 10, 0, 0)
        ctx.masses(list(s.mass_per_type))
        g = ctx.pair_compute(s, 3, 6)
    _check_fp64(g, r)
    # the Tersoff asymmetry matters too
    o2 = oracle.Oracle(3, BNC, s.elements, asym, s.elements, tap_flag=1, threads=8)
    o2.set_system(s)
    o2.build_list()
    assert abs(r["energy"] - o2.compute(1, 0)["energy"]) > 1e-6
    assert np.abs(g["eatom"] - r["eatom"]).max() <= 1e-10 * np.abs(r["eatom"]).max()
    assert np.abs(g["vatom"] - r["vatom"]).max() <= 1e-10 * np.abs(r["vatom"]).max()
    # the asymmetry matters: the symmetric file gives a different energy
    assert abs(r["energy"] - oracle_for(s).compute(1, 0)["energy"]) > 1e-6


@pytest.mark.parametrize("name", sorted(CASES))
def test_mixed_force_error(name):
    s = CASES[name]()
    r = oracle_for(s).compute(1, 2)
    with ctx_for(s, precision=capi.LMFF_MIXED) as ctx:
        g = ctx.pair_compute(s, 1, 2)
    rms_err = math.sqrt(((g["f"] - r["f"]) ** 2).sum(1).mean())
    assert rms_err <= 1e-5, rms_err
    assert abs(g["energy"] - r["energy"]) <= 1e-6 * abs(r["energy"])


def _periodic_z_stack():
    """12 alternating sheets 3.35 A apart in a box periodic in z (40.2 A): stencil rows wrap in z."""
    s = S.layered_stack(15, 9, ["G", "BN"] * 6, seed=23, jitter=0.07)
    s.boxlo[2], s.boxhi[2] = -1.675, 12 * 3.35 - 1.675
    s.periodic = (1, 1, 1)
    return s


def _bilayers_far_apart():
    """Two bilayers 40 A apart in a non-periodic z: 16 z-bins, empty rows of bins between them."""
    s = S.layered_stack(15, 9, ["G", "BN", "G", "BN"], seed=24, jitter=0.07)
    s.x[s.mol >= 3, 2] += 40.0
    s.boxhi[2] += 40.0
    return s


@pytest.mark.parametrize("geometry", ["trilayer", "periodic_z", "bilayers_far_apart"])
def test_neighbor_sets_bit_exact(geometry):
    s = {"trilayer": lambda: S.layered_stack(15, 9, ["G", "BN", "G"], seed=21, jitter=0.07),
         "periodic_z": _periodic_z_stack, "bilayers_far_apart": _bilayers_far_apart}[geometry]()
    print(f"neighbor sets, {geometry}: {s.n} atoms", flush=True)
    o = oracle_for(s)
    offs, ent, sh = o.neighbor_list()
    want = {(i, int(j), *map(int, sh[e])) for i in range(s.n)
            for e, j in zip(range(offs[i], offs[i + 1]), ent[offs[i]:offs[i + 1]]) if s.mol[j] != s.mol[i]}
    with ctx_for(s) as ctx:
        g = ctx.pair_compute(s, 1, 2)
        # every atom's ILP row: the rows of a half-pair system hold each pair once (from the atom
        # in the lower z bin; within one z bin from both atoms), so the entries with their mirror
        # images (j, i, -shift) must make up the reference's full list exactly
        got = set()
        for i in range(s.n):
            for j, sx, sy, sz in ctx.ilp_list(0, i).tolist():
                got.add((i, j, sx, sy, sz))
                got.add((j, i, -sx, -sy, -sz))
    assert got == want
    # the short lists (Tersoff, normals) through the forces of every atom
    _check_fp64(g, o.compute(1, 2))


def test_normals_match_oracle():
    s = ribbons()
    nref, cref, _, _ = oracle_for(s).normals()
    with ctx_for(s) as ctx:
        ctx.pair_compute(s, 1, 0)
        nv, cnt = ctx.normals(0)
    assert (cnt == cref).all()
    np.testing.assert_allclose(np.abs((nv * nref).sum(1)), 1.0, atol=1e-13)


def test_empty_and_ragged_systems():
    s = S.layered_stack(15, 9, ["G", "G"], seed=1)
    with ctx_for(s) as ctx:
        e = S.System(np.zeros((0, 3)), np.zeros(0, np.int32), np.zeros(0, np.int64), np.zeros(0, np.int64),
                     s.boxlo.copy(), s.boxhi.copy(), s.periodic, ["C"])
        g = ctx.pair_compute(e, 3, 6)
        assert g["energy"] == 0.0 and g["f"].shape == (0, 3)
        # ragged: two systems of different size resident together evaluate like each alone
        a = S.layered_stack(15, 9, ["G", "G"], seed=2)
        b = S.layered_stack(17, 10, ["G", "G"], seed=3)
        ia = ctx.add_system(a)
        ib = ctx.add_system(b)
        ctx.evaluate(1, 2)
        ta, tb = ctx.thermo(ia), ctx.thermo(ib)
        _, _, fa = ctx.get(ia)
        _, _, fb = ctx.get(ib)
    for sysm, t, f in ((a, ta, fa), (b, tb, fb)):
        r = oracle_for(sysm).compute(1, 2)
        assert abs(t["pe"] - r["energy"]) <= 1e-10 * abs(r["energy"])
        assert np.abs(f - r["f"]).max() / math.sqrt((r["f"] ** 2).sum(1).mean()) <= 1e-10


def test_resident_step_matches_oracle_nve():
    s = S.layered_stack(15, 9, ["G", "BN"], seed=31)
    o = oracle_for(s)
    x_ref, v_ref, th = o.run_nve(s.x, s.v, s.mass_per_type, 0.001, 40, rebuild_every=10, thermo_every=10)
    with ctx_for(s) as ctx:
        sid = ctx.add_system(s)
        ctx.integrator(sid, capi.LMFF_NVE, dt=0.001)
        ctx.evaluate(1, 2)
        rec = ctx.run(40, 10, use_graph=True)[0]
        x, v, f = ctx.get(sid)
    L = (s.boxhi - s.boxlo)
    d = x - x_ref
    d[:, :2] -= L[:2] * np.round(d[:, :2] / L[:2])
    assert np.abs(d).max() < 1e-9
    assert np.abs(v - v_ref).max() < 1e-7
    pe = np.array([r["pe"] for r in rec])
    np.testing.assert_allclose(pe, th[1:, 1], rtol=1e-10)


def test_pairs_entering_from_beyond_the_split_between_rebuilds():
    """3000 K and one list build for 60 steps: atoms move well over 0.3 A, so pairs that were
    beyond rc + 0.6 A at the build come within rc before the next build (the ILP rows' back
    segment must then be walked).  Positions and energies follow the oracle run with the same
    schedule."""
    s = S.layered_stack(15, 9, ["G", "BN"], seed=51, temperature=3000.0)
    o = oracle_for(s)
    x_ref, v_ref, th = o.run_nve(s.x, s.v, s.mass_per_type, 0.001, 60, rebuild_every=60, thermo_every=20)
    ctx = capi.Context(0)
    try:
        capi.setup_for(ctx, s, every=60, delay=0, check=0)
        sid = ctx.add_system(s)
        ctx.integrator(sid, capi.LMFF_NVE, dt=0.001)
        ctx.evaluate(1, 2)
        rec = ctx.run(60, 20, use_graph=True)[0]
        x, v, f = ctx.get(sid)
    finally:
        ctx.close()
    L = (s.boxhi - s.boxlo)
    d = x - x_ref
    d[:, :2] -= L[:2] * np.round(d[:, :2] / L[:2])
    disp = s.x - x_ref
    disp[:, :2] -= L[:2] * np.round(disp[:, :2] / L[:2])
    assert np.sqrt((disp ** 2).sum(1)).max() > 0.5
    assert np.abs(d).max() < 1e-8
    pe = np.array([r["pe"] for r in rec])
    np.testing.assert_allclose(pe, th[1:, 1], rtol=1e-10)


def test_fp64_nve_drift_10k_matches_oracle():
    """10,000 FP64 NVE steps of the 4,800-atom AB bilayer (dt 1 fs, rebuild every 10 steps)
    against the oracle's record of the same run (tests/golden/nve_drift_c1.json).  While
    roundoff differences are small the two runs are one trajectory; later they separate (the
    dynamics is chaotic), so the total-energy error is compared by its statistics."""
    import json

    with open(os.path.join(GOLDEN, "nve_drift_c1.json")) as f:
        fx = json.load(f)
    s = S.config_c1()
    with ctx_for(s) as ctx:
        sid = ctx.add_system(s)
        ctx.integrator(sid, capi.LMFF_NVE, dt=fx["dt_ps"])
        ctx.evaluate(1, 2)
        th0 = ctx.thermo(sid)
        rec = ctx.run(fx["nsteps"], fx["thermo_every"], use_graph=True)[0]
    e_gpu = np.array([th0["pe"] + th0["ke"]] + [r["pe"] + r["ke"] for r in rec]) / s.n
    e_orc = (np.array(fx["pe"]) + np.array(fx["ke"])) / s.n
    assert len(e_gpu) == len(e_orc)
    early = np.abs(e_gpu[:5] - e_orc[:5])
    print("first 200 steps |dE| per atom:", early.tolist())
    assert early.max() < 1e-8
    d_gpu, d_orc = e_gpu - e_gpu[0], e_orc - e_orc[0]
    half = len(e_orc) // 2
    sd = d_orc[half:].std()
    print("late mean gpu/oracle:", d_gpu[half:].mean(), d_orc[half:].mean(), "std:", d_gpu[half:].std(), sd)
    assert abs(d_gpu[half:].mean() - d_orc[half:].mean()) < 3 * sd + 1e-6
    assert 0.5 < d_gpu[half:].std() / sd < 2.0
    assert abs(np.abs(d_gpu).max() / np.abs(d_orc).max() - 1.0) < 0.5


def test_graph_replay_equals_eager_and_is_deterministic():
    s = S.layered_stack(15, 9, ["G", "BN"], seed=41)
    outs = []
    for graph in (True, False, True):
        with ctx_for(s) as ctx:
            sid = ctx.add_system(s)
            ctx.integrator(sid, capi.LMFF_NVE, dt=0.001)
            ctx.evaluate(0, 0)
            ctx.run(35, 0, use_graph=graph)
            outs.append(ctx.get(sid))
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("style", [capi.LMFF_NVE, capi.LMFF_LANGEVIN])
def test_alone_equals_in_mix_bitwise(style):
    systems = [S.twisted_bilayer(3, 2, seed=5), S.layered_stack(15, 9, ["G", "G"], seed=6),
               S.twisted_bilayer(4, 1, seed=7)]
    alone = []
    for s in systems:
        with ctx_for(s) as ctx:
            sid = ctx.add_system(s)
            ctx.integrator(sid, style, dt=0.001, t_target=300.0, damp=0.1, seed=99)
            ctx.evaluate(0, 0)
            ctx.run(30, 10)
            alone.append(ctx.get(sid))
    with ctx_for(systems[0]) as ctx:
        ids = []
        for s in systems:
            ids.append(ctx.add_system(s))
            ctx.integrator(ids[-1], style, dt=0.001, t_target=300.0, damp=0.1, seed=99)
        ctx.evaluate(0, 0)
        ctx.run(30, 10)
        mixed = [ctx.get(i) for i in ids]
    for a, m in zip(alone, mixed):
        for p, q in zip(a, m):
            assert np.array_equal(p, q)


def test_langevin_drives_temperature():
    s = S.layered_stack(15, 9, ["G", "G"], seed=8, temperature=50.0)
    with ctx_for(s) as ctx:
        sid = ctx.add_system(s)
        ctx.integrator(sid, capi.LMFF_LANGEVIN, dt=0.001, t_target=600.0, damp=0.05, seed=3)
        ctx.evaluate(0, 0)
        rec = ctx.run(600, 100)[0]
    assert rec[0]["temp"] < rec[-1]["temp"]
    assert 400 < np.mean([r["temp"] for r in rec[-2:]]) < 800


def test_error_paths():
    s = S.layered_stack(15, 9, ["G", "G"], seed=9)
    with capi.Context(0) as ctx:
        ctx.settings(1, 16.0, 1)
        with pytest.raises(capi.LmffError) as e:
            ctx.coeff_tersoff(os.path.join(POT, "missing.tersoff"), ["C(O)"])
        assert e.value.code == 2
        with pytest.raises(capi.LmffError) as e:
            ctx.coeff_tersoff(BNC, ["C(O)"])  # no C(O) entries in BNC.tersoff
        assert e.value.code == 3
    with ctx_for(s) as ctx:
        small_box = s.copy()
        small_box.boxhi[0] = 30.0
        with pytest.raises(capi.LmffError) as e:
            ctx.pair_compute(small_box)
        assert e.value.code == 9
        bad = s.copy()
        bad.x[5, 0] = np.nan
        with pytest.raises(capi.LmffError) as e:
            ctx.pair_compute(bad)
        assert e.value.code == 7
        # an atom pushed into its layer gets more than three neighbours within 2 A
        squeezed = s.copy()
        squeezed.x[0, :2] = squeezed.x[1, :2] + np.array([0.3, 0.2])
        squeezed.x[0, 2] = squeezed.x[1, 2]
        with pytest.raises(capi.LmffError) as e:
            ctx.pair_compute(squeezed)
        assert e.value.code in (5, 6)
        # the context stays usable after an error
        g = ctx.pair_compute(s, 1, 2)
        assert np.isfinite(g["energy"])


def test_full_size_c2_properties():
    """228,800-atom graphene/hBN bilayer: size-independent checks at the benchmark size."""
    s = S.config_c2()
    with ctx_for(s) as ctx:
        g = ctx.pair_compute(s, 1, 2)
    with ctx_for(s, precision=capi.LMFF_MIXED) as ctx:
        m = ctx.pair_compute(s, 1, 2)
    assert np.abs(g["f"].sum(0)).max() < 1e-8 * s.n
    assert -7.9 < g["energy"] / s.n < -7.5  # the 4,800-atom cut of the same stack gives -7.716
    rms_err = math.sqrt(((m["f"] - g["f"]) ** 2).sum(1).mean())
    assert rms_err <= 1e-5
    assert abs(m["energy"] - g["energy"]) <= 1e-6 * abs(g["energy"])


def test_pair_compute_list_reuse_follows_the_oracle():
    """LAMMPS' loop through the Pair boundary: positions uploaded every step, the list built
    once (new_list=1) and reused (new_list=0) for 30 steps at 3000 K, so atoms move past the ILP
    rows' split margin and the far flag must be raised from the uploaded positions.  Velocity
    Verlet on the host; positions and energies follow the oracle's run with the same schedule."""
    s = S.layered_stack(15, 9, ["G", "BN"], seed=81, temperature=3000.0)
    o = oracle_for(s)
    x_ref, v_ref, th = o.run_nve(s.x, s.v, s.mass_per_type, 0.001, 30, rebuild_every=30, thermo_every=10)
    dt, ftm2v = 0.001, 1.0 / 1.0364269e-4
    m = s.masses()[:, None]
    x, v = s.x.copy(), s.v.copy()
    pe = []
    with ctx_for(s) as ctx:
        f = ctx.pair_compute(s, 1, 0, new_list=1, x=x)["f"]
        for step in range(1, 31):
            v += 0.5 * dt * ftm2v / m * f
            x += dt * v
            r = ctx.pair_compute(s, 1, 0, new_list=0, x=x)
            f = r["f"]
            v += 0.5 * dt * ftm2v / m * f
            if step % 10 == 0:
                pe.append(r["energy"])
    L = s.boxhi - s.boxlo
    d = x - x_ref
    d[:, :2] -= L[:2] * np.round(d[:, :2] / L[:2])
    disp = x - s.x
    assert np.sqrt((disp ** 2).sum(1)).max() > 0.5
    assert np.abs(d).max() < 1e-8
    np.testing.assert_allclose(pe, th[1:, 1], rtol=1e-10)


def test_pair_compute_fresh_lists_with_reordered_atoms():
    """LAMMPS reneighbouring through the Pair boundary: a fresh list every 10 steps, each time
    with the atoms handed over in a new order (as atom sorting and exchange reorder them), the
    other steps reusing the list.  Fresh calls for the same atom count and box reload the
    system in place; positions and energies follow the oracle."""
    s = S.layered_stack(15, 9, ["G", "BN"], seed=82, temperature=3000.0)
    o = oracle_for(s)
    x_ref, v_ref, th = o.run_nve(s.x, s.v, s.mass_per_type, 0.001, 30, rebuild_every=10, thermo_every=10)
    dt, ftm2v = 0.001, 1.0 / 1.0364269e-4
    rng = np.random.default_rng(5)
    x, v = s.x.copy(), s.v.copy()
    t = s.copy()
    pe = []
    with ctx_for(s) as ctx:
        f = None
        for step in range(0, 31):
            if step > 0:
                v += 0.5 * dt * ftm2v / t.masses()[:, None] * f
                x += dt * v
            fresh = step % 10 == 0
            if fresh:
                perm = rng.permutation(s.n)
                x, v = x[perm], v[perm]
                t.type, t.tag, t.mol = t.type[perm], t.tag[perm], t.mol[perm]
            r = ctx.pair_compute(t, 1, 0, new_list=int(fresh), x=x)
            f = r["f"]
            if step > 0:
                v += 0.5 * dt * ftm2v / t.masses()[:, None] * f
            if step % 10 == 0 and step > 0:
                pe.append(r["energy"])
    order = np.argsort(t.tag)
    L = s.boxhi - s.boxlo
    d = x[order] - x_ref
    d[:, :2] -= L[:2] * np.round(d[:, :2] / L[:2])
    assert np.abs(d).max() < 1e-8
    np.testing.assert_allclose(pe, th[1:, 1], rtol=1e-10)


def test_image_flags_set_get_and_range():
    """lmff_system_set_image / get_image: a round trip in the caller's atom order, kept through
    a short run in which no atom crosses the box, and -512..511 enforced."""
    s = S.layered_stack(15, 9, ["G", "BN"], seed=83, temperature=10.0)
    rng = np.random.default_rng(2)
    img = rng.integers(-40, 40, size=(s.n, 3)).astype(np.int32)
    img[:, 2] = 0
    with ctx_for(s) as ctx:
        sid = ctx.add_system(s)
        ctx.integrator(sid, capi.LMFF_NVE, dt=0.001)
        ctx.set_image(sid, img)
        np.testing.assert_array_equal(ctx.get_image(sid), img)
        x0, _, _ = ctx.get(sid)
        ctx.run(20, 20)
        x1, _, _ = ctx.get(sid)
        lo, hi = s.boxlo[:2] + 0.2, s.boxhi[:2] - 0.2
        inside = np.all((x0[:, :2] > lo) & (x0[:, :2] < hi), axis=1)
        assert inside.sum() > s.n // 2
        np.testing.assert_array_equal(ctx.get_image(sid)[inside], img[inside])
        bad = img.copy()
        bad[0, 0] = 512
        with pytest.raises(capi.LmffError):
            ctx.set_image(sid, bad)
