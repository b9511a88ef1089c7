"""Pins the CPU oracle (oracle/lmff_oracle.c) before it is trusted as the parity checker.

Every check here is independent of the oracle's own formulas: closed forms, an all-pairs
numpy evaluation from the published functional forms (tests/refimpl.py), finite
differences, the strain derivative of the energy, brute-force neighbor sets and
invariants.  CPU only.
"""
import math
import sys
import os

import numpy as np
import pytest

from conftest import POT
import refimpl
from lmff import structures as S

CT = os.path.join(POT, "C.opt.tersoff")
BNC = os.path.join(POT, "BNC.tersoff")
ILP = os.path.join(POT, "BNCH.ILP")


def make_oracle(s, tap=1, threads=1):
    import oracle

    if s.elements == ["C"]:
        o = oracle.Oracle(1, CT, ["C(O)"], ILP, ["C"], tap_flag=tap, threads=threads)
    else:
        o = oracle.Oracle(len(s.elements), BNC, s.elements, ILP, s.elements, tap_flag=tap, threads=threads)
    o.set_system(s)
    o.build_list()
    return o


def small(layers=("G", "G"), seed=3, jitter=0.05, nx=15, ny=9):
    return S.layered_stack(nx, ny, list(layers), seed=seed, jitter=jitter)


# ---------------------------------------------------------------- Tap
def test_tap_closed_form(oracle_lib):
    xs = np.linspace(0.0, 0.999, 50)
    vals = np.array([oracle_lib.tap(16.0 * x, 16.0) for x in xs])
    coef = np.polyfit(xs, vals, 7)
    np.testing.assert_allclose(coef, [20, -70, 84, -35, 0, 0, 0, 1], atol=1e-7)
    assert oracle_lib.tap(0.0, 16.0) == 1.0
    assert oracle_lib.tap(16.0, 16.0) == 0.0 and oracle_lib.tap(17.0, 16.0) == 0.0
    # Tap' = Tap'' = Tap''' = 0 at x = 1 and dTap is the derivative
    for r in (1.0, 5.0, 11.3, 15.99):
        h = 1e-5
        fd = (oracle_lib.tap(r + h, 16.0) - oracle_lib.tap(r - h, 16.0)) / (2 * h)
        assert abs(fd - oracle_lib.dtap(r, 16.0)) < 1e-9
    assert abs(oracle_lib.dtap(16.0 - 1e-9, 16.0)) < 1e-20 * 1e9
    assert oracle_lib.tap(16.0 * (1 - 1e-3), 16.0) < 1e-8


# ---------------------------------------------------------------- parameter files
def test_ilp_parameters_parsed(oracle_lib):
    o = oracle_lib.Oracle(3, BNC, ["C", "B", "N"], ILP, ["C", "B", "N"])
    import ctypes as C

    out = np.zeros(16)
    n = oracle_lib.lib().orc_ilp_param(o.h, 0, out.ctypes.data_as(C.POINTER(C.c_double)))
    assert n == 9  # C/B/N pairs only; H lines are skipped
    # C-C row: z0 alpha delta eps(meV) C(meV) d sR reff C6(meV A^6) S rcut
    assert out[2] == 3.205843 and out[3] == 7.511126 and out[4] == 1.235334
    assert math.isclose(out[5], 1.528338e-5 * 1e-3, rel_tol=1e-15)       # epsilon in eV
    assert math.isclose(out[6], 37.530428e-3, rel_tol=1e-15)             # C in eV
    assert math.isclose(out[10], 25.714535, rel_tol=1e-15)               # C6 in eV A^6
    assert math.isclose(out[13], 1.0 / 1.235334**2, rel_tol=1e-15)       # delta^-2
    assert math.isclose(out[14], 7.511126 / 3.205843, rel_tol=1e-15)     # lambda
    assert math.isclose(out[15], 0.7954443 * 3.681440, rel_tol=1e-15)    # seff


def test_tersoff_parameters_parsed(oracle_lib):
    o = oracle_lib.Oracle(3, BNC, ["C", "B", "N"], ILP, ["C", "B", "N"])
    import ctypes as C

    out = np.zeros(22)
    n = oracle_lib.lib().orc_tersoff_param(o.h, 0, out.ctypes.data_as(C.POINTER(C.c_double)))
    assert n == 27
    # first BNC entry is N B B: n = 0.72751, R + D = 2.0
    assert out[9] == 0.72751 and out[13] == 1.95 and out[14] == 0.05
    assert math.isclose(out[17], 2.0**2, rel_tol=1e-15)
    nn = 0.72751
    assert math.isclose(out[18], (2 * nn * 1e-16) ** (-1 / nn), rel_tol=1e-14)
    assert math.isclose(out[19], (2 * nn * 1e-8) ** (-1 / nn), rel_tol=1e-14)
    assert o.cutmax == pytest.approx(2.1)


@pytest.mark.parametrize("bad", ["missing", "duplicate", "short"])
def test_bad_potential_files(oracle_lib, tmp_path, bad):
    lines = [l for l in open(BNC) if l.strip() and not l.startswith("#")]
    if bad == "missing":
        lines = lines[1:]
    elif bad == "duplicate":
        lines = lines + lines[:1]
    else:
        lines[0] = " ".join(lines[0].split()[:10]) + "\n"
    p = tmp_path / "x.tersoff"
    p.write_text("".join(lines))
    with pytest.raises(oracle_lib.OracleError):
        oracle_lib.Oracle(3, str(p), ["C", "B", "N"], ILP, ["C", "B", "N"])


# ---------------------------------------------------------------- closed forms
def test_tersoff_perfect_graphene_closed_form(oracle_lib):
    a = 2.46
    s = S.layered_stack(15, 9, ["G"], jitter=0.0, a=a, b=a * math.sqrt(3.0), temperature=None)
    o = make_oracle(s)
    r = o.compute(1, 2)
    P = refimpl.tersoff_params(CT, {"C(O)"})[("C(O)", "C(O)", "C(O)")]
    bond = a / math.sqrt(3.0)
    g = P["gamma"] * (1 + P["c"] ** 2 / P["d"] ** 2 - P["c"] ** 2 / (P["d"] ** 2 + (-0.5 - P["h"]) ** 2))
    zeta = 2 * g
    b = (1 + (P["beta"] * zeta) ** P["n"]) ** (-0.5 / P["n"])
    e_atom = 1.5 * (P["A"] * math.exp(-P["lam1"] * bond) - b * P["B"] * math.exp(-P["lam2"] * bond))
    assert r["e_tersoff"] / s.n == pytest.approx(e_atom, rel=1e-12)
    assert r["e_vdw"] == 0.0 and r["e_rep"] == 0.0  # one layer: no interlayer term
    assert np.abs(r["f"]).max() < 1e-10


def test_ilp_flat_bilayer_normals_are_z(oracle_lib):
    a = 2.46
    s = S.layered_stack(15, 9, ["G", "G"], jitter=0.0, a=a, b=a * math.sqrt(3.0), temperature=None)
    o = make_oracle(s)
    nrm, cnt, dn, dri = o.normals()
    assert (cnt == 3).all()
    np.testing.assert_allclose(np.abs(nrm[:, 2]), 1.0, atol=1e-15)
    np.testing.assert_allclose(nrm[:, :2], 0.0, atol=1e-15)
    np.testing.assert_allclose(dri, 0.0, atol=0)  # three neighbours: n does not depend on x_i
    # AB stacking binds more strongly than AA
    ab = o.compute(1, 0)
    s2 = S.layered_stack(15, 9, ["G", "G"], jitter=0.0, a=a, b=a * math.sqrt(3.0), temperature=None)
    s2.x[s.mol == 2, 1] -= a * math.sqrt(3.0) / 3.0
    s2.x[:, 1] %= s2.boxhi[1]
    aa = make_oracle(s2).compute(1, 0)
    e_ab = ab["e_vdw"] + ab["e_rep"]
    e_aa = aa["e_vdw"] + aa["e_rep"]
    assert e_ab < e_aa < 0
    assert -0.030 < e_ab / s.n < -0.015  # ILP binding of AB graphene ~ -23 meV/atom


# ---------------------------------------------------------------- independent brute force
@pytest.mark.parametrize("layers,seed", [(("G", "G"), 3), (("G", "BN"), 4), (("BN", "G", "BN"), 5)])
def test_energy_matches_independent_numpy(oracle_lib, layers, seed):
    s = small(layers, seed=seed, nx=15, ny=9 if len(layers) < 3 else 9)
    o = make_oracle(s)
    r = o.compute(1, 2)
    el = s.elements if s.elements != ["C"] else ["C"]
    tfile, telem = (CT, ["C(O)"]) if s.elements == ["C"] else (BNC, s.elements)
    et = refimpl.tersoff_energy(s, tfile, telem)
    ev, er = refimpl.ilp_energy(s, ILP, el)
    assert r["e_tersoff"] == pytest.approx(et, rel=1e-12)
    assert r["e_vdw"] == pytest.approx(ev, rel=1e-12)
    assert r["e_rep"] == pytest.approx(er, rel=1e-12)
    nrm, cnt, _, _ = o.normals()
    nref, cref = refimpl.normals(s, ILP, el)
    assert (cnt == cref).all()
    np.testing.assert_allclose(np.abs((nrm * nref).sum(1)), 1.0, atol=1e-12)


@pytest.mark.parametrize("layers,seed", [(("G", "G"), 3), (("G", "BN"), 4)])
def test_peratom_energy_halves_match_independent_numpy(oracle_lib, layers, seed):
    """Every atom's energy against tests/refimpl.py, which gives half of each tallied pair
    energy to each atom as Pair::ev_tally does (LMFF/pair.cpp:939-941): the Tersoff repulsive
    pair, 1/2 b_ij f_A of each ordered pair, the ILP van der Waals half pair and the repulsion of
    each ordered pair.  A split that gives each term to its first atom keeps every total and
    must fail."""
    s = small(layers, seed=seed, nx=15, ny=9)
    o = make_oracle(s)
    r = o.compute(3, 0)
    el = s.elements
    tfile, telem = (CT, ["C(O)"]) if s.elements == ["C"] else (BNC, s.elements)
    et, eat = refimpl.tersoff_energy(s, tfile, telem, peratom=True)
    ev, er, eai = refimpl.ilp_energy(s, ILP, el, peratom=True)
    ref = eat + eai
    scale = np.abs(ref).max()
    assert np.abs(r["eatom"] - ref).max() <= 1e-11 * scale
    assert r["eatom"].sum() == pytest.approx(r["energy"], rel=1e-12)
    _, eat1 = refimpl.tersoff_energy(s, tfile, telem, peratom=True, split="first")
    _, _, eai1 = refimpl.ilp_energy(s, ILP, el, peratom=True, split="first")
    wrong = eat1 + eai1
    assert wrong.sum() == pytest.approx(ref.sum(), rel=1e-12)
    assert np.abs(r["eatom"] - wrong).max() > 1e-4 * scale


def test_asymmetric_tables_follow_tag_parity(oracle_lib, tmp_path):
    """With parameter rows that differ between (i, j) and (j, i) (tests/asym_tables.py), the
    reference computes each unordered pair's van der Waals and Tersoff repulsive energy with the
    row of the atom that the tag-parity rule selects (pair_tersoff.cpp:150-158,
    pair_ilp_graphene_hbn.cpp:488-496).  The oracle against refimpl's restatement of that rule,
    per term and per atom; averaging both rows instead must fail."""
    import asym_tables
    import oracle

    tasym, asym = asym_tables.write(tmp_path)
    s = small(("G", "BN"), seed=6, nx=15, ny=9)
    o = oracle.Oracle(3, tasym, s.elements, asym, s.elements, tap_flag=1, threads=os.cpu_count() or 1)
    o.set_system(s)
    o.build_list()
    r = o.compute(3, 0)
    et, eat = refimpl.tersoff_energy(s, tasym, s.elements, peratom=True)
    ev, er, eai = refimpl.ilp_energy(s, asym, s.elements, peratom=True)
    assert r["e_tersoff"] == pytest.approx(et, rel=1e-12)
    assert r["e_vdw"] == pytest.approx(ev, rel=1e-12)
    assert r["e_rep"] == pytest.approx(er, rel=1e-12)
    ref = eat + eai
    assert np.abs(r["eatom"] - ref).max() <= 1e-11 * np.abs(ref).max()
    # averaging the two rows is a different answer for both terms
    et_avg = refimpl.tersoff_energy(s, tasym, s.elements, parity=False)
    ev_avg, _ = refimpl.ilp_energy(s, asym, s.elements, parity=False)
    assert abs(et_avg - et) > 1e3 * 1e-12 * abs(et)
    assert abs(ev_avg - ev) > 1e3 * 1e-12 * abs(ev)
    assert abs(r["e_tersoff"] - et_avg) > 1e-9 and abs(r["e_vdw"] - ev_avg) > 1e-9


def _fd_force(o, x0, atom, comp, h=1e-3):
    e = []
    for k in (-2, -1, 1, 2):
        x = x0.copy()
        x[atom, comp] += k * h
        o.put_x(x)
        e.append(o.compute(1, 0)["energy"])
    o.put_x(x0)
    return -(e[0] - 8 * e[1] + 8 * e[2] - e[3]) / (12 * h)


@pytest.mark.parametrize("layers", [("G", "G"), ("G", "BN")])
def test_forces_are_minus_energy_gradient(oracle_lib, layers):
    s = small(layers, seed=11, jitter=0.08)
    o = make_oracle(s)
    x0 = o.get_x()
    f = o.compute(1, 0)["f"]
    rng = np.random.default_rng(0)
    for atom in rng.choice(s.n, 6, replace=False):
        for comp in range(3):
            fd = _fd_force(o, x0, atom, comp)
            assert abs(fd - f[atom, comp]) < 2e-7 * max(1.0, abs(f[atom, comp])), (atom, comp, fd, f[atom, comp])


def test_forces_with_two_neighbour_normals(oracle_lib):
    """Armchair-edged ribbons (open in x) give 2-neighbour normals whose d n/d x_i is nonzero."""
    s = small(("G", "G"), seed=12, jitter=0.06, nx=12, ny=9)
    s.boxhi[0] += 30.0  # vacuum along x: the layers become ribbons with open edges
    s.x[:, 0] += 5.0
    o = make_oracle(s)
    nrm, cnt, dn, dri = o.normals()
    assert (cnt == 2).sum() > 0
    x0 = o.get_x()
    f = o.compute(1, 0)["f"]
    edge = np.nonzero(cnt == 2)[0]
    for atom in list(edge[:4]) + [int(np.nonzero(cnt == 3)[0][0])]:
        for comp in range(3):
            fd = _fd_force(o, x0, atom, comp)
            assert abs(fd - f[atom, comp]) < 2e-7 * max(1.0, abs(f[atom, comp])), (atom, comp, fd, f[atom, comp])


def test_virial_is_strain_derivative(oracle_lib):
    s = small(("G", "BN"), seed=13, jitter=0.05)
    o = make_oracle(s)
    r = o.compute(1, 2)
    x0 = o.get_x()
    for d in range(3):
        es = []
        for eps in (-2e-6, -1e-6, 1e-6, 2e-6):
            s2 = s.copy()
            s2.x = x0.copy()
            s2.x[:, d] *= 1 + eps
            if s.periodic[d]:
                s2.boxlo[d] *= 1 + eps
                s2.boxhi[d] *= 1 + eps
            es.append(make_oracle(s2).compute(1, 0)["energy"])
        dEde = (es[0] - 8 * es[1] + 8 * es[2] - es[3]) / (12 * 1e-6)
        assert r["virial"][d] == pytest.approx(-dEde, rel=2e-6, abs=1e-3)


def test_peratom_sums(oracle_lib):
    s = small(("G", "BN"), seed=14)
    o = make_oracle(s)
    r = o.compute(3, 6)
    assert r["eatom"].sum() == pytest.approx(r["energy"], rel=1e-13)
    np.testing.assert_allclose(r["vatom"].sum(0), r["virial"], rtol=1e-11, atol=1e-9)
    np.testing.assert_allclose(r["f"].sum(0), 0.0, atol=1e-10)


def test_translation_and_point_symmetry(oracle_lib):
    s = small(("G", "BN"), seed=15)
    r0 = make_oracle(s).compute(1, 2)
    s2 = s.copy()
    s2.x = s.x + np.array([3.3, -7.1, 1.25])
    s2.x[:, :2] %= (s.boxhi - s.boxlo)[:2]
    r1 = make_oracle(s2).compute(1, 2)
    assert r1["energy"] == pytest.approx(r0["energy"], rel=1e-12)
    np.testing.assert_allclose(r1["f"], r0["f"], atol=1e-9)
    # rotation by 180 degrees about z maps the periodic box onto itself
    s3 = s.copy()
    L = s.boxhi - s.boxlo
    s3.x[:, 0] = (L[0] - s.x[:, 0]) % L[0]
    s3.x[:, 1] = (L[1] - s.x[:, 1]) % L[1]
    r3 = make_oracle(s3).compute(1, 2)
    assert r3["energy"] == pytest.approx(r0["energy"], rel=1e-12)
    np.testing.assert_allclose(r3["f"][:, :2], -r0["f"][:, :2], atol=1e-9)
    np.testing.assert_allclose(r3["f"][:, 2], r0["f"][:, 2], atol=1e-9)


def test_neighbor_list_matches_brute_force(oracle_lib):
    s = small(("G", "BN"), seed=16, jitter=0.3)
    o = make_oracle(s)
    offs, ent, sh = o.neighbor_list()
    L = s.boxhi - s.boxlo
    cutsq = 18.0**2
    x = s.x
    for i in np.random.default_rng(1).choice(s.n, 40, replace=False):
        got = sorted((int(j), *map(int, sh[e])) for e, j in zip(range(offs[i], offs[i + 1]), ent[offs[i]:offs[i + 1]]))
        want = []
        for sx in (-1, 0, 1):
            for sy in (-1, 0, 1):
                # the reference's operation order: x_i - (x_j + s*L), no fused multiply-add
                gx = x[:, 0] + sx * L[0]
                gy = x[:, 1] + sy * L[1]
                dx = x[i, 0] - gx
                dy = x[i, 1] - gy
                dz = x[i, 2] - x[:, 2]
                rsq = dx * dx + dy * dy + dz * dz
                for j in np.nonzero(rsq <= cutsq)[0]:
                    if j == i and sx == 0 and sy == 0:
                        continue
                    want.append((int(j), sx, sy, 0))
        assert got == sorted(want)


def test_nve_energy_error_is_second_order(oracle_lib):
    """Velocity Verlet conserves a shadow energy: the total-energy error is bounded and
    scales as dt^2 (halving dt divides it by ~4).  A force that is not the gradient of the
    energy would instead show a drift that does not shrink this way."""
    s = small(("G", "G"), seed=17, jitter=0.02)
    errs = []
    for dt in (0.001, 0.0005):
        o = make_oracle(s, threads=4)
        nst = int(round(0.1 / dt))
        x, v, th = o.run_nve(s.x, s.v, s.mass_per_type, dt, nst, rebuild_every=10, thermo_every=nst // 20)
        etot = th[:, 1] + th[:, 2]
        errs.append(np.abs(etot - etot[0]).max())
        assert 100 < 2 * th[-1, 2] / ((3 * s.n - 3) * S.BOLTZ) < 500
    assert errs[0] / s.n < 5e-4
    assert 3.0 < errs[0] / errs[1] < 5.0


def test_reference_data_file_parses(oracle_lib):
    path = os.path.join(os.path.dirname(__file__), "golden", "inf0.data")
    s = S.read_lammps_data(path)
    assert s.n == 5016 and sorted(set(s.mol.tolist())) == [1, 2, 3, 4, 5, 6]
    import oracle

    o = oracle.Oracle(6, CT, ["C(O)"] * 6, ILP, ["C"] * 6, tap_flag=0)
    o.set_system(s)
    o.build_list()
    r = o.compute(1, 2)
    assert -8.1 < r["energy"] / s.n < -7.9  # C(O) graphene at the deck's 1.43876 A bond


def test_nve_drift_fixture_is_the_current_oracle(oracle_lib):
    """The committed 10,000-step NVE record (tests/golden/make_nve_drift.py) starts as the
    current oracle does: the first 100 steps are recomputed and must agree."""
    import json

    with open(os.path.join(os.path.dirname(__file__), "golden", "nve_drift_c1.json")) as f:
        fx = json.load(f)
    s = S.config_c1()
    assert s.n == fx["natoms"]
    o = make_oracle(s, threads=8)
    _, _, th = o.run_nve(s.x, s.v, s.mass_per_type, fx["dt_ps"], 100, rebuild_every=10, thermo_every=fx["thermo_every"])
    k = len(th)
    assert k == 3 and fx["step"][:k] == [int(v) for v in th[:, 0]]
    np.testing.assert_allclose(th[:, 1], fx["pe"][:k], rtol=1e-11)
    np.testing.assert_allclose(th[:, 2], fx["ke"][:k], rtol=1e-11)
    # the record itself: bounded total-energy error, no trend after the initial transient
    e = (np.array(fx["pe"]) + np.array(fx["ke"])) / s.n
    assert np.abs(e - e[0]).max() < 5e-4
    late = e[len(e) // 2:]
    slope = np.polyfit(np.array(fx["step"][len(e) // 2:], dtype=float), late, 1)[0]
    assert abs(slope) * 5000 < 3 * late.std() + 1e-6


class _Harmonic:
    """Stands in for the oracle: E = k/2 |x - c|^2 per atom, F = -k (x - c)."""

    def __init__(self, c, k=2.0):
        self.c, self.k, self.x = np.asarray(c, dtype=float), k, None
        # box record as Oracle keeps it (type, tag, mol, lo, hi, periodic): an open box
        self._keep = (None, None, None, np.full(3, -100.0), np.full(3, 100.0), np.zeros(3, dtype=np.int32))

    def put_x(self, x):
        self.x = np.array(x, dtype=float)

    def pbc(self):
        pass

    def build_list(self):
        pass

    def get_x(self):
        return self.x.copy()

    def compute(self, eflag=1, vflag=0):
        d = self.x - self.c
        return {"f": -self.k * d, "energy": 0.5 * self.k * float((d * d).sum())}


def test_fire_restatement_first_iterations_by_hand():
    """Iteration 1 starts from v = 0 (v.f = 0: the 'not downhill' branch, no dt change inside the
    initial delay) and takes x1 = x0 + dt^2 ftm2v/m f0; iteration 2 has v.f > 0 and mixes
    v = (1-a) v' + a |v'| f/|f| with a = alpha0 (no dt growth before delaystep)."""
    import oracle

    ftm2v, m, dt, k = 1.0 / 1.0364269e-4, 12.0, 0.001, 2.0
    x0 = np.array([[0.03, -0.02, 0.01]])
    h = _Harmonic(np.zeros((1, 3)), k)
    r1 = oracle.fire_minimize(h, x0, [m], dt, 0.0, 0.0, 1, dmax=10.0)
    f0 = -k * x0
    v1 = dt * ftm2v / m * f0
    x1 = x0 + dt * v1
    assert np.allclose(r1["x"], x1, rtol=1e-14, atol=0)
    r2 = oracle.fire_minimize(h, x0, [m], dt, 0.0, 0.0, 2, dmax=10.0)
    f1 = -k * x1
    vp = v1 + dt * ftm2v / m * f1
    a = 0.25
    v2 = (1 - a) * vp + a * np.linalg.norm(v1) / np.linalg.norm(f1) * f1
    x2 = x1 + dt * v2
    assert np.allclose(r2["x"], x2, rtol=1e-13, atol=0)
    # with a small dmax the second step is dtv = dmax / max|v1_c| instead of dt
    dmax = 1e-6
    r3 = oracle.fire_minimize(h, x0, [m], dt, 0.0, 0.0, 2, dmax=dmax)
    dtv = dmax / np.abs(v1).max()
    assert dt * np.abs(v1).max() > dmax
    vp = v1 + dtv * ftm2v / m * f1
    v2 = (1 - a) * vp + a * np.linalg.norm(v1) / np.linalg.norm(f1) * f1
    assert np.allclose(r3["x"], x1 + dtv * v2, rtol=1e-13, atol=0)


def test_fire_restatement_converges_to_the_minimum():
    import oracle

    rng = np.random.default_rng(3)
    c = rng.normal(size=(20, 3))
    h = _Harmonic(c, 5.0)
    r = oracle.fire_minimize(h, c + rng.normal(scale=0.2, size=c.shape), np.full(20, 12.0), 0.001, 0.0, 1e-8, 20000)
    assert r["stop"] == 3
    assert np.abs(r["x"] - c).max() < 1e-8
    r = oracle.fire_minimize(h, c + 0.1, np.full(20, 12.0), 0.001, 1e-12, 0.0, 20000)
    assert r["stop"] == 2 and r["e_final"] < 1e-8


def test_cg_restatement_on_harmonic_wells():
    """On E = k/2 |x - c|^2 the first quadratic line search is exact: from h = f it tries
    alpha = min(1, dmax / max|h|), the secant model fits (relerr 0), and alpha0 = 1/k lands on
    c; the next force check stops on ftol after 1 iteration and 2 evaluations.  On an
    anisotropic well (k per coordinate) Polak-Ribiere CG reaches the minimum."""
    import oracle

    rng = np.random.default_rng(5)
    c = rng.normal(size=(10, 3))
    h = _Harmonic(c, 5.0)
    r = oracle.cg_minimize(h, c + 0.03 * rng.uniform(-1, 1, size=c.shape), 0.0, 1e-10, 100, 1000)
    assert (r["stop"], r["iterations"], r["evaluations"]) == (3, 1, 2)
    assert np.abs(r["x"] - c).max() < 1e-13

    class Aniso(_Harmonic):
        def compute(self, eflag=1, vflag=0):
            d = self.x - self.c
            kk = np.array([1.0, 4.0, 9.0])
            return {"f": -kk * d, "energy": 0.5 * float((kk * d * d).sum())}

    a = Aniso(c)
    r = oracle.cg_minimize(a, c + 0.05 * rng.uniform(-1, 1, size=c.shape), 0.0, 1e-9, 200, 10000)
    assert r["stop"] == 3 and np.abs(r["x"] - c).max() < 1e-8


# ---------------------------------------------------------------- locality of the force
def test_patch_with_a_24_A_rim_reproduces_the_full_system(oracle_lib):
    """The multilayer benchmarks spot-check the GPU on a cut-out patch (bench.patch_baseline):
    an atom's force depends only on atoms within rc + two normal-neighbour hops, so the oracle
    on the patch plus a 24 A rim (open x, y boundaries) must give the interior atoms the forces
    of the full periodic system."""
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench

    s = S.layered_stack(60, 40, ["G", "BN", "G", "BN"], seed=31, jitter=0.05)
    full = make_oracle(s, threads=os.cpu_count() or 1).compute(1, 0)["f"]
    _, err = bench.patch_baseline(s, full, seed=5, frac=0.04)
    assert err["atoms_checked"] > 1000
    assert err["rel_max_over_rms"] < 1e-12
    # a rim shorter than the interaction range does not: the check can fail
    _, short = bench.patch_baseline(s, full, seed=5, frac=0.04, halo=6.0)
    assert short["rel_max_over_rms"] > 1e-4


# ---------------------------------------------------------------- the fused LMFF CPU path
@pytest.mark.parametrize("case", ["bilayer", "hetero_trilayer", "ribbons", "asymmetric", "null_types"])
def test_lmff_fused_restatement_matches_the_stock_oracle(oracle_lib, case, tmp_path):
    """orc_compute_lmff (the reference's fused LMFF CPU path, LMFF/pair_lmff.cpp:1121-1613: one
    short/normal pass, one fused Tersoff-repulsive + ILP pass that pushes each pair force to both
    atoms and applies the normal-derivative forces once per atom through dprodni) is the CPU
    baseline's path.  It must give the stock oracle's forces and energies up to rounding, also
    with two-neighbour normals, asymmetric parameter rows and types left out of one term."""
    import asym_tables
    import oracle

    if case == "bilayer":
        s = small(("G", "G"), seed=11)
        o = make_oracle(s, threads=4)
    elif case == "hetero_trilayer":
        s = small(("G", "BN", "G"), seed=12)
        o = make_oracle(s, threads=4)
    elif case == "ribbons":
        s = S.layered_stack(12, 9, ["G", "G"], seed=13, jitter=0.06)
        s.boxhi[0] += 30.0
        s.x[:, 0] += 5.0
        o = make_oracle(s, threads=4)
    elif case == "asymmetric":
        s = small(("G", "BN"), seed=14)
        tf, fi = asym_tables.write(tmp_path)
        o = oracle.Oracle(3, tf, s.elements, fi, s.elements, threads=4)
        o.set_system(s)
        o.build_list()
    else:
        # types 1 and 2 of a C-only bilayer: Tersoff only for type 2, ILP for both
        s = small(("G", "G"), seed=15)
        s.type = np.where(s.mol == 1, 1, 2).astype(np.int32)
        s.elements = ["C", "C"]
        o = oracle.Oracle(2, CT, ["NULL", "C(O)"], ILP, ["C", "C"], threads=4)
        o.set_system(s)
        o.build_list()
    r = o.compute(1, 0)
    q = o.compute_lmff(1)
    rms = np.sqrt((r["f"] ** 2).sum(1).mean())
    assert np.abs(q["f"] - r["f"]).max() <= 1e-12 * rms
    for k in ("energy", "e_vdw", "e_rep", "e_tersoff"):
        assert q[k] == pytest.approx(r[k], rel=1e-12, abs=1e-12)


# ---------------------------------------------------------------- per-atom virial split
def test_peratom_virial_matches_an_autograd_restatement(oracle_lib):
    """The oracle's per-atom virial (every tally, incl. the split of the normal-derivative
    forces between k and j) against tests/refimpl_virial.py: forces by automatic
    differentiation of the published energy forms, the reference's tally calls on top.  An
    open jittered bilayer flake, so edge atoms with two-neighbour normals take part."""
    import refimpl_virial

    s = S.layered_stack(6, 4, ["G", "G"], seed=21, jitter=0.08)
    s.periodic = (0, 0, 0)
    r = make_oracle(s, threads=os.cpu_count() or 1).compute(0, 4)
    w = refimpl_virial.peratom_virial(s, CT, "C(O)", ILP, "C")
    scale = np.abs(w).max()
    assert np.abs(r["vatom"] - w).max() <= 1e-10 * scale
    # the split is what is pinned: sharing the k-forces' virial with i instead of j differs
    wi = refimpl_virial.peratom_virial(s, CT, "C(O)", ILP, "C", k_partner="i")
    assert np.abs(r["vatom"] - wi).max() > 1e-6 * scale
