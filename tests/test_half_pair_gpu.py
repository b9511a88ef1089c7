"""The half-pair ILP kernel on the GPU.  It evaluates every interlayer pair once, from the atom
in the lower z bin (from the lower atom index within one z bin), and sums the partner's terms
in exact 64-bit fixed point; the pull kernel evaluates every pair from both of its atoms.  Both
are checked here against each other (FP64 1e-12 of the RMS force, mixed 1e-6 eV/A RMS) and the
half-pair kernel against the oracle (1e-10) on geometries that reach each of its paths:
partner z bins above the atom (stacks, bilayers, a periodic z), pairs within one z bin
(corrugated sheets, whose atoms of two molecules share z bins), periodic images inside a
window (a twisted bilayer in a small cell) and the back segment of the lists (atoms moved past
the build's split margin).  LMFF_HALF=0 in the environment of a new context keeps the pull
kernel."""
import math
import os

import numpy as np
import pytest

from lmff import capi, structures as S
from test_gpu_parity import _check_fp64, ctx_for, oracle_for

pytestmark = pytest.mark.gpu


def _evaluate(s, precision, half, eflag=1, vflag=2, x=None):
    old = os.environ.get("LMFF_HALF")
    os.environ["LMFF_HALF"] = "1" if half else "0"
    try:
        with ctx_for(s, precision=precision) as ctx:
            r = ctx.pair_compute(s, eflag, vflag)
            if x is not None:
                # the same list, positions moved (the list is reused)
                r = ctx.pair_compute(s, eflag, vflag, new_list=0, x=x)
            return r
    finally:
        if old is None:
            del os.environ["LMFF_HALF"]
        else:
            os.environ["LMFF_HALF"] = old


def _corrugated():
    """An AB bilayer bent along x (amplitude 2 A, one wavelength per box): each sheet spans
    several z bins, so atoms of the two molecules share z bins (dz = 0 pairs) and meet at
    negative offsets too."""
    s = S.layered_stack(30, 20, ["G", "BN"], seed=74)
    lx = s.boxhi[0] - s.boxlo[0]
    s.x[:, 2] += 2.0 * np.sin(2.0 * math.pi * (s.x[:, 0] - s.boxlo[0]) / lx)
    return s


def _periodic_z():
    s = S.layered_stack(15, 9, ["G", "BN"] * 6, seed=75, jitter=0.05)
    s.boxlo[2], s.boxhi[2] = -1.675, 12 * 3.35 - 1.675
    s.periodic = (1, 1, 1)
    return s


def _open_box():
    """A G/BN bilayer in a box open in x and y as well (no periodic images anywhere): windows
    stop at the grid's edges and the edge bins hold the atoms beyond them."""
    s = S.layered_stack(24, 14, ["G", "BN"], seed=78, jitter=0.04)
    s.periodic = (0, 0, 0)
    return s


CASES = {
    "bilayer": lambda: S.layered_stack(30, 20, ["G", "BN"], seed=71),
    "ten_layer": lambda: S.layered_stack(15, 9, ["G" if k % 2 == 0 else "BN" for k in range(10)], seed=72),
    "periodic_z": _periodic_z,
    "corrugated": _corrugated,
    # 2,016 atoms in a 45 x 59 A cell: a window (18 x 9 bins) wraps onto the cell's own bins
    "twisted_small_cell": lambda: S.twisted_bilayer(4, 1, seed=73, min_atoms=2000, max_atoms=8000),
    "open_box": _open_box,
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_half_pair_equals_pull(name):
    s = CASES[name]()
    print(f"half-pair vs pull ILP kernel, {name}: {s.n} atoms", flush=True)
    a = _evaluate(s, capi.LMFF_FP64, True)
    b = _evaluate(s, capi.LMFF_FP64, False)
    rms = math.sqrt((b["f"] ** 2).sum(1).mean())
    assert np.abs(a["f"] - b["f"]).max() <= 1e-12 * rms
    # the kernels add in different orders: bitwise equal forces would mean one path ran twice
    assert not np.array_equal(a["f"], b["f"])
    for k in ("energy", "e_vdw", "e_rep"):
        assert abs(a[k] - b[k]) <= 1e-12 * abs(b[k]), k
    assert np.abs(a["virial"] - b["virial"]).max() <= 1e-12 * np.abs(b["virial"]).max()
    am = _evaluate(s, capi.LMFF_MIXED, True)
    bm = _evaluate(s, capi.LMFF_MIXED, False)
    assert math.sqrt(((am["f"] - bm["f"]) ** 2).sum(1).mean()) <= 1e-6
    assert abs(am["energy"] - bm["energy"]) <= 1e-7 * abs(bm["energy"])


@pytest.mark.parametrize("name", ["corrugated", "periodic_z", "twisted_small_cell", "open_box"])
def test_half_pair_vs_oracle(name):
    s = CASES[name]()
    print(f"half-pair ILP kernel vs the oracle, {name}: {s.n} atoms", flush=True)
    r = oracle_for(s).compute(1, 2)
    g = _evaluate(s, capi.LMFF_FP64, True)
    _check_fp64(g, r)


def test_half_pair_back_segment_vs_oracle():
    """The middle sheet slid by 0.5 A after the list build raises the far flag: entries of the
    back segment (beyond rc + 0.6 A at the build) are tested and the pairs that came within rc
    run through the queue of the half-pair kernel."""
    s = S.layered_stack(30, 20, ["G", "BN", "G"], seed=76)
    x1 = s.x.copy()
    x1[s.mol == 2] += np.array([0.4, 0.3, 0.0])
    g = _evaluate(s, capi.LMFF_FP64, True, x=x1)
    moved = s.copy()
    moved.x = x1
    r = oracle_for(moved).compute(1, 2)
    _check_fp64(g, r)


def test_half_pair_fixed_point_range_falls_back_to_the_pull_kernel():
    """A flat AB bilayer with one atom of the upper sheet pushed down to 1.7 A above an atom of the
    lower one: that pair's ILP terms (|C| about 7) pass the range of the fixed-point partner sums.
    The kernel flags it, and the evaluation is repeated with the pull kernel (which the context
    keeps from then on): the result equals the pull kernel's and the oracle's."""
    s = S.layered_stack(30, 20, ["G", "G"], jitter=0.0, seed=77, temperature=None)
    top, bot = np.where(s.mol == 2)[0], np.where(s.mol == 1)[0]
    d2 = ((s.x[top, None, :2] - s.x[None, bot, :2]) ** 2).sum(-1)
    k = top[np.argmin(d2.min(1))]
    assert d2.min() < 1e-12  # an atom of the upper sheet right above one of the lower sheet
    s.x[k, 2] -= 1.65
    print(f"half-pair kernel past its fixed-point range: {s.n} atoms, one pair 1.7 A apart", flush=True)
    a = _evaluate(s, capi.LMFF_FP64, True)
    b = _evaluate(s, capi.LMFF_FP64, False)
    rms = math.sqrt((b["f"] ** 2).sum(1).mean())
    assert np.abs(a["f"] - b["f"]).max() <= 1e-12 * rms
    assert abs(a["energy"] - b["energy"]) <= 1e-12 * abs(b["energy"])
    _check_fp64(a, oracle_for(s).compute(1, 2))


def _with_env(key, value, fn):
    old = os.environ.get(key)
    os.environ[key] = value
    try:
        return fn()
    finally:
        if old is None:
            del os.environ[key]
        else:
            os.environ[key] = old


@pytest.mark.parametrize("name", ["ten_layer", "corrugated", "twisted_small_cell", "open_box"])
def test_half_pair_window_beyond_lds(name):
    """LMFF_HALF_WCAP=64 keeps 64 atoms of every window in the kernel's LDS; the partner sums of
    every later slot go straight to the global fixed-point sums (the kernel's OVF variant, which
    splits each row segment where its window slots reach 64).  The partner sums are integers
    either way; the atom's own sums are floating-point sums over the lanes, which the split
    regroups, so the results equal the default's to rounding (FP64 1e-12 of the RMS force, mixed
    1e-6 eV/A RMS) and the oracle's to the FP64 bar.  Periodic images inside a window and
    windows cut by the edges of an open box are among the cases."""
    s = CASES[name]()
    print(f"half-pair windows beyond the LDS, {name}: {s.n} atoms", flush=True)
    a = _evaluate(s, capi.LMFF_FP64, True)
    b = _with_env("LMFF_HALF_WCAP", "64", lambda: _evaluate(s, capi.LMFF_FP64, True))
    rms = math.sqrt((a["f"] ** 2).sum(1).mean())
    assert np.abs(a["f"] - b["f"]).max() <= 1e-12 * rms
    for k in ("energy", "e_vdw", "e_rep"):
        assert abs(a[k] - b[k]) <= 1e-12 * abs(a[k]), k
    assert np.abs(a["virial"] - b["virial"]).max() <= 1e-12 * np.abs(a["virial"]).max()
    _check_fp64(b, oracle_for(s).compute(1, 2))
    am = _evaluate(s, capi.LMFF_MIXED, True)
    bm = _with_env("LMFF_HALF_WCAP", "64", lambda: _evaluate(s, capi.LMFF_MIXED, True))
    assert math.sqrt(((am["f"] - bm["f"]) ** 2).sum(1).mean()) <= 1e-6
    assert abs(am["energy"] - bm["energy"]) <= 1e-7 * abs(am["energy"])


def _resident(s, x=None, steps=0):
    """Forces and potential energy of a resident system, after an optional move of its atoms (the
    next list build re-fits the bins of the open z side), then an optional NVE run."""
    with ctx_for(s) as ctx:
        sid = ctx.add_system(s)
        ctx.integrator(sid, capi.LMFF_NVE, dt=0.001)
        ctx.evaluate(1, 2)
        if x is not None:
            ctx.set(sid, x=x)
            ctx.evaluate(1, 2)
        pe = ctx.thermo(sid)["pe"]
        f = ctx.get(sid)[2]
        if steps:
            ctx.run(steps, steps, use_graph=True)
        return f, pe, ctx.get(sid)[0]


def test_half_pair_bins_follow_a_moving_stack():
    """A resident 4-layer stack (one z bin per sheet, centred on the sheets) whose top sheet then
    moves up by 4 A while its bottom sheet ripples by 1.5 A: the sheets are no longer evenly
    spaced, so the next list build fits the z bins to the atoms' extent instead (six bins), and
    the forces equal those of a context that starts from the moved positions and the oracle's;
    30 Verlet steps (captured steps with the new tile count) follow the same trajectory."""
    s = S.layered_stack(15, 9, ["G", "BN", "G", "BN"], seed=79)
    x1 = s.x.copy()
    x1[s.mol == 4, 2] += 4.0
    lx = s.boxhi[0] - s.boxlo[0]
    bot = s.mol == 1
    x1[bot, 2] += 1.5 * np.sin(2.0 * math.pi * (s.x[bot, 0] - s.boxlo[0]) / lx)
    moved = s.copy()
    moved.x = x1
    print(f"resident stack moved after set-up: {s.n} atoms", flush=True)
    fa, pea, xa = _resident(s, x1, steps=30)
    fb, peb, xb = _resident(moved, steps=30)
    r = oracle_for(moved).compute(1, 2)
    rms = math.sqrt((r["f"] ** 2).sum(1).mean())
    assert np.abs(fa - fb).max() <= 1e-12 * rms
    assert np.abs(fa - r["f"]).max() <= 1e-10 * rms
    assert abs(pea - r["energy"]) <= 1e-10 * abs(r["energy"])
    assert abs(pea - peb) <= 1e-12 * abs(peb)
    assert np.abs(xa - xb).max() <= 1e-9
