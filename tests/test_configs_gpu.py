"""The multilayer configurations on the GPU against the oracle.

  * the 10-layer alternating graphene/hBN stack and the 4-layer stack, each at a size the
    oracle evaluates in seconds (every atom compared): FP64 forces, energies, virial and
    per-atom tallies within 1e-10, mixed precision within RMS 1e-5 eV/A;
  * the full 2,704,000-atom 10-layer stack (260x260 cells per layer) and the full 19,998,784-atom
    4-layer stack (1118x1118): the GPU's FP64 forces on a seeded 1 % patch of the box against the
    oracle on that patch plus a 24 A rim (an atom's force depends only on atoms within
    rc + two normal-neighbour hops < 24 A; tests/test_oracle.py pins the cut-out), plus
    size-independent checks: the total force vanishes, the energy per atom lies in the band of
    these stacks, and mixed precision stays within RMS 1e-5 eV/A of FP64.
Tolerances as in test_gpu_parity.py."""
import math
import os
import sys

import numpy as np
import pytest

from lmff import capi, structures as S
from test_gpu_parity import _check_fp64, ctx_for, oracle_for

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def alternating(nx, ny, layers, seed):
    return S.layered_stack(nx, ny, ["G" if k % 2 == 0 else "BN" for k in range(layers)], seed=seed)


@pytest.mark.parametrize("layers,seed", [(10, 41), (4, 42)])
def test_stack_every_atom_vs_oracle(layers, seed):
    s = alternating(15, 9, layers, seed)
    print(f"{layers}-layer stack, {s.n} atoms, FP64 and mixed vs the oracle", flush=True)
    r = oracle_for(s).compute(3, 6)
    with ctx_for(s) as ctx:
        g = ctx.pair_compute(s, 3, 6)
    _check_fp64(g, r)
    assert np.abs(g["eatom"] - r["eatom"]).max() <= 1e-10 * np.abs(r["eatom"]).max()
    assert np.abs(g["vatom"] - r["vatom"]).max() <= 1e-10 * np.abs(r["vatom"]).max()
    with ctx_for(s, precision=capi.LMFF_MIXED) as ctx:
        m = ctx.pair_compute(s, 1, 2)
    assert math.sqrt(((m["f"] - r["f"]) ** 2).sum(1).mean()) <= 1e-5
    assert abs(m["energy"] - r["energy"]) <= 1e-6 * abs(r["energy"])


def _full_size(layers, nx, ny, seed):
    sys.path.insert(0, REPO)
    import bench

    s = alternating(nx, ny, layers, seed)
    print(f"{layers}-layer stack, {s.n} atoms: FP64 and mixed on the GPU", flush=True)
    with ctx_for(s) as ctx:
        g = ctx.pair_compute(s, 1, 2)
    with ctx_for(s, precision=capi.LMFF_MIXED) as ctx:
        m = ctx.pair_compute(s, 1, 0)
    print("oracle on a seeded 1 % patch", flush=True)
    _, err = bench.patch_baseline(s, g["f"], seed=20240 + layers)
    return s, g, m, err


def _size_independent(s, g, m, err):
    rms = math.sqrt((g["f"] ** 2).sum(1).mean())
    assert err["atoms_checked"] >= 0.009 * s.n
    assert err["rel_max_over_rms"] <= 1e-10
    # Newton's third law over the periodic system: the forces sum to zero (to rounding)
    assert np.abs(g["f"].sum(0)).max() <= 1e-9 * rms * math.sqrt(s.n)
    # jittered graphene/hBN stacks at 3.35 A: -7.8 .. -7.6 eV per atom (Tersoff cohesion plus a
    # few tens of meV of ILP binding)
    assert -7.8 < g["energy"] / s.n < -7.6
    assert g["e_vdw"] < 0 < g["e_rep"]
    d = m["f"] - g["f"]
    assert math.sqrt((d ** 2).sum(1).mean()) <= 1e-5


def test_ten_layer_stack_2704000_atoms():
    s, g, m, err = _full_size(10, 260, 260, 12347)
    assert s.n == 2704000
    _size_independent(s, g, m, err)


def test_four_layer_stack_19998784_atoms():
    s, g, m, err = _full_size(4, 1118, 1118, 12349)
    assert s.n == 19998784
    _size_independent(s, g, m, err)
