"""FIRE minimisation on the GPU (min_style fire; MinFire::iterate, LMFF/min_fire.cpp:104-460)
against the oracle's CPU restatement of the same algorithm on the oracle's forces."""
import numpy as np
import pytest

from lmff import capi, structures as S

pytestmark = pytest.mark.gpu


def oracle_for(s, threads=8):
    import os

    import oracle
    from conftest import POT

    o = oracle.Oracle(3, os.path.join(POT, "BNC.tersoff"), s.elements, os.path.join(POT, "BNCH.ILP"), s.elements,
                      threads=threads)
    o.set_system(s)
    return o


def gpu_minimize(systems, etol, ftol, maxiter, **fire):
    with capi.Context(0) as ctx:
        capi.setup_for(ctx, systems[0])
        sids = [ctx.add_system(s) for s in systems]
        for sid in sids:
            ctx.integrator(sid, capi.LMFF_NVE, dt=0.001)
        st = ctx.minimize(etol, ftol, maxiter, nsys=len(systems), **fire)
        xs = [ctx.get(sid)[0] for sid in sids]
    return st, xs


def wrapped_diff(a, b, s):
    d = a - b
    L = s.boxhi - s.boxlo
    for k in range(3):
        if s.periodic[k]:
            d[:, k] -= L[k] * np.round(d[:, k] / L[k])
    return np.abs(d).max()


def test_fire_follows_the_oracle_to_the_force_tolerance():
    import oracle

    s = S.layered_stack(15, 9, ["G", "BN"], seed=71, jitter=0.08)
    ref = oracle.fire_minimize(oracle_for(s), s.x, s.masses(), 0.001, 0.0, 0.1, 400)
    st, xs = gpu_minimize([s], 0.0, 0.1, 400)
    g = st[0]
    print("oracle", {k: v for k, v in ref.items() if k not in ("x", "f")}, "gpu", g)
    assert ref["stop"] == 3 and g["stop"] == ref["stop"]
    assert g["iterations"] == ref["iterations"] and g["evaluations"] == ref["evaluations"]
    assert abs(g["e_initial"] - ref["e_initial"]) < 1e-10 * abs(ref["e_initial"])
    assert abs(g["e_final"] - ref["e_final"]) < 1e-9 * abs(ref["e_final"])
    assert g["fnorm2_final"] < 0.1 and g["e_final"] < g["e_initial"]
    assert wrapped_diff(xs[0], ref["x"], s) < 1e-7


def test_fire_energy_tolerance_and_packed_systems():
    a = S.layered_stack(15, 9, ["G", "BN"], seed=72, jitter=0.05)
    b = S.layered_stack(15, 9, ["G", "BN"], seed=73, jitter=0.03)
    alone = [gpu_minimize([s], 1e-6, 0.0, 600) for s in (a, b)]
    packed, xs = gpu_minimize([a, b], 1e-6, 0.0, 600)
    for (st, x1), p, x2, s in zip(alone, packed, xs, (a, b)):
        q = st[0]
        assert q["stop"] == 2, q
        assert abs(q["e_final"] - q["e_next_to_last"]) < 1e-6 * 0.5 * (abs(q["e_final"]) + abs(q["e_next_to_last"]))
        assert p["stop"] == q["stop"] and p["iterations"] == q["iterations"]
        assert abs(p["e_final"] - q["e_final"]) < 1e-9 * abs(q["e_final"])
        assert wrapped_diff(x2, x1[0], s) < 1e-7


def test_fire_parameter_errors():
    s = S.layered_stack(15, 9, ["G", "BN"], seed=74)
    with capi.Context(0) as ctx:
        capi.setup_for(ctx, s)
        ctx.add_system(s)
        with pytest.raises(capi.LmffError, match="tmax has to be larger than tmin"):
            ctx.minimize(0.0, 0.1, 10, tmax=0.01)
        with pytest.raises(capi.LmffError, match="dtgrow"):
            ctx.minimize(0.0, 0.1, 10, dtgrow=0.9)


def gpu_cg(s, etol, ftol, maxiter, maxeval, skin=2.0):
    with capi.Context(0) as ctx:
        capi.setup_for(ctx, s, skin=skin)
        sid = ctx.add_system(s)
        st = ctx.minimize_cg(etol, ftol, maxiter, maxeval)
        x = ctx.get(sid)[0]
    return st, x


@pytest.mark.parametrize("skin,etol,maxiter", [(2.0, 1e-5, 1000), (0.2, 1e-8, 60)])
def test_cg_follows_the_oracle(skin, etol, maxiter):
    """minimize etol 1e-5 1000 10000 as the reference's decks use it; then a 0.2 A skin, so that
    lists are rebuilt (and atoms wrapped) in the middle of line searches."""
    import os

    import oracle
    from conftest import POT

    s = S.layered_stack(15, 9, ["G", "BN"], seed=71, jitter=0.08)
    o = oracle.Oracle(3, os.path.join(POT, "BNC.tersoff"), s.elements, os.path.join(POT, "BNCH.ILP"), s.elements,
                      skin=skin, threads=8)
    o.set_system(s)
    ref = oracle.cg_minimize(o, s.x, etol, 1e-5, maxiter, 10000, skin=skin)
    g, x = gpu_cg(s, etol, 1e-5, maxiter, 10000, skin=skin)
    print("oracle", {k: v for k, v in ref.items() if k not in ("x", "f")}, "gpu", g)
    assert g["stop"] == ref["stop"]
    assert g["iterations"] == ref["iterations"] and g["evaluations"] == ref["evaluations"]
    assert abs(g["e_final"] - ref["e_final"]) < 1e-9 * abs(ref["e_final"])
    assert g["e_final"] < g["e_initial"]
    assert wrapped_diff(x, ref["x"], s) < 1e-7


def test_cg_needs_one_system():
    a = S.layered_stack(15, 9, ["G", "BN"], seed=75)
    with capi.Context(0) as ctx:
        capi.setup_for(ctx, a)
        ctx.add_system(a)
        ctx.add_system(a.copy())
        with pytest.raises(capi.LmffError, match="exactly one system"):
            ctx.minimize_cg(1e-5, 1e-5, 10, 100)
