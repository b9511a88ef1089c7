"""The device thermostat and the job server against the oracle.

  * fix langevin on the device (Philox4x32-10 noise keyed by seed, step and tag) against the
    numpy restatement of FixLangevin::post_force in tests/refimpl_langevin.py driven by the
    oracle's forces: one step within 1e-12 (A, A/ps), a 100-step run with neighbor rebuilds
    within 1e-9 A and 1e-7 A/ps (the bar of the NVE trajectory test);
  * NVE jobs through the packed job server against the oracle's own Verlet run of each job."""
import numpy as np
import pytest

import refimpl_langevin as L
from lmff import capi, serving, structures as S
from test_gpu_parity import ctx_for, oracle_for

pytestmark = pytest.mark.gpu

T, DAMP, DT, SEED = 300.0, 0.1, 0.001, 0x1234_5678_9ABC


def _oracle_forces(s):
    o = oracle_for(s)

    def forces(x):
        o.put_x(x)
        o.pbc()
        o.build_list()
        return o.compute(0, 0)["f"]

    return forces


def _device_run(s, nsteps, step0=0):
    with ctx_for(s) as ctx:
        sid = ctx.add_system(s)
        ctx.integrator(sid, capi.LMFF_LANGEVIN, dt=DT, t_target=T, damp=DAMP, seed=SEED)
        if step0:
            ctx.set_step(sid, step0)
        ctx.run(nsteps, nsteps, use_graph=True)
        x, v, _ = ctx.get(sid)
    return x, v


def _mod_box(d, s):
    L_ = s.boxhi - s.boxlo
    for k in range(3):
        if s.periodic[k]:
            d[:, k] -= L_[k] * np.round(d[:, k] / L_[k])
    return d


@pytest.mark.parametrize("step0", [0, 5_000_000_123])
def test_one_langevin_step_matches_the_restatement(step0):
    s = S.layered_stack(15, 9, ["G", "BN"], seed=61)
    print(f"one Langevin step of {s.n} atoms from step {step0}", flush=True)
    x, v = _device_run(s, 1, step0)
    xr, vr = L.verlet_langevin(s.x, s.v, s.masses(), s.tag, _oracle_forces(s), 1, SEED, T, DAMP, DT, step0=step0)
    assert np.abs(_mod_box(x - xr, s)).max() <= 1e-12
    assert np.abs(v - vr).max() <= 1e-12
    # what a wrong noise amplitude would do to the same step
    _, vw = L.verlet_langevin(s.x, s.v, s.masses(), s.tag, _oracle_forces(s), 1, SEED, T, DAMP, DT, step0=step0,
                              kb_factor=48.0)
    assert np.abs(v - vw).max() > 1e-4


def test_hundred_langevin_steps_match_the_restatement():
    s = S.layered_stack(15, 9, ["G", "BN"], seed=62)
    print(f"100 Langevin steps of {s.n} atoms", flush=True)
    x, v = _device_run(s, 100)
    xr, vr = L.verlet_langevin(s.x, s.v, s.masses(), s.tag, _oracle_forces(s), 100, SEED, T, DAMP, DT)
    assert np.abs(_mod_box(x - xr, s)).max() <= 1e-9
    assert np.abs(v - vr).max() <= 1e-7


def test_server_nve_jobs_follow_the_oracle():
    """Three twisted-bilayer NVE jobs of different sizes and lengths, two queued at the start and
    one arriving later, served packed in FP64; each job's final state against the oracle's
    Verlet run of that job alone (rebuild every 10 steps, the server's tick)."""
    jobs = serving.tbg_job_stream(3, rate_hz=200.0, steps=30, seed=19, max_atoms=5000)
    for j, t, n in zip(jobs, [0.0, 0.0, 0.004], [30, 20, 40]):
        j.arrival, j.steps, j.style = t, n, capi.LMFF_NVE
    srv = serving.Server(precision=capi.LMFF_FP64, hbm_budget_bytes=4e9, tick=10)
    try:
        st = srv.run(jobs)
    finally:
        srv.close()
    assert st["jobs"] == 3
    for j in jobs:
        s = j.system
        print(f"job {j.job_id}: {s.n} atoms, {j.steps} steps", flush=True)
        o = oracle_for(s)
        xr, vr, th = o.run_nve(s.x, s.v, s.mass_per_type, 0.001, j.steps, rebuild_every=10, thermo_every=10)
        assert np.abs(_mod_box(j.result["x"] - xr, s)).max() < 1e-9
        assert np.abs(j.result["v"] - vr).max() < 1e-7
        assert j.result["pe"] == pytest.approx(th[-1, 1], rel=1e-10)
