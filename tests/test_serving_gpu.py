"""The job server on the GPU: a small stream of twisted-bilayer jobs of uneven size and length,
packed into one launch per step, gives every job exactly (bitwise) the result it gets alone."""
import numpy as np
import pytest

from lmff import capi, serving

pytestmark = pytest.mark.gpu


def test_stream_results_equal_alone_bitwise():
    jobs = serving.tbg_job_stream(5, rate_hz=200.0, steps=40, seed=11, max_atoms=6000)
    # three jobs queued at the start (packed from the first tick), two arriving later, uneven
    # lengths so that jobs retire while others keep running
    for j, t, n in zip(jobs, [0.0, 0.0, 0.0, 0.003, 0.006], [40, 20, 60, 30, 40]):
        j.arrival, j.steps = t, n
    alone = {j.job_id: serving.run_alone(j, precision=capi.LMFF_MIXED, tick=10) for j in jobs}
    srv = serving.Server(precision=capi.LMFF_MIXED, hbm_budget_bytes=4e9, tick=10)
    try:
        st = srv.run(jobs)
    finally:
        srv.close()
    assert st["jobs"] == len(jobs)
    assert st["max_resident"] >= 2
    for j in jobs:
        a = alone[j.job_id]
        assert np.array_equal(j.result["x"], a["x"])
        assert np.array_equal(j.result["v"], a["v"])
        assert np.array_equal(j.result["image"], a["image"])
        assert j.result["e_interlayer"] == a["e_interlayer"]
        assert j.result["e_interlayer"] < 0
    assert st["latency_p99_s"] >= st["latency_p50_s"] > 0


def test_hbm_estimate_covers_what_a_job_allocates():
    """The server admits by lmff_system_bytes: for bilayer jobs the estimate must cover what the
    library actually allocates for them (without being far above it)."""
    from lmff import structures as S

    for s in (S.layered_stack(30, 20, ["G", "BN"], seed=3), S.twisted_bilayer(5, 3, seed=4, target_atoms=12000)):
        with capi.Context(0) as ctx:
            capi.setup_for(ctx, s)
            est = ctx.system_bytes(s.n, 2)
            b0, _ = ctx.memory()
            sid = ctx.add_system(s)
            ctx.integrator(sid, capi.LMFF_LANGEVIN, dt=0.001, t_target=300.0, damp=1.0, seed=5)
            ctx.run(20, 10)
            b1, _ = ctx.memory()
        used = b1 - b0
        print(s.n, "atoms: estimate", est, "allocated", used)
        # per-context tables (parameters, cells, chunk sums) are the only bytes outside the estimate
        assert used <= est + 4 * 2 ** 20
        assert est <= 1.2 * used + 4 * 2 ** 20
