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
        assert j.result["e_interlayer"] == a["e_interlayer"]
        assert j.result["e_interlayer"] < 0
    assert st["latency_p99_s"] >= st["latency_p50_s"] > 0
