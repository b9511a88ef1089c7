"""Scheduler logic of the job server on the CPU, with a fake device context (no GPU):
admission by HBM budget in arrival order, retirement at tick boundaries, one packed run per
tick covering every resident job, and latency bookkeeping."""
import numpy as np
import pytest

from lmff import serving, structures as S


class FakeCtx:
    def __init__(self, bytes_per_atom=1000.0):
        self.bpa = bytes_per_atom
        self.systems = {}
        self.next = 0
        self.runs = []

    def system_bytes(self, natoms, nlayers=2):
        return int(natoms * self.bpa)

    def add_system(self, s):
        sid = self.next
        self.next += 1
        self.systems[sid] = s
        return sid

    def integrator(self, sid, *a, **k):
        assert sid in self.systems

    def remove_system(self, sid):
        del self.systems[sid]

    def get(self, sid):
        s = self.systems[sid]
        return s.x, np.zeros_like(s.x), np.zeros_like(s.x)

    def get_image(self, sid):
        return np.zeros((self.systems[sid].n, 3), dtype=np.int32)

    def run(self, nsteps, thermo_every, use_graph=True, nsys=1):
        assert nsys == len(self.systems) and nsteps == thermo_every
        self.runs.append(sorted(self.systems))
        return [[{"e_vdw": -1.0, "e_rep": 0.5, "pe": -7.0, "temp": 300.0}] for _ in range(nsys)]

    def profile_read(self):
        # device time of the last run (ms), per-kernel-family times and counts
        return 0.0, np.zeros(7), np.zeros(7, dtype=np.int64)

    def close(self):
        pass


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        self.t += 0.001
        return self.t


def jobs_of(sizes, steps, arrivals):
    out = []
    for k, (n, st, a) in enumerate(zip(sizes, steps, arrivals)):
        s = S.layered_stack(15, 9, ["G", "G"], seed=k)
        s.x = s.x[: n] if n < s.n else s.x
        out.append(serving.Job(job_id=k, system=s, steps=st, arrival=a))
    return out


def test_budget_limits_residency_and_everything_finishes():
    fake = FakeCtx(bytes_per_atom=1000.0)
    jobs = jobs_of([1080] * 6, [30, 20, 40, 10, 30, 20], [0.0, 0.0, 0.0, 0.001, 0.002, 0.002])
    srv = serving.Server(ctx=fake, hbm_budget_bytes=1080 * 1000 * 2.5, tick=10)
    st = srv.run(jobs, clock=Clock())
    assert st["jobs"] == 6
    assert st["max_resident"] <= 2
    assert all(len(r) <= 2 for r in fake.runs)
    assert all(j.done_steps == j.steps for j in jobs)
    assert all(j.finished_at >= j.admitted_at >= j.arrival for j in jobs)
    assert st["atom_steps"] == sum(j.natoms * j.steps for j in jobs)
    # admission is FIFO by arrival
    admitted = sorted(jobs, key=lambda j: j.admitted_at)
    assert [j.job_id for j in admitted] == sorted(j.job_id for j in jobs)
    assert st["latency_p99_s"] >= st["latency_p50_s"] > 0


def test_oversized_job_still_runs_alone():
    fake = FakeCtx(bytes_per_atom=1e9)
    jobs = jobs_of([1080, 1080], [10, 10], [0.0, 0.0])
    st = serving.Server(ctx=fake, hbm_budget_bytes=1.0, tick=10).run(jobs, clock=Clock())
    assert st["jobs"] == 2 and st["max_resident"] == 1


def test_job_stream_is_seeded_and_in_range():
    a = serving.tbg_job_stream(6, rate_hz=10.0, steps=100, seed=5, max_atoms=8000)
    b = serving.tbg_job_stream(6, rate_hz=10.0, steps=100, seed=5, max_atoms=8000)
    assert [j.natoms for j in a] == [j.natoms for j in b]
    assert all(2000 <= j.natoms <= 8000 for j in a)
    arr = [j.arrival for j in a]
    assert arr == sorted(arr) and arr[0] > 0
    for j in a:
        m, n = j.system.meta["m"], j.system.meta["n"]
        assert 1 <= n < m <= 12


def test_cancel_queued_and_running_jobs_and_snapshots():
    """on_tick hook: after the first tick, cancel one running and one queued job and take a
    snapshot of another running one; cancelled jobs free their slots and are not counted."""
    fake = FakeCtx(bytes_per_atom=1000.0)
    jobs = jobs_of([1080] * 5, [40, 40, 40, 20, 20], [0.0] * 5)
    srv = serving.Server(ctx=fake, hbm_budget_bytes=1080 * 1000 * 2.5, tick=10)   # two at a time
    snaps = []

    def on_tick(server, now):
        if len(server.log) == 1:
            running = sorted(j.job_id for j in server.resident.values())
            assert running == [0, 1]
            server.cancel(1)        # running
            server.cancel(3)        # still queued
            snaps.append(server.snapshot(0))
            with pytest.raises(KeyError):
                server.snapshot(2)

    st = srv.run(jobs, clock=Clock(), on_tick=on_tick)
    status = {j.job_id: j.status for j in jobs}
    assert status == {0: "done", 1: "cancelled", 2: "done", 3: "cancelled", 4: "done"}
    assert st["jobs"] == 3 and st["cancelled"] == 2
    assert jobs[1].done_steps == 10 and jobs[3].done_steps == 0 and jobs[1].result is None
    assert snaps[0]["step"] == 10 and snaps[0]["x"].shape == (1080, 3)
    assert not fake.systems and srv.reserved == 0.0
    assert st["atom_steps"] == sum(j.natoms * j.done_steps for j in jobs)
