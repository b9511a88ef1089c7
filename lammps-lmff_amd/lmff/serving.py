"""Continuous-batching job server for independent bilayer MD jobs on one GPU.

Jobs (e.g. twisted-bilayer-graphene cells of uneven size) arrive over time.  Resident jobs
are packed into one atom array (one context, one launch per kernel per step covering all of
them).  Every `tick` steps -- the neighbor-rebuild interval, so that each job's rebuilds and
hence its trajectory are the same as when it runs alone -- the server retires finished
jobs, admits arrived jobs while the HBM budget allows, and advances every resident job by
one tick.  Per-job state never leaves the device between ticks.

A job's result (final interlayer energy, thermo trace, latency) is identical bit for bit
to running that job alone with the same tick, because every kernel works per atom, the
per-system reductions never straddle systems, and the Langevin noise is keyed by
(job seed, job step, atom tag).
"""
from __future__ import annotations

import heapq
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import capi, structures as S


@dataclass
class Job:
    job_id: int
    system: S.System
    steps: int
    arrival: float                 # seconds after the server starts
    seed: int = 1
    t_target: float = 300.0
    damp: float = 1.0
    style: int = capi.LMFF_LANGEVIN
    # filled in by the server
    admitted_at: float | None = None
    finished_at: float | None = None
    done_steps: int = 0
    sys_id: int = -1
    trace: list = field(default_factory=list)
    result: dict | None = None
    status: str = "pending"        # pending -> queued -> running -> done | cancelled

    @property
    def natoms(self) -> int:
        return self.system.n


def tbg_job_stream(njobs: int = 512, rate_hz: float = 50.0, steps: int = 2000, seed: int = 2024,
                   min_atoms: int = 2000, max_atoms: int = 80000, sizes: str = "log-uniform"):
    """Seeded stream of twisted-bilayer jobs with Poisson arrivals.

    Each job draws a commensurate twist (m, n), 1 <= n < m <= 12.  sizes="log-uniform": a
    target size log-uniform in [min_atoms, max_atoms] (small cells are the common case), the
    moire cell replicated to reach it; sizes="minimal": the smallest replication with at least
    min_atoms atoms and every side >= 36.5 A (the size set by the twist alone)."""
    if sizes not in ("log-uniform", "minimal"):
        raise ValueError(f"unknown size distribution {sizes!r}")
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for k in range(njobs):
        t += rng.exponential(1.0 / rate_hz)
        while True:
            m = int(rng.integers(2, 13))
            n = int(rng.integers(1, m))
            target = int(round(math.exp(rng.uniform(math.log(min_atoms), math.log(max_atoms)))))
            if sizes == "minimal":
                target = min_atoms
            try:
                s = S.twisted_bilayer(m, n, min_atoms=min_atoms, max_atoms=max_atoms, target_atoms=target,
                                      seed=int(rng.integers(1 << 30)))
                break
            except ValueError:
                continue
        jobs.append(Job(job_id=k, system=s, steps=steps, arrival=t, seed=int(rng.integers(1, 1 << 62))))
    return jobs


class Server:
    """Admits jobs by HBM budget and runs them packed, one launch per kernel per step."""

    def __init__(self, precision=capi.LMFF_MIXED, hbm_budget_bytes: float = 200e9, tick: int = 10,
                 max_resident: int = 10 ** 9, device: int = 0, dt: float = 0.001, ctx=None):
        if ctx is None:
            ctx = capi.Context(device)
            capi.setup_graphene(ctx, precision=precision, every=tick)
        self.ctx = ctx
        self.budget = float(hbm_budget_bytes)
        self.tick = int(tick)
        self.max_resident = max_resident
        self.dt = dt
        self.resident: dict[int, Job] = {}
        self.reserved = 0.0
        self.log = []
        self._cancel: set[int] = set()
        self.cancelled: list[Job] = []

    def cancel(self, job_id: int):
        """Requests cancellation: a queued job never starts, a running one leaves at the next tick
        boundary (its slot and HBM reservation are freed, no result is kept)."""
        self._cancel.add(job_id)

    def snapshot(self, job_id: int):
        """Positions, velocities and image flags of a running job now (caller's atom order)."""
        for job in self.resident.values():
            if job.job_id == job_id:
                x, v, _ = self.ctx.get(job.sys_id)
                return {"step": job.done_steps, "x": x, "v": v, "image": self.ctx.get_image(job.sys_id)}
        raise KeyError(f"job {job_id} is not running")

    def close(self):
        self.ctx.close()

    def _job_bytes(self, job: Job) -> float:
        return float(self.ctx.system_bytes(job.natoms, 2))

    def _admit(self, job: Job, now: float):
        sid = self.ctx.add_system(job.system)
        self.ctx.integrator(sid, job.style, dt=self.dt, t_target=job.t_target, damp=job.damp, seed=job.seed)
        job.sys_id = sid
        job.admitted_at = now
        job.status = "running"
        self.resident[sid] = job
        self.reserved += self._job_bytes(job)

    def _collect(self, job: Job):
        """The finished job's result, read while the packed layout is still intact (a removal
        marks it for a re-pack, which the next read would pay for every remaining job)."""
        x, v, _ = self.ctx.get(job.sys_id)
        img = self.ctx.get_image(job.sys_id)
        last = job.trace[-1] if job.trace else {}
        job.result = {"e_interlayer": last.get("e_vdw", float("nan")) + last.get("e_rep", float("nan")),
                      "pe": last.get("pe"), "temp": last.get("temp"), "x": x, "v": v, "image": img}

    def _retire(self, job: Job, now: float):
        self.ctx.remove_system(job.sys_id)
        self.reserved -= self._job_bytes(job)
        del self.resident[job.sys_id]
        job.finished_at = now
        job.status = "done"

    def _drop(self, job: Job, now: float):
        self.ctx.remove_system(job.sys_id)
        self.reserved -= self._job_bytes(job)
        del self.resident[job.sys_id]
        job.finished_at = now
        job.status = "cancelled"
        self.cancelled.append(job)

    def run(self, jobs, clock=None, time_scale: float = 1.0, on_tick=None):
        """Serves `jobs` (sorted by arrival).  Arrival times are in seconds of wall clock
        (scaled by time_scale).  on_tick(server, now), if given, is called after every tick
        (it may cancel jobs or take snapshots).  Returns summary statistics."""
        clock = clock or time.perf_counter
        pending = sorted(jobs, key=lambda j: j.arrival)
        queue = []   # arrived, not admitted (FIFO)
        t0 = clock()
        finished = []
        atom_steps = 0
        gpu_s = 0.0      # device time of the ticks (HIP events around each lmff_run)
        admit_s = 0.0    # host time spent admitting (staging + the re-pack at the next tick)
        pi = 0
        while pi < len(pending) or queue or self.resident:
            now = (clock() - t0) / time_scale
            while pi < len(pending) and pending[pi].arrival <= now:
                pending[pi].status = "queued"
                queue.append(pending[pi])
                pi += 1
            # cancellations: queued jobs never start, running ones leave now
            if self._cancel:
                for job in [j for j in queue if j.job_id in self._cancel]:
                    queue.remove(job)
                    job.status = "cancelled"
                    job.finished_at = now
                    self.cancelled.append(job)
                for job in [j for j in self.resident.values() if j.job_id in self._cancel]:
                    self._drop(job, now)
                self._cancel -= {j.job_id for j in self.cancelled}
            # retire finished jobs: read every result first, then remove them all (one re-pack at
            # the next tick instead of one per retired job)
            done = [j for j in self.resident.values() if j.done_steps >= j.steps]
            for job in done:
                self._collect(job)
            for job in done:
                self._retire(job, now)
                finished.append(job)
            # admit in arrival order while the budget allows
            while queue and len(self.resident) < self.max_resident:
                job = queue[0]
                need = self._job_bytes(job)
                if self.reserved + need > self.budget and self.resident:
                    break
                queue.pop(0)
                ta = time.perf_counter()
                self._admit(job, now)
                admit_s += time.perf_counter() - ta
            if not self.resident:
                if pi < len(pending):
                    wait = pending[pi].arrival - now
                    if wait > 0:
                        time.sleep(min(wait * time_scale, 0.05))
                continue
            nsys = len(self.resident)
            order = sorted(self.resident)  # packed order = admission order = sys id order
            rec = self.ctx.run(self.tick, self.tick, use_graph=True, nsys=nsys)
            gpu_s += self.ctx.profile_read()[0] * 1e-3
            for k, sid in enumerate(order):
                job = self.resident[sid]
                job.done_steps += self.tick
                job.trace.append(rec[k][-1])
                atom_steps += job.natoms * self.tick
            self.log.append((now, nsys, sum(j.natoms for j in self.resident.values())))
            if on_tick is not None:
                on_tick(self, (clock() - t0) / time_scale)
        t_end = (clock() - t0) / time_scale
        lat = np.array([j.finished_at - j.arrival for j in finished])
        return {
            "jobs": len(finished),
            "cancelled": len(self.cancelled),
            "makespan_s": t_end,
            "atom_steps": atom_steps,
            "atom_steps_per_s": atom_steps / t_end if t_end > 0 else float("nan"),
            "latency_p50_s": float(np.percentile(lat, 50)) if len(lat) else float("nan"),
            "latency_p99_s": float(np.percentile(lat, 99)) if len(lat) else float("nan"),
            "max_resident": max((n for _, n, _ in self.log), default=0),
            "ticks": len(self.log),
            "gpu_busy_s": gpu_s,
            "admit_s": admit_s,
        }


def run_alone(job: Job, precision=capi.LMFF_MIXED, tick: int = 10, dt: float = 0.001):
    """Runs one job by itself (the sequential baseline and the alone-vs-mix reference)."""
    ctx = capi.Context(0)
    try:
        capi.setup_graphene(ctx, precision=precision, every=tick)
        sid = ctx.add_system(job.system)
        ctx.integrator(sid, job.style, dt=dt, t_target=job.t_target, damp=job.damp, seed=job.seed)
        trace = []
        done = 0
        while done < job.steps:
            rec = ctx.run(tick, tick, use_graph=True)
            trace.append(rec[0][-1])
            done += tick
        x, v, _ = ctx.get(sid)
        last = trace[-1]
        return {"e_interlayer": last["e_vdw"] + last["e_rep"], "pe": last["pe"], "temp": last["temp"], "x": x,
                "v": v, "image": ctx.get_image(sid), "trace": trace}
    finally:
        ctx.close()
