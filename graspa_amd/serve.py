"""Served stream of GCMC / Widom jobs with continuous batching.

Jobs (a framework, its adsorbates, T, P and a cycle count) arrive over time.
The server keeps a fixed set of batch slots on one GPU; every job that has
arrived is admitted into a free slot as soon as one is free and the resident
footprint of the admitted jobs (framework, guest store, structure factors,
tables) stays under the memory budget.  Between launches of the batch kernel
the server retires finished jobs and admits waiting ones, so jobs that arrive
while others run share the device with them.  Each job draws its moves from its
own counter-based stream (seed, stream), so its trajectory is the same alone
and in any mix.

This is the host side of the path only: job construction, admission and the
launch loop; every move runs in the device kernel (graspa_batch_launch).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import engine


@dataclass
class Job:
    """One adsorption job: a System (deck) and how long to run it."""
    job_id: int
    system: object
    cycles: int
    seed: int
    stream: int = 0
    arrival: float = 0.0          # seconds after the server starts
    capacity: int = 1024          # molecules per species
    # filled by the server
    admitted: Optional[float] = None
    finished: Optional[float] = None
    stats: Optional[engine.BatchStats] = None
    cycle_seconds: Optional[np.ndarray] = None
    trace: list = field(default_factory=list)


def footprint_bytes(system, capacity=1024):
    """Device memory a job holds while resident: framework copies and screen
    cells, the guest store, force field, Ewald tables and structure factors,
    and the real-space table."""
    n_fw = len(system.fw_type)
    guests = sum(capacity * c.molsize for c in system.components)
    T = system.ntypes
    nvec = system.nvec if not system.no_charges else 0
    per_fw = 32 + 4 + 16 + 16 + 36 + 32      # SoA FP64, type, fixed-point copies, sorted copy, cell slots
    per_guest = 32 + 12 + 16
    ewald = nvec * (16 * 3 + 8 + 16 + 8)
    return int(n_fw * per_fw + guests * per_guest + T * T * 48 + ewald + 64 * 1024 + 2 * 1024 * 1024)


class Server:
    """Continuous-batching GCMC server on one GPU."""

    def __init__(self, n_slots=64, precision=engine.GRASPA_MIXED, memory_budget=8 << 30, moves_per_launch=512,
                 trace_cap=0, cycle_cap=8192, device=0):
        self.n_slots = n_slots
        self.precision = precision
        self.budget = memory_budget
        self.mpl = moves_per_launch
        self.trace_cap = trace_cap
        self.cycle_cap = cycle_cap
        self.device = device
        self.batch = engine.Batch([], [], [], trace_cap=trace_cap, cycle_cap=cycle_cap, n_slots=n_slots)
        self.slots: List[Optional[Job]] = [None] * n_slots
        self._boxes: List[Optional[engine.Box]] = [None] * n_slots
        self._params = [None] * n_slots
        self.resident = 0
        self.launches = 0
        self.moves = 0

    def _admit(self, job, slot, now):
        box = engine.Box(job.system, precision=self.precision, device=self.device, capacity=job.capacity)
        box.init_ewald()
        params = engine.make_mc_params(job.system, box)
        self.batch.set_box(slot, box, params, job.seed, job.stream)
        self.batch.set_cycles(slot, job.cycles, production=True)
        self.slots[slot] = job
        self._boxes[slot] = box
        self._params[slot] = params
        self.resident += footprint_bytes(job.system, job.capacity)
        job.admitted = now

    def _retire(self, slot, now):
        job = self.slots[slot]
        job.stats = self.batch.stats(slot)
        job.cycle_seconds = self.batch.cycle_times(slot, min(job.cycles, self.cycle_cap))
        if self.trace_cap:
            job.trace = self.batch.trace(slot)
        job.finished = now
        self.batch.set_box(slot, None)
        self._boxes[slot].close()
        self._boxes[slot] = None
        self._params[slot] = None
        self.slots[slot] = None
        self.resident -= footprint_bytes(job.system, job.capacity)

    def run(self, jobs, time_scale=1.0):
        """Serves `jobs` (sorted by arrival, seconds scaled by time_scale)
        until all have finished; returns them with their statistics."""
        pending = sorted(jobs, key=lambda j: j.arrival)
        done = []
        t0 = time.perf_counter()
        while pending or any(s is not None for s in self.slots):
            now = time.perf_counter() - t0
            # admission: arrived jobs, free slot, memory budget
            while pending and pending[0].arrival * time_scale <= now:
                free = [i for i, s in enumerate(self.slots) if s is None]
                need = footprint_bytes(pending[0].system, pending[0].capacity)
                if need > self.budget:
                    raise ValueError("job %d needs %d bytes, more than the memory budget" % (pending[0].job_id, need))
                if not free or self.resident + need > self.budget:
                    break
                self._admit(pending.pop(0), free[0], now)
            if all(s is None for s in self.slots):
                # nothing resident: wait for the next arrival
                time.sleep(max(0.0, min(0.01, pending[0].arrival * time_scale - now)))
                continue
            before = self.batch.total_moves()
            self.batch.launch(self.mpl)
            self.launches += 1
            self.moves += self.batch.total_moves() - before
            now = time.perf_counter() - t0
            for i, job in enumerate(self.slots):
                if job is not None and self.batch.slot_done(i):
                    self._retire(i, now)
                    done.append(job)
        self.elapsed = time.perf_counter() - t0
        return done

    def close(self):
        for i, b in enumerate(self._boxes):
            if b is not None:
                b.close()
                self._boxes[i] = None
        self.batch.close()


def poisson_jobs(n_jobs, rate, seed=4000, cycles_lo=50, cycles_hi=5000, system_fn=None):
    """Jobs drawn from the 256-box distribution, Poisson arrivals at `rate`
    jobs/s, lengths log-uniform in [cycles_lo, cycles_hi] cycles."""
    from . import synth
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for j in range(n_jobs):
        t += rng.exponential(1.0 / rate)
        cyc = int(round(10 ** rng.uniform(np.log10(cycles_lo), np.log10(cycles_hi))))
        s = system_fn(j) if system_fn else synth.config3_box(j, seed_base=seed)
        jobs.append(Job(job_id=j, system=s, cycles=cyc, seed=seed + j, stream=j, arrival=t))
    return jobs
