"""Served stream of GCMC / Widom jobs with continuous batching.

Jobs (a framework, its adsorbates, T, P and a cycle count) arrive over time.
The server keeps a fixed set of batch slots on one GPU; every job that has
arrived is admitted into a free slot as soon as one is free and the resident
footprint of the admitted jobs (the boxes' device memory as the engine
reports it) stays under the memory budget.  Launches are queued
asynchronously and the boxes of arriving jobs are built while a launch runs;
between launches the server retires finished jobs and admits the built ones,
so jobs that arrive while others run share the device with them.  Each job draws its moves from its
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
    prepared: bool = False
    footprint: int = 0            # device bytes of its box (graspa_box_footprint)
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
    """Continuous-batching GCMC server on one GPU.

    Each launch of the batch kernel is queued asynchronously; while it runs,
    the host builds the boxes of jobs that have arrived (framework upload,
    screen cells, structure factors on the box's own stream) so that they are
    ready to enter the batch when the launch ends.  A job is admitted into a
    free slot when its box's real device footprint (graspa_box_footprint)
    fits the memory budget beside the resident and the prepared boxes."""

    def __init__(self, n_slots=64, precision=engine.GRASPA_MIXED, memory_budget=8 << 30, moves_per_launch=512,
                 trace_cap=0, cycle_cap=8192, device=0, prepare_ahead=4, time_slice_us=0.0):
        self.n_slots = n_slots
        self.precision = precision
        self.budget = memory_budget
        self.mpl = moves_per_launch
        self.trace_cap = trace_cap
        self.cycle_cap = cycle_cap
        self.device = device
        self.prepare_ahead = prepare_ahead
        self.batch = engine.Batch([], [], [], trace_cap=trace_cap, cycle_cap=cycle_cap, n_slots=n_slots)
        if time_slice_us > 0:
            # every box runs about this long per launch, at its own pace
            self.batch.set_time_slice(time_slice_us)
        self.slots: List[Optional[Job]] = [None] * n_slots
        self._boxes: List[Optional[engine.Box]] = [None] * n_slots
        self._params = [None] * n_slots
        self._fp = [0] * n_slots
        self.ready = []               # (job, box, params, footprint) built and waiting for a slot
        self.resident = 0             # bytes of the boxes in slots
        self.prepared = 0             # bytes of the boxes built and waiting
        self.launches = 0
        self.moves = 0
        self.peak_bytes = 0
        self.max_concurrent = 0       # most jobs resident in one launch

    def _prepare(self, job):
        """Builds the job's box (host set-up, uploads and S(k) on the box's stream)."""
        est = footprint_bytes(job.system, job.capacity)
        if est > self.budget:
            raise ValueError("job %d needs about %d bytes, more than the memory budget" % (job.job_id, est))
        box = engine.Box(job.system, precision=self.precision, device=self.device, capacity=job.capacity)
        box.init_ewald()
        params = engine.make_mc_params(job.system, box)
        fp = box.footprint()
        self.ready.append((job, box, params, fp))
        self.prepared += fp
        job.prepared = True

    def _admit(self, now):
        """Moves prepared boxes into free slots while their footprint fits."""
        while self.ready:
            free = [i for i, s in enumerate(self.slots) if s is None]
            job, box, params, fp = self.ready[0]
            if not free or self.resident + fp > self.budget:
                break
            self.ready.pop(0)
            self.prepared -= fp
            slot = free[0]
            self.batch.set_box(slot, box, params, job.seed, job.stream)
            self.batch.set_cycles(slot, job.cycles, production=True)
            self.slots[slot] = job
            self._boxes[slot] = box
            self._params[slot] = params
            self._fp[slot] = fp
            self.resident += fp
            job.admitted = now
            job.footprint = fp
        self.peak_bytes = max(self.peak_bytes, self.resident + self.prepared)

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
        self.resident -= self._fp[slot]
        self._fp[slot] = 0

    def _can_prepare(self, job):
        est = footprint_bytes(job.system, job.capacity)
        n_free = sum(1 for s in self.slots if s is None)
        return (len(self.ready) < n_free + self.prepare_ahead
                and self.resident + self.prepared + est <= self.budget)

    def run(self, jobs, time_scale=1.0):
        """Serves `jobs` (sorted by arrival, seconds scaled by time_scale)
        until all have finished; returns them with their statistics."""
        pending = sorted(jobs, key=lambda j: j.arrival)
        done = []
        t0 = time.perf_counter()
        while pending or self.ready or any(s is not None for s in self.slots):
            now = time.perf_counter() - t0
            self._admit(now)
            if all(s is None for s in self.slots):
                if pending and pending[0].arrival * time_scale <= now:
                    # the device is idle: build the next box now
                    self._prepare(pending.pop(0))
                    continue
                if self.ready:
                    raise ValueError("job %d does not fit the memory budget" % self.ready[0][0].job_id)
                # nothing resident: wait for the next arrival
                time.sleep(max(0.0, min(0.01, pending[0].arrival * time_scale - now)))
                continue
            before = self.batch.total_moves()
            self.max_concurrent = max(self.max_concurrent, sum(1 for x in self.slots if x is not None))
            self.batch.launch_async(self.mpl)
            # while the launch runs: build the boxes of arrived jobs
            while pending and pending[0].arrival * time_scale <= time.perf_counter() - t0 and \
                    self._can_prepare(pending[0]):
                self._prepare(pending.pop(0))
            self.batch.wait()
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
        for _, box, _, _ in self.ready:
            box.close()
        self.ready = []
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
