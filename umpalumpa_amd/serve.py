"""Continuous-batching server for FR reconstruction jobs on one GPU.

Jobs arrive with a slice count, box, blob radius and symmetry group.  The server admits them
first come, first served while their device footprint (volume, weights, traverse spaces, blob
table) fits the HBM budget, keeps the admitted volumes resident, and on every tick inserts the
next chunk of slices of every admitted job with one batch step (``umpa_fr_batch_step``): one
persistent launch per group of jobs sharing settings, so jobs that arrive while others run are
inserted by the same launches.  A job's voxels receive its spaces in index order whatever else
is in flight, so its result is the same alone and in the mix (bit for bit under the exact
strategy).  Ticks are enqueued ahead of the GPU (up to ``depth`` in flight) and completions
are noticed through stream tickets, so the host loop never waits on an idle GPU.  By default
each tick serves the admitted jobs in order of remaining work (in seconds, at the measured
insertion rate) minus twice the time they have waited, up to a work quantum that fills the GPU:
short jobs overtake long ones, which keeps their latency near their isolated runtime, until a
long job has waited long enough to take its turn, which bounds the worst latency.

The reference has no serving layer; its harness runs one reconstruction at a time
(examples/fourier_reconstruction/fr.cpp:9-76).
"""
from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field

import numpy as np

from . import fr as F


@dataclass
class Job:
    """A reconstruction request.  ``slices(c)`` returns the device tensor of chunk c
    ([m, Hy, Hx] complex64, or ([m, Hy, Hx, 2] float16, [m] float32 scales) for FP16)."""
    job_id: int
    box: int
    n_slices: int
    radius: float
    symmetry: str
    rotations: np.ndarray          # [n_slices, 3, 3] float64
    slices: object                 # callable chunk index -> tensor(s)
    arrival: float = 0.0           # seconds after the server's start
    # filled by the server
    admitted: float | None = None
    done: float | None = None
    volume: object = None
    weights: object = None
    _spaces: object = None
    _table: object = None
    _chunk_spaces: list = field(default_factory=list)
    _next_chunk: int = 0
    _last_ticket: int = 0
    _order: int | None = None

    def _k(self) -> int:
        if self._order is None:
            self._order = len(F.symmetry_operators(self.symmetry))
        return self._order

    def chunk_cost(self, c: int, chunk: int) -> float:
        """Relative insertion work of chunk c: spaces x slice area."""
        m = min(chunk, self.n_slices - c * chunk)
        return float(self._k() * m * (self.box + 1) ** 2)

    def remaining_cost(self, chunk: int) -> float:
        m = max(0, self.n_slices - self._next_chunk * chunk)
        return float(self._k() * m * (self.box + 1) ** 2)

    def footprint(self, chunk: int) -> int:
        """Device bytes the job holds while admitted."""
        return 12 * (self.box + 1) ** 3 + 112 * self.n_slices * self._k() + 4 * 10000


class Server:
    """Admission by HBM budget, resident volumes, chunked insertion with continuous batching."""

    def __init__(self, device: int = 0, chunk: int = 512, hbm_budget: float = 200e9,
                 strategy: int = F.FR_FAST_MATH, depth: int = 3, alpha: float = 15.0,
                 max_active: int = 64, policy: str = "aged", tick_work: float = 4 * 512 * 257 ** 2,
                 age_weight: float = 2.0):
        import torch
        self.torch = torch
        self.dev = torch.device("cuda", device)
        self.chunk = chunk
        self.budget = hbm_budget
        self.strategy = strategy
        self.depth = depth
        self.alpha = alpha
        self.max_active = max_active
        # "srpt": each tick takes the admitted jobs with the least remaining work first, until
        # the tick holds `tick_work` (enough to fill the GPU), so short jobs overtake long ones;
        # "oldest": the same quantum filled in arrival order, which bounds the worst latency
        # (the longest job is never overtaken, later jobs fill what it leaves of the GPU);
        # "aged": the same quantum by remaining work in seconds (at the measured insertion
        # rate) minus `age_weight` times the time the job has waited, so a long job that has
        # waited is no longer overtaken by every newcomer;
        # "fair": every admitted job advances by one chunk per tick
        if policy not in ("srpt", "oldest", "aged", "fair"):
            raise ValueError(policy)
        self.policy = policy
        self.tick_work = tick_work
        self.age_weight = age_weight
        self._rate = None          # work units per second, measured from completed ticks
        self.batch = F.FRBatch(device)
        self._tables = {}
        self.used = 0
        self.total_updates = 0

    def _settings(self, radius):
        return F.FRSettings(F.FR_LOOKUP, 0, F.FR_PRECISE, self.alpha, radius)

    def _table(self, box, radius):
        key = (box, radius)
        if key not in self._tables:
            s = self._settings(radius)
            desc = F.fr_desc(box // 2 + 1, box, 1, box + 1, 1, 10000)
            self._tables[key] = self.torch.from_numpy(F.fill_blob_table(10000, desc, s)).to(
                self.dev)
        return self._tables[key]

    def _admit(self, job: Job, now: float):
        torch = self.torch
        s = self._settings(job.radius)
        sym = F.symmetry_operators(job.symmetry)
        k = len(sym)
        # spaces chunk by chunk, projection indices relative to the chunk's slices
        recs = []
        for c0 in range(0, job.n_slices, self.chunk):
            rot = job.rotations[c0:c0 + self.chunk]
            recs.append(F.generate_traverse_spaces(rot, job.box, s, sym))
            job._chunk_spaces.append((sum(len(r) for r in recs[:-1]), len(recs[-1])))
        sp = np.concatenate(recs)
        with torch.cuda.stream(self.batch.stream):
            job._spaces = torch.from_numpy(sp.view(np.uint8)).to(self.dev, non_blocking=False)
            job.volume = torch.zeros((job.box + 1,) * 3, dtype=torch.complex64, device=self.dev)
            job.weights = torch.zeros((job.box + 1,) * 3, dtype=torch.float32, device=self.dev)
        job._table = self._table(job.box, job.radius)
        job.admitted = now
        self.used += job.footprint(self.chunk)

    def _batch_job(self, job: Job, c: int):
        data = job.slices(c)
        off, cnt = job._chunk_spaces[c]
        sp = job._spaces[off * 112:(off + cnt) * 112]
        s = self._settings(job.radius)
        if isinstance(data, tuple):
            return F.BatchJob(data[0], job.volume, job.weights, sp, s, job._table,
                              scales=data[1]), cnt
        return F.BatchJob(data, job.volume, job.weights, sp, s, job._table), cnt

    def run(self, jobs, realtime: bool = True, count_updates: bool = False):
        """Serves `jobs` (sorted by arrival).  With `realtime`, a job becomes visible at its
        arrival time on the host clock; otherwise all are present at once.  Returns the jobs
        with admitted/done times (seconds from the start)."""
        pending = collections.deque(sorted(jobs, key=lambda j: j.arrival))
        active: list[Job] = []
        inflight = collections.deque()   # (ticket, work) of enqueued ticks
        busy_since, done_work = None, 0.0
        t0 = time.perf_counter()
        self.batch.set_count_updates(count_updates)
        self.total_updates = 0
        finished = []
        while pending or active:
            now = time.perf_counter() - t0
            # admission: FIFO within the budget and the slot limit
            while pending and (not realtime or pending[0].arrival <= now) and \
                    len(active) < self.max_active and \
                    self.used + pending[0].footprint(self.chunk) <= self.budget:
                j = pending.popleft()
                self._admit(j, now)
                active.append(j)
            if pending and not active and (not realtime or pending[0].arrival <= now) and \
                    pending[0].footprint(self.chunk) > self.budget:
                raise ValueError(f"job {pending[0].job_id} needs more than the HBM budget")
            # completions
            while inflight and self.batch.query(inflight[0][0]):
                done_work += inflight.popleft()[1]
                if busy_since is not None and now > busy_since:
                    self._rate = done_work / (now - busy_since)
            if not inflight:
                busy_since, done_work = None, 0.0
            for j in list(active):
                if j._next_chunk == len(j._chunk_spaces) and self.batch.query(j._last_ticket):
                    j.done = time.perf_counter() - t0
                    active.remove(j)
                    self.used -= j.footprint(self.chunk)
                    finished.append(j)
            ready = [j for j in active if j._next_chunk < len(j._chunk_spaces)]
            if ready and self.policy in ("srpt", "oldest", "aged"):
                if self.policy == "oldest":
                    ready.sort(key=lambda j: (j.arrival, j.job_id))
                elif self.policy == "aged" and self._rate:
                    ready.sort(key=lambda j: (j.remaining_cost(self.chunk) / self._rate -
                                              self.age_weight * (now - j.arrival), j.arrival))
                else:
                    ready.sort(key=lambda j: (j.remaining_cost(self.chunk), j.arrival))
                picked, work = [], 0.0
                for j in ready:
                    if picked and work >= self.tick_work:
                        break
                    picked.append(j)
                    work += j.chunk_cost(j._next_chunk, self.chunk)
                ready = picked
            if ready and len(inflight) < self.depth:
                bjobs, ends = [], []
                for j in ready:
                    bj, cnt = self._batch_job(j, j._next_chunk)
                    bjobs.append(bj)
                    ends.append(cnt)
                self.batch.step(bjobs, [0] * len(bjobs), ends, self.strategy)
                if count_updates:    # (the counting variant waits for each step)
                    self.total_updates += self.batch.last_updates()
                ticket = self.batch.record()
                if busy_since is None:
                    busy_since = time.perf_counter() - t0
                inflight.append((ticket, sum(j.chunk_cost(j._next_chunk, self.chunk)
                                             for j in ready)))
                for j in ready:
                    j._next_chunk += 1
                    j._last_ticket = ticket
            elif not ready and not pending:
                time.sleep(0)
            elif not ready or len(inflight) >= self.depth:
                time.sleep(20e-6)
        self.batch.synchronize()
        return finished


def latency_stats(jobs):
    lat = np.array([j.done - j.arrival for j in jobs])
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "max_s": float(lat.max())}
