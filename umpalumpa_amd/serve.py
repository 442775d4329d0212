"""Continuous-batching server for FR reconstruction jobs on one GPU.

Jobs arrive with a slice count, box, blob radius and symmetry group.  The server admits them
first come, first served while their device footprint (volume, weights, traverse spaces, a ring
of slice chunks, blob table) fits the HBM budget, keeps the admitted volumes resident, and on
every tick inserts the next traverse spaces of the admitted jobs, in priority order up to a
quantum of about 10 ms of work (one job at most half of it while others are ready), with one
batch step (``umpa_fr_batch_step``): one launch per group of jobs sharing settings, so jobs
that arrive while others run are inserted by the same launches.  A job's voxels receive its
spaces in index order whatever else is in flight, so its result is the same alone and in the
mix (bit for bit under the exact strategy).  Ticks are enqueued ahead of the GPU (up to
``depth`` in flight) and completions are noticed through stream events, so the host loop
never waits on an idle GPU.  The loop only plans ticks: each lane's stream work (ring copies,
the step's work-item schedule, the launches) runs on a thread of its own, so a step over large
volumes (hundreds of thousands of work items, milliseconds of host work) does not hold up
admissions or the other lane.  Small jobs (under about 20 ms of work) run on a second lane, a
batch with its own stream, whose launches take CUs as the main lane's bounded launches free
them, instead of queueing behind its ticks.  By default each tick serves the jobs with the
least remaining work first ("srpt"), so short jobs overtake long ones; "aged" lets a long job
that has waited take its turn.

Device memory comes from torch's caching allocator.  Finished jobs keep their volumes (the
results) until the caller drops them, and when an allocation fails the allocator frees every
cached block, each free synchronising the device: a multi-second stall of the loop.  A
long-running caller releases finished volumes and empties the cache at quiet moments (bench.py
does so before each timed run).

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
    ([m, Hy, Hx] complex64, or ([m, Hy, Hx, 2] float16, [m] float32 scales) for FP16).  The
    server calls it from the serving loop (chunk 0, for its shape) and from its lane threads
    (each chunk it copies into the job's ring), there with the loop's current stream as the
    current stream; the copies follow that stream's work."""
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
    ready: float | None = None     # admission finished (spaces resident)
    started: float | None = None   # first tick submitted
    done: float | None = None
    volume: object = None
    weights: object = None
    _spaces: object = None         # all traverse spaces, chunk after chunk
    _table: object = None
    _chunk_spaces: list = field(default_factory=list)   # (first space, spaces) per chunk
    _cursor: int = 0               # spaces already enqueued
    _filled: int = 0               # chunks already copied into the ring
    _ring: object = None           # (slices, scales or None) of `_slots` chunk slots
    _slots: int = 1
    _last: object = None           # future of the last tick that holds the job
    _order: int | None = None
    _slice_bytes: int | None = None
    _prep: object = None           # future of the host-side admission work
    _lane: object = None           # the server lane the job runs on

    def _k(self) -> int:
        if self._order is None:
            self._order = len(F.symmetry_operators(self.symmetry))
        return self._order

    def n_spaces(self) -> int:
        return self._k() * self.n_slices

    def remaining_cost(self, chunk: int = 0) -> float:
        """Relative insertion work not yet enqueued: spaces x slice area."""
        return float((self.n_spaces() - self._cursor) * (self.box + 1) ** 2)

    def space_cost(self) -> float:
        """Relative insertion work of one traverse space of this job."""
        return float((self.box + 1) ** 2)

    def slice_bytes(self) -> int:
        if self._slice_bytes is None:
            data = self.slices(0)
            self._slice_bytes = (data[0][0].numel() * 2 + 4) if isinstance(data, tuple) \
                else data[0].numel() * 8
        return self._slice_bytes

    def ring_slots(self, chunk: int, ring_bytes: float) -> int:
        """Chunk slots of the job's slice ring: as many as fit ring_bytes, 1 to 8, no more
        than the job has chunks."""
        n_chunks = -(-self.n_slices // chunk)
        return max(1, min(8, n_chunks, int(ring_bytes // (chunk * self.slice_bytes()))))

    def footprint(self, chunk: int, ring_bytes: float = 512e6) -> int:
        """Device bytes the job holds while admitted: volume and weights, traverse spaces, the
        slice ring."""
        return 12 * (self.box + 1) ** 3 + 112 * self.n_spaces() + 4 * 10000 + \
            self.ring_slots(chunk, ring_bytes) * chunk * self.slice_bytes()


class _Lane:
    """One batch stream of the server with its ticks in flight.  Its stream work is enqueued
    by one worker thread, in submission order; only that thread calls into the batch while a
    run is in progress."""

    def __init__(self, device, depth, wave_items):
        import concurrent.futures
        self.batch = F.FRBatch(device)
        if wave_items:
            self.batch.set_wave_items(wave_items)
        self.depth = depth
        self.worker = concurrent.futures.ThreadPoolExecutor(max_workers=1)
        # (future, work) of submitted ticks; a future's result is (event recorded after the
        # tick's launches, voxel-updates counted)
        self.inflight = collections.deque()


class Server:
    """Admission by HBM budget, resident volumes, chunked insertion with continuous batching."""

    def __init__(self, device: int = 0, chunk: int = 512, hbm_budget: float = 200e9,
                 strategy: int = F.FR_FAST_MATH, depth: int = 2, alpha: float = 15.0,
                 max_active: int = 256, policy: str = "srpt",
                 tick_work: float = 4 * 512 * 257 ** 2,
                 age_weight: float = 2.0, tick_seconds: float | None = 1e-2,
                 share: float | None = 0.5, min_tick_work: float = 512 * 257 ** 2,
                 fast_seconds: float | None = 2e-2, wave_items: int = 2,
                 ring_bytes: float = 512e6, poll_seconds: float = 2e-4):
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
        # "fair": every admitted job advances by its ring window (up to 8 chunks) per tick
        if policy not in ("srpt", "oldest", "aged", "fair"):
            raise ValueError(policy)
        self.policy = policy
        self.tick_work = tick_work
        self.age_weight = age_weight
        # a tick's work: tick_seconds of the measured insertion rate (tick_work before the
        # rate is known, or with tick_seconds None), at least min_tick_work: the rate is
        # measured on the host clock, so ticks too short for the host to keep up would lower
        # it and shrink the next ticks further (at 2 ms, 26 spaces per tick and 40x less
        # throughput on 12 jobs of the serving mix).  With other jobs ready, one job takes at
        # most `share` of the tick, so a long job's chunk is spread over ticks and short jobs
        # get into every tick (None: whole chunks)
        self.tick_seconds = tick_seconds
        self.share = share
        self.min_tick_work = min_tick_work
        # Two lanes, each a batch on its own stream (default priority): jobs whose whole work
        # takes under fast_seconds at the measured rate (under tick_work before it is known)
        # run on the fast lane, every other job on the main lane.  Launches are bounded
        # (wave_items work items per wavefront, then the workgroup ends), so the fast lane's
        # launches take CUs as the main lane's workgroups end instead of waiting behind its
        # ticks in flight.  fast_seconds None: one lane.  A job stays on its lane, so its
        # spaces still reach its voxels in order.
        self.fast_seconds = fast_seconds
        # Each admitted job copies its slice chunks into a ring of up to 8 chunk slots (at most
        # ring_bytes) and its traverse spaces name the slot, so one step covers as many
        # consecutive chunks as the ring holds: a job advances by up to 8 chunks per tick
        # instead of one.  The copies are exact, so results do not change.
        self.ring_bytes = ring_bytes
        self.poll_seconds = poll_seconds
        self._rate = None          # work units per second, measured from completed ticks
        self._main = _Lane(device, depth, wave_items)
        # the fast lane's jobs are small by construction: all their chunks are enqueued at
        # once (up to 64 steps in flight)
        self._fast = _Lane(device, 64, wave_items) if fast_seconds else None
        self.batch = self._main.batch
        # admission off the serving loop: traverse spaces are made by worker threads (the
        # native generator releases the interpreter lock) and uploaded on a stream of their
        # own, which the job's lane waits for; the loop never blocks on either
        import concurrent.futures
        self._workers = concurrent.futures.ThreadPoolExecutor(max_workers=4)
        self._copy = torch.cuda.Stream(self.dev)
        self._tables = {}
        self.used = 0
        self.total_updates = 0

    def close(self) -> None:
        """Waits for the work in flight and ends the server's threads (admission workers and
        lane threads); the server takes no further jobs."""
        for ln in (self._main, self._fast):
            if ln is not None:
                ln.worker.shutdown(wait=True)
                ln.batch.synchronize()
        self._workers.shutdown(wait=True)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    def _settings(self, radius):
        return F.FRSettings(F.FR_LOOKUP, 0, F.FR_PRECISE, self.alpha, radius)

    def _table(self, box, radius):
        # the table samples the blob over [0, r^2] and does not depend on the box: jobs of
        # one radius share it, and with it a launch per tick
        key = radius
        if key not in self._tables:
            s = self._settings(radius)
            desc = F.fr_desc(box // 2 + 1, box, 1, box + 1, 1, 10000)
            self._tables[key] = self.torch.from_numpy(F.fill_blob_table(10000, desc, s)).to(
                self.dev)
        return self._tables[key]

    def _spaces_host(self, job: Job):
        """Host side of admission (a worker thread): the job's traverse spaces chunk by chunk
        (projection indices relative to the chunk's slices) in pinned memory."""
        s = self._settings(job.radius)
        sym = F.symmetry_operators(job.symmetry)
        recs, chunks, off = [], [], 0
        for c, c0 in enumerate(range(0, job.n_slices, self.chunk)):
            # projection indices name the chunk's slot in the job's ring
            r = F.generate_traverse_spaces(job.rotations[c0:c0 + self.chunk], job.box, s, sym,
                                           first_index=(c % job._slots) * self.chunk)
            chunks.append((off, len(r)))
            off += len(r)
            recs.append(r)
        host = self.torch.from_numpy(np.concatenate(recs).view(np.uint8)).pin_memory()
        return chunks, host

    def _is_small(self, job: Job) -> bool:
        """Does the job go to the fast lane: under fast_seconds of work at the measured rate
        (under tick_work before the rate is known)?"""
        if self._fast is None:
            return False
        work = job.remaining_cost(self.chunk)
        return work <= self.tick_work if not self._rate else \
            work / self._rate <= self.fast_seconds

    def _admit(self, job: Job, now: float):
        torch = self.torch
        job._lane = self._fast if self._is_small(job) else self._main
        job._slots = job.ring_slots(self.chunk, self.ring_bytes)
        data = job.slices(0)
        first = data[0] if isinstance(data, tuple) else data
        n = job._slots * self.chunk
        with torch.cuda.stream(job._lane.batch.stream):
            job.volume = torch.zeros((job.box + 1,) * 3, dtype=torch.complex64, device=self.dev)
            job.weights = torch.zeros((job.box + 1,) * 3, dtype=torch.float32, device=self.dev)
            job._ring = (torch.empty((n,) + tuple(first.shape[1:]), dtype=first.dtype,
                                     device=self.dev),
                         torch.empty(n, dtype=torch.float32, device=self.dev)
                         if isinstance(data, tuple) else None)
        job._table = self._table(job.box, job.radius)
        job._prep = self._workers.submit(self._spaces_host, job)
        job.admitted = now
        self.used += job.footprint(self.chunk, self.ring_bytes)

    def _ready(self, job: Job) -> bool:
        """Has the job's admission finished?  Completes it (the upload) when the host work
        is done."""
        if job._spaces is not None:
            return True
        if not job._prep.done():
            return False
        chunks, host = job._prep.result()
        torch = self.torch
        with torch.cuda.stream(self._copy):
            dev = host.to(self.dev, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(self._copy)
        stream = job._lane.batch.stream
        stream.wait_event(ev)
        dev.record_stream(stream)
        job._chunk_spaces = chunks
        job._spaces = dev
        job._prep = None
        job.ready = time.perf_counter() - self._t0
        return True

    def _window_end(self, job: Job) -> int:
        """End of the spaces one step may cover from the job's cursor: those of the chunks the
        ring holds together with the cursor's chunk."""
        per = job._k() * self.chunk   # spaces of a full chunk
        return min(job.n_spaces(), (job._cursor // per + job._slots) * per)

    def _fill(self, job: Job, end: int, stream):
        """Copies the chunks with spaces below `end` into the ring (chunk c to slot c mod
        slots), on `stream`: the steps that read a slot's previous chunk precede the copy
        there."""
        torch = self.torch
        per = job._k() * self.chunk
        ring, ring_sc = job._ring
        while job._filled <= (end - 1) // per:
            c = job._filled
            o = (c % job._slots) * self.chunk
            data = job.slices(c)
            with torch.cuda.stream(stream):
                if isinstance(data, tuple):
                    m = data[0].shape[0]
                    ring[o:o + m].copy_(data[0], non_blocking=True)
                    ring_sc[o:o + m].copy_(data[1], non_blocking=True)
                else:
                    m = data.shape[0]
                    ring[o:o + m].copy_(data, non_blocking=True)
            for t in (data if isinstance(data, tuple) else (data,)):
                t.record_stream(stream)
            job._filled += 1

    @staticmethod
    def _done(fut) -> bool:
        """Has a submitted tick run on the GPU?  (Raises what the tick's thread raised.)"""
        return fut.done() and fut.result()[0].query()

    def _retire(self, lane) -> float:
        """Drops the lane's completed ticks (in order); returns the work they held."""
        work = 0.0
        while lane.inflight and self._done(lane.inflight[0][0]):
            fut, w = lane.inflight.popleft()
            self.total_updates += fut.result()[1]
            work += w
        return work

    def _enqueue(self, lane, plan, caller, count_updates):
        """The lane thread's part of a tick: ring copies, the step, an event after it."""
        torch = self.torch
        stream = lane.batch.stream
        with torch.cuda.stream(caller):
            # the slices' producers (on the caller's stream) precede the copies into the rings
            stream.wait_stream(caller)
            bjobs = []
            for j, _, e in plan:
                self._fill(j, e, stream)
                bjobs.append(F.BatchJob(j._ring[0], j.volume, j.weights, j._spaces,
                                        self._settings(j.radius), j._table, scales=j._ring[1]))
            lane.batch.step(bjobs, [b for _, b, _ in plan], [e for _, _, e in plan],
                            self.strategy)
            n = lane.batch.last_updates() if count_updates else 0
            ev = torch.cuda.Event()
            ev.record(stream)
        return ev, n

    def _tick(self, lane, ready, now, count_updates):
        """Submits one tick of `lane` over its ready jobs (in priority order) to the lane's
        thread; returns the work it holds."""
        # the tick: jobs in priority order, each with the spaces from its cursor to the end of
        # its ring window that fit its share of the quantum (the whole quantum when it is
        # alone), until the quantum is full; "fair", and the fast lane, give every ready job
        # its whole window
        quantum = self.tick_work if not (self.tick_seconds and self._rate) \
            else max(self._rate * self.tick_seconds, self.min_tick_work)
        cap = quantum if len(ready) == 1 or self.share is None else self.share * quantum
        whole = self.policy == "fair" or lane is self._fast
        plan, work = [], 0.0
        for j in ready:
            if plan and work >= quantum and not whole:
                break
            left = self._window_end(j) - j._cursor
            if whole:
                n_sp = left
            else:
                room = max(quantum - work, 0.0)
                n_sp = max(1, min(left, int(min(cap, room if plan else cap) / j.space_cost())))
            plan.append((j, j._cursor, j._cursor + n_sp))
            work += n_sp * j.space_cost()
        fut = lane.worker.submit(self._enqueue, lane, plan,
                                 self.torch.cuda.current_stream(self.dev), count_updates)
        lane.inflight.append((fut, work))
        for j, _, e in plan:
            if j.started is None:
                j.started = time.perf_counter() - self._t0
            j._last = fut
            j._cursor = e
        return work

    def run(self, jobs, realtime: bool = True, count_updates: bool = False):
        """Serves `jobs` (sorted by arrival).  With `realtime`, a job becomes visible at its
        arrival time on the host clock; otherwise all are present at once.  Returns the jobs
        with admitted/done times (seconds from the start)."""
        pending = collections.deque(sorted(jobs, key=lambda j: j.arrival))
        active: list[Job] = []
        lanes = [ln for ln in (self._fast, self._main) if ln is not None]
        busy_since, done_work = None, 0.0   # main lane: the insertion rate
        main = self._main
        n_main = 0                          # admitted main-lane jobs
        t0 = self._t0 = time.perf_counter()
        for ln in lanes:
            ln.batch.set_count_updates(count_updates)
        self.total_updates = 0
        finished = []
        while pending or active:
            now = time.perf_counter() - t0
            # admission: FIFO within the budget and the slot limit (of main-lane jobs: a
            # backlog of long jobs does not keep short ones out)
            while pending and (not realtime or pending[0].arrival <= now) and \
                    (n_main < self.max_active or self._is_small(pending[0])) and \
                    self.used + pending[0].footprint(self.chunk, self.ring_bytes) <= self.budget:
                j = pending.popleft()
                self._admit(j, now)
                active.append(j)
                n_main += j._lane is main
            if pending and not active and (not realtime or pending[0].arrival <= now) and \
                    pending[0].footprint(self.chunk, self.ring_bytes) > self.budget:
                raise ValueError(f"job {pending[0].job_id} needs more than the HBM budget")
            # completions
            w = self._retire(main)
            if w:
                done_work += w
                if busy_since is not None and now > busy_since:
                    self._rate = done_work / (now - busy_since)
            if not main.inflight:
                busy_since, done_work = None, 0.0
            if self._fast is not None:
                self._retire(self._fast)
            for j in list(active):
                if self._ready(j) and j._cursor == j.n_spaces() and \
                        (j._last is None or self._done(j._last)):
                    j.done = time.perf_counter() - t0
                    active.remove(j)
                    n_main -= j._lane is main
                    self.used -= j.footprint(self.chunk, self.ring_bytes)
                    j._ring = j._spaces = None   # the volume and weights stay: the result
                    finished.append(j)
            ticked = False
            for lane in lanes:
                ready = [j for j in active if j._lane is lane and j._spaces is not None and
                         j._cursor < j.n_spaces()]
                if not ready:
                    continue
                if len(lane.inflight) >= lane.depth:
                    continue
                if self.policy == "oldest":
                    ready.sort(key=lambda j: (j.arrival, j.job_id))
                elif self.policy == "aged" and self._rate:
                    ready.sort(key=lambda j: (j.remaining_cost(self.chunk) / self._rate -
                                              self.age_weight * (now - j.arrival), j.arrival))
                elif self.policy != "fair":
                    ready.sort(key=lambda j: (j.remaining_cost(self.chunk), j.arrival))
                self._tick(lane, ready, now, count_updates)
                ticked = True
                if lane is main and busy_since is None:
                    busy_since = time.perf_counter() - t0
                while lane is self._fast and len(lane.inflight) < lane.depth:
                    ready = [j for j in ready if j._cursor < j.n_spaces()]
                    if not ready:
                        break
                    self._tick(lane, ready, now, count_updates)
            if ticked:
                continue
            # idle: sleep long enough for the lane threads to take the interpreter lock (a
            # loop that takes it back at once keeps them waiting for the interpreter's switch
            # interval, 5 ms, every time they return from a native call)
            time.sleep(self.poll_seconds)
        for ln in lanes:
            for fut, _ in ln.inflight:
                fut.result()
            ln.batch.synchronize()
            self._retire(ln)
        return finished


def latency_stats(jobs):
    lat = np.array([j.done - j.arrival for j in jobs])
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "max_s": float(lat.max())}
