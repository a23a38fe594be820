"""Continuous batching of drug-screening jobs at the level of cells.

The reference has no serving layer: each MPI rank runs its share of (drug sample, conc)
jobs to completion, one cell at a time (DrugSimulation/main.cpp:206-208,
CEP7Cell/modules/drugs_herg.cpp:113-335).  Here a job is a set of cells of one model with
their drug / IC50 / conductance parameters, a BCL and a beat count, and the server keeps
one *resident* batch per (model, precision, BCL, dt) -- a pool -- whose cells the jobs share:

* a job is admitted, in arrival order within its pool, as soon as the pool has as many free
  cells as the job has; admission happens at the pool's next beat boundary, while the jobs
  already in the pool keep running (``CellBatch.restart_cells`` starts the new cells from
  their model's initial states inside the live batch, ``park_cells`` gives a finished job's
  cells back);
* every launch advances every busy cell of the pool by one beat, so small jobs share
  wavefronts with other jobs instead of each holding a launch of its own;
* pools are created when their first job arrives, sized under a device-memory budget (each
  cell's bytes from ``device_bytes``); an idle pool is closed when another pool needs its
  memory;
* each pool runs on its own HIP stream and is polled without blocking, so pools of different
  models and BCLs overlap on the device;
* per job: latency from arrival to biomarkers on the host, and the cell-beats done.

A job's result does not depend on what else runs: every cell is integrated independently of
its neighbours, and a restarted cell computes exactly what the same cell computes in a batch
of its own.  Pools of one BCL keep the cells of a wavefront on beats of one length, so no
cell waits on a longer beat than its own.

Launches of different pools overlap only on different hardware queues: HIP maps streams onto
``GPU_MAX_HW_QUEUES`` queues (4 unless set, some taken by the runtime), and a beat kernel
holds its queue for a beat of device time.  A serving process sets ``GPU_MAX_HW_QUEUES``
above the number of pools in its environment before the first HIP call.
"""
from __future__ import annotations

import heapq
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import MODEL_ORD_CIPA, MODEL_TT06, PREC_MIXED, CellBatch, device_bytes

BCL_CHOICES = (500.0, 1000.0, 2000.0)
# cells per pool and model (capacity of a resident batch; the memory budget may cut it)
DEFAULT_CAPACITY = {MODEL_ORD_CIPA: 4096, MODEL_TT06: 8192}


@dataclass
class Job:
    """One screening request.  ``cells`` holds the per-cell arrays of ``CellBatch.set_cells``
    (celltype, conc, herg, gscale, bcl); all cells of a job share one BCL."""
    jid: int
    model: int
    cells: dict
    nbeats: int
    nwin: int = 0                  # biomarker window (0: all beats)
    precision: int = PREC_MIXED
    dt: float = 0.0                # 0: model default (ORd 0.1 ms, TT06 0.02 ms)
    arrival: float = 0.0           # host clock (time.perf_counter) at submission
    # filled by the server
    t_admit: float = 0.0
    t_done: float = 0.0
    beats_done: int = 0
    biomarkers: np.ndarray | None = None
    states: np.ndarray | None = None
    status: np.ndarray | None = None
    slots: np.ndarray | None = field(default=None, repr=False)

    @property
    def ncells(self) -> int:
        return len(self.cells["bcl"])

    @property
    def bcl(self) -> float:
        return float(self.cells["bcl"][0])

    @property
    def latency(self) -> float:
        return self.t_done - self.arrival

    def step_dt(self) -> float:
        if self.dt > 0:
            return self.dt
        return 0.1 if self.model == MODEL_ORD_CIPA else 0.02

    def key(self) -> tuple:
        return (self.model, self.precision, self.bcl, self.step_dt())


def _runs(idx: np.ndarray):
    """(first, count, offset) of the runs of consecutive values in a sorted index array."""
    if len(idx) == 0:
        return
    br = np.nonzero(np.diff(idx) != 1)[0] + 1
    starts = np.concatenate([[0], br])
    ends = np.concatenate([br, [len(idx)]])
    for s, e in zip(starts, ends):
        yield int(idx[s]), int(e - s), int(s)


class Pool:
    """A resident batch of ``capacity`` cells of one (model, precision, BCL, dt).  Parked
    cells are free; a job takes the lowest free cells, so busy cells stay packed into few
    wavefronts."""

    def __init__(self, key: tuple, capacity: int, factory, nbytes: int):
        self.key = key
        model, precision, bcl, dt = key
        self.capacity = capacity
        self.nbytes = nbytes
        b = factory(model, capacity, precision)
        # placeholder cells of the pool's BCL: the biomarker Ca trace is sized for it
        b.set_cells(celltype=np.zeros(capacity), conc=np.zeros(capacity), bcl=np.full(capacity, bcl))
        b.set_run_config(dt=dt, nbeats_total=1, nwin=1)
        b.init_consts()
        b.park_cells(np.arange(capacity))
        self.batch = b
        self.free = list(range(capacity))          # a heap of free cells
        self.jobs: list[Job] = []
        self.pending: deque[Job] = deque()
        self.inflight = False
        self.launches = 0

    def busy(self) -> bool:
        return bool(self.jobs or self.pending or self.inflight)

    def _admit(self, now: float) -> None:
        # first come, first served within the pool: the head waits for free cells
        while self.pending and len(self.free) >= self.pending[0].ncells:
            job = self.pending.popleft()
            slots = np.sort(np.array([heapq.heappop(self.free) for _ in range(job.ncells)], dtype=np.int64))
            c = job.cells
            for first, count, off in _runs(slots):
                sl = slice(off, off + count)
                self.batch.set_cells(celltype=c["celltype"][sl], conc=c["conc"][sl], herg=c["herg"][sl],
                                     gscale=c["gscale"][sl], bcl=c["bcl"][sl], first=first)
            nwin = job.nwin if job.nwin > 0 else job.nbeats
            self.batch.restart_cells(slots, job.nbeats, min(nwin, job.nbeats))
            job.slots = slots
            job.t_admit = now
            self.jobs.append(job)

    def poll(self) -> tuple[list[Job], int]:
        """Retire a finished launch, hand back the cells of the jobs it completed, admit waiting
        jobs and launch the next beat.  Returns (jobs completed, cell-beats done)."""
        done, cb = [], 0
        if self.inflight:
            if self.batch.query():
                return done, cb
            self.inflight = False
            for job in self.jobs:
                job.beats_done += 1
                cb += job.ncells
            done = [j for j in self.jobs if j.beats_done >= j.nbeats]
            if done:
                bm = self.batch.biomarker_array()
                st = self.batch.states()
                status = self.batch.status()
                now = time.perf_counter()
                for j in done:
                    j.biomarkers, j.states, j.status = bm[j.slots], st[j.slots], status[j.slots]
                    j.t_done = now
                freed = np.concatenate([j.slots for j in done])
                self.batch.park_cells(freed)
                for s in freed.tolist():
                    heapq.heappush(self.free, s)
                self.jobs = [j for j in self.jobs if j.beats_done < j.nbeats]
        self._admit(time.perf_counter())
        if self.jobs:
            self.batch.run_beats_async(1)
            self.inflight = True
            self.launches += 1
        return done, cb

    def close(self) -> None:
        if self.inflight:
            self.batch.sync()
            self.inflight = False
        self.batch.close()


class Server:
    """Admission + launch loop over the pools.  ``batch_factory`` builds a pool's batch (the
    default is the HIP ``CellBatch``; tests pass a host stand-in to exercise the scheduling
    alone); ``cell_bytes(model, bcl, dt)`` gives a cell's device bytes."""

    def __init__(self, mem_budget: int, capacity: dict | None = None, batch_factory=None, cell_bytes=None):
        self.mem_budget = int(mem_budget)
        self.capacity = {**DEFAULT_CAPACITY, **(capacity or {})}
        self.pools: dict[tuple, Pool] = {}
        self.waiting: deque[Job] = deque()      # jobs whose pool is not created yet
        self.finished: list[Job] = []
        self.mem_used = 0
        self.cell_beats = 0
        self._factory = batch_factory or CellBatch
        self._cell_bytes = cell_bytes or (lambda m, bcl, dt: -(-device_bytes(m, 1024, bcl, dt) // 1024))

    # ---- request side
    def submit(self, job: Job, now: float | None = None) -> None:
        job.arrival = time.perf_counter() if now is None else now
        if len(set(np.asarray(job.cells["bcl"], dtype=float).tolist())) != 1:
            raise ValueError(f"job {job.jid}: all cells of a job share one BCL")
        cap = self.capacity[job.model]
        if job.ncells > cap or job.ncells * self._cell_bytes(job.model, job.bcl, job.step_dt()) > self.mem_budget:
            raise ValueError(f"job {job.jid} ({job.ncells} cells) does not fit a pool "
                             f"({cap} cells, {self.mem_budget} B)")
        pool = self.pools.get(job.key())
        if pool is not None:
            pool.pending.append(job)
        else:
            self.waiting.append(job)

    # ---- pools
    def _pool_for(self, job: Job) -> Pool | None:
        key = job.key()
        if key in self.pools:
            return self.pools[key]
        per_cell = self._cell_bytes(job.model, job.bcl, job.step_dt())
        cap = min(self.capacity[job.model], (self.mem_budget - self.mem_used) // per_cell)
        if cap < job.ncells:
            # close idle pools until the new one fits
            for k, p in list(self.pools.items()):
                if not p.busy():
                    p.close()
                    self.mem_used -= p.nbytes
                    del self.pools[k]
            cap = min(self.capacity[job.model], (self.mem_budget - self.mem_used) // per_cell)
            if cap < job.ncells:
                return None
        cap = int(cap)
        pool = Pool(key, cap, self._factory, cap * per_cell)
        self.pools[key] = pool
        self.mem_used += pool.nbytes
        return pool

    def poll(self) -> list[Job]:
        """One scheduling round over every pool.  Returns the jobs completed in this round."""
        still = deque()
        while self.waiting:
            job = self.waiting.popleft()
            pool = self._pool_for(job)
            if pool is None:
                still.append(job)
            else:
                pool.pending.append(job)
        self.waiting = still
        done = []
        for pool in list(self.pools.values()):
            d, cb = pool.poll()
            self.cell_beats += cb
            done += d
        self.finished += done
        return done

    @property
    def active(self) -> list[Job]:
        return [j for p in self.pools.values() for j in p.jobs]

    @property
    def pending(self) -> list[Job]:
        return list(self.waiting) + [j for p in self.pools.values() for j in p.pending]

    def idle(self) -> bool:
        return not self.waiting and not any(p.busy() for p in self.pools.values())

    def drain(self, sleep: float = 1e-4) -> None:
        while not self.idle():
            self.poll()
            if not self.idle():
                time.sleep(sleep)

    def close(self) -> None:
        for p in self.pools.values():
            p.close()
        self.pools.clear()
        self.mem_used = 0


def latency_percentiles(jobs: list[Job]) -> dict:
    lat = np.array([j.latency for j in jobs]) * 1e3
    if lat.size == 0:
        return {"p50": None, "p99": None, "n": 0}
    return {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)), "n": int(lat.size)}


def synthetic_jobs(seed: int, njobs: int, max_cells: int = 400, beats=(10, 500),
                   tt06_fraction: float = 0.5, precision: int = PREC_MIXED,
                   bcl_choices=BCL_CHOICES) -> list[Job]:
    """Screening requests of the serving workload: cells log-uniform in [1, max_cells], beats
    uniform in [beats[0], beats[1]], model ORd-CiPA or TT06 with the given probability, one
    BCL per job drawn from bcl_choices, drug parameters drawn as in the population workloads."""
    from . import workload
    rng = np.random.default_rng(seed)
    jobs = []
    for j in range(njobs):
        n = int(np.floor(10 ** rng.uniform(0, np.log10(max_cells))))
        n = max(1, min(n, max_cells))
        nb = int(rng.integers(beats[0], beats[1] + 1))
        s = int(rng.integers(0, 2**31 - 1))
        bcl = float(rng.choice(np.asarray(bcl_choices, dtype=float)))
        if rng.uniform() < tt06_fraction:
            pop = workload.tt_population(s, n, bcl=bcl)
            cells = {"celltype": pop["celltype"], "conc": pop["conc"],
                     "herg": np.concatenate([pop["ic50"], pop["hill"]], axis=1),
                     "gscale": pop["gscale"], "bcl": pop["bcl"]}
            jobs.append(Job(j, MODEL_TT06, cells, nb, precision=precision))
        else:
            pop = workload.ord_population(s, 1, 1, n, bcl_choices=(bcl,))
            cells = {k: pop[k][:n] for k in ("celltype", "conc", "herg", "gscale", "bcl")}
            jobs.append(Job(j, MODEL_ORD_CIPA, cells, nb, precision=precision))
    return jobs


def run_poisson(server: Server, jobs: list[Job], rate: float, seed: int = 0, timeout: float = 600.0):
    """Submit `jobs` with exponential inter-arrival times of mean 1/rate (s) while the server
    runs; returns (finished jobs, wall seconds from first arrival to last result)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=len(jobs))
    t0 = time.perf_counter()
    due = t0 + np.cumsum(gaps) - gaps[0]
    k = 0
    while k < len(jobs) or not server.idle():
        now = time.perf_counter()
        while k < len(jobs) and due[k] <= now:
            server.submit(jobs[k], now=float(due[k]))
            k += 1
        server.poll()
        if now - t0 > timeout:
            raise TimeoutError("serving run exceeded its time limit")
        time.sleep(1e-4)
    return server.finished, time.perf_counter() - t0
