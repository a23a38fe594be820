"""Continuous-batching request loop for drug-screening jobs.

The reference has no serving layer: each MPI rank runs its share of (drug sample, conc)
jobs to completion, one cell at a time (DrugSimulation/main.cpp:206-208,
CEP7Cell/modules/drugs_herg.cpp:113-335).  Here a job is a set of cells of one model with
their drug / IC50 / conductance parameters and a beat count; the server

* admits jobs in arrival order as device memory (a byte budget, each job sized with
  ``device_bytes``) and launch slots free up,
* keeps every admitted job's cells resident on the device between beats (one
  ``CellBatch`` per job, each with its own HIP stream),
* advances each admitted job ``chunk_beats`` beats per launch with ``run_beats_async``
  and polls completion without blocking, so launches of different jobs overlap on the
  GPU and a job that arrives while others run starts as soon as it is admitted,
* reports per-job latency (arrival -> biomarkers on the host) and the cell-beats done.

A job's result does not depend on what else runs: its cells live in their own batch, and
inside a batch every cell is integrated independently of its neighbours.

Launches of different jobs overlap only on different hardware queues: HIP maps streams onto
``GPU_MAX_HW_QUEUES`` queues (4 unless set, some taken by the runtime), and a beat kernel holds
its queue for up to a beat of device time.  A serving process sets ``GPU_MAX_HW_QUEUES`` well
above ``max_active`` (e.g. 24 for 16) in its environment before the first HIP call.  Device
memory of finished jobs is recycled by the library (``release_cached_memory`` returns it):
freeing it would wait for every job's running kernel.
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import MODEL_ORD_CIPA, MODEL_TT06, PREC_MIXED, CellBatch, device_bytes


@dataclass
class Job:
    """One screening request.  ``cells`` holds the per-cell arrays of ``CellBatch.set_cells``
    (celltype, conc, herg, gscale, bcl)."""
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
    inflight: int = 0
    biomarkers: np.ndarray | None = None
    states: np.ndarray | None = None
    status: np.ndarray | None = None
    bytes: int = 0
    batch: object = field(default=None, repr=False)

    @property
    def ncells(self) -> int:
        return len(self.cells["bcl"])

    @property
    def latency(self) -> float:
        return self.t_done - self.arrival

    def step_dt(self) -> float:
        if self.dt > 0:
            return self.dt
        return 0.1 if self.model == MODEL_ORD_CIPA else 0.02


class Server:
    """Admission + launch loop.  ``batch_factory`` builds the per-job batch (the default is
    the HIP ``CellBatch``; tests pass a host stand-in to exercise the scheduling alone)."""

    def __init__(self, mem_budget: int, max_active: int = 16, chunk_beats: int = 1,
                 batch_factory=None, sizer=None):
        self.mem_budget = int(mem_budget)
        self.max_active = int(max_active)
        self.chunk_beats = int(chunk_beats)
        self.pending: deque[Job] = deque()
        self.active: list[Job] = []
        self.finished: list[Job] = []
        self.mem_used = 0
        self.cell_beats = 0
        self._factory = batch_factory or CellBatch
        self._sizer = sizer or (lambda j: device_bytes(j.model, j.ncells, float(np.max(j.cells["bcl"])),
                                                       j.step_dt()))

    # ---- request side
    def submit(self, job: Job, now: float | None = None) -> None:
        job.arrival = time.perf_counter() if now is None else now
        job.bytes = self._sizer(job)
        if job.bytes > self.mem_budget:
            raise ValueError(f"job {job.jid} needs {job.bytes} B, above the budget {self.mem_budget} B")
        self.pending.append(job)

    # ---- scheduling
    def _admit(self) -> None:
        # first come, first served: the queue head waits for memory, later jobs wait behind it
        while self.pending and len(self.active) < self.max_active:
            job = self.pending[0]
            if self.mem_used + job.bytes > self.mem_budget:
                break
            self.pending.popleft()
            b = self._factory(job.model, job.ncells, job.precision)
            c = job.cells
            b.set_cells(celltype=c["celltype"], conc=c["conc"], herg=c["herg"], gscale=c["gscale"],
                        bcl=c["bcl"])
            nwin = job.nwin if job.nwin > 0 else job.nbeats
            b.set_run_config(dt=job.step_dt(), nbeats_total=job.nbeats, nwin=min(nwin, job.nbeats))
            b.init_consts()
            job.batch = b
            job.t_admit = time.perf_counter()
            self.mem_used += job.bytes
            self.active.append(job)
            self._launch(job)

    def _launch(self, job: Job) -> None:
        k = min(self.chunk_beats, job.nbeats - job.beats_done)
        job.batch.run_beats_async(k)
        job.inflight = k

    def _finish(self, job: Job) -> None:
        b = job.batch
        job.biomarkers = b.biomarker_array()
        job.states = b.states()
        job.status = b.status()
        job.t_done = time.perf_counter()
        b.close()
        job.batch = None
        self.mem_used -= job.bytes
        self.finished.append(job)

    def poll(self) -> list[Job]:
        """One scheduling round: retire finished launches, relaunch or complete their jobs,
        admit waiting jobs.  Returns the jobs completed in this round."""
        done = []
        still = []
        for job in self.active:
            if job.batch.query():
                still.append(job)
                continue
            job.beats_done += job.inflight
            self.cell_beats += job.inflight * job.ncells
            if job.beats_done >= job.nbeats:
                self._finish(job)
                done.append(job)
            else:
                self._launch(job)
                still.append(job)
        self.active = still
        self._admit()
        return done

    def idle(self) -> bool:
        return not self.pending and not self.active

    def drain(self, sleep: float = 1e-4) -> None:
        while not self.idle():
            self.poll()
            if self.active:
                time.sleep(sleep)


def latency_percentiles(jobs: list[Job]) -> dict:
    lat = np.array([j.latency for j in jobs]) * 1e3
    if lat.size == 0:
        return {"p50": None, "p99": None, "n": 0}
    return {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)), "n": int(lat.size)}


def synthetic_jobs(seed: int, njobs: int, max_cells: int = 400, beats=(10, 500),
                   tt06_fraction: float = 0.5, precision: int = PREC_MIXED) -> list[Job]:
    """Screening requests of the serving workload: cells log-uniform in [1, max_cells], beats
    uniform in [beats[0], beats[1]], model ORd-CiPA or TT06 with the given probability, drug
    parameters drawn as in the population workloads."""
    from . import workload
    rng = np.random.default_rng(seed)
    jobs = []
    for j in range(njobs):
        n = int(np.floor(10 ** rng.uniform(0, np.log10(max_cells))))
        n = max(1, min(n, max_cells))
        nb = int(rng.integers(beats[0], beats[1] + 1))
        s = int(rng.integers(0, 2**31 - 1))
        if rng.uniform() < tt06_fraction:
            pop = workload.tt_population(s, n)
            cells = {"celltype": pop["celltype"], "conc": pop["conc"],
                     "herg": np.concatenate([pop["ic50"], pop["hill"]], axis=1),
                     "gscale": pop["gscale"], "bcl": pop["bcl"]}
            jobs.append(Job(j, MODEL_TT06, cells, nb, precision=precision))
        else:
            pop = workload.ord_population(s, 1, 1, n, bcl_choices=(1000.0,))
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
