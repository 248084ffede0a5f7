"""Serving loop: a stream of independent simulation jobs on one GPU.

Jobs arrive at seeded Poisson times.  The loop admits waiting jobs in arrival order while their
device-memory estimate fits the budget, keeps every admitted job's lattice resident, advances all
of them together in chunks of steps (``Batch``: jobs whose step is the fused kernel alone advance
in batched launches, one launch per kernel variant and step covering the lines of all of them,
replayed from captured HIP graphs; the others launch their own steps on their own streams, so
their work overlaps on the device), and retires a job as soon as its steps are done, which frees
its memory for the next arrivals (continuous batching).  It reports
each job's completion latency (completion - arrival) and the aggregate lattice updates per second.

The mix follows the serving workload: D2Q9 or D3Q19 MRT cases, 2D sides log-uniform in [64, 512],
3D sides log-uniform in [32, 128], step counts uniform in [500, 5000], FP32 storage, about 30 %
with a salinity field and 10 % with an immersed-boundary body (a cylinder or sphere of 200-1800
markers).  FP16 storage is not offered: it is switched off in the library.
"""
from __future__ import annotations

import concurrent.futures as cf
import dataclasses
import math
import time
from typing import List, Optional

import numpy as np

from . import Batch, Simulation, _abi, cases, estimate_device_bytes


@dataclasses.dataclass
class Job:
    jid: int
    case: "cases.Case"
    steps: int
    arrival: float             # seconds after the start of the run
    bytes: int = 0
    done_steps: int = 0
    start: Optional[float] = None
    finish: Optional[float] = None
    sim: Optional[Simulation] = None

    @property
    def cells(self) -> int:
        return self.case.n_cells

    @property
    def latency(self) -> float:
        return self.finish - self.arrival


def _log_uniform(u: float, lo: int, hi: int) -> int:
    return int(round(math.exp(math.log(lo) + u * (math.log(hi) - math.log(lo)))))


def _job_case(rng: np.random.Generator, jid: int, storage: int) -> "cases.Case":
    three_d = rng.random() < 0.5
    kind = rng.random()
    salinity = kind < 0.3
    ibm = 0.3 <= kind < 0.4
    seed = cases.SEED_BASE + 1000 + jid
    if three_d:
        nx, ny, nz = (_log_uniform(rng.random(), 32, 128) for _ in range(3))
        if ibm:
            m = int(rng.integers(200, 1801))
            d = max(6.0, min(nx, ny, nz) / 4.0)
            # light gravity: the body drifts a few cells in 5000 steps and stays clear of the walls
            c = cases.falling_sphere(n=(nx, ny, nz), diameter=d, markers=m, storage=storage, tau=0.7, gravity=5e-6,
                                     density_ratio=3.0, method=_abi.IB_STRESS, kernel=_abi.IB_KERNEL_REFERENCE)
            c.body.pos = (nx / 2.0, ny / 2.0, nz / 2.0)
            return c.replace(name=f"job{jid}_sphere_{nx}x{ny}x{nz}")
        c = cases.lock_exchange(nx=nx, ny=ny, nz=nz, storage=storage, seed=seed,
                                scalar_lattice=7 if salinity else 0)
        if not salinity:
            c = c.replace(scalar=_abi.SCALAR_NONE, scalar_lattice=0, gravity=0.0, init={
                k: v for k, v in c.init.items() if k != "S"})
        return c.replace(name=f"job{jid}_{'le' if salinity else 'box'}_{nx}x{ny}x{nz}")
    nx, nz = _log_uniform(rng.random(), 64, 512), _log_uniform(rng.random(), 64, 512)
    if ibm:
        m = int(rng.integers(200, 1801))
        return cases.cylinder_2d(n=(nx, nz), markers=m, storage=storage, gravity=5e-6).replace(
            name=f"job{jid}_cyl_{nx}x{nz}")
    c = cases.lattice_taylor_green(nx=nx, nz=nz, storage=storage, seed=seed)
    if salinity:
        X, _, Z = c.coords()
        S = 0.5 * (1.0 - np.tanh((X - X.mean()) / 3.0))
        c = c.replace(scalar=_abi.SCALAR_SALINITY, scalar_lattice=0, taus=0.52, beta_s=1e-3, gravity=1e-2,
                      init=dict(c.init, S=S))
    return c.replace(name=f"job{jid}_tg_{nx}x{nz}")


def serving_mix(n_jobs: int = 256, mean_gap_s: float = 0.02, steps=(500, 5000), seed: int = cases.SEED_BASE + 5,
                storage: int = _abi.F32) -> List[Job]:
    """The seeded job stream: exponential inter-arrival gaps of mean `mean_gap_s`."""
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for j in range(n_jobs):
        t += float(rng.exponential(mean_gap_s))
        c = _job_case(rng, j, storage)
        jobs.append(Job(jid=j, case=c, steps=int(rng.integers(steps[0], steps[1] + 1)), arrival=t))
    return jobs


class Server:
    """Continuous batching under a device-memory budget.

    Admitted jobs are set up (host-side lattice set-up and upload, ``ipl_create``) by `workers`
    threads while the batch steps on the device, so set-up of new arrivals overlaps the running
    jobs' work; a job joins the batch at the first chunk boundary after its set-up ends.  A
    finished job leaves the batch at once and is released (``ipl_destroy``: its graph, stream and
    device memory) by a separate thread, also while the batch steps; its bytes return to the
    budget when the release has ended."""

    def __init__(self, budget_bytes: float = 200e9, chunk: int = 400, library: str = None, workers: int = 4,
                 batched: bool = True, max_lines: int = 2048):
        self.budget = int(budget_bytes)
        self.chunk = chunk
        self.library = library
        self.workers = workers
        self.batched = batched  # batched launches (else every job launches its own steps)
        # jobs of more than max_lines workgroups per step launch their own steps on their own streams
        # (their launches are large enough alone); the smaller ones share batched launches (None: all)
        self.max_lines = max_lines

    def _create(self, job: Job):
        t = time.perf_counter()
        sim = Simulation(job.case, self.library)
        try:
            sim.set_shared(True)  # it steps beside the other jobs (as a batch member)
            sim.prepare()  # the job's step graphs, captured here rather than in the stepping loop
        except Exception:
            sim.close()
            raise
        return sim, time.perf_counter() - t

    @staticmethod
    def _release(sim: Simulation) -> float:
        t = time.perf_counter()
        sim.close()
        return time.perf_counter() - t

    def run(self, jobs: List[Job], log=None, on_done=None) -> dict:
        """Serves `jobs` to completion.  `on_done(job)` is called for each finished job while its
        simulation is still open (to read its fields), `log(str)` with one line per finished job."""
        pending = sorted(jobs, key=lambda j: j.arrival)
        for j in pending:
            j.bytes = estimate_device_bytes(j.case, self.library)
            if j.bytes > self.budget:
                raise ValueError(f"job {j.jid} needs {j.bytes} B, more than the whole budget")
        active: List[Job] = []
        captures: List[int] = []
        plain_jobs = []  # (job id, why the fused path refused it)
        creating = {}
        releasing = {}
        used = 0
        done: List[Job] = []
        busy = create = close = 0.0
        t0 = time.perf_counter()
        with Batch(self.library, batched=self.batched, max_lines=self.max_lines) as batch, cf.ThreadPoolExecutor(max_workers=self.workers) as pool, \
                cf.ThreadPoolExecutor(max_workers=1) as closer:
            try:
                while pending or active or creating:
                    for fut in [f for f in releasing if f.done()]:
                        j = releasing.pop(fut)
                        close += fut.result()
                        used -= j.bytes
                    now = time.perf_counter() - t0
                    # admit in arrival order while the memory estimate fits
                    while pending and pending[0].arrival <= now and used + pending[0].bytes <= self.budget:
                        j = pending.pop(0)
                        used += j.bytes
                        creating[pool.submit(self._create, j)] = j
                    for fut in [f for f in creating if f.done()]:
                        j = creating.pop(fut)
                        j.sim, dt = fut.result()
                        create += dt
                        if getattr(j.sim, "path", _abi.PATH_FUSED) == _abi.PATH_PLAIN:
                            plain_jobs.append((j.jid, j.sim.path_reason))
                            if log:
                                log(f"[serve] job {j.jid} {j.case.name} runs on the plain path: {j.sim.path_reason}")
                        j.start = time.perf_counter() - t0
                        batch.add(j.sim)
                        active.append(j)
                    if not active:
                        nxt = pending[0].arrival - (time.perf_counter() - t0) if pending else 0.05
                        waits = list(creating) + list(releasing)
                        if waits:
                            # an arrived job held back by the budget waits for a set-up or release
                            cf.wait(waits, timeout=0.05 if nxt <= 0 else min(nxt, 0.05), return_when=cf.FIRST_COMPLETED)
                        elif nxt > 0:
                            time.sleep(min(nxt, 0.05))
                        continue
                    counts = [min(self.chunk, j.steps - j.done_steps) for j in active]
                    tb = time.perf_counter()
                    batch.step(counts)
                    busy += time.perf_counter() - tb
                    now = time.perf_counter() - t0
                    for j, k in zip(list(active), counts):
                        j.done_steps += k
                        if j.done_steps >= j.steps:
                            j.finish = now
                            captures.append(j.sim.graph_captures())
                            if on_done:
                                on_done(j)
                            batch.remove(j.sim)
                            releasing[closer.submit(self._release, j.sim)] = j
                            j.sim = None
                            active.remove(j)
                            done.append(j)
                            if log:
                                log(f"[serve] job {j.jid} {j.case.name} done: latency {j.latency:.3f} s")
                for fut in list(releasing):
                    close += fut.result()
                    used -= releasing.pop(fut).bytes
                counters = batch.counters() if hasattr(batch, "counters") else {}
            finally:
                for fut in releasing:
                    fut.exception()
                for fut, j in creating.items():  # after an error: release what the workers set up
                    try:
                        fut.result()[0].close()
                    except Exception:
                        pass
                for j in active:
                    if j.sim is not None:
                        j.sim.close()
        makespan = max(j.finish for j in done)
        lat = np.array([j.latency for j in done])
        updates = float(sum(j.cells * j.steps for j in done))
        return {"jobs": len(done), "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
                "mean_s": float(lat.mean()), "makespan_s": float(makespan), "busy_s": float(busy),
                "setup_s": float(create), "close_s": float(close),
                "aggregate_mlups": updates / makespan / 1e6, "busy_mlups": updates / busy / 1e6,
                "cell_updates": updates, "budget_gb": self.budget / 1e9, "chunk": self.chunk, "batched": self.batched,
                "setup_workers": self.workers, "max_graph_captures_per_job": int(max(captures)),
                "plain_path_jobs": len(plain_jobs), "plain_path_reasons": sorted({r for _, r in plain_jobs}),
                "batch_counters": counters}
