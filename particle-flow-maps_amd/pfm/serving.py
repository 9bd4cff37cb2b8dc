"""Serving: a stream of simulation jobs driven with continuous batching.

The reference runs one scene per process from its main loop
(3D/run_3D.py:618-871, 2D/run_2D.py:676-976).  A server instead takes jobs of
mixed 2D/3D resolutions as they arrive and keeps every admitted job's state
(particles, flow maps, face velocities, the multigrid hierarchy) resident in
device memory until it finishes:

* Admission.  A job is admitted when a slot is free and its device footprint
  (``footprint_bytes``, an upper bound of what its ``Simulation`` allocates)
  fits in what the running jobs leave of the memory budget.  Pending jobs are
  scanned in arrival order and the first that fits is taken (first fit), so a
  large job at the head does not hold back small ones behind it.
* Overlap.  Each slot is a host thread with its own HIP stream.  A job's steps
  are issued on its slot's stream; the library's calls release the GIL and
  its host synchronisations (the CFL max-speed read, the MGPCG convergence
  test) wait on that stream only, so the device interleaves the kernels of all
  running jobs while each job keeps the reference's step order.
* Results.  Every kernel of the step is deterministic (no float atomics, fixed
  reduction orders), so a job's fields are bit-identical whether it runs alone
  or in the mix.
* Multigrid hierarchies allocated by the library are released when the
  server is idle or stops (``hipFree`` synchronises the whole device, which
  would stall the other slots mid-step).
* Batched 2D lane (``batch_2d=True``).  2D jobs are not given a slot of their
  own: the admitted 2D jobs waiting for a lane form one batch
  (``pfm.batch.Batch2D``) that advances all of them together, one persistent
  launch per phase of the step, while other lanes and the 3D slots keep
  running.  A job that has done its steps is disabled in its batch (its
  fields stay as they are) and reported; the batch's memory counts against
  the budget until the whole batch has finished.  A batch's result for a job
  does not depend on the other jobs of the batch, so it is the same alone and
  in the mix; it uses the fp32 V-cycle.

Per-step latency is the host wall time from issuing a step to its completion
on the job's stream; ``ServeReport`` gives p50/p99 over every step of every
job and jobs/s over the stream.
"""
from __future__ import annotations

import collections
import dataclasses
import math
import queue
import threading
import time
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib as L
from . import core
from .simulation import Simulation, StepConfig


@dataclasses.dataclass(frozen=True)
class Job:
    """One simulation request: a scene of `res` cells run for `steps` steps."""

    id: int
    res: tuple
    steps: int
    seed: int
    arrival: float = 0.0  # seconds after the start of the stream

    @property
    def dim(self) -> int:
        return len(self.res)

    @property
    def ppc(self) -> int:
        return 4 if self.dim == 2 else 8


def _log_uniform_res(rng, lo, hi, mult=8):
    v = math.exp(rng.uniform(math.log(lo), math.log(hi)))
    return int(min(hi, max(lo, mult * round(v / mult))))


def job_stream(n_jobs: int, rate: float, seed: int = 3, frac_2d: float = 0.7, res_2d=(128, 1024),
               res_3d=(64, 192), steps=(20, 200)) -> List[Job]:
    """A seeded job stream: Poisson arrivals at `rate` jobs/s; 2D jobs with
    probability `frac_2d` at a square resolution log-uniform in `res_2d`, else
    3D cubes log-uniform in `res_3d` (both rounded to multiples of 8); step
    counts uniform in `steps` (inclusive)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for i in range(n_jobs):
        t += float(rng.exponential(1.0 / rate)) if rate > 0 else 0.0
        if rng.uniform() < frac_2d:
            n = _log_uniform_res(rng, *res_2d)
            res = (n, n)
        else:
            n = _log_uniform_res(rng, *res_3d)
            res = (n, n, n)
        k = int(rng.integers(steps[0], steps[1] + 1))
        jobs.append(Job(id=i, res=res, steps=k, seed=int(rng.integers(0, 2**31 - 1)), arrival=t))
    return jobs


def step_config(job: Job, mg_precision: int = 1) -> StepConfig:
    """The knobs of a served job: the reference schedule (reinit every 12 steps
    in 2D, 20 with a short reinit every step in 3D), CFL 0.5, MGPCG to a 1e-6
    relative residual, seeded jitter, cell-major particle numbering."""
    if job.dim == 2:
        return StepConfig(particles_per_cell=4, reinit_every=12, reinit_every_grad_m=12, CFL=0.5, tol=1e-6,
                          tol_mode="rel", seed=job.seed % (2**31), particle_order=1, mg_precision=mg_precision)
    return StepConfig(particles_per_cell=8, reinit_every=20, reinit_every_grad_m=1, CFL=0.5, tol=1e-6,
                      tol_mode="rel", seed=job.seed % (2**31), particle_order=1, write_imp_grad_m=False,
                      mg_precision=mg_precision)


_SLACK = 2 << 20  # allocation granularity allowance per device allocation


def mg_bytes(res, base_level: int = 3, precision: int = 0) -> int:
    """Upper bound of the device bytes pfm_mg_create allocates: per level r
    (fp32 on level 0, the storage type below), z and its ping-pong copy; x, p,
    its ping-pong copy and Ap on level 0, the right-hand-side copy in mixed
    precision, FP64 partials and the scalar block."""
    n_levels = max(1, int(math.log2(min(res))) - base_level + 1)
    cells = [math.prod(max(1, n >> l) for n in res) for l in range(n_levels)]
    esz = 4 if precision == 0 else 2
    b = 4 * cells[0] + esz * sum(cells[1:]) + 2 * esz * sum(cells)
    b += (16 + (4 if precision else 0)) * cells[0]
    return b + (3 * n_levels + 8) * _SLACK + (1 << 20)


def binning_bytes(res, n: int) -> int:
    """Device bytes of a Binning (core.Binning): the P2G workspace, cell starts,
    the permutation and the cell of every particle."""
    import ctypes

    g = L.make_grid(tuple(res), 1.0 / res[1])
    ws = int(L.load(require_gpu=False).pfm_p2g_workspace_size(ctypes.byref(g), int(n)))
    if ws == 0:
        raise L.PfmError(f"pfm_p2g_workspace_size failed: {L.last_error()}")
    return ws + 4 * (math.prod(res) + 1) + 8 * max(n, 1) + 3 * _SLACK


def footprint_bytes(job: Job, mg_precision: int = 1, base_level: int = 3) -> int:
    """Device bytes a served job holds while it runs (an upper bound): the
    tensors Simulation allocates, its Binning and MGPCG hierarchy, and the
    central-vector buffer of the initial condition."""
    d, res = job.dim, tuple(job.res)
    cells = math.prod(res)
    n = cells * job.ppc
    # particle SoA vectors of d floats: pos, init_imp, T_init/T_grad_m/init_C
    # (d each), and in 2D imp, T and C (d each) and imp_grad_m
    vecs = 2 + 3 * d + ((1 + 2 * d + 1) if d == 2 else 0)
    particles = 4 * d * n * vecs
    faces = sum(math.prod(r + (1 if a == c else 0) for c, r in enumerate(res)) for a in range(d))
    grid = 4 * faces * (2 + 2 * d)  # u, p2g weight, psi and T grids
    ic = 4 * d * cells
    n_tensors = d * (2 + 2 * d) + vecs * d + 2
    return (particles + grid + ic + n_tensors * _SLACK + binning_bytes(res, n)
            + mg_bytes(res, base_level, mg_precision))


class Admission:
    """First-fit admission by slots and device-memory footprint (host logic).

    ``used`` counts the footprints of running jobs.  A finished job may leave
    part of its footprint allocated until the server frees it (``retained``);
    that part still counts against the budget."""

    def __init__(self, budget_bytes: int, slots: int, footprint=footprint_bytes, takes_slot=None):
        if slots < 1:
            raise ValueError("need at least one slot")
        self.budget = int(budget_bytes)
        self.slots = int(slots)
        self.footprint = footprint
        # jobs for which takes_slot(job) is False (the batched 2D lane) are
        # limited by memory only
        self.takes_slot = takes_slot or (lambda job: True)
        self.used = 0
        self.retained = 0
        self.slots_used = 0
        self.running: Dict[int, int] = {}

    def admit_from(self, pending: "collections.deque[Job]") -> Tuple[List[Job], List[Job]]:
        """Remove from `pending` and return (admitted, rejected) jobs, in
        arrival order.  A job whose footprint exceeds the whole budget can never
        run and is rejected."""
        admitted, rejected = [], []
        for job in list(pending):
            slot = self.takes_slot(job)
            if slot and self.slots_used >= self.slots:
                continue
            fp = self.footprint(job)
            if fp > self.budget:
                pending.remove(job)
                rejected.append(job)
            elif self.used + self.retained + fp <= self.budget:
                pending.remove(job)
                self.running[job.id] = fp
                self.used += fp
                self.slots_used += 1 if slot else 0
                admitted.append(job)
        return admitted, rejected

    def blocked_by_retained(self, pending) -> bool:
        """Whether a pending job would be admitted if the retained bytes were freed."""
        if self.retained == 0:
            return False
        return any(self.used + self.footprint(j) <= self.budget and
                   (not self.takes_slot(j) or self.slots_used < self.slots) for j in pending)

    def release(self, job: Job, retain: int = 0) -> None:
        fp = self.running.pop(job.id)
        self.used -= fp
        self.slots_used -= 1 if self.takes_slot(job) else 0
        self.retained += min(int(retain), fp)

    def release_batch(self, jobs: Sequence[Job]) -> None:
        """Release every job of a finished batch at once, keeping all their
        footprints as retained: the batch's device memory stays allocated as
        one object until the server drops it."""
        for job in jobs:
            self.release(job, retain=self.running.get(job.id, 0))

    def free_retained(self) -> None:
        self.retained = 0


def build_scene(job: Job, mg_precision: int = 1, device="cuda") -> Simulation:
    """The job's scene with its seeded initial condition, ready to step.

    2D: 8-32 Gaussian point vortices at uniform positions in [0.2, 0.8]^2 with
    circulations N(0, 1) and core radius 1.5 dx (the point-vortex field of
    2D/init_conditions.py:5-9,29-60), then the initial G2P of 2D/run_2D.py:731-738.
    3D: one or two vortex rings (add_vortex_ring, 3D/init_conditions.py:10-22)
    of major radius in [0.1, 0.2], core 0.02, unit circulation, centred in
    [0.35, 0.65]^2 x [0.3, 0.5], normal along z.  Both are projected."""
    sim = Simulation(job.res, step_config(job, mg_precision), device=device)
    if job.dim == 2:
        centers, weights, rad = point_vortex_ic(job)
        sim.init_leapfrog_2d(centers, weights, rad=rad)
        sim.initialize_2d_particles()
    else:
        rng = np.random.default_rng(job.seed)
        rings = []
        for _ in range(int(rng.integers(1, 3))):
            c = [float(rng.uniform(0.35, 0.65)), float(rng.uniform(0.35, 0.65)), float(rng.uniform(0.3, 0.5))]
            rings.append(dict(w=1.0, rad=float(rng.uniform(0.1, 0.2)), delta=0.02, c=c, unit_x=[1.0, 0.0, 0.0],
                              unit_y=[0.0, 1.0, 0.0]))
        sim.init_rings_3d(rings)
    return sim


@dataclasses.dataclass
class JobResult:
    job: Job
    admitted: float
    started: float
    finished: float
    step_ms: List[float]
    poisson_iters: List[int]
    kinetic_energy: float
    fields: Optional[List[torch.Tensor]] = None  # final face velocities (keep_fields=True)
    error: Optional[str] = None


@dataclasses.dataclass
class ServeReport:
    results: List[JobResult]
    slots: int
    mg_precision: int
    wall_s: float
    batch_2d: bool = False

    def summary(self) -> dict:
        ok = [r for r in self.results if r.error is None]
        steps = np.array([t for r in ok for t in r.step_ms], dtype=np.float64)
        t0 = min(r.job.arrival for r in self.results) if self.results else 0.0
        t1 = max(r.finished for r in self.results) if self.results else 0.0
        span = max(t1 - t0, 1e-9)
        return {
            "jobs": len(self.results),
            "failed": len(self.results) - len(ok),
            "steps": int(steps.size),
            "jobs_per_s": len(ok) / span,
            "makespan_s": span,
            "step_ms_p50": float(np.percentile(steps, 50)) if steps.size else None,
            "step_ms_p99": float(np.percentile(steps, 99)) if steps.size else None,
            "queue_wait_s_mean": float(np.mean([r.admitted - r.job.arrival for r in self.results]))
            if self.results else None,
            "slots": self.slots,
            "mg_precision": {0: "fp32", 1: "fp16", 2: "bf16"}[self.mg_precision],
            "batch_2d": self.batch_2d,
        }


def batch_step_config() -> StepConfig:
    """The knobs of the batched 2D lane: those of a served 2D job (the seed is
    per scene) with the fp32 V-cycle."""
    return StepConfig(particles_per_cell=4, reinit_every=12, reinit_every_grad_m=12, CFL=0.5, tol=1e-6,
                      tol_mode="rel", particle_order=1, write_imp_grad_m=False)


def point_vortex_ic(job: Job):
    """The seeded 2D initial condition of build_scene: centres, weights, core radius."""
    rng = np.random.default_rng(job.seed)
    nv = int(rng.integers(8, 33))
    centers = rng.uniform(0.2, 0.8, size=(nv, 2))
    weights = rng.normal(0.0, 1.0, size=nv)
    return centers, weights, 1.5 / job.res[1]


class Server:
    """Runs job streams on one GPU with `slots` concurrently resident jobs
    (and, with batch_2d, `batch_lanes` batches of 2D jobs)."""

    def __init__(self, slots: int = 8, mem_budget: Optional[int] = None, mg_precision: int = 1, device="cuda",
                 keep_fields: bool = False, batch_2d: bool = False, batch_lanes: int = 2, max_batch: int = 64,
                 batch_window_s: float = 0.02):
        L.load()  # fails loudly without the HIP library or a GPU
        self.device = torch.device(device)
        if mem_budget is None:
            free, _total = torch.cuda.mem_get_info(self.device)
            mem_budget = int(0.85 * free)
        self.slots = int(slots)
        self.mg_precision = int(mg_precision)
        self.keep_fields = keep_fields
        self.batch_2d = bool(batch_2d)
        self.batch_lanes = int(batch_lanes)
        self.max_batch = int(max_batch)
        self.batch_window_s = float(batch_window_s)
        # 2D jobs of the batched lane do not take a slot
        self.admission = Admission(mem_budget, self.slots,
                                   footprint=lambda j: footprint_bytes(j, 0 if self._batched(j) else self.mg_precision),
                                   takes_slot=lambda j: not self._batched(j))

    def _batched(self, job: Job) -> bool:
        return self.batch_2d and job.dim == 2

    def _run_batch(self, group, t0):
        """One batch of 2D jobs from initial condition to the last job's last
        step; returns [(JobResult, ...)] and the batch (its memory is freed by
        the caller when the device is idle)."""
        from .batch import Batch2D

        jobs = [j for j, _ in group]
        # no HIP graphs here: capturing one while the server's other threads
        # make synchronous HIP calls (a job's solver set-up, the allocator)
        # fails both sides on ROCm, so a served batch replays its launches
        # directly (the same kernels, the same results)
        b = Batch2D([j.res for j in jobs], batch_step_config(), seeds=[j.seed % (2**31) for j in jobs],
                    device=self.device, graphs=False)
        for k, j in enumerate(jobs):
            c, w, rad = point_vortex_ic(j)
            b.set_point_vortices(k, c, w, rad)
        b.project()
        b.init_particles()
        stream = torch.cuda.current_stream(self.device)
        stream.synchronize()
        started = time.perf_counter() - t0
        left = [j.steps for j in jobs]
        step_ms = [[] for _ in jobs]
        iters = [[] for _ in jobs]
        out = [None] * len(jobs)
        while any(n > 0 for n in left):
            a = time.perf_counter()
            b.step()
            stream.synchronize()
            ms = 1e3 * (time.perf_counter() - a)
            info = b.scene_info()
            ns = b.state.n_solves
            for k, j in enumerate(jobs):
                if left[k] <= 0:
                    continue
                step_ms[k].append(ms)
                iters[k].extend(int(x) for x in info["iters"][k][:ns])
                left[k] -= 1
                if left[k] == 0:
                    u = [b.u_view(k, a_) for a_ in range(2)]
                    ke = float(core.kinetic_energy(j.res, u).item())
                    fields = [x.clone() for x in u] if self.keep_fields else None
                    b.set_enabled(k, False)
                    out[k] = JobResult(j, group[k][1], started, time.perf_counter() - t0, step_ms[k], iters[k], ke,
                                       fields)
        return out, b

    def _run_job(self, job: Job, admitted: float, t0: float):
        started = time.perf_counter() - t0
        sim = build_scene(job, self.mg_precision, self.device)
        stream = torch.cuda.current_stream(self.device)
        stream.synchronize()
        step_ms = []
        for _ in range(job.steps):
            a = time.perf_counter()
            sim.step()
            stream.synchronize()
            step_ms.append(1e3 * (time.perf_counter() - a))
        ke = sim.kinetic_energy()
        fields = [u.clone() for u in sim.u] if self.keep_fields else None
        stream.synchronize()
        res = JobResult(job, admitted, started, time.perf_counter() - t0, step_ms, list(sim.poisson_iters), ke,
                        fields)
        return res, sim.solver

    def run(self, jobs: Sequence[Job], realtime: bool = True, progress=None) -> ServeReport:
        """Serve `jobs` (arrival times honoured when `realtime`, else all
        pending at once) and return when every job has finished.  `progress`,
        if given, is called with each JobResult as it completes."""
        jobs = sorted(jobs, key=lambda j: (j.arrival, j.id))
        work: "queue.Queue" = queue.Queue()
        done: "queue.Queue" = queue.Queue()
        t0 = time.perf_counter()

        def worker():
            s = torch.cuda.Stream(device=self.device)
            with torch.cuda.stream(s):
                while True:
                    item = work.get()
                    if item is None:
                        return
                    job, admitted = item
                    try:
                        done.put(self._run_job(job, admitted, t0))
                    except Exception as e:  # reported per job; the server keeps going
                        now = time.perf_counter() - t0
                        done.put((JobResult(job, admitted, now, now, [], [], float("nan"), None, repr(e)), None))

        bwork: "queue.Queue" = queue.Queue()

        def batch_worker():
            s = torch.cuda.Stream(device=self.device)
            with torch.cuda.stream(s):
                while True:
                    item = bwork.get()
                    if item is None:
                        return
                    group = [item]
                    # the 2D jobs admitted within the collection window join this batch
                    deadline = time.perf_counter() + self.batch_window_s
                    while len(group) < self.max_batch:
                        try:
                            nxt = bwork.get(timeout=max(0.0, deadline - time.perf_counter()))
                        except queue.Empty:
                            break
                        if nxt is None:
                            bwork.put(None)
                            break
                        group.append(nxt)
                    # one item per batch: the main thread releases the batch's jobs and
                    # takes the batch object together, so the batch's memory is counted
                    # until the server itself drops it
                    try:
                        out, b = self._run_batch(group, t0)
                        done.put((out, b))
                        del out, b  # this thread keeps no reference to the batch
                    except Exception as e:  # reported per job; the server keeps going
                        now = time.perf_counter() - t0
                        done.put(([JobResult(job, admitted, now, now, [], [], float("nan"), None, repr(e))
                                   for job, admitted in group], None))

        threads = [threading.Thread(target=worker, daemon=True) for _ in range(self.slots)]
        if self.batch_2d:
            threads += [threading.Thread(target=batch_worker, daemon=True) for _ in range(self.batch_lanes)]
        for th in threads:
            th.start()
        arrivals = collections.deque(jobs)
        pending: "collections.deque[Job]" = collections.deque()
        results: List[JobResult] = []
        retired = []  # MGPCG objects of finished jobs; dropping them frees the hierarchy
        try:
            while len(results) < len(jobs):
                now = time.perf_counter() - t0
                while arrivals and (not realtime or arrivals[0].arrival <= now):
                    pending.append(arrivals.popleft())
                if retired and (not self.admission.running or self.admission.blocked_by_retained(pending)):
                    retired.clear()
                    self.admission.free_retained()
                admitted, rejected = self.admission.admit_from(pending)
                for job in admitted:
                    (bwork if self._batched(job) else work).put((job, time.perf_counter() - t0))
                for job in rejected:
                    results.append(JobResult(job, now, now, now, [], [], float("nan"), None,
                                             "rejected: footprint exceeds the memory budget"))
                timeout = 0.05
                if arrivals and realtime:
                    timeout = max(1e-4, min(timeout, arrivals[0].arrival - now))
                try:
                    r, solver = done.get(timeout=timeout)
                except queue.Empty:
                    continue
                if isinstance(r, list):
                    # a whole batch: its memory stays counted until the batch object is freed
                    self.admission.release_batch([x.job for x in r])
                    finished = r
                else:
                    keep = mg_bytes(r.job.res, 3, self.mg_precision) if solver is not None else 0
                    self.admission.release(r.job, retain=keep)
                    finished = [r]
                if solver is not None:
                    retired.append(solver)
                del solver
                for x in finished:
                    results.append(x)
                    if progress is not None:
                        progress(x)
        finally:
            for _ in range(self.slots):
                work.put(None)
            for _ in range(self.batch_lanes if self.batch_2d else 0):
                bwork.put(None)
            for th in threads:
                th.join()
            retired.clear()
            self.admission.free_retained()
        return ServeReport(sorted(results, key=lambda r: r.job.id), self.slots, self.mg_precision,
                           time.perf_counter() - t0, self.batch_2d)
