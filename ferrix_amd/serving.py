"""Serving loop: solve requests that arrive over time, with continuous
batching on one GPU.

A request is a linear static job (one mesh, 1-4 load cases sharing K), an
NLGEOM static job (Newton-Raphson with the reference's increment control,
``static_step.rs:47-267``), or a C3D20 block.  The server

* admits an arrived request when its estimated device memory fits what the
  device has free: hipMemGetInfo's free bytes plus the released slot blocks
  the context keeps for reuse, less a reserve and the estimates of admitted
  requests whose memory is not allocated yet (FIFO; the head of the queue
  waits until memory frees up).  A fixed budget over the estimates can be
  given instead;
* uploads its mesh (host symbolic phase + copies) on an admission thread --
  the host work overlaps with the running batch -- then queues its
  assembly, penalty BCs and admission to the running :class:`Batch` with no
  host wait (``fx_batch_add_system``): it joins the next iteration chunk
  while the other systems keep going;
* runs NLGEOM requests on worker threads, each with its own context and
  stream (``fx_static_step``: increments, element and assembly kernels on
  the worker's stream); every Newton iteration's linear solve joins the
  running batch through the step's solver seam (``fx_step_def.solve``) and
  advances with the other systems' iterations;
* retires a request when its result is on the host, freeing its memory.

Latency is completion minus arrival.  The serial comparison
(:func:`serve_serial`) takes the same stream one request at a time with the
reference's pattern (one FP64 ``fx_solve`` per load case, one
``fx_static_step`` per NLGEOM step).
"""
from __future__ import annotations

import collections
import queue
import threading
import time
from concurrent.futures import Future, ThreadPoolExecutor
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import _lib as L
from . import meshgen
from .batch import System, admit, assemble
from .device import Batch, Context, DeviceMesh, SolveOptions, run_job
from .inp import ConvergenceError, Project, deck_stem
from .model import Job


@dataclass
class Request:
    rid: int
    arrival: float            # seconds after the stream starts
    kind: str                 # "linear" | "nlgeom" | "c3d20"
    jobs: List[Job]

    @property
    def n_dof(self) -> int:
        return self.jobs[0].n_dof

    @property
    def n_rhs(self) -> int:
        return len(self.jobs)

    def est_bytes(self) -> int:
        """Device memory estimate before upload: stored entries per DOF row
        (about 45 for C3D4, 160 for C3D20) times the operator copies and
        element contributions, plus the vectors."""
        per_row = 160 if self.kind == "c3d20" else 45
        return int(self.n_dof * (per_row * 40 + 8 * 24 * self.n_rhs))


def request_from_deck(rid: int, arrival: float, deck_path: str,
                      default_initial_temperature: Optional[float] = None) -> Request:
    """A request given as an ``.inp`` deck, read with the reference's keyword
    surface and quirks (``ferrix_amd.inp``, ``parser.rs``).  A deck of one
    linear ``*STATIC`` step joins the batch like a generated request (its
    loads are the load case; a C3D20 mesh keeps the lane-group operator).
    A deck with NLGEOM or several steps runs through the step driver like an
    NLGEOM request, each step from the previous step's displacement, as
    ``main.rs:49-63`` runs them.  default_initial_temperature: the
    project-wide temperature before ``*INITIAL CONDITIONS``
    (``Project::default_initial_temperature``; 0.0 unless given)."""
    p = Project.from_jobname(deck_path)
    if default_initial_temperature is not None:
        p.default_initial_temperature = default_initial_temperature
    job = Job.from_project(p, name=deck_stem(deck_path))
    single_linear = len(job.steps) == 1 and not job.steps[0].nlgeom
    c3d20 = bool(np.any(np.diff(np.asarray(job.mesh.el_ptr)) == 20))
    kind = ("c3d20" if c3d20 else "linear") if single_linear else "nlgeom"
    return Request(rid, arrival, kind, [job])


def make_requests(n: int, rate: float, seed: int = 5, arrival_seed: int = 55) -> List[Request]:
    """Requests of the serving mix with Poisson arrivals at `rate` per second
    (rate <= 0: everything arrives at t = 0)."""
    specs = meshgen.request_specs(n, seed)
    rng = np.random.default_rng(arrival_seed)
    t = 0.0
    out = []
    for k, sp in enumerate(specs):
        if rate > 0:
            t += float(rng.exponential(1.0 / rate))
        out.append(Request(k, t if rate > 0 else 0.0, sp["kind"], meshgen.request_jobs(sp)))
    return out


@dataclass
class ServeResult:
    latency: Dict[int, float] = field(default_factory=dict)
    x: Dict[int, np.ndarray] = field(default_factory=dict)       # linear: (n_rhs, n_dof); NLGEOM: (1, n_dof)
    solves: Dict[int, int] = field(default_factory=dict)         # linear solves per request
    failures: Dict[int, str] = field(default_factory=dict)
    makespan: float = 0.0
    peak_bytes: int = 0          # largest sum of admitted requests' estimates
    min_free: float = float("inf")  # least free + reusable device bytes an admission saw
    max_admitted: int = 0        # most requests admitted and not yet finished at once
    batched_nl_solves: int = 0   # Newton iterations' solves of NLGEOM requests done in the batch

    def dof_solves(self, reqs: List[Request]) -> int:
        by = {r.rid: r for r in reqs}
        return sum(by[k].n_dof * (by[k].n_rhs if by[k].kind != "nlgeom" else 1) * self.solves[k]
                   for k in self.x)


def _static_opts(precision: int) -> SolveOptions:
    if precision == L.FX_PREC_MIXED_POLY:
        return SolveOptions(precision=precision, stop_mode=L.FX_STOP_REL_R2, tol=1e-10, max_iter=400_000)
    return SolveOptions(precision=precision, stop_mode=L.FX_STOP_REL_R2, tol=1e-10, max_iter=100_000,
                        inner_tol=1e-3, max_refine=100)


class _BatchSolve:
    """One Newton iteration's linear solve handed from an NLGEOM worker to
    the serving loop, which adds it to the running batch and sets `done`
    when its result is in `x`."""

    def __init__(self, dm: DeviceMesh, rhs: np.ndarray):
        self.dm, self.rhs = dm, rhs
        self.x: Optional[np.ndarray] = None
        self.iterations = 0
        self.status = 0
        self.done = threading.Event()


class _NlWorker:
    """One thread with its own context: NLGEOM requests through the step
    driver.  With a solve queue, every Newton iteration's system goes to the
    serving loop's batch (the step's solver seam) and the worker waits for
    its result; without one, the step solves with its own PCG."""

    def __init__(self, precision: int, solves: Optional["queue.Queue"] = None):
        self.precision = precision
        self.solves = solves
        self.local = threading.local()
        self.lock = threading.Lock()
        self.closed = False  # set by the serving loop when it stops serving solves

    def close(self):
        """No more batched solves: queued ones are released with a failure."""
        with self.lock:
            self.closed = True
        while self.solves is not None and not self.solves.empty():
            job = self.solves.get_nowait()
            job.status = -L.FX_ERR_INVALID
            job.done.set()

    def _solver(self, dm: DeviceMesh, rhs: np.ndarray):
        job = _BatchSolve(dm, rhs)
        with self.lock:
            if self.closed:
                raise ConvergenceError("the server stopped before the solve")
            self.solves.put(job)
        job.done.wait()
        if job.status != 1:
            e = ConvergenceError(f"batched linear solve ended with status {job.status}")
            e.status = L.FX_ERR_NOT_CONVERGED if job.status in (-L.FX_ERR_NOT_CONVERGED, 3) else L.FX_ERR_BREAKDOWN
            raise e
        return job.x, job.iterations

    def __call__(self, req: Request):
        ctx = getattr(self.local, "ctx", None)
        if ctx is None:
            ctx = self.local.ctx = Context(0)
        job = req.jobs[0]
        dm = DeviceMesh(ctx, job.mesh, job.node_temp)
        try:
            so = _static_opts(L.FX_PREC_MIXED if self.precision == L.FX_PREC_MIXED_POLY else self.precision)
            u, reps = run_job(ctx, job, solve_opts=so, direct_opts=so, mesh=dm,
                              solver=self._solver if self.solves is not None else None)
            return u, sum(r.linear_solves for r in reps)
        finally:
            dm.close()


def serve_batched(ctx: Context, reqs: List[Request], precision: int = L.FX_PREC_MIXED,
                  mem_budget: Optional[int] = None, mem_reserve: int = 4 << 30, chunk_iters: int = 16,
                  nl_workers: int = 4, prep_context: bool = True, nl_in_batch: bool = True,
                  log_every: int = 0) -> ServeResult:
    """Continuous batching over the request stream (see the module doc).
    mem_budget: None admits by the device's free memory (less mem_reserve);
    a number admits while the admitted requests' estimates stay under it.
    prep_context: the admission thread uploads and assembles on a context of
    its own (its own stream), so its synchronisations do not wait behind the
    batch's queued chunks; False: on the batch's context.  nl_in_batch:
    NLGEOM requests' linear solves join the batch (False: each step solves
    with its own PCG on the worker's stream)."""
    res = ServeResult()
    bt = Batch(ctx, _static_opts(precision))
    pctx = Context(0) if prep_context else ctx
    pool = ThreadPoolExecutor(max_workers=nl_workers)
    prep = ThreadPoolExecutor(max_workers=1)
    nl_queue: "queue.Queue[_BatchSolve]" = queue.Queue()
    worker = _NlWorker(precision, nl_queue if nl_in_batch else None)
    pending = collections.deque(sorted(reqs, key=lambda r: r.arrival))
    waiting: collections.deque = collections.deque()
    running: Dict[int, tuple] = {}                   # slot -> (request, System, bytes)
    nl_slots: Dict[int, _BatchSolve] = {}            # slot -> a Newton iteration's solve
    nl: Dict[int, tuple] = {}                        # rid -> (future, request, bytes)
    staged: collections.deque = collections.deque()  # (future System, request, bytes), admission order
    used = 0
    inflight = 0  # estimates of admitted requests whose memory is not allocated yet

    def fits(need: int) -> bool:
        if mem_budget is not None:
            return used + need <= mem_budget
        free, _total, cached = ctx.memory()
        res.min_free = min(res.min_free, free + cached)
        return need <= free + cached - mem_reserve - inflight

    def prepare(r: Request) -> System:
        # the host symbolic phase and the copies; assembly and penalty are
        # queued with the admission (fx_batch_add_system, no host wait)
        return System(r.rid, r.jobs, DeviceMesh(pctx, r.jobs[0].mesh, r.jobs[0].node_temp))

    t0 = time.perf_counter()
    logged = 0
    try:
        while pending or waiting or running or nl or staged:
            if log_every and len(res.latency) >= logged + log_every:
                logged = len(res.latency) // log_every * log_every
                print(f"# served {len(res.latency)}/{len(reqs)} at {time.perf_counter() - t0:.1f} s", flush=True)
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            # admission: FIFO under the free memory (or the budget)
            while waiting:
                r = waiting[0]
                need = r.est_bytes()
                if (running or nl or staged) and not fits(need):
                    break
                waiting.popleft()
                used += need
                res.peak_bytes = max(res.peak_bytes, used)
                if r.kind == "nlgeom":
                    nl[r.rid] = (pool.submit(worker, r), r, need)
                    inflight += need  # allocated by the worker as its step runs; counted until it ends
                    res.max_admitted = max(res.max_admitted, len(running) + len(nl) + len(staged))
                    continue
                staged.append((prep.submit(prepare, r), r, need))
                inflight += need
                res.max_admitted = max(res.max_admitted, len(running) + len(nl) + len(staged))
            while staged and staged[0][0].done():
                fut, r, need = staged.popleft()
                inflight -= need  # uploaded: now in the device's used memory
                try:
                    s = fut.result()
                except Exception as e:  # noqa: BLE001 -- a failed request is reported, the server goes on
                    res.failures[r.rid] = f"{type(e).__name__}: {e}"
                    res.latency[r.rid] = time.perf_counter() - t0 - r.arrival
                    used -= need
                    continue
                running[admit(bt, s)] = (r, s, need)
            # Newton iterations' solves of the NLGEOM workers join the batch
            while not nl_queue.empty():
                job = nl_queue.get_nowait()
                nl_slots[bt.add(job.dm, job.rhs[None, :])] = job
            if running or nl_slots:
                bt.run(max_iters=chunk_iters)
                done = bt.finished()
                for slot in (k for k in done if k in nl_slots):
                    st, rep = bt.poll(slot)
                    job = nl_slots.pop(slot)
                    if st == 1:
                        job.x = bt.get(slot)[0]
                    job.status, job.iterations = st, rep.iterations
                    res.batched_nl_solves += 1
                    bt.remove(slot)  # before the worker goes on with its mesh
                    job.done.set()
                for slot in (k for k in done if k in running):
                    st, rep = bt.poll(slot)
                    r, s, need = running.pop(slot)
                    if st == 1:
                        res.x[r.rid] = bt.get(slot)
                    else:
                        res.failures[r.rid] = f"status {-st}"
                    res.solves[r.rid] = 1
                    res.latency[r.rid] = time.perf_counter() - t0 - r.arrival
                    bt.remove(slot)
                    if pctx is ctx:
                        s.dm.close()
                    else:  # on the admission thread: its stream sync does not stall this loop
                        prep.submit(s.dm.close)
                    used -= need
            for rid in [k for k, v in nl.items() if v[0].done()]:
                fut, r, need = nl.pop(rid)
                inflight -= need
                try:
                    u, solves = fut.result()
                    res.x[rid] = u[None, :]
                    res.solves[rid] = solves
                except Exception as e:  # noqa: BLE001 -- a failed request is reported, the server goes on
                    res.failures[rid] = f"{type(e).__name__}: {e}"
                res.latency[rid] = time.perf_counter() - t0 - r.arrival
                used -= need
            if not running and not nl_slots and not waiting:
                # idle: wait for the next arrival, a worker's solve or an admission
                nxt = pending[0].arrival - (time.perf_counter() - t0) if pending else 0.0
                if nl and nl_queue.empty():
                    time.sleep(0.0002)
                else:
                    time.sleep(min(max(nxt, 0.0), 0.002) if pending or staged else 0.0)
        res.makespan = time.perf_counter() - t0
    finally:
        worker.close()
        for job in nl_slots.values():  # solves the loop left in the batch (an error stopped it)
            job.status = -L.FX_ERR_INVALID
            job.done.set()
        prep.shutdown(wait=True)
        pool.shutdown(wait=True)
        bt.close()
        if pctx is not ctx:
            pctx.close()  # after the batch; destroys any mesh a failure left behind
    return res


def serve_serial(ctx: Context, reqs: List[Request], precision: int = L.FX_PREC_FP64, log_every: int = 0) -> ServeResult:
    """The same stream one request at a time, in arrival order: one
    fx_solve per load case, one step driver call per NLGEOM step."""
    res = ServeResult()
    so = _static_opts(precision)
    t0 = time.perf_counter()
    for r in sorted(reqs, key=lambda q: q.arrival):
        now = time.perf_counter() - t0
        if r.arrival > now:
            time.sleep(r.arrival - now)
        job = r.jobs[0]
        dm = DeviceMesh(ctx, job.mesh, job.node_temp)
        try:
            if r.kind == "nlgeom":
                u, reps = run_job(ctx, job, solve_opts=so, direct_opts=so, mesh=dm)
                res.x[r.rid] = u[None, :]
                res.solves[r.rid] = sum(q.linear_solves for q in reps)
            else:
                s = System(r.rid, r.jobs, dm)
                assemble(s)
                res.x[r.rid] = np.array([dm.solve(b, so)[0] for b in s.rhs])
                res.solves[r.rid] = 1
        except Exception as e:  # noqa: BLE001
            res.failures[r.rid] = f"{type(e).__name__}: {e}"
        finally:
            dm.close()
        res.latency[r.rid] = time.perf_counter() - t0 - r.arrival
        if log_every and len(res.latency) % log_every == 0:
            print(f"# served {len(res.latency)}/{len(reqs)} at {time.perf_counter() - t0:.1f} s", flush=True)
    res.makespan = time.perf_counter() - t0
    return res
