"""Serving loop: solve requests that arrive over time, with continuous
batching on one GPU.

A request is a linear static job (one mesh, 1-4 load cases sharing K), an
NLGEOM static job (Newton-Raphson with the reference's increment control,
``static_step.rs:47-267``), or a C3D20 block.  The server

* admits an arrived request when its estimated device memory fits the
  budget (FIFO; the head of the queue waits until memory frees up);
* uploads its mesh (host symbolic phase + copies), assembles K and applies
  the penalty BCs on an admission thread -- the host work overlaps with the
  running batch -- and adds the system to the running :class:`Batch`: it
  joins the next iteration chunk while the other systems keep going;
* runs NLGEOM requests on worker threads, each with its own context and
  stream (``fx_static_step``), so their Newton solves overlap with the
  batch on the device;
* retires a request when its result is on the host, freeing its memory.

Latency is completion minus arrival.  The serial comparison
(:func:`serve_serial`) takes the same stream one request at a time with the
reference's pattern (one FP64 ``fx_solve`` per load case, one
``fx_static_step`` per NLGEOM step).
"""
from __future__ import annotations

import collections
import threading
import time
from concurrent.futures import Future, ThreadPoolExecutor
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import _lib as L
from . import meshgen
from .batch import System, assemble
from .device import Batch, Context, DeviceMesh, SolveOptions, run_job
from .inp import Project, deck_stem
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
    specs = meshgen.cfg5_request_specs(n, seed)
    rng = np.random.default_rng(arrival_seed)
    t = 0.0
    out = []
    for k, sp in enumerate(specs):
        if rate > 0:
            t += float(rng.exponential(1.0 / rate))
        out.append(Request(k, t if rate > 0 else 0.0, sp["kind"], meshgen.cfg5_jobs(sp)))
    return out


@dataclass
class ServeResult:
    latency: Dict[int, float] = field(default_factory=dict)
    x: Dict[int, np.ndarray] = field(default_factory=dict)       # linear: (n_rhs, n_dof); NLGEOM: (1, n_dof)
    solves: Dict[int, int] = field(default_factory=dict)         # linear solves per request
    failures: Dict[int, str] = field(default_factory=dict)
    makespan: float = 0.0
    peak_bytes: int = 0

    def dof_solves(self, reqs: List[Request]) -> int:
        by = {r.rid: r for r in reqs}
        return sum(by[k].n_dof * (by[k].n_rhs if by[k].kind != "nlgeom" else 1) * self.solves[k]
                   for k in self.x)


def _static_opts(precision: int) -> SolveOptions:
    return SolveOptions(precision=precision, stop_mode=L.FX_STOP_REL_R2, tol=1e-10, max_iter=100_000,
                        inner_tol=1e-3, max_refine=100)


class _NlWorker:
    """One thread with its own context: NLGEOM requests through the step
    driver."""

    def __init__(self, precision: int):
        self.precision = precision
        self.local = threading.local()

    def __call__(self, req: Request):
        ctx = getattr(self.local, "ctx", None)
        if ctx is None:
            ctx = self.local.ctx = Context(0)
        job = req.jobs[0]
        dm = DeviceMesh(ctx, job.mesh, job.node_temp)
        try:
            so = _static_opts(self.precision)
            u, reps = run_job(ctx, job, solve_opts=so, direct_opts=so, mesh=dm)
            return u, sum(r.linear_solves for r in reps)
        finally:
            dm.close()


def serve_batched(ctx: Context, reqs: List[Request], precision: int = L.FX_PREC_MIXED,
                  mem_budget: int = 200 << 30, chunk_iters: int = 16, nl_workers: int = 2,
                  prep_context: bool = True) -> ServeResult:
    """Continuous batching over the request stream (see the module doc).
    prep_context: the admission thread uploads and assembles on a context of
    its own (its own stream), so its synchronisations do not wait behind the
    batch's queued chunks; False: on the batch's context."""
    res = ServeResult()
    bt = Batch(ctx, _static_opts(precision))
    pctx = Context(0) if prep_context else ctx
    pool = ThreadPoolExecutor(max_workers=nl_workers)
    prep = ThreadPoolExecutor(max_workers=1)
    worker = _NlWorker(precision)
    pending = collections.deque(sorted(reqs, key=lambda r: r.arrival))
    waiting: collections.deque = collections.deque()
    running: Dict[int, tuple] = {}                   # slot -> (request, System, bytes)
    nl: Dict[int, tuple] = {}                        # rid -> (future, request, bytes)
    staged: collections.deque = collections.deque()  # (future System, request, bytes), admission order
    used = 0

    def prepare(r: Request) -> System:
        s = System(r.rid, r.jobs, DeviceMesh(pctx, r.jobs[0].mesh, r.jobs[0].node_temp))
        assemble(s)
        return s

    t0 = time.perf_counter()
    try:
        while pending or waiting or running or nl or staged:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            # admission: FIFO under the memory budget
            while waiting:
                r = waiting[0]
                need = r.est_bytes()
                if (running or nl or staged) and used + need > mem_budget:
                    break
                waiting.popleft()
                used += need
                res.peak_bytes = max(res.peak_bytes, used)
                if r.kind == "nlgeom":
                    nl[r.rid] = (pool.submit(worker, r), r, need)
                    continue
                staged.append((prep.submit(prepare, r), r, need))
            while staged and staged[0][0].done():
                fut, r, need = staged.popleft()
                try:
                    s = fut.result()
                except Exception as e:  # noqa: BLE001 -- a failed request is reported, the server goes on
                    res.failures[r.rid] = f"{type(e).__name__}: {e}"
                    res.latency[r.rid] = time.perf_counter() - t0 - r.arrival
                    used -= need
                    continue
                running[bt.add(s.dm, s.rhs)] = (r, s, need)
            if running:
                bt.run(max_iters=chunk_iters)
                for slot in list(running):
                    st, rep = bt.poll(slot)
                    if st == 0:
                        continue
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
                try:
                    u, solves = fut.result()
                    res.x[rid] = u[None, :]
                    res.solves[rid] = solves
                except Exception as e:  # noqa: BLE001 -- a failed request is reported, the server goes on
                    res.failures[rid] = f"{type(e).__name__}: {e}"
                res.latency[rid] = time.perf_counter() - t0 - r.arrival
                used -= need
            if not running and not waiting:
                # idle: wait for the next arrival, a worker or an admission
                nxt = pending[0].arrival - (time.perf_counter() - t0) if pending else 0.0
                time.sleep(min(max(nxt, 0.0), 0.002) if nl or pending or staged else 0.0)
        res.makespan = time.perf_counter() - t0
    finally:
        prep.shutdown(wait=True)
        pool.shutdown(wait=True)
        bt.close()
        if pctx is not ctx:
            pctx.close()  # after the batch; destroys any mesh a failure left behind
    return res


def serve_serial(ctx: Context, reqs: List[Request], precision: int = L.FX_PREC_FP64) -> ServeResult:
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
    res.makespan = time.perf_counter() - t0
    return res
