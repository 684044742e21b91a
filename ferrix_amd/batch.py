"""Host driver for many independent linear static systems on one GPU.

A *system* is one mesh's resident K (after assembly and the penalty BCs,
``static_step.rs:101-198``) with 1-4 load cases that share it.  Two ways to
solve a population of systems:

* :func:`solve_serial` -- one :meth:`DeviceMesh.solve` per load case, one
  after the other: the reference's pattern (one ``Solver::solve`` per step,
  ``solvers/mod.rs:18-44``) on the GPU, the FP64 comparison arm;
* :func:`solve_batched` -- continuous batching on one :class:`Batch`:
  systems are admitted while the running set's resident operator bytes stay
  under a working-set budget, advanced together by shared kernel launches,
  and retired as they finish, which frees room for the next ones.
"""
from __future__ import annotations

import queue
import threading
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import _lib as L
from .device import Batch, DeviceMesh, SolveOptions
from .model import Job


@dataclass
class System:
    """A resident mesh plus its load cases (jobs sharing mesh and BCs)."""

    key: int
    jobs: List[Job]
    dm: Optional[DeviceMesh] = None
    rhs: Optional[np.ndarray] = None  # (n_rhs, n_dof) after the penalty
    penalty: float = 0.0

    @property
    def n_dof(self) -> int:
        return self.jobs[0].n_dof

    @property
    def n_rhs(self) -> int:
        return len(self.jobs)


def upload(ctx, systems: List[System]):
    for s in systems:
        if s.dm is None:
            s.dm = DeviceMesh(ctx, s.jobs[0].mesh, s.jobs[0].node_temp)


def admit(bt: Batch, s: System) -> int:
    """Assembly, penalty and admission of one system with no host wait
    (Batch.add_system): the same arithmetic as assemble() + Batch.add."""
    st0 = s.jobs[0].steps[0]
    return bt.add_system(s.dm, [(j.steps[0].load_dof, j.steps[0].load_val) for j in s.jobs], st0.bc_dof,
                         st0.bc_val)


def assemble(s: System):
    """Linear K at u = 0, penalty BCs of the first load case's step on K and
    every load case's right-hand side (all load cases share the BC set).
    The first right-hand side goes through fx_apply_penalty; the others get
    the same per-record update rhs[d] += penalty * (target - 0)."""
    s.dm.assemble(False)
    st0 = s.jobs[0].steps[0]
    rows = []
    for k, j in enumerate(s.jobs):
        st = j.steps[0]
        b = np.zeros(j.n_dof)
        np.add.at(b, st.load_dof, st.load_val)
        if k == 0:
            b, s.penalty = s.dm.apply_penalty(st0.bc_dof, st0.bc_val, b)
        elif s.penalty > 0.0:
            # one addition per record, in record order (np.add.at is unbuffered)
            np.add.at(b, np.asarray(st0.bc_dof), s.penalty * (np.asarray(st0.bc_val, np.float64) - 0.0))
        rows.append(b)
    s.rhs = np.array(rows)


@dataclass
class PopulationResult:
    x: Dict[int, np.ndarray] = field(default_factory=dict)     # key -> (n_rhs, n_dof)
    seconds: float = 0.0                                         # wall, assembly + solve
    latency: Dict[int, float] = field(default_factory=dict)     # key -> s from start to result on host
    iterations: Dict[int, int] = field(default_factory=dict)      # key -> slowest load case
    iters_rhs: Dict[int, List[int]] = field(default_factory=dict)  # serial: key -> per load case
    failures: Dict[int, str] = field(default_factory=dict)
    spmv_stats: tuple = ()  # batched: (launches, timed launches, timed ms, algorithmic bytes of all launches)
    inner_tune: tuple = ()  # polynomial batch: (shape, ms one launch, ms split) of fx_batch_autotune
    first_run_at: float = 0.0  # batched: s from the start to the first run call (the admissions before it)
    collect_s: float = 0.0     # batched: host s spent collecting finished systems (poll, get, remove)

    def dof_solves(self, systems: List[System]) -> int:
        return sum(s.n_dof * s.n_rhs for s in systems if s.key in self.x)


def solve_serial(systems: List[System], opts: SolveOptions) -> PopulationResult:
    """Assemble and solve each load case with fx_solve, one at a time."""
    res = PopulationResult()
    t0 = time.perf_counter()
    for s in systems:
        assemble(s)
        xs, its = [], []
        for b in s.rhs:
            x, rep = s.dm.solve(b, opts)
            xs.append(x)
            its.append(rep.iterations)
        res.x[s.key] = np.array(xs)
        res.iterations[s.key] = max(its)
        res.iters_rhs[s.key] = its
        res.latency[s.key] = time.perf_counter() - t0
    res.seconds = time.perf_counter() - t0
    return res


def operator_bytes(s: System, opts: SolveOptions) -> int:
    """Resident bytes one SpMV pass of this system streams: K values in the
    working precision plus one 4-byte column index per 3x3 block."""
    vs = 4 if opts.precision in (L.FX_PREC_MIXED, L.FX_PREC_MIXED_POLY) else 8
    return s.dm.nnz_struct * vs + (s.dm.nnz_struct // 9) * 4


def solve_batched(ctx, systems: List[System], opts: SolveOptions, ws_bytes: int = 0, chunk_iters: int = 256,
                  time_spmv: bool = False, assemble_ahead: bool = False,
                  device_admission: bool = True, inner_shape="tune", overlap: bool = True) -> PopulationResult:
    """Continuous batching over `systems` in order.  ws_bytes > 0 caps the
    summed operator bytes of the running systems (at least one always runs);
    0 admits everything at once.  Each Batch.run call advances the running
    systems by up to chunk_iters iterations (the library overlaps the host's
    bookkeeping of one 16-iteration chunk with the next); finished systems
    are collected and new ones admitted between calls.  overlap: run()
    returns with up to three chunks still running on the GPU, which keeps
    iterating while this loop collects and admits (Batch.set_overlap); the
    results are the same bit for bit.

    assemble_ahead: the systems' meshes belong to another context than the
    batch's (their own stream).  A host thread assembles them in order while
    the batch runs, and each pass admits the ones that are ready, so the GPU
    iterates the first systems while later ones are assembled.  Otherwise
    each system is assembled just before its admission.

    device_admission (without assemble_ahead): each system's assembly,
    penalty and admission are queued on the device with no host wait
    (fx_batch_add_system); False: the host reads max_diag and the
    penalised right-hand sides back per system first (assemble()).  The
    results are the same bit for bit.

    inner_shape (polynomial batch): "tune" times the Chebyshev step's two
    shapes on the running systems once, after the first run call, and keeps
    the faster (fx_batch_autotune); 0 / 1 sets one (one launch per step, one
    per load-case count).  Every shape gives the same bits."""
    if assemble_ahead and any(s.dm.ctx is ctx for s in systems):
        raise ValueError("assemble_ahead: the meshes must belong to another context than the batch")
    res = PopulationResult()
    bt = Batch(ctx, opts)
    if time_spmv:
        bt.time_spmv(True)
    if overlap:
        bt.set_overlap(True)
    poly = opts.precision == L.FX_PREC_MIXED_POLY
    tune = poly and inner_shape == "tune"
    if poly and not tune:
        bt.set_inner_shape(int(inner_shape))
    running: Dict[int, System] = {}   # slot -> system
    ws = 0
    ready: "queue.Queue[Optional[System]]" = queue.Queue()
    stop = threading.Event()
    failed: List[BaseException] = []

    def produce():
        try:
            for s in systems:
                if stop.is_set():
                    break
                assemble(s)
                ready.put(s)
        except BaseException as e:  # re-raised by the admitting thread
            failed.append(e)
        finally:
            ready.put(None)

    worker = threading.Thread(target=produce, daemon=True) if assemble_ahead else None
    qi = 0
    exhausted = False

    def fetch(block: bool) -> Optional[System]:
        """The next system to admit (assembled when assemble_ahead), or None
        if none is ready now or none is left."""
        nonlocal qi, exhausted
        if worker is None:
            if qi < len(systems):
                qi += 1
                return systems[qi - 1]
            exhausted = True
            return None
        try:
            s = ready.get(block=block)
        except queue.Empty:
            return None
        if s is None:
            exhausted = True
            if failed:
                raise failed[0]
        return s

    nxt: Optional[System] = None
    t0 = time.perf_counter()
    if worker is not None:
        worker.start()
    try:
        while True:
            while True:
                if nxt is None and not exhausted:
                    nxt = fetch(block=not running)  # wait only when nothing runs
                if nxt is None:
                    break
                need = operator_bytes(nxt, opts)
                if running and ws_bytes > 0 and ws + need > ws_bytes:
                    break
                if worker is None and device_admission:
                    slot = admit(bt, nxt)
                else:
                    if worker is None:
                        assemble(nxt)
                    slot = bt.add(nxt.dm, nxt.rhs)
                running[slot] = nxt
                ws += need
                nxt = None
            if not running:
                if exhausted and nxt is None:
                    break
                continue
            if not res.first_run_at:
                res.first_run_at = time.perf_counter() - t0
            n_run = bt.run(max_iters=chunk_iters)
            if tune and n_run > 0:
                res.inner_tune = bt.autotune()
                tune = False
            if n_run == len(running):
                continue  # none finished: nothing to collect, no memory freed
            tc = time.perf_counter()
            for slot in bt.finished():
                st, rep = bt.poll(slot)
                s = running.pop(slot)
                ws -= operator_bytes(s, opts)
                if st == 1:
                    res.x[s.key] = bt.get(slot)
                else:
                    res.failures[s.key] = f"status {-st}"
                res.iterations[s.key] = rep.iterations
                res.latency[s.key] = time.perf_counter() - t0
                bt.remove(slot)
            res.collect_s += time.perf_counter() - tc
        res.seconds = time.perf_counter() - t0
        res.spmv_stats = bt.spmv_stats()
    finally:
        stop.set()
        if worker is not None:
            while worker.is_alive():  # the worker's last put must not block
                try:
                    ready.get(timeout=0.1)
                except queue.Empty:
                    pass
            worker.join()
        bt.close()
    return res
