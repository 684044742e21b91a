"""Flat job description handed to the C ABI (and to the CPU oracle in tests).

A :class:`Job` is a mesh (:class:`~ferrix_amd.inp.MeshArrays`), the nodal
temperature field and a list of :class:`StepSpec`.  Loads and boundary
conditions are stored exactly as the reference keeps them
(``src/solver/step/boundary_conds.rs:10-101``): one record per (node, dof)
in deck order, duplicates kept, each with its amplitude and origin step.
Records whose node is absent from the mesh are dropped at conversion, which
is what ``static_step.rs:104`` and ``:148`` do at evaluation time.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from .inp import Amplitude, MeshArrays, Project, initial_node_temperatures

# solver selection (src/solver/solvers/mod.rs:48-55)
SOLVER_DEFAULT, SOLVER_DIRECT, SOLVER_ITERATIVE = 0, 1, 2
_SOLVER_CODE = {"DEFAULT": SOLVER_DEFAULT, "DIRECT": SOLVER_DIRECT, "ITERATIVE": SOLVER_ITERATIVE}

# PCG stopping rules
STOP_ABS_SQRT_RZ = 0  # sqrt(r.z) < tol (iterative.rs:63, tol 1e-8 from static_step.rs:204)
STOP_REL_R2 = 1  # ||r||_2 / ||b||_2 <= tol


@dataclass
class AmplitudeTable:
    """All amplitudes of a step packed for the C ABI.  Amplitude 0 is the
    default ramp (``amplitude.rs:48-55``)."""

    total_time: List[int] = field(default_factory=lambda: [0])
    shift_x: List[float] = field(default_factory=lambda: [0.0])
    shift_y: List[float] = field(default_factory=lambda: [0.0])
    ptr: List[int] = field(default_factory=lambda: [0, 0])
    has_data: List[int] = field(default_factory=lambda: [0])
    times: List[float] = field(default_factory=list)
    values: List[float] = field(default_factory=list)

    def add(self, a: Amplitude) -> int:
        if a.data is None and not a.total_time and a.shift_x == 0.0 and a.shift_y == 0.0:
            return 0
        self.total_time.append(int(a.total_time))
        self.shift_x.append(float(a.shift_x))
        self.shift_y.append(float(a.shift_y))
        if a.data is not None:
            t, v = a.data
            n = min(len(t), len(v)) if len(t) else 0
            self.times.extend(float(x) for x in t[:n])
            self.values.extend(float(x) for x in v[:n])
            self.has_data.append(1)
        else:
            self.has_data.append(0)
        self.ptr.append(len(self.times))
        return len(self.total_time) - 1


@dataclass
class StepSpec:
    """One ``*STATIC`` step as arrays.

    ``load_dof``/``bc_dof`` are global DOF indices (3*node_index + dof);
    ``*_amp`` index into ``amps``; ``*_origin`` is the origin step."""

    load_dof: np.ndarray
    load_val: np.ndarray
    load_amp: np.ndarray
    load_origin: np.ndarray
    bc_dof: np.ndarray
    bc_val: np.ndarray
    bc_amp: np.ndarray
    bc_origin: np.ndarray
    amps: AmplitudeTable
    initial_time_increment: float = 1.0
    time_period: float = 1.0
    min_time_increment: float = 1e-5
    max_time_increment: float = 1.0
    max_increments: int = 10_000
    nlgeom: bool = False
    solver: int = SOLVER_DEFAULT
    # linear-solve settings; reference values: abs sqrt(r.z) < 1e-8, 1000 its
    stop_mode: int = STOP_ABS_SQRT_RZ
    tol: float = 1e-8
    max_iter: int = 1000

    @staticmethod
    def build(loads, bcs, *, inc=None, nlgeom=False, solver=SOLVER_DEFAULT, **kw) -> "StepSpec":
        """loads / bcs: iterables of (global_dof, value, Amplitude, origin)."""
        amps = AmplitudeTable()

        def pack(recs):
            recs = list(recs)
            d = np.array([r[0] for r in recs], np.int64)
            v = np.array([r[1] for r in recs], np.float64)
            a = np.array([amps.add(r[2]) if r[2] is not None else 0 for r in recs], np.int32)
            o = np.array([r[3] for r in recs], np.int32)
            return d, v, a, o

        ld, lv, la, lo = pack(loads)
        bd, bv, ba, bo = pack(bcs)
        s = StepSpec(ld, lv, la, lo, bd, bv, ba, bo, amps, nlgeom=nlgeom, solver=solver, **kw)
        if inc is not None:
            s.initial_time_increment = inc.initial_time_increment
            s.time_period = inc.time_period
            s.min_time_increment = inc.min_time_increment
            s.max_time_increment = inc.max_time_increment
            s.max_increments = inc.max_iterations
        return s


@dataclass
class Job:
    """Mesh + temperatures + steps.  ``name`` labels the workload."""

    mesh: MeshArrays
    node_temp: np.ndarray
    steps: List[StepSpec]
    name: str = "job"

    @property
    def n_dof(self) -> int:
        return self.mesh.n_dof

    @staticmethod
    def from_project(p: Project, name: str = "deck") -> "Job":
        mesh = p.arrays()
        temps = initial_node_temperatures(p)
        steps = []
        for st in p.steps:
            def conv(recs):
                out = []
                for r in recs:
                    j = p.get_index_for_node_id(r.node_id)
                    if j is None:
                        continue
                    out.append((3 * j + r.dof, r.value, r.amplitude, r.origin_step))
                return out

            steps.append(StepSpec.build(conv(st.loads), conv(st.bcs), inc=st.increment_data,
                                        nlgeom=st.nlgeom, solver=_SOLVER_CODE[st.solver]))
        return Job(mesh, temps, steps, name)


def with_solver(step: StepSpec, *, solver: Optional[int] = None, stop_mode: Optional[int] = None,
                tol: Optional[float] = None, max_iter: Optional[int] = None) -> StepSpec:
    """Copy of ``step`` with linear-solve settings replaced."""
    import copy

    s = copy.copy(step)
    if solver is not None:
        s.solver = solver
    if stop_mode is not None:
        s.stop_mode = stop_mode
    if tol is not None:
        s.tol = tol
    if max_iter is not None:
        s.max_iter = max_iter
    return s
