"""Python host layer over the C ABI: the GPU counterparts of FerriX's
assembly, solver and step seams.

* :class:`Context` -- one HIP device and stream (``fx_ctx``); ``solve_csr``
  solves a caller's CSR matrix (``Solver::solve``, ``solvers/mod.rs:18-44``).
* :class:`DeviceMesh` -- a resident mesh (``fx_mesh``): ``assemble``
  (``Assembler::assemble``, ``src/solver/assembler.rs:37-151``),
  ``internal_force`` (``:157-241``), ``stress_strain``
  (``static_step.rs:271-379``), ``apply_penalty``
  (``src/solver/step/static_step.rs:183-195``), ``solve``
  (``Solver::solve``, ``src/solver/solvers/mod.rs:18-44``) and
  ``static_step`` (``StaticStep::solve``, ``static_step.rs:47-267``).
* :class:`Batch` -- many resident systems (meshes x up to 4 load cases)
  solved by one set of kernel launches, with systems joining while others
  run (``fx_batch``).
* :func:`run_job` -- every step of a :class:`~ferrix_amd.model.Job` in
  order, like ``main.rs:49-63`` and the reference's parity tests.

Errors come back as the :mod:`ferrix_amd.inp` exception classes that mirror
``FerrixError``.
"""
from __future__ import annotations

import ctypes as C
import weakref
from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np

from . import _lib as L
from .inp import ConvergenceError, FerrixError, InvalidModelState, NumericalError
from .model import Job, StepSpec, SOLVER_ITERATIVE, STOP_REL_R2


class HipError(FerrixError):
    pass


class NoDeviceError(FerrixError):
    pass


_ERR = {
    L.FX_ERR_INVALID: InvalidModelState,
    L.FX_ERR_OOM: InvalidModelState,
    L.FX_ERR_HIP: HipError,
    L.FX_ERR_SINGULAR_JACOBIAN: NumericalError,
    L.FX_ERR_NOT_CONVERGED: ConvergenceError,
    L.FX_ERR_BREAKDOWN: NumericalError,
    L.FX_ERR_NO_DEVICE: NoDeviceError,
}


def _p(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


@dataclass
class SolveOptions:
    """``fx_solve_opts``; defaults are the reference's PCG settings
    (``static_step.rs:204``: absolute sqrt(r.z) < 1e-8, 1000 iterations)."""

    precision: int = L.FX_PREC_FP64
    stop_mode: int = L.FX_STOP_ABS_SQRT_RZ
    tol: float = 1e-8
    max_iter: int = 1000
    inner_tol: float = 1e-3
    max_refine: int = 50
    spmv_group: int = 0
    poly_degree: int = 0  # MIXED_POLY batches: Chebyshev steps per outer iteration (0: 15)

    def c(self) -> L.fx_solve_opts:
        return L.fx_solve_opts(self.precision, self.stop_mode, self.tol, self.max_iter, self.inner_tol,
                               self.max_refine, self.spmv_group, self.poly_degree)


# The reference's DEFAULT/DIRECT solver is a sparse Cholesky; the GPU path
# serves it with FP64 PCG to a tight relative residual.
DIRECT_STANDIN = SolveOptions(stop_mode=L.FX_STOP_REL_R2, tol=1e-12, max_iter=200_000)


class Context:
    def __init__(self, device: int = 0):
        self.lib = L.load()
        h = C.c_void_p()
        rc = self.lib.fx_ctx_create(device, C.byref(h))
        if rc != L.FX_OK:
            raise _ERR.get(rc, FerrixError)(self.lib.fx_last_error(None).decode())
        self.h = h
        self._meshes = weakref.WeakSet()
        self._batches = weakref.WeakSet()

    def check(self, rc: int):
        if rc != L.FX_OK:
            msg = self.lib.fx_last_error(self.h).decode()
            raise _ERR.get(rc, FerrixError)(msg)

    def memory(self):
        """(free, total, pool) device bytes: hipMemGetInfo and the
        released batch-slot blocks this context keeps for reuse."""
        fr, tot, pool = C.c_int64(0), C.c_int64(0), C.c_int64(0)
        self.check(self.lib.fx_ctx_memory(self.h, C.byref(fr), C.byref(tot), C.byref(pool)))
        return fr.value, tot.value, pool.value

    def measure_fp64_peak(self, reps: int = 10):
        """(FP64 VALU, FP64 MFMA) TFLOP/s measured on this device by
        register-only kernels (fx_measure_fp64_peak)."""
        v, m = C.c_double(), C.c_double()
        self.check(self.lib.fx_measure_fp64_peak(self.h, reps, C.byref(v), C.byref(m)))
        return v.value, m.value

    def solve_csr(self, row_ptr, col, val, rhs, opts: Optional[SolveOptions] = None):
        """``Solver::solve(k_global, f_global, ...)`` on the caller's CSR
        (``solvers/mod.rs:18-44``; fx_solve_csr): Jacobi-PCG from the
        matrix's own diagonal, x0 = 0.  Returns (x, fx_solve_report)."""
        rp = _c(row_ptr, np.int64)
        ci = _c(col, np.int64)
        v = _c(val, np.float64)
        b = _c(rhs, np.float64)
        n = len(rp) - 1
        if len(b) != n:
            raise InvalidModelState(f"solve_csr: rhs has {len(b)} entries for {n} rows")
        if len(ci) < rp[-1] or len(v) < rp[-1]:
            raise InvalidModelState("solve_csr: fewer column indices or values than row_ptr[n]")
        x = np.zeros(n)
        rep = L.fx_solve_report()
        o = (opts or SolveOptions()).c()
        self.check(self.lib.fx_solve_csr(self.h, n, _p(rp), _p(ci), _p(v), _p(b), _p(x), C.byref(o),
                                         C.byref(rep)))
        return x, rep

    def close(self):
        if getattr(self, "h", None):
            for b in list(self._batches):
                b.close()
            for m in list(self._meshes):
                m.close()
            self.lib.fx_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceMesh:
    def __init__(self, ctx: Context, mesh, node_temp: np.ndarray):
        self.ctx = ctx
        self.lib = ctx.lib
        self.n_dof = mesh.n_dof
        keep = [_c(mesh.xyz, np.float64), _c(node_temp, np.float64), _c(mesh.el_ptr, np.int64),
                _c(mesh.conn, np.int32), _c(mesh.mat, np.int32), _c(mesh.lut_n_e, np.int32),
                _c(mesh.lut_t_e, np.float64), _c(mesh.lut_v_e, np.float64), _c(mesh.lut_n_nu, np.int32),
                _c(mesh.lut_t_nu, np.float64), _c(mesh.lut_v_nu, np.float64)]
        xyz, temp, el_ptr, conn, mat, ne, te, ve, nn_, tn, vn = keep
        mt = L.fx_material_table(len(ne), _p(ne), _p(te), _p(ve), _p(nn_), _p(tn), _p(vn))
        h = C.c_void_p()
        ctx.check(self.lib.fx_mesh_upload(ctx.h, mesh.n_nodes, _p(xyz), _p(temp), mesh.n_el, _p(el_ptr),
                                          _p(conn), _p(mat), C.byref(mt), C.byref(h)))
        self.h = h
        ctx._meshes.add(self)
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        ctx.check(self.lib.fx_mesh_info(h, C.byref(a), C.byref(b), C.byref(c)))
        self.nnz_struct = b.value
        self.device_bytes = c.value

    def close(self):
        if getattr(self, "h", None) and getattr(self.ctx, "h", None):
            self.lib.fx_mesh_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state --------------------------------------------------------------
    def set_u(self, u):
        u = _c(u, np.float64)
        assert u.shape == (self.n_dof,)
        self.ctx.check(self.lib.fx_state_set_u(self.h, _p(u)))

    def zero_u(self):
        self.ctx.check(self.lib.fx_state_set_u(self.h, None))

    def profile_spmv(self, precision: int, reps: int = 50, group: int = 0):
        """(ms per launch, algorithmic bytes per launch) of the PCG SpMV."""
        ms, by = C.c_double(), C.c_double()
        self.ctx.check(self.lib.fx_profile_spmv(self.h, precision, reps, group, C.byref(ms), C.byref(by)))
        return ms.value, by.value

    def profile_element(self, reps: int = 20, mfma: bool = False):
        """(ms per launch, FLOP per launch) of the element-K kernels
        (fx_profile_element): the VALU ones, or C3D20 on FP64 MFMA."""
        ms, fl = C.c_double(0.0), C.c_double(0.0)
        self.ctx.check(self.lib.fx_profile_element(self.h, reps, 1 if mfma else 0, C.byref(ms), C.byref(fl)))
        return ms.value, fl.value

    def autotune_spmv(self, precision: int, reps: int = 20):
        """Times the PCG SpMV for every lane-group size and keeps the
        fastest for this mesh's later solves; returns (group, ms per
        launch)."""
        g, ms = C.c_int32(), C.c_double()
        self.ctx.check(self.lib.fx_mesh_autotune_spmv(self.h, precision, reps, C.byref(g), C.byref(ms)))
        return g.value, ms.value

    def autotune_elements(self, reps: int = 20, allow_mfma: bool = False):
        """Times the element-K kernels (C3D4 at 64..512 elements per
        workgroup; C3D20 on VALU and, if allowed, on MFMA) and keeps the
        fastest for this mesh; returns (C3D4 workgroup size, C3D20 on MFMA
        (-1: no C3D20 elements), ms per launch of the chosen kernels)."""
        b, mf, ms = C.c_int32(), C.c_int32(), C.c_double()
        self.ctx.check(self.lib.fx_mesh_autotune_elements(self.h, reps, 1 if allow_mfma else 0, C.byref(b),
                                                          C.byref(mf), C.byref(ms)))
        return b.value, mf.value, ms.value

    def get_u(self) -> np.ndarray:
        u = np.zeros(self.n_dof)
        self.ctx.check(self.lib.fx_state_get_u(self.h, _p(u)))
        return u

    # -- assembly seam -------------------------------------------------------
    def assemble(self, nlgeom: bool = False) -> Tuple[float, int]:
        md = C.c_double()
        nnz = C.c_int64()
        self.ctx.check(self.lib.fx_assemble(self.h, int(nlgeom), C.byref(md), C.byref(nnz)))
        self.nnz = nnz.value
        return md.value, nnz.value

    def internal_force(self, nlgeom: bool = False) -> np.ndarray:
        f = np.zeros(self.n_dof)
        self.ctx.check(self.lib.fx_internal_force(self.h, int(nlgeom), _p(f)))
        return f

    def stress_strain(self):
        """Nodal (stress, strain), each (n_nodes, 6), at the resident u
        (StaticStep::calculate_stress_strain, static_step.rs:271-379)."""
        n = self.n_dof // 3
        S = np.zeros((n, 6))
        E = np.zeros((n, 6))
        self.ctx.check(self.lib.fx_stress_strain(self.h, _p(S), _p(E)))
        return S, E

    def reaction_forces(self, dofs, f_ext=None, nlgeom: bool = False) -> np.ndarray:
        """RF at the constrained DOFs: f_int(u) - f_ext there, at the
        resident u (the reference writes no RF; with a converged solution
        they balance the applied loads, sum RF = -sum f_ext per direction)."""
        f = self.internal_force(nlgeom)
        dofs = np.asarray(dofs, np.int64)
        rf = f[dofs]
        if f_ext is not None:
            rf = rf - np.asarray(f_ext, np.float64)[dofs]
        return rf

    def matrix(self):
        """Filtered CSR (row_ptr int64, col int32, values float64)."""
        rp = np.zeros(self.n_dof + 1, np.int64)
        col = np.zeros(max(1, self.nnz), np.int32)
        val = np.zeros(max(1, self.nnz), np.float64)
        self.ctx.check(self.lib.fx_matrix_get(self.h, _p(rp), _p(col), _p(val)))
        k = int(rp[-1])
        return rp, col[:k], val[:k]

    def apply_penalty(self, dofs, targets, rhs):
        dofs = _c(dofs, np.int64)
        targets = _c(targets, np.float64)
        rhs = np.array(rhs, dtype=np.float64, copy=True)
        pen = C.c_double()
        self.ctx.check(self.lib.fx_apply_penalty(self.h, len(dofs), _p(dofs), _p(targets), _p(rhs), C.byref(pen)))
        return rhs, pen.value

    # -- solver seam ---------------------------------------------------------
    def solve(self, rhs, opts: Optional[SolveOptions] = None):
        rhs = _c(rhs, np.float64)
        x = np.zeros(self.n_dof)
        o = (opts or SolveOptions()).c()
        rep = L.fx_solve_report()
        self.ctx.check(self.lib.fx_solve(self.h, _p(rhs), _p(x), C.byref(o), C.byref(rep)))
        return x, rep

    # -- step driver ---------------------------------------------------------
    def static_step(self, step: StepSpec, step_id: int, timer: L.fx_timer, direct_opts: SolveOptions = None,
                    solve_opts: Optional[SolveOptions] = None, on_increment=None, solver=None) -> L.fx_step_report:
        """StaticStep::solve.  on_increment(step_id, inc_id, global_time), if
        given, runs after every converged increment with its displacement
        resident (the reference's writer.write_increment); an exception
        raised there stops the step and is raised again here.
        solver(mesh, rhs) -> (x, iterations), if given, solves every Newton
        iteration's system (K resident on `mesh`, assembled and penalised;
        rhs the penalised residual) in place of the step's own PCG
        (fx_step_def.solve, the Solver seam); an exception raised there stops
        the step and is raised again here."""
        sd, keep = step_def(step, step_id, direct_opts or DIRECT_STANDIN, solve_opts)
        raised = []
        if solver is not None:
            def solve_cb(_user, _mesh, n, rhs, x, its):
                try:
                    b = np.ctypeslib.as_array(rhs, shape=(n,)).copy()
                    xs, k = solver(self, b)
                    np.ctypeslib.as_array(x, shape=(n,))[:] = xs
                    its[0] = int(k)
                    return 0
                except BaseException as e:  # carried across the C frames, raised below
                    raised.append(e)
                    return int(getattr(e, "status", L.FX_ERR_NOT_CONVERGED))
            keep["solve"] = L.fx_solve_fn(solve_cb)
            sd.solve = C.cast(keep["solve"], C.c_void_p)
        if on_increment is not None:
            def hook(_user, sid, inc, t):
                try:
                    on_increment(sid, inc, t)
                    return 0
                except BaseException as e:  # carried across the C frames, raised below
                    raised.append(e)
                    return 1
            keep["hook"] = L.fx_increment_fn(hook)
            sd.on_increment = C.cast(keep["hook"], C.c_void_p)
        rep = L.fx_step_report()
        rc = self.lib.fx_static_step(self.h, C.byref(sd), C.byref(timer), C.byref(rep))
        if raised:
            raise raised[0]
        try:
            self.ctx.check(rc)
        except FerrixError as e:
            e.report = rep  # the step's counters up to the error
            raise
        del keep
        return rep


class Batch:
    """Batched PCG over resident systems (``fx_batch``).  Per load case the
    result equals :meth:`DeviceMesh.solve` with the same options (bitwise
    for FP64)."""

    def __init__(self, ctx: Context, opts: Optional[SolveOptions] = None):
        self.ctx = ctx
        self.lib = ctx.lib
        o = (opts or SolveOptions()).c()
        h = C.c_void_p()
        ctx.check(self.lib.fx_batch_create(ctx.h, C.byref(o), C.byref(h)))
        self.h = h
        self._n = {}
        ctx._batches.add(self)

    def add(self, mesh: DeviceMesh, rhs) -> int:
        """Admits mesh's resident K with rhs of shape (n_dof,) or
        (n_rhs, n_dof); returns the slot."""
        rhs = _c(np.atleast_2d(rhs), np.float64)
        assert rhs.shape[1] == mesh.n_dof
        slot = C.c_int32()
        self.ctx.check(self.lib.fx_batch_add(self.h, mesh.h, rhs.shape[0], _p(rhs), C.byref(slot)))
        self._n[slot.value] = rhs.shape
        return slot.value

    def add_system(self, mesh: DeviceMesh, load_cases, bc_dof, bc_target) -> int:
        """Assembles mesh's linear K, builds one load vector per entry of
        load_cases (a (load_dof, load_val) record pair each, summed per DOF
        in record order), applies the penalty BCs (records bc_dof, bc_target
        in deck order) to K and to every load vector, and admits the system,
        with no host wait (fx_batch_add_system); returns the slot.  The same
        results as DeviceMesh.assemble + apply_penalty + add on the summed
        vectors."""
        cases = [(_c(d, np.int64), _c(v, np.float64)) for d, v in load_cases]
        assert cases and all(d.shape == v.shape for d, v in cases)
        lc = _c(np.concatenate([np.full(len(d), k, np.int32) for k, (d, _) in enumerate(cases)]), np.int32)
        ld = _c(np.concatenate([d for d, _ in cases]), np.int64)
        lv = _c(np.concatenate([v for _, v in cases]), np.float64)
        bd = _c(bc_dof, np.int64)
        bv = _c(bc_target, np.float64)
        assert bd.shape == bv.shape
        slot = C.c_int32()
        self.ctx.check(self.lib.fx_batch_add_system(self.h, mesh.h, len(cases), len(ld), _p(lc), _p(ld), _p(lv),
                                                    len(bd), _p(bd), _p(bv), C.byref(slot)))
        self._n[slot.value] = (len(cases), mesh.n_dof)
        return slot.value

    def run(self, max_iters: int = 0) -> int:
        """Advances the running systems; returns how many still run."""
        n = C.c_int32()
        self.ctx.check(self.lib.fx_batch_run(self.h, max_iters, C.byref(n)))
        return n.value

    def poll(self, slot: int):
        """(state, fx_solve_report): state 0 running, 1 converged,
        -fx_status on an error."""
        st = C.c_int32()
        rep = L.fx_solve_report()
        self.ctx.check(self.lib.fx_batch_poll(self.h, slot, C.byref(st), C.byref(rep)))
        return st.value, rep

    def get(self, slot: int) -> np.ndarray:
        x = np.empty(self._n[slot])
        self.ctx.check(self.lib.fx_batch_get(self.h, slot, _p(x)))
        return x

    def result(self, slot: int) -> np.ndarray:
        """Solutions of a finished slot; raises like DeviceMesh.solve when
        its solve stopped on an error."""
        st, rep = self.poll(slot)
        if st < 0:
            raise _ERR.get(-st, FerrixError)(f"batch slot {slot}: solve stopped with status {-st}")
        if st == 0:
            raise InvalidModelState(f"batch slot {slot} is still running")
        return self.get(slot)

    def remove(self, slot: int):
        self.ctx.check(self.lib.fx_batch_remove(self.h, slot))
        self._n.pop(slot, None)

    def autotune(self, reps: int = 10):
        """Polynomial batch: times the Chebyshev step as one launch and as one
        launch per load-case count on the running systems and keeps the
        faster; returns (shape, ms one launch, ms split)."""
        sh, a, b = C.c_int32(), C.c_double(), C.c_double()
        self.ctx.check(self.lib.fx_batch_autotune(self.h, reps, C.byref(sh), C.byref(a), C.byref(b)))
        return sh.value, a.value, b.value

    def set_inner_shape(self, shape: int):
        """Polynomial batch: 0 one launch per Chebyshev step, 1 one per load-case count."""
        self.ctx.check(self.lib.fx_batch_set_inner_shape(self.h, int(shape)))

    def set_overlap(self, on: bool):
        """on: a budgeted run() returns with up to three iteration chunks still
        running on the GPU, which keeps iterating while the caller collects
        and admits systems (fx_batch_set_overlap)."""
        self.ctx.check(self.lib.fx_batch_set_overlap(self.h, 1 if on else 0))

    def finished(self) -> List[int]:
        """Slots whose solves have finished and are still in the batch."""
        n = C.c_int32()
        self.ctx.check(self.lib.fx_batch_finished(self.h, None, 0, C.byref(n)))
        if n.value == 0:
            return []
        out = np.zeros(n.value, np.int32)
        self.ctx.check(self.lib.fx_batch_finished(self.h, _p(out), n.value, C.byref(n)))
        return [int(k) for k in out[:n.value]]

    def time_spmv(self, on: bool = True):
        """Bracket every later SpMV launch with HIP events."""
        self.ctx.check(self.lib.fx_batch_time_spmv(self.h, int(on)))

    def spmv_stats(self):
        """(launches, timed launches, timed ms, algorithmic bytes of all launches)."""
        a, b = C.c_int64(), C.c_int64()
        c, d = C.c_double(), C.c_double()
        self.ctx.check(self.lib.fx_batch_spmv_stats(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)))
        return a.value, b.value, c.value, d.value

    def close(self):
        if getattr(self, "h", None) and getattr(self.ctx, "h", None):
            self.lib.fx_batch_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def step_options(step: StepSpec, precision=L.FX_PREC_FP64) -> SolveOptions:
    """Linear-solve options a StepSpec asks for (ITERATIVE settings)."""
    return SolveOptions(precision=precision, stop_mode=step.stop_mode, tol=step.tol, max_iter=step.max_iter)


def step_def(step: StepSpec, step_id: int, direct_opts: SolveOptions, solve_opts: Optional[SolveOptions]):
    a = step.amps
    keep = dict(
        ld=_c(step.load_dof, np.int64), lv=_c(step.load_val, np.float64), la=_c(step.load_amp, np.int32),
        lo=_c(step.load_origin, np.int32), bd=_c(step.bc_dof, np.int64), bv=_c(step.bc_val, np.float64),
        ba=_c(step.bc_amp, np.int32), bo=_c(step.bc_origin, np.int32),
        tt=_c(a.total_time, np.int32), sx=_c(a.shift_x, np.float64), sy=_c(a.shift_y, np.float64),
        hd=_c(a.has_data, np.int32), ptr=_c(a.ptr, np.int64), times=_c(a.times or [0.0], np.float64),
        values=_c(a.values or [0.0], np.float64))
    amps = L.fx_amplitudes(len(a.total_time), _p(keep["tt"]), _p(keep["sx"]), _p(keep["sy"]), _p(keep["hd"]),
                           _p(keep["ptr"]), _p(keep["times"]), _p(keep["values"]))
    so = (solve_opts or step_options(step)).c()
    sd = L.fx_step_def(step_id, int(step.nlgeom), int(step.solver), step.initial_time_increment, step.time_period,
                       step.min_time_increment, step.max_time_increment, step.max_increments,
                       len(keep["ld"]), _p(keep["ld"]), _p(keep["lv"]), _p(keep["la"]), _p(keep["lo"]),
                       len(keep["bd"]), _p(keep["bd"]), _p(keep["bv"]), _p(keep["ba"]), _p(keep["bo"]),
                       amps, so, direct_opts.c())
    return sd, keep


def run_job(ctx: Context, job: Job, precision=L.FX_PREC_FP64, direct_opts: Optional[SolveOptions] = None,
            solve_opts: Optional[SolveOptions] = None, mesh: Optional[DeviceMesh] = None, writer=None,
            solver=None):
    """All steps of ``job`` from u = 0; returns (u, [fx_step_report]).
    writer: a result writer (``ferrix_amd.vtk.VtkWriter``) driven as
    main.rs:45-66 drives the reference's: init, one write per converged
    increment (U, NT, and nodal S and E from fx_stress_strain), finish --
    also after a failed step."""
    dm = mesh or DeviceMesh(ctx, job.mesh, job.node_temp)
    dm.set_u(np.zeros(job.n_dof))
    timer = L.fx_timer(0.0, 0.0, 1.0)
    reps = []
    hook = None
    if writer is not None:
        writer.init()

        def hook(step_id, inc_id, global_time):
            stress, strain = dm.stress_strain()
            writer.write_increment(step_id, inc_id, global_time, dm.get_u(), job.node_temp, stress, strain)
    try:
        for k, st in enumerate(job.steps):
            so = solve_opts or step_options(st, precision)
            do = direct_opts or SolveOptions(precision=precision, stop_mode=STOP_REL_R2, tol=1e-12, max_iter=200_000)
            reps.append(dm.static_step(st, k + 1, timer, direct_opts=do, solve_opts=so, on_increment=hook,
                                       solver=solver))
    finally:
        if writer is not None:
            writer.finish()
    return dm.get_u(), reps
