"""CPU oracle for the FerriX static-step hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this module.  It wraps ``fx_oracle.c`` (element kernels,
triplet assembly, triplet->CSR, Jacobi-PCG) and restates the step driver of
``src/solver/step/static_step.rs:47-267`` in Python:

* increment loop with the clamp of the last increment (``:71-84``), the
  cut-back ``dt /= 2`` (``:251-258``) and the timer that is not rolled back;
* NR loop of at most 15 iterations (``:99``): f_ext from amplitude-scaled
  loads (``:102-108``), f_int (``:127``), residual and BC-masked norms, the
  relative criterion ``< 1e-3`` with the ``1e-6`` floor (``:139-171``);
  tangent assembly, penalty ``1e6 * max_diag`` per BC record (``:184-195``),
  solve, ``u += du``.
* The thermal force (``:111-118``) is exactly zero because the nodal
  temperature field never changes (``state.rs:47-78``); the oracle checks
  that with ``fxo_element_thermal`` in its unit tests and skips it here.

DIRECT (the reference default, faer sparse Cholesky, ``direct.rs:19-55``) is
stood in for by SuperLU from scipy (an exact sparse factorisation); PCG is the
C restatement.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

OK, ERR_INVALID, ERR_OOM, ERR_SINGULAR, ERR_NOT_CONVERGED = 0, 1, 2, 4, 5


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "libfx_oracle.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libfx_oracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        P = C.c_void_p
        i64, i32, dbl = C.c_int64, C.c_int, C.c_double
        L.fxo_element_flat.argtypes = [i32, P, P, P, P, P, P, P, P, P, P, i32, P, P]
        L.fxo_element_thermal_flat.argtypes = [i32, P, P, P, P, P, P, P, P, P, i32, i32, P, P, dbl, P]
        L.fxo_assemble.argtypes = [i64, P, P, P, P, P, P, P, P, P, P, P, P, dbl, P, P, P, P, P, P]
        L.fxo_internal_force.argtypes = [i64, P, P, P, P, P, P, P, P, P, P, P, P, P, i64, P, P]
        L.fxo_stress_strain.argtypes = [i64, i64, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P]
        L.fxo_triplets_to_csr.argtypes = [i64, i64, P, P, P, P, P, P, P]
        L.fxo_pcg.argtypes = [i64, P, P, P, P, P, dbl, i64, i32, i32, P, P]
        L.fxo_spmv.argtypes = [i64, P, P, P, P, P, i32]
        L.fxo_jacobi.argtypes = [i64, P, P, P, P]
        L.fxo_pcg_time_iters.argtypes = [i64, P, P, P, P, i64, i32]
        L.fxo_pcg_time_iters.restype = dbl
        L.fxo_pcg_par.argtypes = [i64, P, P, P, P, P, dbl, i64, i32, i32, P, P]
        L.fxo_pcg_par_time.argtypes = [i64, P, P, P, P, dbl, i32, i64, i32, P]
        L.fxo_pcg_par_time.restype = dbl
        L.fxo_lut_interp.argtypes = [i32, P, P, dbl]
        L.fxo_lut_interp.restype = dbl
        L.fxo_shape_c3d20.argtypes = [dbl, dbl, dbl, P, P]
        L.fxo_shape_c3d4.argtypes = [dbl, dbl, dbl, P, P]
        L.fxo_gauss_points.argtypes = [i32, P, P]
        L.fxo_d_matrix.argtypes = [dbl, dbl, P]
        L.fxo_fnv1a_pattern.argtypes = [i64, P, i64, P]
        L.fxo_fnv1a_pattern.restype = C.c_uint64
        _LIB = L
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def pattern_fnv1a(row_ptr, col) -> str:
    """FNV-1a fingerprint of a CSR pattern: row_ptr as int64 LE, then col as
    int32 LE (the golden fixtures' pattern hash; a test helper)."""
    rp = np.ascontiguousarray(row_ptr, np.int64)
    ci = np.ascontiguousarray(col, np.int32)
    return f"{int(lib().fxo_fnv1a_pattern(len(rp) - 1, _p(rp), len(ci), _p(ci))):016x}"


class OracleError(Exception):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


# ---------------------------------------------------------------------------
# Element-level
# ---------------------------------------------------------------------------
def gauss_points(nn):
    c = np.zeros(81)
    w = np.zeros(27)
    g = lib().fxo_gauss_points(nn, _p(c), _p(w))
    return c[: 3 * g].reshape(g, 3), w[:g]


def shape_c3d20(xi, eta, zeta):
    N = np.zeros(20)
    dN = np.zeros(60)
    lib().fxo_shape_c3d20(xi, eta, zeta, _p(N), _p(dN))
    return N, dN.reshape(3, 20)


def shape_c3d4(xi, eta, zeta):
    N = np.zeros(4)
    dN = np.zeros(12)
    lib().fxo_shape_c3d4(xi, eta, zeta, _p(N), _p(dN))
    return N, dN.reshape(3, 4)


def d_matrix(E, nu):
    D = np.zeros(36)
    lib().fxo_d_matrix(E, nu, _p(D))
    return D.reshape(6, 6)


def lut_interp(ts, vs, temp):
    t = np.ascontiguousarray(ts, np.float64)
    v = np.ascontiguousarray(vs, np.float64)
    return lib().fxo_lut_interp(len(t), _p(t), _p(v), temp)


def _mat_args(mesh):
    return (mesh.lut_n_e, mesh.lut_t_e, mesh.lut_v_e, mesh.lut_n_nu, mesh.lut_t_nu, mesh.lut_v_nu)


def element(mesh, X, Tn, m=0, ucfg=None, uel=None, want_k=True, want_f=False):
    """K_el ((3nn)^2) and/or f_el (3nn) of one element with coordinates X
    (nn x 3) using material m of ``mesh``."""
    nn = X.shape[0]
    X = np.ascontiguousarray(X, np.float64)
    Tn = np.ascontiguousarray(Tn, np.float64)
    K = np.zeros((3 * nn, 3 * nn)) if want_k else None
    f = np.zeros(3 * nn) if want_f else None
    uc = None if ucfg is None else np.ascontiguousarray(ucfg, np.float64)
    ue = None if uel is None else np.ascontiguousarray(uel, np.float64)
    rc = lib().fxo_element_flat(nn, _p(X), _p(uc), _p(ue), _p(Tn), *[_p(a) for a in _mat_args(mesh)], m,
                                _p(K), _p(f))
    if rc != OK:
        raise OracleError(rc, "Singular Jacobian" if rc == ERR_SINGULAR else f"element rc={rc}")
    return K, f


def thermal_force(mesh, X, Ti, Tc, alpha_t, alpha_v, t_ref, m=0):
    """element.rs:453-520 for one element with an expansion LUT."""
    nn = X.shape[0]
    f = np.zeros(3 * nn)
    ta = np.ascontiguousarray(alpha_t, np.float64)
    va = np.ascontiguousarray(alpha_v, np.float64)
    rc = lib().fxo_element_thermal_flat(nn, _p(np.ascontiguousarray(X, np.float64)),
                                        _p(np.ascontiguousarray(Ti, np.float64)),
                                        _p(np.ascontiguousarray(Tc, np.float64)),
                                        *[_p(a) for a in _mat_args(mesh)], m, len(ta), _p(ta), _p(va), t_ref, _p(f))
    if rc != OK:
        raise OracleError(rc, "thermal force failed")
    return f


# ---------------------------------------------------------------------------
# Global assembly
# ---------------------------------------------------------------------------
@dataclass
class CSR:
    n: int
    row_ptr: np.ndarray
    col: np.ndarray
    val: np.ndarray

    @property
    def nnz(self):
        return int(self.row_ptr[-1])

    def to_scipy(self):
        import scipy.sparse as sp

        return sp.csr_matrix((self.val, self.col, self.row_ptr), shape=(self.n, self.n))


def assemble_triplets(mesh, Tnode, u=None, drop_tol=1e-12):
    """assembler.rs:37-151 -> (rows, cols, vals, max_diag)."""
    cap = int(np.sum((3 * np.diff(mesh.el_ptr)) ** 2))
    tr = np.zeros(cap, np.int32)
    tc = np.zeros(cap, np.int32)
    tv = np.zeros(cap, np.float64)
    nt = np.zeros(1, np.int64)
    md = np.zeros(1, np.float64)
    bad = np.zeros(1, np.int64)
    uu = None if u is None else np.ascontiguousarray(u, np.float64)
    rc = lib().fxo_assemble(mesh.n_el, _p(mesh.el_ptr), _p(mesh.conn), _p(mesh.mat), _p(mesh.xyz),
                            _p(np.ascontiguousarray(Tnode, np.float64)), _p(uu),
                            *[_p(a) for a in _mat_args(mesh)], drop_tol, _p(nt), _p(tr), _p(tc), _p(tv),
                            _p(md), _p(bad))
    if rc != OK:
        raise OracleError(rc, f"assembly failed rc={rc} element={int(bad[0])}")
    n = int(nt[0])
    return tr[:n], tc[:n], tv[:n], float(md[0])


def triplets_to_csr(n, rows, cols, vals) -> CSR:
    rp = np.zeros(n + 1, np.int64)
    col = np.zeros(max(1, len(rows)), np.int32)
    val = np.zeros(max(1, len(rows)), np.float64)
    nnz = np.zeros(1, np.int64)
    rows = np.ascontiguousarray(rows, np.int32)
    cols = np.ascontiguousarray(cols, np.int32)
    vals = np.ascontiguousarray(vals, np.float64)
    rc = lib().fxo_triplets_to_csr(n, len(rows), _p(rows), _p(cols), _p(vals), _p(rp), _p(col), _p(val),
                                   _p(nnz))
    if rc != OK:
        raise OracleError(rc, "triplets_to_csr failed")
    k = int(nnz[0])
    return CSR(n, rp, col[:k].copy(), val[:k].copy())


def internal_force(mesh, Tnode, u, ucfg=None):
    f = np.zeros(mesh.n_dof)
    bad = np.zeros(1, np.int64)
    rc = lib().fxo_internal_force(mesh.n_el, _p(mesh.el_ptr), _p(mesh.conn), _p(mesh.mat), _p(mesh.xyz),
                                  _p(np.ascontiguousarray(Tnode, np.float64)),
                                  _p(np.ascontiguousarray(u, np.float64)),
                                  _p(None if ucfg is None else np.ascontiguousarray(ucfg, np.float64)),
                                  *[_p(a) for a in _mat_args(mesh)], mesh.n_dof, _p(f), _p(bad))
    if rc != OK:
        raise OracleError(rc, f"internal force failed rc={rc} element={int(bad[0])}")
    return f


def stress_strain(mesh, Tnode, u):
    """Nodal stress and strain (n_nodes x 6 each, Voigt xx yy zz xy yz zx),
    averaged over the elements of each node (static_step.rs:271-379)."""
    n = len(mesh.xyz) if mesh.xyz.ndim == 2 else len(mesh.xyz) // 3
    S = np.zeros((n, 6))
    E = np.zeros((n, 6))
    bad = np.zeros(1, np.int64)
    rc = lib().fxo_stress_strain(n, mesh.n_el, _p(mesh.el_ptr), _p(mesh.conn), _p(mesh.mat), _p(mesh.xyz),
                                 _p(np.ascontiguousarray(Tnode, np.float64)),
                                 _p(np.ascontiguousarray(u, np.float64)),
                                 *[_p(a) for a in _mat_args(mesh)], _p(S), _p(E), _p(bad))
    if rc != OK:
        raise OracleError(rc, f"stress recovery failed rc={rc} element={int(bad[0])}")
    return S, E


def assemble_csr(mesh, Tnode, u=None, bc_dofs=None, penalty_scale=1.0e6):
    """Triplets (+ penalty triplets appended in BC order when bc_dofs is
    given and max_diag > 0) -> CSR.  Returns (CSR, max_diag, penalty)."""
    r, c, v, md = assemble_triplets(mesh, Tnode, u)
    pen = 0.0
    if bc_dofs is not None and md > 0.0:
        pen = md * penalty_scale
        bd = np.asarray(bc_dofs, np.int32)
        r = np.concatenate([r, bd])
        c = np.concatenate([c, bd])
        v = np.concatenate([v, np.full(len(bd), pen)])
    return triplets_to_csr(mesh.n_dof, r, c, v), md, pen


# ---------------------------------------------------------------------------
# Solvers
# ---------------------------------------------------------------------------
def pcg(K: CSR, b, tol=1e-8, max_iter=1000, stop_mode=0, nthreads=0):
    """iterative.rs:26-121 (stop_mode 0) or ||r||/||b|| <= tol (mode 1).
    Returns (x, iterations, final criterion, converged)."""
    x = np.zeros(K.n)
    it = np.zeros(1, np.int64)
    fr = np.zeros(1, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    rc = lib().fxo_pcg(K.n, _p(K.row_ptr), _p(K.col), _p(K.val), _p(b), _p(x), tol, max_iter, stop_mode,
                       nthreads, _p(it), _p(fr))
    return x, int(it[0]), float(fr[0]), rc == OK


def pcg_par(K: CSR, b, tol=1e-8, max_iter=1000, stop_mode=0, nthreads=0):
    """The same PCG with the reference's rayon structure (parallel dots,
    updates and Jacobi apply, per-iteration allocations): the CPU baseline's
    solver.  Returns (x, iterations, final criterion, converged)."""
    x = np.zeros(K.n)
    it = np.zeros(1, np.int64)
    fr = np.zeros(1, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    rc = lib().fxo_pcg_par(K.n, _p(K.row_ptr), _p(K.col), _p(K.val), _p(b), _p(x), tol, max_iter, stop_mode,
                           nthreads, _p(it), _p(fr))
    return x, int(it[0]), float(fr[0]), rc == OK


def pcg_par_time(K: CSR, b, tol, stop_mode, max_iter, nthreads):
    """Wall seconds and iterations of pcg_par to `tol` (tol = 0: exactly
    `max_iter` iterations); seconds < 0 when it did not converge."""
    it = np.zeros(1, np.int64)
    b = np.ascontiguousarray(b, np.float64)
    sec = lib().fxo_pcg_par_time(K.n, _p(K.row_ptr), _p(K.col), _p(K.val), _p(b), tol, stop_mode, max_iter,
                                 nthreads, _p(it))
    return float(sec), int(it[0])


def spmv(K: CSR, x, nthreads=0):
    y = np.zeros(K.n)
    lib().fxo_spmv(K.n, _p(K.row_ptr), _p(K.col), _p(K.val), _p(np.ascontiguousarray(x, np.float64)), _p(y),
                   nthreads)
    return y


def jacobi(K: CSR):
    d = np.zeros(K.n)
    lib().fxo_jacobi(K.n, _p(K.row_ptr), _p(K.col), _p(K.val), _p(d))
    return d


def direct_solve(K: CSR, b):
    """Stand-in for direct.rs:19-55 (faer Cholesky): SuperLU on the same
    matrix.  Raises NumericalError-like OracleError on a singular matrix."""
    import scipy.sparse.linalg as spla

    A = K.to_scipy().tocsc()
    try:
        lu = spla.splu(A, permc_spec="MMD_AT_PLUS_A")
    except RuntimeError as e:
        raise OracleError(ERR_SINGULAR, f"Cholesky factorization failed: {e}. Check for rigid body modes!")
    return lu.solve(np.asarray(b, np.float64))


# ---------------------------------------------------------------------------
# Amplitude / time (amplitude.rs:39-98, time.rs:8-63)
# ---------------------------------------------------------------------------
class SolverTime:
    def __init__(self):
        self.glob, self.local, self.local_max = 0.0, 0.0, 1.0

    def new_step(self, max_time):
        self.local = 0.0
        self.local_max = max_time

    def new_increment(self, dt):
        self.glob += dt
        self.local += dt


def _interp(t, times, values, sx, sy):
    if len(times) == 0:
        return 0.0
    tt = t - sx
    idx = int(np.searchsorted(np.asarray(times), tt, side="left"))
    if idx == 0:
        return values[0] + sy
    if idx >= len(times):
        return values[-1] + sy
    t0, t1, v0, v1 = times[idx - 1], times[idx], values[idx - 1], values[idx]
    slope = (v1 - v0) / (t1 - t0)
    return v0 + slope * (tt - t0) + sy


def amp_y(amps, a, timer: SolverTime, origin, current_step):
    if amps.has_data[a]:
        times = amps.times[amps.ptr[a]:amps.ptr[a + 1]]
        vals = amps.values[amps.ptr[a]:amps.ptr[a + 1]]
        t = timer.glob if amps.total_time[a] else timer.local
        return _interp(t, times, vals, amps.shift_x[a], amps.shift_y[a])
    if current_step > origin:
        return 1.0
    return timer.local / timer.local_max


# ---------------------------------------------------------------------------
# Static step (static_step.rs:47-267)
# ---------------------------------------------------------------------------
@dataclass
class StepLog:
    increments: List[dict] = field(default_factory=list)
    linear_solves: int = 0
    u_committed: Optional[np.ndarray] = None  # displacement of the last converged increment
    pcg_iterations: List[int] = field(default_factory=list)
    nnz: List[int] = field(default_factory=list)


def static_step(job, step_index, u, timer: SolverTime, solver_override=None, log: Optional[StepLog] = None,
                nthreads=0):
    """Runs step ``step_index`` (0-based; step_id = index + 1) from
    displacement ``u``; returns the displacement at the end of the step."""
    st = job.steps[step_index]
    step_id = step_index + 1
    mesh, T = job.mesh, job.node_temp
    n = mesh.n_dof
    timer.new_step(st.time_period)
    log = log if log is not None else StepLog()
    solver = st.solver if solver_override is None else solver_override
    current_time = 0.0
    n_inc = 0
    dt = st.initial_time_increment
    u_start = np.array(u, dtype=np.float64, copy=True)
    bc_dofs = np.asarray(st.bc_dof, np.int64)
    while current_time < st.time_period:
        n_inc += 1
        if n_inc > st.max_increments:
            raise OracleError(ERR_NOT_CONVERGED, f"Maximum increment count of {st.max_increments} exceeded.")
        if current_time + dt > st.time_period:
            dt = st.time_period - current_time
        timer.new_increment(dt)
        u_curr = u_start.copy()
        converged = False
        inc = dict(dt=dt, iters=[], rel=[])
        for it in range(15):
            f_ext = np.zeros(n)
            for d, v, a, o in zip(st.load_dof, st.load_val, st.load_amp, st.load_origin):
                f_ext[d] += amp_y(st.amps, a, timer, o, step_id) * v
            f_int = internal_force(mesh, T, u_curr, u_curr if st.nlgeom else None)
            R = f_ext - f_int
            rc, fc = R.copy(), f_ext.copy()
            rc[bc_dofs] = 0.0
            fc[bc_dofs] = 0.0
            res_norm = math.sqrt(float(np.sum(rc * rc)))
            f_norm = math.sqrt(float(np.sum(fc * fc)))
            rel = res_norm / f_norm if f_norm > 1e-6 else res_norm
            inc["rel"].append(rel)
            if rel < 1e-3:
                converged = True
                break
            r_, c_, v_, md = assemble_triplets(mesh, T, u_curr if st.nlgeom else None)
            if md > 0.0:
                pen = md * 1.0e6
                r_ = np.concatenate([r_, bc_dofs.astype(np.int32)])
                c_ = np.concatenate([c_, bc_dofs.astype(np.int32)])
                v_ = np.concatenate([v_, np.full(len(bc_dofs), pen)])
                for d, v, a, o in zip(st.bc_dof, st.bc_val, st.bc_amp, st.bc_origin):
                    target = amp_y(st.amps, a, timer, o, step_id) * v
                    R[d] += pen * (target - u_curr[d])
            K = triplets_to_csr(n, r_, c_, v_)
            log.nnz.append(K.nnz)
            if solver == 2:
                du, its, _, ok = pcg(K, R, st.tol, st.max_iter, st.stop_mode, nthreads)
                log.pcg_iterations.append(its)
                if not ok:
                    raise OracleError(ERR_NOT_CONVERGED, f"PCG did not converge after {st.max_iter} iterations")
            else:
                du = direct_solve(K, R)
            log.linear_solves += 1
            u_curr = u_curr + du
            inc["iters"].append(it)
        inc["converged"] = converged
        log.increments.append(inc)
        if converged:
            current_time += dt
            u_start = u_curr
            log.u_committed = u_start
        else:
            dt /= 2.0
            if dt < st.min_time_increment:
                raise OracleError(ERR_NOT_CONVERGED, "Minimum time increment reached. Convergence failed.")
    return u_start


def run_job(job, solver_override=None, nthreads=0):
    """All steps in order (main.rs:49-63); returns (u, [StepLog])."""
    u = np.zeros(job.n_dof)
    timer = SolverTime()
    logs = []
    for k in range(len(job.steps)):
        lg = StepLog()
        u = static_step(job, k, u, timer, solver_override, lg, nthreads)
        logs.append(lg)
    return u, logs
