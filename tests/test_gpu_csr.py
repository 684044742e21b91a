"""GPU: Solver::solve on the caller's CSR (fx_solve_csr), the seam's own
contract (solvers/mod.rs:18-44, called at static_step.rs:198-204 with the
matrix the step just built).

The matrix handed in is the oracle's CSR of each system (triplets -> CSR
with the penalty BC entries, static_step.rs:183-198), so the GPU solves
exactly the matrix the reference's solver receives.  Bars: under the
reference's stop rule the oracle's PCG iteration count within 2 iterations
(the SpMV row sums and dots associate differently; at a relative residual of
1e-12 within 10%), the solution within 1e-8 relative of the oracle's PCG and
of its direct solve.
"""
import numpy as np
import pytest

import oracle as O
from ferrix_amd import _lib as L
from ferrix_amd import meshgen
from ferrix_amd.device import SolveOptions
from ferrix_amd.inp import ConvergenceError, InvalidModelState

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def system(job):
    """The oracle's penalised CSR and right-hand side of the job's first
    step at u = 0 (the matrix and residual the reference's solver gets)."""
    st = job.steps[0]
    csr, md, pen = O.assemble_csr(job.mesh, job.node_temp, None, st.bc_dof)
    b = np.zeros(job.n_dof)
    np.add.at(b, st.load_dof, st.load_val)
    np.add.at(b, np.asarray(st.bc_dof), pen * np.asarray(st.bc_val, np.float64))
    return csr, b


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_solve_csr_reference_rule_cantilever(ctx):
    """sqrt(r.z) < 1e-8, at most 1000 iterations (iterative.rs:26-121,
    static_step.rs:204) on the 20x4x4 C3D4 cantilever."""
    csr, b = system(meshgen.c3d4_cantilever())
    x, rep = ctx.solve_csr(csr.row_ptr, csr.col, csr.val, b, SolveOptions())
    xo, it_o, _, ok = O.pcg(csr, b, 1e-8, 1000, 0)
    assert ok
    say("iterations gpu", rep.iterations, "oracle", it_o, "group", rep.spmv_group)
    assert abs(rep.iterations - it_o) <= 2
    assert rep.final_criterion < 1e-8
    xd = O.direct_solve(csr, b)
    assert rel(x, xo) < 1e-8
    assert rel(x, xd) < 1e-8


@pytest.mark.parametrize("group", [0, 2, 64])
def test_solve_csr_c3d20_block(ctx, group):
    """A jittered 10x4x3 C3D20 block (156 entries per row) to a relative
    residual of 1e-12, every lane-group size giving the direct solution."""
    csr, b = system(meshgen.c3d20_block(seed=7, nx=10, ny=4, nz=3))
    opts = SolveOptions(stop_mode=L.FX_STOP_REL_R2, tol=1e-12, max_iter=100_000, spmv_group=group)
    x, rep = ctx.solve_csr(csr.row_ptr, csr.col, csr.val, b, opts)
    xo, it_o, _, ok = O.pcg(csr, b, 1e-12, 100_000, 1)
    assert ok
    say("group", rep.spmv_group, "iterations gpu", rep.iterations, "oracle", it_o)
    # at 1e-12 the finite-precision iteration count depends on the summation
    # order of the row sums and dots: within 10% of the oracle's
    assert abs(rep.iterations - it_o) <= max(2, it_o // 10)
    assert rel(x, O.direct_solve(csr, b)) < 1e-8


def test_solve_csr_uses_the_matrix_passed_in(ctx):
    """Two different matrices through the same context give their own
    solutions: nothing of an earlier call's matrix is kept."""
    c1, b1 = system(meshgen.c3d4_cantilever())
    scaled = O.CSR(c1.n, c1.row_ptr, c1.col, 2.0 * c1.val)
    opts = SolveOptions(stop_mode=L.FX_STOP_REL_R2, tol=1e-12, max_iter=100_000)
    x1, _ = ctx.solve_csr(c1.row_ptr, c1.col, c1.val, b1, opts)
    x2, _ = ctx.solve_csr(scaled.row_ptr, scaled.col, scaled.val, b1, opts)
    assert rel(x2, 0.5 * x1) < 1e-9


def test_solve_csr_jacobi_guard(ctx):
    """Rows are scaled by 1/K_ii only when |K_ii| > 1e-9, else by 1
    (preconditioner.rs:44-48): diag(1e-10, 2, 3) with a coupling between
    rows 1 and 2.  The GPU takes the oracle's iterations and reaches the
    exact solution."""
    rp = np.array([0, 1, 3, 5], np.int64)
    col = np.array([0, 1, 2, 1, 2], np.int64)
    val = np.array([1e-10, 2.0, 0.5, 0.5, 3.0])
    A = np.array([[1e-10, 0, 0], [0, 2.0, 0.5], [0, 0.5, 3.0]])
    b = np.array([1e-10, 1.0, 2.0])
    inv = O.jacobi(O.CSR(3, rp, col.astype(np.int32), val))
    assert np.array_equal(inv, np.array([1.0, 0.5, 1.0 / 3.0]))
    x, rep = ctx.solve_csr(rp, col, val, b, SolveOptions(stop_mode=L.FX_STOP_REL_R2, tol=1e-14, max_iter=100))
    _, it_o, _, ok = O.pcg(O.CSR(3, rp, col.astype(np.int32), val), b, 1e-14, 100, 1)
    assert ok and abs(rep.iterations - it_o) <= 1
    assert rel(x, np.linalg.solve(A, b)) < 1e-12


def test_solve_csr_not_converged(ctx):
    csr, b = system(meshgen.c3d4_cantilever())
    with pytest.raises(ConvergenceError):
        ctx.solve_csr(csr.row_ptr, csr.col, csr.val, b, SolveOptions(max_iter=5))


def test_solve_csr_rejects_malformed_matrices(ctx):
    rp = np.array([0, 2, 3], np.int64)
    b = np.ones(2)
    with pytest.raises(InvalidModelState):  # column out of range
        ctx.solve_csr(rp, np.array([0, 2, 1], np.int64), np.ones(3), b)
    with pytest.raises(InvalidModelState):  # not ascending
        ctx.solve_csr(rp, np.array([1, 0, 1], np.int64), np.ones(3), b)
    with pytest.raises(InvalidModelState):  # row_ptr decreasing
        ctx.solve_csr(np.array([0, 2, 1], np.int64), np.array([0, 1], np.int64), np.ones(2), b)
    with pytest.raises(InvalidModelState):  # mixed precision needs a mesh
        ctx.solve_csr(np.array([0, 1, 2], np.int64), np.array([0, 1], np.int64), np.ones(2), b,
                      SolveOptions(precision=L.FX_PREC_MIXED))
