"""Polynomial-preconditioned mixed-precision batches (FX_PREC_MIXED_POLY):
flexible CG in FP64 preconditioned by Chebyshev steps on the Jacobi-scaled
operator stored in FP16, with FP64 iterates.

Bars:
* every load case within 1e-8 relative of the direct solve of the oracle's
  penalised matrix (solved to ||b - K x||/||b|| <= 1e-10);
* bit for bit alone vs in a mix, whatever was admitted before or later, and
  whatever the run chunking (a slot's arithmetic never depends on its
  neighbours or on the host's pacing);
* both layouts of the FP16 operator: sliced ELL (C3D4 meshes) and node
  blocks with lane groups (C3D20 meshes, or C3D4 uploaded with FX_SELL=0);
* the reference's stop rule (sqrt(r.z) < 1e-8 on the FP64 residual) and the
  not-converged error of the reference's solver;
* single systems through fx_solve / fx_static_step (a one-slot batch kept
  with the mesh): bit for bit the system alone in a batch, and the FP64
  steps' increments, Newton iterations and solves;
* every inner-step shape and row width gives the same bits.
"""
import numpy as np
import pytest

import oracle as O
from ferrix_amd import _lib as L
from ferrix_amd import meshgen
from ferrix_amd.device import Batch, DeviceMesh, SolveOptions
from ferrix_amd.inp import ConvergenceError, InvalidModelState

from test_gpu_batch import population, small_specs, system

pytestmark = pytest.mark.gpu

POLY = SolveOptions(precision=L.FX_PREC_MIXED_POLY, stop_mode=L.FX_STOP_REL_R2, tol=1e-10, max_iter=100_000)


def say(*a):
    print(*a, flush=True)


def oracle_errors(jobs, x, b):
    csr, _, _ = O.assemble_csr(jobs[0].mesh, jobs[0].node_temp, None, jobs[0].steps[0].bc_dof)
    out = []
    for r in range(b.shape[0]):
        xd = O.direct_solve(csr, b[r])
        out.append(float(np.linalg.norm(x[r] - xd) / np.linalg.norm(xd)))
    return out


def solve_alone(ctx, sys, opts):
    out = []
    for dm, b in sys:
        bt = Batch(ctx, opts)
        s = bt.add(dm, b)
        bt.run()
        out.append((bt.result(s), bt.poll(s)[1]))
        bt.close()
    return out


def test_poly_batch_alone_vs_mix_and_oracle(ctx):
    specs = small_specs(5, seed=31)
    sys = population(ctx, specs)
    solo = solve_alone(ctx, sys, POLY)
    bt = Batch(ctx, POLY)
    slots = [bt.add(dm, b) for dm, b in sys[:2]]
    bt.run(max_iters=40)  # the first two run a while before the others join
    slots += [bt.add(dm, b) for dm, b in sys[2:]]
    bt.run()
    for k, (s, spec) in enumerate(zip(slots, specs)):
        st, rep = bt.poll(s)
        assert st == 1
        x = bt.get(s)
        assert np.array_equal(x, solo[k][0]), f"system {k}: mix differs from alone"
        assert rep.iterations == solo[k][1].iterations
        errs = oracle_errors(meshgen.population_mesh(spec), x, sys[k][1])
        say("system", k, "n_dof", spec["n_dof"], "inner steps", rep.iterations, "outer", rep.outer_spmv,
            "rel err", max(errs))
        assert max(errs) < 1e-8
        assert rep.outer_spmv * 15 == rep.iterations


@pytest.mark.parametrize("sell", [True, False])
def test_poly_operator_layouts(ctx, sell):
    """The sliced-ELL FP16 operator and the node-block lane-group one
    (FX_SELL=0) each meet the bar; their summation orders differ, so their
    results agree to the tolerance, not bitwise."""
    specs = small_specs(3, seed=5)
    sys = population(ctx, specs, sell=sell)
    bt = Batch(ctx, POLY)
    slots = [bt.add(dm, b) for dm, b in sys]
    bt.run()
    for k, (s, spec) in enumerate(zip(slots, specs)):
        assert bt.poll(s)[0] == 1
        errs = oracle_errors(meshgen.population_mesh(spec), bt.result(s), sys[k][1])
        say("sell", sell, "system", k, "rel err", max(errs))
        assert max(errs) < 1e-8


def test_poly_c3d20_block(ctx):
    """A jittered 10x4x3 C3D20 block (node-block operator, 27 blocks per node
    row) beside two C3D4 systems in one batch."""
    job = meshgen.c3d20_block(seed=7, nx=10, ny=4, nz=3)
    dm20, b20 = system(ctx, [job])
    specs = small_specs(2, seed=8)
    sys = population(ctx, specs)
    bt = Batch(ctx, POLY)
    s20 = bt.add(dm20, b20)
    slots = [bt.add(dm, b) for dm, b in sys]
    bt.run()
    assert bt.poll(s20)[0] == 1
    errs = oracle_errors([job], bt.result(s20), b20)
    say("c3d20 rel err", errs, "report", bt.poll(s20)[1].iterations)
    assert max(errs) < 1e-8
    for k, s in enumerate(slots):
        assert max(oracle_errors(meshgen.population_mesh(specs[k]), bt.result(s), sys[k][1])) < 1e-8


@pytest.mark.parametrize("degree", [1, 8, 32])
def test_poly_degrees(ctx, degree):
    specs = small_specs(2, seed=13)
    sys = population(ctx, specs)
    opts = SolveOptions(precision=L.FX_PREC_MIXED_POLY, stop_mode=L.FX_STOP_REL_R2, tol=1e-10, max_iter=200_000,
                        poly_degree=degree)
    bt = Batch(ctx, opts)
    slots = [bt.add(dm, b) for dm, b in sys]
    bt.run()
    for k, s in enumerate(slots):
        st, rep = bt.poll(s)
        assert st == 1 and rep.iterations == degree * rep.outer_spmv
        assert max(oracle_errors(meshgen.population_mesh(specs[k]), bt.result(s), sys[k][1])) < 1e-8


def test_poly_chunking_does_not_change_results(ctx):
    sys = population(ctx, small_specs(4, seed=17))
    out = []
    for step in (0, 15, 100):
        bt = Batch(ctx, POLY)
        slots = [bt.add(dm, b) for dm, b in sys]
        calls = 0
        while bt.run(max_iters=step) > 0:
            calls += 1
            assert calls < 100_000
        out.append([(bt.result(s), bt.poll(s)[1].iterations) for s in slots])
        bt.close()
    for k in range(len(sys)):
        for other in out[1:]:
            assert np.array_equal(out[0][k][0], other[k][0]), f"system {k}"
            assert out[0][k][1] == other[k][1]


def test_poly_reference_stop_rule(ctx):
    """sqrt(r.z) < 1e-8 with z = D^-1 r on the FP64 residual (iterative.rs:63)."""
    specs = small_specs(2, seed=23)
    sys = population(ctx, specs)
    opts = SolveOptions(precision=L.FX_PREC_MIXED_POLY, stop_mode=L.FX_STOP_ABS_SQRT_RZ, tol=1e-8, max_iter=100_000)
    bt = Batch(ctx, opts)
    slots = [bt.add(dm, b) for dm, b in sys]
    bt.run()
    for k, s in enumerate(slots):
        st, rep = bt.poll(s)
        assert st == 1 and rep.final_criterion < 1e-8
        dm, b = sys[k]
        x = bt.result(s)
        jobs = meshgen.population_mesh(specs[k])
        csr, _, _ = O.assemble_csr(jobs[0].mesh, jobs[0].node_temp, None, jobs[0].steps[0].bc_dof)
        inv = O.jacobi(csr)
        for r in range(b.shape[0]):
            res = b[r] - O.spmv(csr, x[r])
            assert np.sqrt(res @ (inv * res)) < 2e-8  # the true residual, to rounding


def test_poly_not_converged(ctx):
    sys = population(ctx, small_specs(2, seed=3))
    bt = Batch(ctx, SolveOptions(precision=L.FX_PREC_MIXED_POLY, stop_mode=L.FX_STOP_REL_R2, tol=1e-10,
                                 max_iter=30))
    slots = [bt.add(dm, b) for dm, b in sys]
    bt.run()
    for s in slots:
        st, rep = bt.poll(s)
        assert st == -L.FX_ERR_NOT_CONVERGED
        assert rep.iterations == 30
        with pytest.raises(ConvergenceError):
            bt.result(s)


@pytest.mark.parametrize("kind", ["c3d4", "c3d20"])
def test_poly_single_solve(ctx, kind):
    """fx_solve in FX_PREC_MIXED_POLY: the system as the one slot of a batch
    kept with the mesh.  Bit for bit the same system alone in a batch, within
    1e-8 of the oracle's direct solve; a second solve reuses the batch, other
    options make a new one, and destroying the mesh frees it."""
    if kind == "c3d4":
        jobs = meshgen.population_mesh(small_specs(3, seed=71)[1])
    else:
        jobs = [meshgen.c3d20_block(seed=5, nx=10, ny=4, nz=3, jitter=True)]
    dm, b = system(ctx, jobs[:1])
    x, rep = dm.solve(b[0], POLY)
    say(kind, "inner steps", rep.iterations, "outer", rep.outer_spmv)
    assert rep.iterations > 0 and rep.final_criterion <= 1e-10
    bt = Batch(ctx, POLY)
    k = bt.add(dm, b)
    bt.run()
    assert np.array_equal(bt.result(k)[0], x)
    bt.close()
    x2, _ = dm.solve(b[0], POLY)
    assert np.array_equal(x2, x)
    looser = SolveOptions(precision=L.FX_PREC_MIXED_POLY, stop_mode=L.FX_STOP_REL_R2, tol=1e-6, max_iter=100_000)
    x3, rep3 = dm.solve(b[0], looser)
    assert rep3.iterations < rep.iterations and np.linalg.norm(x3 - x) < 1e-4 * np.linalg.norm(x)
    assert max(oracle_errors(jobs, x[None, :], b[:1])) < 1e-8
    dm.close()


def test_poly_static_steps(ctx):
    """fx_static_step with its linear solves in FX_PREC_MIXED_POLY: the
    linear cantilever and the NLGEOM beam take the FP64 steps' increments,
    Newton iterations and linear solves, U within 1e-8 of theirs."""
    from ferrix_amd.device import run_job
    for job in (meshgen.c3d4_cantilever(), meshgen.c3d20_nlgeom_beam(nx=8, ny=2, nz=2, tip_load=20000.0)):
        u64, r64 = run_job(ctx, job, precision=L.FX_PREC_FP64)
        up, rp = run_job(ctx, job, precision=L.FX_PREC_MIXED_POLY)
        e = np.linalg.norm(up - u64) / np.linalg.norm(u64)
        say("n_dof", job.n_dof, "NR", r64[0].nr_iterations, rp[0].nr_iterations, "solves", r64[0].linear_solves,
            rp[0].linear_solves, "U rel diff", e)
        for f in ("increments", "cutbacks", "nr_iterations", "linear_solves"):
            assert getattr(r64[0], f) == getattr(rp[0], f), f
        assert e < 1e-8


def test_poly_inner_shapes_bitwise(ctx):
    """The Chebyshev step as one launch (shape 0), as one launch per
    load-case count (shape 1), and the shape fx_batch_autotune picks after
    timing both on the running systems mid-solve: the same bits."""
    specs = small_specs(8, seed=41)
    sys = population(ctx, specs)
    counts = sorted({b.shape[0] for _, b in sys})
    say("load-case counts", counts)
    assert len(counts) >= 2
    out, tuned = [], None
    for shape in (0, 1, "tune"):
        bt = Batch(ctx, POLY)
        if shape != "tune":
            bt.set_inner_shape(shape)
        else:
            bt.set_overlap(True)  # the tuner first takes up the chunks left in flight
        slots = [bt.add(dm, b) for dm, b in sys]
        if shape == "tune":
            assert bt.run(max_iters=45) > 0
            tuned = bt.autotune(reps=5)
            say("autotune: shape", tuned[0], "ms one launch", tuned[1], "ms split", tuned[2])
            assert tuned[0] in (0, 1) and tuned[1] > 0 and tuned[2] > 0
        assert bt.run() == 0
        out.append([(bt.result(s), bt.poll(s)[1].iterations) for s in slots])
        bt.close()
    for k in range(len(sys)):
        for other in out[1:]:
            assert np.array_equal(out[0][k][0], other[k][0]), f"system {k}"
            assert out[0][k][1] == other[k][1]
    assert max(oracle_errors(meshgen.population_mesh(specs[0]), out[1][0][0], sys[0][1])) < 1e-8


@pytest.mark.parametrize("overlap", [False, True])
def test_poly_overlap_driver_bitwise(ctx, overlap):
    """A driver loop of budgeted runs that collects finished systems (copied
    out while later chunks still run, with overlap), drops a running system
    and admits it again later, and admits the others into freed slots:
    every result and iteration count equals the system solved alone."""
    specs = small_specs(6, seed=43)
    sys = population(ctx, specs)
    solo = []
    for dm, b in sys:
        bt = Batch(ctx, POLY)
        s = bt.add(dm, b)
        assert bt.run() == 0
        solo.append((bt.result(s), bt.poll(s)[1].iterations))
        bt.close()
    bt = Batch(ctx, POLY)
    bt.set_overlap(overlap)
    todo, running, got = list(range(len(sys))), {}, {}
    dropped, passes = None, 0
    while running or todo:
        while todo and len(running) < 3:
            k = todo.pop(0)
            running[bt.add(*sys[k])] = k
        bt.run(max_iters=30)
        passes += 1
        if passes == 2:  # a system still running leaves, and joins again at the end
            live = [s for s in running if bt.poll(s)[0] == 0]
            if live:
                dropped = running.pop(live[0])
                bt.remove(live[0])
                todo.append(dropped)
        for s in bt.finished():
            k = running.pop(s)
            got[k] = (bt.result(s), bt.poll(s)[1].iterations)
            bt.remove(s)
    bt.close()
    say("overlap", overlap, "passes", passes, "dropped", dropped)
    assert dropped is not None and passes > 3
    for k in range(len(sys)):
        assert np.array_equal(got[k][0], solo[k][0]), f"system {k}"
        assert got[k][1] == solo[k][1], f"system {k}"


def test_poly_autotune_needs_poly_batch(ctx):
    bt = Batch(ctx, SolveOptions(precision=L.FX_PREC_FP64, stop_mode=L.FX_STOP_REL_R2, tol=1e-10))
    with pytest.raises(InvalidModelState, match="polynomial batch"):
        bt.autotune()
    with pytest.raises(InvalidModelState, match="shape must be 0 or 1"):
        bt.set_inner_shape(2)
    bt.close()


def test_poly_rows_per_workgroup_bitwise(ctx):
    """A batch whose running systems hold more than 2 M node rows runs the
    Chebyshev step with 512 rows per workgroup; a system alone runs it with
    256.  The system's solutions are the same bits either way."""
    specs = meshgen.population_specs(n_meshes=40, seed=61, dof_lo=1.5e5, dof_hi=2.0e5)
    sys = population(ctx, specs)
    rows = sum(dm.n_dof // 3 for dm, _ in sys)
    say("node rows in the batch", rows)
    assert rows > 2 << 20
    bt = Batch(ctx, POLY)
    slots = [bt.add(dm, b) for dm, b in sys]
    assert bt.run() == 0
    mixed = [(bt.result(s), bt.poll(s)[1].iterations) for s in slots[:2]]
    bt.close()
    for (dm, b), (x_mix, it_mix) in zip(sys[:2], mixed):
        alone = Batch(ctx, POLY)
        k = alone.add(dm, b)
        alone.run()
        assert np.array_equal(alone.result(k), x_mix)
        assert alone.poll(k)[1].iterations == it_mix
        alone.close()
    for dm, _ in sys:
        dm.close()
