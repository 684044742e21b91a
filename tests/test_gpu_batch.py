"""Batched solves (fx_batch): many resident systems, up to 4 load cases each,
advanced by one set of kernel launches.

Bars:
* FP64 on the lane-group operator (the single-solve SpMV): every batched
  load case is BIT FOR BIT the fx_solve result of the same system and
  options (same lane group);
* FP64 on the sliced-ELL operator (short-row meshes in a batch): within
  1e-9 relative of fx_solve (different summation order in the SpMV rows);
* any precision: bit for bit alone vs in a mix, admitted at any time;
* mixed precision: bit for bit alone vs in a mix (a slot's arithmetic does
  not depend on its neighbours), and within 1e-8 relative of the direct
  solve of the oracle's matrix;
* the reference's stop rule and error behaviour (not converged) per slot.
"""
import dataclasses
import os

import numpy as np
import pytest

import oracle as O
from ferrix_amd import meshgen
from ferrix_amd import _lib as L
from ferrix_amd.device import Batch, DeviceMesh, SolveOptions
from ferrix_amd.inp import ConvergenceError

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def system(ctx, jobs):
    """Resident K (assembled, penalty applied) of jobs sharing one mesh and
    BC set, with one right-hand side per job."""
    j0 = jobs[0]
    dm = DeviceMesh(ctx, j0.mesh, j0.node_temp)
    dm.assemble(False)
    rhs = []
    for j in jobs:
        st = j.steps[0]
        b = np.zeros(j.n_dof)
        np.add.at(b, st.load_dof, st.load_val)
        rhs.append(b)
    st = j0.steps[0]
    out = []
    pen = None
    for b in rhs:
        if pen is None:
            bp, pen = dm.apply_penalty(st.bc_dof, st.bc_val, b)
        else:  # K already carries the penalty: the right-hand side only (u = 0)
            bp = b.copy()
            for d, t in zip(st.bc_dof, st.bc_val):
                bp[d] += pen * t
        out.append(bp)
    return dm, np.array(out)


def population(ctx, specs, sell=True):
    """Resident systems; sell=False uploads them without the sliced-ELL
    operator (FX_SELL=0), so batches run the single-solve SpMV."""
    old = os.environ.get("FX_SELL")
    os.environ["FX_SELL"] = "1" if sell else "0"
    try:
        return [system(ctx, meshgen.population_mesh(s)) for s in specs]
    finally:
        if old is None:
            del os.environ["FX_SELL"]
        else:
            os.environ["FX_SELL"] = old


def small_specs(n, seed=11):
    return meshgen.population_specs(n_meshes=n, seed=seed, dof_lo=300, dof_hi=6000)


REL = SolveOptions(stop_mode=L.FX_STOP_REL_R2, tol=1e-10, max_iter=100_000, spmv_group=8)


def test_fp64_batch_equals_solve_bitwise(ctx):
    sys = population(ctx, small_specs(6), sell=False)
    bt = Batch(ctx, REL)
    slots = [bt.add(dm, b) for dm, b in sys]
    assert bt.run() == 0
    for (dm, b), s in zip(sys, slots):
        st, rep = bt.poll(s)
        assert st == 1
        x = bt.get(s)
        for r in range(b.shape[0]):
            xs, rs = dm.solve(b[r], REL)
            say("slot", s, "rhs", r, "n_dof", dm.n_dof, "iters", rs.iterations)
            assert np.array_equal(x[r], xs), f"slot {s} rhs {r}: max diff {np.max(np.abs(x[r] - xs))}"


def test_fp64_reference_rule_batch(ctx):
    """Default options: sqrt(r.z) < 1e-8, 1000 iterations (the reference's
    PCG); the slot's iteration count is its slowest load case's."""
    sys = population(ctx, small_specs(3, seed=5), sell=False)
    bt = Batch(ctx, SolveOptions(spmv_group=8))
    slots = [bt.add(dm, b) for dm, b in sys]
    bt.run()
    for (dm, b), s in zip(sys, slots):
        st, rep = bt.poll(s)
        if st < 0:
            assert -st == L.FX_ERR_NOT_CONVERGED
            continue
        its = []
        x = bt.get(s)
        for r in range(b.shape[0]):
            xs, rs = dm.solve(b[r], SolveOptions(spmv_group=8))
            its.append(rs.iterations)
            assert np.array_equal(x[r], xs)
        assert rep.iterations == max(its)


def test_fp64_batch_sell_matches_solve(ctx):
    sys = population(ctx, small_specs(6, seed=13), sell=True)
    bt = Batch(ctx, REL)
    slots = [bt.add(dm, b) for dm, b in sys]
    assert bt.run() == 0
    for (dm, b), s in zip(sys, slots):
        x = bt.result(s)
        _, rep = bt.poll(s)
        its = []
        for r in range(b.shape[0]):
            xs, rs = dm.solve(b[r], REL)
            its.append(rs.iterations)
            err = np.linalg.norm(x[r] - xs) / np.linalg.norm(xs)
            say("slot", s, "rhs", r, "n_dof", dm.n_dof, "rel diff", err)
            assert err < 1e-9
        assert abs(rep.iterations - max(its)) <= max(2, max(its) // 50)


@pytest.mark.parametrize("overlap", [False, True])
def test_continuous_admission_and_removal(ctx, overlap):
    """Systems admitted while others run, one removed mid-solve and its slot
    reused: every finished result equals the system solved alone (overlap:
    the budgeted runs return with chunks still in flight)."""
    sys = population(ctx, small_specs(5, seed=23))
    solo = []
    for dm, b in sys:
        bt = Batch(ctx, REL)
        s = bt.add(dm, b)
        bt.run()
        solo.append(bt.result(s))
        bt.close()
    bt = Batch(ctx, REL)
    bt.set_overlap(overlap)
    s0 = bt.add(*sys[0])
    s1 = bt.add(*sys[1])
    bt.run(max_iters=20)
    s2 = bt.add(*sys[2])
    bt.run(max_iters=13)
    bt.remove(s1)  # dropped while running
    s3 = bt.add(*sys[3])
    assert s3 == s1  # the slot is reused
    bt.run(max_iters=7)
    s4 = bt.add(*sys[4])
    assert bt.run() == 0
    for k, s in [(0, s0), (2, s2), (3, s3), (4, s4)]:
        assert np.array_equal(bt.result(s), solo[k]), f"system {k}"


@pytest.mark.parametrize("spmv_group", [4, 32])
def test_mixed_batch_alone_vs_mix(ctx, spmv_group):
    opts = SolveOptions(precision=L.FX_PREC_MIXED, stop_mode=L.FX_STOP_REL_R2, tol=1e-10, max_iter=100_000,
                        inner_tol=1e-3, spmv_group=spmv_group)
    specs = small_specs(4, seed=31)
    sys = population(ctx, specs)
    solo = []
    for dm, b in sys:
        bt = Batch(ctx, opts)
        s = bt.add(dm, b)
        bt.run()
        solo.append(bt.result(s))
        bt.close()
    bt = Batch(ctx, opts)
    slots = [bt.add(dm, b) for dm, b in sys[:2]]
    bt.run(max_iters=9)
    slots += [bt.add(dm, b) for dm, b in sys[2:]]
    bt.run()
    for k, (s, spec) in enumerate(zip(slots, specs)):
        st, rep = bt.poll(s)
        assert st == 1 and rep.refinements >= 2
        x = bt.get(s)
        assert np.array_equal(x, solo[k]), f"system {k}"
        # against the direct solve of the oracle's penalised matrix
        jobs = meshgen.population_mesh(spec)
        csr, _, _ = O.assemble_csr(jobs[0].mesh, jobs[0].node_temp, None, jobs[0].steps[0].bc_dof)
        b = sys[k][1]
        for r in range(b.shape[0]):
            xd = O.direct_solve(csr, b[r])
            err = np.linalg.norm(x[r] - xd) / np.linalg.norm(xd)
            say("system", k, "rhs", r, "rel err", err, "inner iters", rep.iterations, "passes", rep.refinements)
            assert err < 1e-8


def test_mixed_batch_matches_fp64_solve(ctx):
    specs = small_specs(2, seed=41)
    sys = population(ctx, specs)
    mixed = SolveOptions(precision=L.FX_PREC_MIXED, stop_mode=L.FX_STOP_REL_R2, tol=1e-10, max_iter=100_000)
    bt = Batch(ctx, mixed)
    slots = [bt.add(dm, b) for dm, b in sys]
    bt.run()
    for (dm, b), s in zip(sys, slots):
        x = bt.result(s)
        for r in range(b.shape[0]):
            x64, _ = dm.solve(b[r], SolveOptions(stop_mode=L.FX_STOP_REL_R2, tol=1e-12, max_iter=100_000))
            assert np.linalg.norm(x[r] - x64) / np.linalg.norm(x64) < 1e-8


def test_mixed_batch_symmetric_and_full_fp32_operators(ctx):
    """Mixed batches on sliced-ELL meshes use the symmetric FP32 operator by
    default (upper blocks stored once, lower blocks read transposed);
    FX_SSYM=0 at fx_batch_add keeps the full FP32 operator.  Both solve to
    1e-8 of the oracle's direct solve, and the symmetric one streams fewer
    algorithmic bytes per iteration."""
    specs = small_specs(3, seed=43)
    sys = population(ctx, specs)
    mixed = SolveOptions(precision=L.FX_PREC_MIXED, stop_mode=L.FX_STOP_REL_R2, tol=1e-10, max_iter=100_000)
    out, per_launch = {}, {}
    old = os.environ.get("FX_SSYM")
    try:
        for mode in ("0", "1"):
            os.environ["FX_SSYM"] = mode
            bt = Batch(ctx, mixed)
            slots = [bt.add(dm, b) for dm, b in sys]
            bt.run()
            out[mode] = [bt.result(s) for s in slots]
            launches, _, _, nbytes = bt.spmv_stats()
            per_launch[mode] = nbytes / launches
            bt.close()
    finally:
        if old is None:
            os.environ.pop("FX_SSYM", None)
        else:
            os.environ["FX_SSYM"] = old
    say("algorithmic bytes per SpMV launch: full", per_launch["0"], "symmetric", per_launch["1"])
    assert per_launch["1"] < 0.8 * per_launch["0"]
    for k, spec in enumerate(specs):
        jobs = meshgen.population_mesh(spec)
        csr, _, _ = O.assemble_csr(jobs[0].mesh, jobs[0].node_temp, None, jobs[0].steps[0].bc_dof)
        b = sys[k][1]
        for r in range(b.shape[0]):
            xd = O.direct_solve(csr, b[r])
            for mode in ("0", "1"):
                assert np.linalg.norm(out[mode][k][r] - xd) / np.linalg.norm(xd) < 1e-8, (k, r, mode)


def test_batch_with_unused_node(ctx):
    """A node that no element uses still gets its three DOFs (mesh.rs:35-55):
    its block row is the diagonal block alone, zero until the penalty of its
    clamp.  Placed inside a slice, that row goes through the sliced-ELL
    operators (FP64, and the symmetric FP32 one of the mixed batch); both
    batches match the oracle's direct solve."""
    job = meshgen.population_mesh(small_specs(1, seed=47)[0])[0]
    m, st = job.mesh, job.steps[0]
    k = m.n_nodes // 2 + 5
    mesh = dataclasses.replace(m, node_ids=np.arange(1, m.n_nodes + 2, dtype=m.node_ids.dtype),
                               xyz=np.insert(m.xyz, k, m.xyz[k] + 0.25, axis=0),
                               conn=np.where(m.conn >= k, m.conn + 1, m.conn).astype(m.conn.dtype))
    temp = np.insert(job.node_temp, k, job.node_temp[k])
    shift = lambda d: np.where(d >= 3 * k, d + 3, d)
    bc_dof = np.concatenate([shift(st.bc_dof), 3 * k + np.arange(3)])
    bc_val = np.concatenate([st.bc_val, np.zeros(3)])
    b = np.zeros(mesh.n_dof)
    np.add.at(b, shift(st.load_dof), st.load_val)
    csr, _, _ = O.assemble_csr(mesh, temp, None, bc_dof)
    for prec in (L.FX_PREC_FP64, L.FX_PREC_MIXED):
        dm = DeviceMesh(ctx, mesh, temp)
        dm.assemble(False)
        bp, _ = dm.apply_penalty(bc_dof, bc_val, b)
        bt = Batch(ctx, SolveOptions(precision=prec, stop_mode=L.FX_STOP_REL_R2, tol=1e-10, max_iter=100_000))
        s = bt.add(dm, bp[None, :])
        bt.run()
        x = bt.result(s)[0]
        xd = O.direct_solve(csr, bp)
        say("precision", prec, "rel err", np.linalg.norm(x - xd) / np.linalg.norm(xd))
        assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-8
        assert np.all(x[3 * k:3 * k + 3] == 0.0)
        bt.close()
        dm.close()


def test_mixed_batch_ill_conditioned_does_not_stall(ctx):
    """Member 216 of the 512-mesh population (26x4x4 cells, 2,025 DOF, two
    load cases; FP64 PCG needs about 680 iterations): the FP32 operator's
    error is large enough here that a residual replacement which kept the
    Krylov direction once left load case 0 on a stalling FP32 Krylov space
    (23,000 iterations).  Each replacement now restarts the direction unless
    the new residual is within 10% of the iterated one."""
    spec = meshgen.population_specs()[216]
    assert (spec["nx"], spec["ny"], spec["nz"]) == (26, 4, 4)
    (dm, b), = population(ctx, [spec])
    bt = Batch(ctx, SolveOptions(precision=L.FX_PREC_MIXED, stop_mode=L.FX_STOP_REL_R2, tol=1e-10, max_iter=6000,
                                 inner_tol=1e-3, max_refine=100))
    s = bt.add(dm, b)
    bt.run()
    st, rep = bt.poll(s)
    say("iterations", rep.iterations, "replacements", rep.refinements, "criterion", rep.final_criterion)
    assert st == 1 and rep.iterations < 6000
    x = bt.result(s)
    jobs = meshgen.population_mesh(spec)
    csr, _, _ = O.assemble_csr(jobs[0].mesh, jobs[0].node_temp, None, jobs[0].steps[0].bc_dof)
    for r in range(b.shape[0]):
        xd = O.direct_solve(csr, b[r])
        assert np.linalg.norm(x[r] - xd) / np.linalg.norm(xd) < 1e-8


@pytest.mark.parametrize("precision", [L.FX_PREC_FP64, L.FX_PREC_MIXED])
def test_run_chunking_does_not_change_results(ctx, precision):
    """fx_batch_run keeps two 16-iteration chunks in flight, learns of
    finished slots from summaries the GPU writes to host memory, and
    rebuilds its block tables lazily.  None of that may change a slot's
    arithmetic: one run to completion, many 5-iteration runs and 37-iteration
    runs give bitwise equal solutions and iteration counts."""
    opts = SolveOptions(precision=precision, stop_mode=L.FX_STOP_REL_R2, tol=1e-10, max_iter=100_000,
                        inner_tol=1e-3)
    sys = population(ctx, small_specs(5, seed=17))
    out = []
    for step in (0, 5, 37):
        bt = Batch(ctx, opts)
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


def test_batch_not_converged(ctx):
    sys = population(ctx, small_specs(2, seed=3))
    bt = Batch(ctx, SolveOptions(max_iter=3, spmv_group=8))
    slots = [bt.add(dm, b) for dm, b in sys]
    bt.run()
    for s in slots:
        st, rep = bt.poll(s)
        assert st == -L.FX_ERR_NOT_CONVERGED
        assert rep.iterations == 3
        with pytest.raises(ConvergenceError):
            bt.result(s)


def test_mesh_destroyed_while_batched(ctx):
    sys = population(ctx, small_specs(2, seed=9))
    bt = Batch(ctx, REL)
    s0 = bt.add(*sys[0])
    bt.add(*sys[1])
    bt.run(max_iters=5)
    sys[1][0].close()  # drops its slot from the batch
    assert bt.run() == 0
    assert bt.poll(s0)[0] == 1


@pytest.mark.parametrize("precision", [L.FX_PREC_FP64, L.FX_PREC_MIXED])
def test_population_driver_budget(ctx, precision):
    """solve_batched admitting everything at once, and under a working-set
    budget that makes systems wait for others to finish (admitted between
    chunks, into reused slots): every system's result is bit for bit the
    same, since a slot's arithmetic does not depend on when it joins."""
    from ferrix_amd.batch import System, operator_bytes, solve_batched, upload

    systems = [System(k, meshgen.population_mesh(sp)) for k, sp in enumerate(small_specs(7, seed=53))]
    upload(ctx, systems)
    opts = SolveOptions(precision=precision, stop_mode=L.FX_STOP_REL_R2, tol=1e-10, max_iter=100_000)
    a = solve_batched(ctx, systems, opts)
    budget = 2 * max(operator_bytes(s, opts) for s in systems)
    c = solve_batched(ctx, systems, opts, ws_bytes=budget, chunk_iters=32)
    for r in (a, c):
        assert not r.failures and sorted(r.x) == list(range(len(systems)))
    for k in range(len(systems)):
        assert np.array_equal(a.x[k], c.x[k]), k
    say("seconds", a.seconds, c.seconds)


@pytest.mark.parametrize("precision", [L.FX_PREC_FP64, L.FX_PREC_MIXED])
def test_population_driver_assembles_ahead(ctx, precision):
    """Meshes on a second context, assembled by a host thread on that
    context's stream while the batch runs (assemble_ahead): every solution and
    iteration count is bit for bit the serial driver's, with everything
    admitted at once and under a working-set budget."""
    from ferrix_amd.batch import System, operator_bytes, solve_batched, upload
    from ferrix_amd.device import Context

    specs = small_specs(7, seed=53)
    ref = [System(k, meshgen.population_mesh(sp)) for k, sp in enumerate(specs)]
    upload(ctx, ref)
    opts = SolveOptions(precision=precision, stop_mode=L.FX_STOP_REL_R2, tol=1e-10, max_iter=100_000)
    a = solve_batched(ctx, ref, opts)
    with pytest.raises(ValueError):
        solve_batched(ctx, ref, opts, assemble_ahead=True)  # meshes on the batch's own context
    mctx = Context(0)
    try:
        systems = [System(k, meshgen.population_mesh(sp)) for k, sp in enumerate(specs)]
        upload(mctx, systems)
        b = solve_batched(ctx, systems, opts, chunk_iters=16, assemble_ahead=True)
        budget = 2 * max(operator_bytes(s, opts) for s in systems)
        c = solve_batched(ctx, systems, opts, ws_bytes=budget, chunk_iters=32, assemble_ahead=True)
        for r in (a, b, c):
            assert not r.failures and sorted(r.x) == list(range(len(systems)))
        for r in (b, c):
            for k in range(len(systems)):
                assert np.array_equal(a.x[k], r.x[k]), k
                assert a.iterations[k] == r.iterations[k], k
        say("seconds", a.seconds, b.seconds, c.seconds)
    finally:
        mctx.close()


def test_mesh_of_another_context_destroyed_while_batched(ctx):
    """A mesh of a second context in a batch: destroying it drops its slot
    (after the batch's stream has finished with it), the other system still
    converges, and closing the mesh's context with the batch alive drops
    the rest."""
    from ferrix_amd.device import Context

    mctx = Context(0)
    bt = Batch(ctx, REL)
    try:
        sys = population(mctx, small_specs(2, seed=9))
        s0 = bt.add(*sys[0])
        bt.add(*sys[1])
        bt.run(max_iters=5)
        sys[1][0].close()
        assert bt.run() == 0
        assert bt.poll(s0)[0] == 1
        x0 = bt.get(s0)
        for r, b in enumerate(sys[0][1]):  # the same system through fx_solve
            x1, _ = sys[0][0].solve(b, REL)
            assert np.max(np.abs(x0[r] - x1)) <= 1e-8 * np.max(np.abs(x1))
        mctx.close()
        assert bt.run() == 0
    finally:
        bt.close()
        mctx.close()


@pytest.mark.parametrize("precision", [L.FX_PREC_FP64, L.FX_PREC_MIXED, L.FX_PREC_MIXED_POLY])
def test_device_admission_bitwise(ctx, precision):
    """fx_batch_add_system (assembly, penalty and admission queued with no
    host wait) against assemble + apply_penalty + fx_batch_add: the same
    solutions bit for bit, and the same iteration counts."""
    from ferrix_amd.batch import System, solve_batched, upload
    specs = small_specs(6, seed=41)
    opts = SolveOptions(precision=precision, stop_mode=L.FX_STOP_REL_R2, tol=1e-10,
                        max_iter=400_000 if precision == L.FX_PREC_MIXED_POLY else 100_000)
    out = []
    for dev in (True, False):
        systems = [System(k, meshgen.population_mesh(sp)) for k, sp in enumerate(specs)]
        upload(ctx, systems)
        out.append(solve_batched(ctx, systems, opts, chunk_iters=16, device_admission=dev))
        for s in systems:
            s.dm.close()
    a, b = out
    assert not a.failures and not b.failures
    for k in b.x:
        assert np.array_equal(a.x[k], b.x[k]), f"system {k}"
        assert a.iterations[k] == b.iterations[k]


def test_device_admission_element_failure(ctx):
    """A mesh whose element K fails (an element collapsed onto a plane)
    admitted without a host wait: its slot finishes with the singular
    Jacobian status, and its neighbours in the batch solve as alone."""
    from ferrix_amd.batch import System, admit
    specs = small_specs(2, seed=43)
    good = [System(k, meshgen.population_mesh(sp)) for k, sp in enumerate(specs)]
    bad_job = meshgen.c3d4_cantilever()
    m = bad_job.mesh
    xyz = m.xyz.copy()
    nodes = m.conn[m.el_ptr[5]:m.el_ptr[6]]
    for nd in nodes[1:]:
        xyz[nd] = xyz[nodes[0]]
    m.xyz = xyz
    bad = System(99, [bad_job])
    for s in good + [bad]:
        s.dm = DeviceMesh(ctx, s.jobs[0].mesh, s.jobs[0].node_temp)
    bt = Batch(ctx, REL)
    slots = [admit(bt, good[0]), admit(bt, bad), admit(bt, good[1])]
    bt.run()
    st_bad, _ = bt.poll(slots[1])
    say("bad slot state", st_bad)
    assert st_bad == -L.FX_ERR_SINGULAR_JACOBIAN
    for s, slot in ((good[0], slots[0]), (good[1], slots[2])):
        assert bt.poll(slot)[0] == 1
        alone = Batch(ctx, REL)
        k = admit(alone, s)
        alone.run()
        assert np.array_equal(bt.result(slot), alone.result(k))
        alone.close()
    bt.close()


def test_foreign_mesh_destroyed_while_batch_runs(ctx):
    """A batch on one context runs meshes of another context while a second
    host thread destroys one of those meshes and then the other context:
    the destroyed mesh's slot leaves the batch (its poll reports an invalid
    slot), the other systems finish with the results they get alone, and
    destroying the batch and the context in either order is clean."""
    import threading
    import time
    from ferrix_amd.device import Context
    specs = small_specs(3, seed=47)
    alone = []
    for sp in specs:
        dm, b = system(ctx, meshgen.population_mesh(sp))
        bt = Batch(ctx, REL)
        k = bt.add(dm, b)
        bt.run()
        alone.append(bt.result(k))
        bt.close()
        dm.close()
    for rnd in range(3):
        other = Context(0)
        sys_ = [system(other, meshgen.population_mesh(sp)) for sp in specs]
        bt = Batch(ctx, REL)
        slots = [bt.add(dm, b) for dm, b in sys_]
        dropped = threading.Event()

        def destroy():
            time.sleep(0.002 * rnd)
            sys_[1][0].close()
            dropped.set()

        th = threading.Thread(target=destroy)
        th.start()
        done = {}
        while len(done) < 2:
            bt.run(max_iters=8)
            for i in (0, 2):
                if i not in done and bt.poll(slots[i])[0] != 0:
                    done[i] = bt.result(slots[i])
        th.join()
        assert dropped.is_set()
        with pytest.raises(Exception):
            bt.poll(slots[1])  # the slot left with its mesh
        for i in (0, 2):
            assert np.array_equal(done[i], alone[i]), (rnd, i)
        if rnd % 2 == 0:
            bt.close()
            other.close()
        else:
            other.close()  # drops the remaining foreign slots first
            bt.close()
        say("round", rnd, "ok")


def test_add_system_load_records(ctx):
    """Load records with repeated DOFs (within and across load cases, in
    scattered order) through fx_batch_add_system: the solutions bit for bit
    those of the dense vectors np.add.at builds, admitted by assemble +
    apply_penalty + fx_batch_add."""
    from ferrix_amd.batch import System, admit, assemble
    sp = small_specs(3, seed=53)[2]
    jobs = meshgen.population_mesh(sp)
    rng = np.random.default_rng(5)
    for j in jobs:
        st = j.steps[0]
        d = np.asarray(st.load_dof, np.int64)
        v = np.asarray(st.load_val, np.float64)
        extra = rng.integers(0, len(d), size=len(d))  # repeats of existing DOFs
        dd = np.concatenate([d, d[extra], rng.integers(0, j.n_dof, size=7)])
        vv = np.concatenate([v, rng.normal(size=len(extra)) * np.abs(v).max(), rng.normal(size=7)])
        perm = rng.permutation(len(dd))
        st.load_dof, st.load_val = dd[perm], vv[perm]
    say("load cases", len(jobs), "records", [len(j.steps[0].load_dof) for j in jobs])
    out = []
    for dev in (True, False):
        s = System(0, jobs)
        s.dm = DeviceMesh(ctx, jobs[0].mesh, jobs[0].node_temp)
        bt = Batch(ctx, REL)
        if dev:
            k = admit(bt, s)
        else:
            assemble(s)
            k = bt.add(s.dm, s.rhs)
        bt.run()
        out.append(bt.result(k))
        bt.close()
        s.dm.close()
    assert np.array_equal(out[0], out[1])
