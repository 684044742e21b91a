"""GPU parity on the benchmark configurations at their full sizes, against
golden fixtures the oracle wrote (tests/golden/make_golden.py):

* the linear C3D20 block: 40x10x10 elements, 18,821 nodes, 56,463 DOF,
  coordinates jittered (seed 2) -- the filtered CSR pattern (nnz and its
  FNV-1a fingerprint), max_diag bit for bit, and the step's displacement
  within 1e-8 of the oracle's direct solve, FP64 and mixed;
* the NLGEOM C3D20 beam: 80x4x4 elements, 7,265 nodes, 21,795 DOF, tip load
  ramped over ten increments (static_step.rs:71-258) -- at 110 kN the
  reference's Picard iteration stops with "Minimum time increment reached":
  the GPU step driver must stop the same way after the same increments,
  cutbacks and linear solves, leaving the displacement of the same last
  converged increment; at 20 kN the step completes and the displacement
  matches the oracle's.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
from ferrix_amd import _lib as L
from ferrix_amd import meshgen
from ferrix_amd.device import DeviceMesh, SolveOptions, run_job
from ferrix_amd.inp import ConvergenceError

from conftest import ROOT

pytestmark = pytest.mark.gpu
GOLD = os.path.join(ROOT, "tests", "golden")


def say(*a):
    print(*a, flush=True)


def golden(name):
    meta = json.load(open(os.path.join(GOLD, f"{name}_oracle.json")))
    arrays = dict(np.load(os.path.join(GOLD, f"{name}_oracle.npz")))
    return meta, arrays


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_block_full_size_pattern_and_max_diag(ctx):
    meta, arr = golden("block")
    job = meshgen.c3d20_block()
    assert (job.mesh.n_nodes, job.mesh.n_el, job.n_dof) == (meta["n_nodes"], meta["n_el"], meta["n_dof"])
    dm = DeviceMesh(ctx, job.mesh, job.node_temp)
    md, nnz = dm.assemble(False)
    say("nnz", nnz, "max_diag", md)
    assert nnz == meta["nnz"] == 8_809_029
    assert md == meta["max_diag"]
    rp, col, _ = dm.matrix()
    assert int(rp.sum()) == int(arr["row_ptr_sum"])
    assert int(col.astype(np.int64).sum()) == int(arr["col_sum"])
    assert O.pattern_fnv1a(rp, col) == meta["pattern_fnv1a"]


@pytest.mark.parametrize("precision", [L.FX_PREC_FP64, L.FX_PREC_MIXED])
def test_block_full_size_step(ctx, precision):
    """One linear step: Jacobi-PCG to ||r||/||b|| <= 1e-10 (the
    configuration's solver setting) against the oracle's direct solve."""
    meta, arr = golden("block")
    job = meshgen.c3d20_block()
    u, reps = run_job(ctx, job, precision=precision)
    e = rel(u, arr["u"])
    say("precision", precision, "pcg", reps[0].pcg_iterations, "rel err", e)
    assert reps[0].linear_solves == meta["linear_solves"]
    assert e < 1e-8


def nlgeom_run(ctx, job, precision):
    """The NLGEOM step through fx_static_step with a hook recording every
    converged increment (inc_id, global time); returns (u, report, error,
    converged increments)."""
    dm = DeviceMesh(ctx, job.mesh, job.node_temp)
    dm.set_u(np.zeros(job.n_dof))
    timer = L.fx_timer(0.0, 0.0, 1.0)
    st = job.steps[0]
    so = SolveOptions(precision=precision, stop_mode=st.stop_mode, tol=st.tol, max_iter=st.max_iter)
    seen = []
    err = None
    try:
        rep = dm.static_step(st, 1, timer, direct_opts=so, solve_opts=so,
                             on_increment=lambda sid, inc, t: seen.append((inc, t)))
    except ConvergenceError as e:
        err, rep = e, e.report
    return dm.get_u(), rep, err, seen


@pytest.mark.parametrize("precision", [L.FX_PREC_FP64, L.FX_PREC_MIXED])
def test_beam_full_size_stops_like_the_oracle(ctx, precision):
    """80x4x4 C3D20 beam, 110 kN: the oracle's Picard iteration converges six
    increments (t = 0.1 ... 0.6), then fails 15 NR iterations at every
    halving of dt until dt < 1e-5 and stops with "Minimum time increment
    reached" (static_step.rs:251-258).  The GPU step stops with the same
    error after the same increments, cutbacks, NR iterations and linear
    solves, and leaves the displacement of the sixth increment."""
    meta, arr = golden("beam")
    assert meta["outcome"].startswith("Minimum time increment reached")
    job = meshgen.c3d20_nlgeom_beam()
    assert job.n_dof == meta["n_dof"] == 21_795
    u, rep, err, seen = nlgeom_run(ctx, job, precision)
    incs = meta["increments"]
    nr = sum(len(r) for r in meta["increment_rel_residuals"])
    conv = [k + 1 for k, i in enumerate(incs) if i[1]]
    say("precision", precision, "increments", rep.increments, "cutbacks", rep.cutbacks, "nr", rep.nr_iterations,
        "solves", rep.linear_solves, "pcg", rep.pcg_iterations, "converged", seen)
    assert err is not None and "Minimum time increment reached" in str(err)
    assert rep.increments == len(incs)
    assert rep.cutbacks == sum(1 for i in incs if not i[1])
    assert rep.nr_iterations == nr
    assert rep.linear_solves == meta["linear_solves"]
    assert [inc for inc, _ in seen] == conv
    assert np.allclose([t for _, t in seen], np.cumsum([incs[k - 1][2] for k in conv]), rtol=0, atol=1e-12)
    e = rel(u, arr["u_last_converged"])
    say("u after the error vs the oracle's last converged increment:", e)
    assert e < 1e-8


@pytest.mark.parametrize("precision", [L.FX_PREC_FP64, L.FX_PREC_MIXED])
def test_beam_converging_load(ctx, precision):
    """The same 80x4x4 beam at a tip load the step completes: the same NR
    history as the oracle and the displacement within 1e-8 of its."""
    meta, arr = golden("beam_converging")
    assert meta["outcome"] == "ok"
    job = meshgen.c3d20_nlgeom_beam(tip_load=meta["tip_load"])
    u, rep, err, seen = nlgeom_run(ctx, job, precision)
    assert err is None
    incs = meta["increments"]
    say("precision", precision, "increments", rep.increments, "nr", rep.nr_iterations, "solves", rep.linear_solves,
        "pcg", rep.pcg_iterations, "tip w", u.reshape(-1, 3)[:, 2].min())
    assert rep.increments == len(incs)
    assert rep.linear_solves == meta["linear_solves"]
    assert rep.nr_iterations == sum(len(r) for r in meta["increment_rel_residuals"])
    e = rel(u, arr["u"])
    say("u vs the oracle:", e)
    assert e < 1e-8
