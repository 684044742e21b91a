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


def test_cfg2_full_size_pattern_and_max_diag(ctx):
    meta, arr = golden("cfg2")
    job = meshgen.cfg2_c3d20_block()
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
def test_cfg2_full_size_step(ctx, precision):
    """One linear step: Jacobi-PCG to ||r||/||b|| <= 1e-10 (the
    configuration's solver setting) against the oracle's direct solve."""
    meta, arr = golden("cfg2")
    job = meshgen.cfg2_c3d20_block()
    u, reps = run_job(ctx, job, precision=precision)
    e = rel(u, arr["u"])
    say("precision", precision, "pcg", reps[0].pcg_iterations, "rel err", e)
    assert reps[0].linear_solves == meta["linear_solves"]
    assert e < 1e-8
