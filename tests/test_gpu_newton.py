"""The device-resident Newton iteration of fx_static_step (fx_newton.hip)
against the host-driven loop (FX_NR_DEVICE=0), static_step.rs:99-210.

Bars:
* displacements bit for bit the host loop's, with the same increments,
  cutbacks, NR iterations, linear solves, PCG iterations and final
  residual ratio, FP64 and mixed precision, linear and NLGEOM steps, the
  front replayed from a captured graph and as plain launches;
* fewer host waits: in FP64 one stream synchronisation per NR iteration
  when the queued iterations cover the solve (the step's first solve and
  the first solve of each increment may need the chunked finish);
* a failing step (the 80x4x4 beam at 110 kN) stops with the host loop's
  error after the same history.
"""
import os
import time

import numpy as np
import pytest

from ferrix_amd import _lib as L
from ferrix_amd import meshgen
from ferrix_amd.device import run_job
from ferrix_amd.inp import ConvergenceError

pytestmark = pytest.mark.gpu

FIELDS = ("increments", "cutbacks", "nr_iterations", "linear_solves", "pcg_iterations", "outer_spmv",
          "last_rel_residual")


def say(*a):
    print(*a, flush=True)


def run(ctx, job, precision, env):
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        t = time.perf_counter()
        err = None
        try:
            u, reps = run_job(ctx, job, precision=precision)
        except ConvergenceError as e:
            err, u, reps = e, None, [e.report]
        return u, reps, err, time.perf_counter() - t
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def compare(ctx, job, precision, label):
    u_h, r_h, e_h, t_h = run(ctx, job, precision, {"FX_NR_DEVICE": "0"})
    u_d, r_d, e_d, t_d = run(ctx, job, precision, {"FX_NR_DEVICE": "1", "FX_NR_GRAPH": "1"})
    u_p, r_p, e_p, t_p = run(ctx, job, precision, {"FX_NR_DEVICE": "1", "FX_NR_GRAPH": "0"})
    for k, (a, b, c) in enumerate(zip(r_h, r_d, r_p)):
        say(f"{label} step {k + 1}: host loop {t_h:.3f} s, {a.host_syncs} waits; device graph {t_d:.3f} s, "
            f"{b.host_syncs} waits; device plain {t_p:.3f} s, {c.host_syncs} waits; nr {a.nr_iterations} "
            f"solves {a.linear_solves} pcg {a.pcg_iterations}")
        for f in FIELDS:
            assert getattr(a, f) == getattr(b, f) == getattr(c, f), (label, k, f)
        assert b.host_syncs < a.host_syncs and c.host_syncs < a.host_syncs
    assert (e_h is None) == (e_d is None) == (e_p is None)
    if e_h is None:
        assert np.array_equal(u_h, u_d) and np.array_equal(u_h, u_p), label
    else:
        assert str(e_h) == str(e_d) == str(e_p)
    return r_h, r_d


@pytest.mark.parametrize("precision", [L.FX_PREC_FP64, L.FX_PREC_MIXED])
def test_linear_step(ctx, precision):
    job = meshgen.c3d4_cantilever()
    r_h, r_d = compare(ctx, job, precision, f"cantilever p{precision}")
    if precision == L.FX_PREC_FP64:
        # iteration 1: front + queued solve (+ its chunked finish); iteration 2: the converged check
        assert r_d[0].host_syncs <= r_d[0].nr_iterations + r_d[0].linear_solves * 12


@pytest.mark.parametrize("precision", [L.FX_PREC_FP64, L.FX_PREC_MIXED])
def test_nlgeom_beam(ctx, precision):
    job = meshgen.c3d20_nlgeom_beam(nx=8, ny=2, nz=2, tip_load=20000.0)
    r_h, r_d = compare(ctx, job, precision, f"beam 8x2x2 p{precision}")
    if precision == L.FX_PREC_FP64:
        assert r_d[0].host_syncs < 5 * r_d[0].nr_iterations


def test_c3d20_linear_block(ctx):
    job = meshgen.c3d20_block(seed=3, nx=10, ny=4, nz=3)
    compare(ctx, job, L.FX_PREC_FP64, "c3d20 10x4x3")


def test_failing_step_full_beam(ctx):
    """80x4x4 C3D20 beam at 110 kN: 'Minimum time increment reached' after
    the same history from both loops (FP64)."""
    job = meshgen.c3d20_nlgeom_beam()
    r_h, r_d = compare(ctx, job, L.FX_PREC_FP64, "beam 80x4x4 110 kN")
    assert r_h[0].cutbacks > 0


@pytest.mark.parametrize("precision", [L.FX_PREC_FP64, L.FX_PREC_MIXED])
def test_converging_full_beam(ctx, precision):
    """The 80x4x4 beam at the tip load its step completes (the beam_converging
    golden's): both loops, timed."""
    import json
    meta = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "beam_converging_oracle.json")))
    job = meshgen.c3d20_nlgeom_beam(tip_load=meta["tip_load"])
    compare(ctx, job, precision, f"beam 80x4x4 {meta['tip_load']:.0f} N p{precision}")
