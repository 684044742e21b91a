"""Serving with continuous batching: requests (linear C3D4 with 1-4 load
cases, NLGEOM C3D4, C3D20 blocks) arrive while others are on the GPU.

Bars: every request's result is BIT FOR BIT the result of the same request
served alone (a slot's arithmetic does not depend on its neighbours; NLGEOM
requests run the deterministic step driver), and within 1e-8 relative of
the same stream served one request at a time with plain FP64 PCG.
"""
import numpy as np
import pytest

from ferrix_amd import meshgen
from ferrix_amd import _lib as L
from ferrix_amd.serving import Request, serve_batched, serve_serial

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def stream():
    specs = meshgen.cfg4_batch_specs(n_meshes=5, seed=17, dof_lo=500, dof_hi=8000)
    reqs = []
    t = 0.0
    for k, sp in enumerate(specs):
        reqs.append(Request(len(reqs), t, "linear", meshgen.cfg4_mesh(sp)))
        t += 0.004
        if k % 2 == 0:
            reqs.append(Request(len(reqs), t, "nlgeom", [meshgen.c3d4_box_nlgeom(sp)]))
            t += 0.004
    reqs.append(Request(len(reqs), t, "c3d20", [meshgen.cfg2_c3d20_block(seed=3, nx=8, ny=2, nz=2)]))
    return reqs


@pytest.mark.parametrize("precision", [L.FX_PREC_MIXED, L.FX_PREC_FP64])
def test_serving_alone_vs_mix(ctx, precision):
    reqs = stream()
    say("requests", [(r.kind, r.n_dof, r.n_rhs) for r in reqs])
    mix = serve_batched(ctx, reqs, precision, chunk_iters=8)
    assert not mix.failures, mix.failures
    ser = serve_serial(ctx, reqs)
    assert not ser.failures, ser.failures
    for r in reqs:
        alone = serve_batched(ctx, [Request(r.rid, 0.0, r.kind, r.jobs)], precision, chunk_iters=8)
        assert np.array_equal(alone.x[r.rid], mix.x[r.rid]), f"request {r.rid} ({r.kind}) differs alone vs mix"
        e = np.linalg.norm(mix.x[r.rid] - ser.x[r.rid], axis=1) / np.linalg.norm(ser.x[r.rid], axis=1)
        say("request", r.rid, r.kind, "solves", mix.solves[r.rid], "rel err vs serial fp64", e.max(),
            "latency ms", 1e3 * mix.latency[r.rid])
        assert e.max() < 1e-8
    assert mix.dof_solves(reqs) > 0
