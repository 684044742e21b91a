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
    specs = meshgen.population_specs(n_meshes=5, seed=17, dof_lo=500, dof_hi=8000)
    reqs = []
    t = 0.0
    for k, sp in enumerate(specs):
        reqs.append(Request(len(reqs), t, "linear", meshgen.population_mesh(sp)))
        t += 0.004
        if k % 2 == 0:
            reqs.append(Request(len(reqs), t, "nlgeom", [meshgen.c3d4_box_nlgeom(sp)]))
            t += 0.004
    reqs.append(Request(len(reqs), t, "c3d20", [meshgen.c3d20_block(seed=3, nx=8, ny=2, nz=2)]))
    return reqs


@pytest.mark.parametrize("precision", [L.FX_PREC_MIXED, L.FX_PREC_FP64, L.FX_PREC_MIXED_POLY])
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
    n_nl = sum(1 for r in reqs if r.kind == "nlgeom")
    assert mix.batched_nl_solves >= n_nl  # every NLGEOM request solved its Newton systems in the batch


def test_serving_nlgeom_own_solver_vs_batch(ctx):
    """NLGEOM requests whose Newton solves join the batch against the same
    requests solving with the step's own FP64 PCG on the worker's stream:
    the same number of linear solves, U within 1e-8."""
    specs = meshgen.population_specs(n_meshes=3, seed=29, dof_lo=500, dof_hi=6000)
    reqs = [Request(k, 0.0, "nlgeom", [meshgen.c3d4_box_nlgeom(sp)]) for k, sp in enumerate(specs)]
    inb = serve_batched(ctx, reqs, L.FX_PREC_FP64, chunk_iters=8)
    own = serve_batched(ctx, reqs, L.FX_PREC_FP64, chunk_iters=8, nl_in_batch=False)
    assert not inb.failures and not own.failures
    assert inb.batched_nl_solves == sum(inb.solves.values()) and own.batched_nl_solves == 0
    for r in reqs:
        e = np.linalg.norm(inb.x[r.rid] - own.x[r.rid]) / np.linalg.norm(own.x[r.rid])
        say("request", r.rid, "n_dof", r.n_dof, "solves", inb.solves[r.rid], own.solves[r.rid], "rel diff", e)
        assert inb.solves[r.rid] == own.solves[r.rid]
        assert e < 1e-8


def test_serving_admits_by_free_memory(ctx):
    """Admission against the device's free memory: a reserve that leaves
    room for about two requests' estimates holds the number admitted at
    once down (the estimates are conservative, so a few small requests may
    share the room), the memory in use at any admission decision stays
    within the room plus one request, and every result still matches the
    unlimited server's."""
    import gc
    gc.collect()  # earlier tests' device objects are released before the measurement
    reqs = stream()
    free, total, pool = ctx.memory()
    assert 0 < free <= total
    big = max(r.est_bytes() for r in reqs)
    room = 2 * big
    res = serve_batched(ctx, reqs, L.FX_PREC_FP64, chunk_iters=8, mem_reserve=free + pool - room)
    unlimited = serve_batched(ctx, reqs, L.FX_PREC_FP64, chunk_iters=8)
    say("free", free, "total", total, "admitted at once", res.max_admitted, "unlimited", unlimited.max_admitted,
        "least free seen", res.min_free, "room", room)
    assert not res.failures
    assert res.max_admitted < unlimited.max_admitted
    assert free + pool - res.min_free <= room + big
    for r in reqs:
        assert np.array_equal(res.x[r.rid], unlimited.x[r.rid]), r.rid


# calculix_parity.rs known answers: node 17's displacement after the last step
DECK_ANSWERS = {
    "C3D4_staticstep": (9.356107e-06, -1.313901e-04, 1.719218e-03),
    "C3D20": (5.752112e-02, -2.497816e-04, -8.074951e-03),
    "C3D4_2staticstep": (2.603480e-04, -1.377070e-04, 1.001299e-03),
}


def test_serving_deck_requests(ctx):
    """Requests that arrive as .inp decks (the reference's own test decks),
    interleaved with generated ones: a one-step linear deck joins the batch,
    a two-step deck runs through the step driver.  Each deck's node 17 meets
    CalculiX's answer at the reference's 1e-6, and every request matches the
    stream served one at a time in plain FP64 to 1e-8."""
    from conftest import deck
    from ferrix_amd.serving import request_from_deck

    gen = stream()[:3]
    reqs = []
    t = 0.0
    for k, name in enumerate(DECK_ANSWERS):
        reqs.append(request_from_deck(len(reqs), t, deck(name), default_initial_temperature=20.0))
        t += 0.003
        r = gen[k]
        reqs.append(Request(len(reqs), t, r.kind, r.jobs))
        t += 0.003
    say("requests", [(r.kind, r.n_dof, r.jobs[0].name) for r in reqs])
    assert [r.kind for r in reqs[0::2]] == ["linear", "c3d20", "nlgeom"]
    mix = serve_batched(ctx, reqs, L.FX_PREC_FP64, chunk_iters=8)
    assert not mix.failures, mix.failures
    ser = serve_serial(ctx, reqs)
    assert not ser.failures, ser.failures
    for r in reqs:
        e = np.linalg.norm(mix.x[r.rid] - ser.x[r.rid], axis=1) / np.linalg.norm(ser.x[r.rid], axis=1)
        say("request", r.rid, r.kind, r.jobs[0].name, "rel err vs serial fp64", e.max())
        assert e.max() < 1e-8
    from ferrix_amd.inp import Project
    for r in reqs[0::2]:
        name = r.jobs[0].name
        idx = Project.from_jobname(deck(name)).get_index_for_node_id(17)
        got = mix.x[r.rid][0][3 * idx:3 * idx + 3]
        say(name, "node 17", got, DECK_ANSWERS[name])
        assert np.all(np.abs(got - np.array(DECK_ANSWERS[name])) < 1e-6)
