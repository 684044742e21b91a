"""C3D20 element K on FP64 MFMA (k_c3d20_mfma, FX_C3D20_MFMA=1): the
product G^T H over the 162 strain rows in the matrix core's order.

Bars (its summation order is not the oracle's, so no bitwise bar):
* one-element meshes (jittered shapes, reference and NLGEOM coordinates):
  the assembled matrix is the element K; max |K_gpu - K_oracle| <= 1e-13
  max |K_oracle|, and K_gpu exactly symmetric;
* a jittered 10x4x3 block: every assembled row within 1e-13 of its largest
  entry, max_diag within 1e-13, and the linear step's U within 1e-8 of the
  oracle's;
* the NLGEOM beam (K rebuilt every Newton iteration): the oracle's number of
  linear solves and U within 1e-7, as the VALU kernel's test asks.

The element autotuner (fx_mesh_autotune_elements) is tested here too: the
C3D4 workgroup size changes no bit of K, and the C3D20 kernel it keeps is
the VALU one unless MFMA is allowed.
"""
import numpy as np
import pytest

import oracle as O
from ferrix_amd import meshgen
from ferrix_amd.device import DeviceMesh, run_job

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture
def mfma(monkeypatch):
    monkeypatch.setenv("FX_C3D20_MFMA", "1")


def dense(rp, col, val, n):
    K = np.zeros((n, n))
    for i in range(n):
        K[i, col[rp[i]:rp[i + 1]]] = val[rp[i]:rp[i + 1]]
    return K


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("nlgeom", [False, True])
def test_single_element_k(ctx, mfma, seed, nlgeom):
    job = meshgen.c3d20_block(seed=seed + 11, nx=1, ny=1, nz=1, jitter=True)
    n = job.n_dof
    u = np.random.default_rng(seed).normal(scale=2.0, size=n) if nlgeom else None
    dm = DeviceMesh(ctx, job.mesh, job.node_temp)
    if u is not None:
        dm.set_u(u)
    dm.assemble(nlgeom)
    Kg = dense(*dm.matrix(), n)
    csr, _, _ = O.assemble_csr(job.mesh, job.node_temp, u)
    Ko = dense(csr.row_ptr, csr.col, csr.val, n)
    err = np.max(np.abs(Kg - Ko)) / np.max(np.abs(Ko))
    say("seed", seed, "nlgeom", nlgeom, "rel err", err)
    assert err <= 1e-13
    assert np.array_equal(Kg, Kg.T)


def test_block_assembly_and_linear_step(ctx, mfma):
    job = meshgen.c3d20_block(seed=7, nx=10, ny=4, nz=3, jitter=True)
    n = job.n_dof
    dm = DeviceMesh(ctx, job.mesh, job.node_temp)
    md, nnz = dm.assemble(False)
    csr, md_o, _ = O.assemble_csr(job.mesh, job.node_temp)
    say("nnz", nnz, csr.nnz, "max_diag", md, md_o)
    assert abs(md - md_o) <= 1e-13 * md_o
    Kg = dense(*dm.matrix(), n)
    Ko = dense(csr.row_ptr, csr.col, csr.val, n)
    row_err = np.max(np.abs(Kg - Ko), axis=1) / np.max(np.abs(Ko), axis=1)
    say("worst row rel err", row_err.max())
    assert row_err.max() <= 1e-13
    u, reps = run_job(ctx, job, mesh=dm)
    u_o, _ = O.run_job(job)
    e = np.linalg.norm(u - u_o) / np.linalg.norm(u_o)
    say("step U rel err", e)
    assert e < 1e-8


def test_nlgeom_beam(ctx, mfma):
    job = meshgen.c3d20_nlgeom_beam(nx=8, ny=2, nz=2, tip_load=20000.0)
    u_o, logs = O.run_job(job)
    u, reps = run_job(ctx, job)
    e = np.linalg.norm(u - u_o) / np.linalg.norm(u_o)
    say("solves", reps[0].linear_solves, logs[0].linear_solves, "U rel err", e)
    assert reps[0].linear_solves == logs[0].linear_solves
    assert e < 1e-7


@pytest.mark.parametrize("block", [64, 256, 512])
def test_c3d4_workgroup_sizes_bitwise(ctx, monkeypatch, block):
    """k_c3d4 at 64 / 256 / 512 elements per workgroup (FX_C3D4_BLOCK at
    upload) assembles the default 128's K bit for bit; the element autotuner
    picks one of the four and leaves K unchanged."""
    job = meshgen.population_mesh(meshgen.population_specs(n_meshes=4, seed=9, dof_lo=2e3, dof_hi=2e4)[1])[0]
    ref = DeviceMesh(ctx, job.mesh, job.node_temp)
    md0, _ = ref.assemble(False)
    K0 = ref.matrix()
    monkeypatch.setenv("FX_C3D4_BLOCK", str(block))
    dm = DeviceMesh(ctx, job.mesh, job.node_temp)
    md1, _ = dm.assemble(False)
    assert md1 == md0
    for a, b in zip(dm.matrix(), K0):
        assert np.array_equal(a, b)
    bs, c20, ms = ref.autotune_elements(reps=3)
    say("block", block, "tuned", bs, c20, f"{ms * 1e3:.1f} us")
    assert bs in (64, 128, 256, 512) and c20 == -1 and ms > 0
    ref.assemble(False)
    for a, b in zip(ref.matrix(), K0):
        assert np.array_equal(a, b)


def test_element_autotune_c3d20(ctx):
    """C3D20: without MFMA allowed the tuner keeps the VALU kernel (K bit for
    bit); allowed, it keeps the faster one (K within 1e-13 of the VALU's)."""
    job = meshgen.c3d20_block(seed=5, nx=10, ny=4, nz=3, jitter=True)
    n = job.n_dof
    dm = DeviceMesh(ctx, job.mesh, job.node_temp)
    dm.assemble(False)
    K0 = dense(*dm.matrix(), n)
    bs, c20, ms = dm.autotune_elements(reps=3, allow_mfma=False)
    assert c20 == 0
    dm.assemble(False)
    assert np.array_equal(dense(*dm.matrix(), n), K0)
    bs, c20, ms = dm.autotune_elements(reps=3, allow_mfma=True)
    say("C3D20 on MFMA:", c20, f"{ms * 1e3:.1f} us per launch")
    assert c20 in (0, 1)
    dm.assemble(False)
    K1 = dense(*dm.matrix(), n)
    assert np.max(np.abs(K1 - K0)) <= 1e-13 * np.max(np.abs(K0))


def test_element_tuning_cache():
    """A second mesh of a timed bucket (the same C3D4 mesh again, on a fresh
    context) takes the stored workgroup size; its K is bit for bit the
    first mesh's."""
    from ferrix_amd.device import Context
    import time
    job = meshgen.population_mesh(meshgen.population_specs(n_meshes=4, seed=9, dof_lo=2e4, dof_hi=4e4)[2])[0]
    c = Context(0)
    try:
        d1 = DeviceMesh(c, job.mesh, job.node_temp)
        t = time.perf_counter()
        b1, _, _ = d1.autotune_elements(reps=10)
        t1 = time.perf_counter() - t
        d2 = DeviceMesh(c, job.mesh, job.node_temp)
        t = time.perf_counter()
        b2, _, _ = d2.autotune_elements(reps=10)
        t2 = time.perf_counter() - t
        say("tuned", b1, f"{t1 * 1e3:.1f} ms; cached", b2, f"{t2 * 1e3:.1f} ms")
        assert b1 == b2
        d1.assemble(False)
        d2.assemble(False)
        for a, b in zip(d1.matrix(), d2.matrix()):
            assert np.array_equal(a, b)
        d1.close()
        d2.close()
    finally:
        c.close()
