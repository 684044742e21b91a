"""Nodal stress/strain recovery (static_step.rs:271-379,
element.rs:526-564) on the GPU against the oracle: BIT FOR BIT (the same
fixed summation order, B's and D's structural zeros skipped exactly)."""
import numpy as np
import pytest

import oracle as O
from ferrix_amd import meshgen
from ferrix_amd.device import DeviceMesh, run_job

pytestmark = pytest.mark.gpu


def check(ctx, job, u, T=None):
    T = job.node_temp if T is None else T
    dm = DeviceMesh(ctx, job.mesh, T)
    dm.set_u(u)
    S, E = dm.stress_strain()
    So, Eo = O.stress_strain(job.mesh, T, u)
    assert np.array_equal(E, Eo), f"strain differs, max {np.max(np.abs(E - Eo))}"
    assert np.array_equal(S, So), f"stress differs, max {np.max(np.abs(S - So))}"
    return S, E


@pytest.mark.parametrize("name,make", [
    ("cantilever_c3d4", meshgen.c3d4_cantilever),
    ("c3d20_jitter", lambda: meshgen.c3d20_block(seed=7, nx=6, ny=3, nz=2)),
    ("c3d20_axis_aligned", lambda: meshgen.c3d20_block(seed=7, nx=4, ny=2, nz=2, jitter=False)),
])
def test_recovery_bitwise(ctx, name, make):
    job = make()
    u = np.random.default_rng(5).normal(scale=1e-2, size=job.n_dof)
    check(ctx, job, u)


def test_recovery_temperature_dependent(ctx):
    job = meshgen.c3d20_block(seed=6, nx=3, ny=2, nz=2)
    m = job.mesh
    m.lut_n_e[0] = 2
    m.lut_t_e[0, :2] = [0.0, 100.0]
    m.lut_v_e[0, :2] = [70000.0, 35000.0]
    rng = np.random.default_rng(3)
    T = rng.uniform(0.0, 100.0, size=len(m.xyz))
    check(ctx, job, rng.normal(scale=1e-3, size=job.n_dof), T)


def test_recovery_after_step(ctx):
    """Stress of the solved cantilever: bitwise the oracle's recovery of the
    GPU displacement, and the bending stress has the expected sign pattern
    (tension on top near the clamp under a downward tip load)."""
    job = meshgen.c3d4_cantilever()
    u, _ = run_job(ctx, job)
    S, E = check(ctx, job, u)
    X = job.mesh.xyz.reshape(-1, 3)
    root = X[:, 0] < 20.0
    top, bottom = root & (X[:, 2] > 15.0), root & (X[:, 2] < 5.0)
    assert S[top, 0].mean() > 0.0 > S[bottom, 0].mean()


def test_reaction_forces_balance_the_load(ctx):
    """RF = f_int(u) at the clamped DOFs, bitwise the oracle's f_int of the
    same u, and in equilibrium with the applied load: the rigid translations
    are in K's null space, so sum f_int = 0 per direction and the reactions
    carry exactly minus the applied total (penalty BCs, converged step)."""
    job = meshgen.c3d4_cantilever()
    st = job.steps[0]
    dm = DeviceMesh(ctx, job.mesh, job.node_temp)
    u, _ = run_job(ctx, job, mesh=dm)
    bc = np.unique(st.bc_dof)
    rf = dm.reaction_forces(bc)
    fo = O.internal_force(job.mesh, job.node_temp, u)
    assert np.array_equal(rf, fo[bc])
    applied = np.zeros(3)
    np.add.at(applied, st.load_dof % 3, st.load_val)
    total = np.array([rf[bc % 3 == d].sum() for d in range(3)])
    assert np.allclose(total, -applied, rtol=1e-6, atol=1e-6 * np.abs(applied).max())
