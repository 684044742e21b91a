"""Pins the CPU oracle (oracle/fx_oracle.c) before it is trusted as the checker.

Every function of the oracle is checked against something other than its
own formula: the reference's unit known answers (c3d20.rs:243-265,
material.rs:283-319, amplitude.rs:108-141), closed forms (Lame D, Gauss
exactness, C3D4 K = V B^T D B with independently derived gradients),
invariants (symmetry, rigid-body null space and rank, patch test,
f_int = K u), brute force (dense numpy solve against PCG), scipy's
duplicate summation for the CSR conversion, and the CalculiX values the
reference's integration tests use (calculix_parity.rs) run end to end.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
from ferrix_amd import meshgen
from ferrix_amd.inp import Project
from ferrix_amd.model import Job, SOLVER_DIRECT

from conftest import DECKS, ROOT, deck

C3D20_LOCAL = np.array([
    [-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1],
    [0, -1, -1], [1, 0, -1], [0, 1, -1], [-1, 0, -1], [0, -1, 1], [1, 0, 1], [0, 1, 1], [-1, 0, 1],
    [-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], dtype=float)  # c3d20.rs:61-84


# ---------------------------------------------------------------- shapes
def test_c3d20_partition_of_unity_reference_points():
    # c3d20.rs:243-252
    for p in [(0.0, 0.0, 0.0), (0.5, -0.3, 0.8)]:
        N, _ = O.shape_c3d20(*p)
        assert abs(N.sum() - 1.0) < 1e-12


def test_c3d20_kronecker_all_nodes():
    # c3d20.rs:254-265 checks nodes 1 and 9; extended to all 20
    for i, p in enumerate(C3D20_LOCAL):
        N, _ = O.shape_c3d20(*p)
        e = np.zeros(20)
        e[i] = 1.0
        assert np.max(np.abs(N - e)) < 1e-12, i


def test_c3d20_derivatives_finite_difference_and_sum():
    rng = np.random.default_rng(0)
    for _ in range(20):
        p = rng.uniform(-1, 1, 3)
        N, dN = O.shape_c3d20(*p)
        assert np.max(np.abs(dN.sum(axis=1))) < 1e-12  # sum N = 1 -> sum dN = 0
        h = 1e-6
        for k in range(3):
            q1, q2 = p.copy(), p.copy()
            q1[k] += h
            q2[k] -= h
            fd = (O.shape_c3d20(*q1)[0] - O.shape_c3d20(*q2)[0]) / (2 * h)
            assert np.max(np.abs(fd - dN[k])) < 1e-8


def test_c3d20_reproduces_serendipity_space():
    """sum_i N_i f(x_i) = f(x) for f in the serendipity space (quadratics
    and the cubic terms x^2 y, x y z ... of the 20-node brick)."""
    rng = np.random.default_rng(1)
    funcs = [lambda x, y, z: 1 + 0 * x, lambda x, y, z: x, lambda x, y, z: y * z, lambda x, y, z: x * x,
             lambda x, y, z: x * y * z, lambda x, y, z: x * x * y, lambda x, y, z: z * z * x]
    for _ in range(10):
        p = rng.uniform(-1, 1, 3)
        N, _ = O.shape_c3d20(*p)
        for f in funcs:
            vals = np.array([f(*c) for c in C3D20_LOCAL])
            assert abs(N @ vals - f(*p)) < 1e-12


def test_c3d4_shape():
    N, dN = O.shape_c3d4(0.25, 0.25, 0.25)
    assert np.allclose(N, 0.25)
    assert np.array_equal(dN, [[-1, 1, 0, 0], [-1, 0, 1, 0], [-1, 0, 0, 1]])


def test_gauss_rules():
    c, w = O.gauss_points(4)
    assert c.shape == (1, 3) and np.allclose(c, 0.25) and w[0] == 1.0 / 6.0
    c, w = O.gauss_points(20)
    assert len(w) == 27
    assert abs(w.sum() - 8.0) < 1e-14
    # 3-point Gauss-Legendre is exact to degree 5 per axis
    val = np.sum(w * c[:, 0] ** 4 * c[:, 1] ** 4 * c[:, 2] ** 4)
    assert abs(val - (2.0 / 5.0) ** 3) < 1e-12
    # xi outermost, zeta innermost (c3d20.rs:40-50)
    assert c[0, 2] < c[1, 2] and c[0, 0] == c[8, 0] and c[9, 0] > c[8, 0]


# ---------------------------------------------------------------- material
def test_d_matrix_lame():
    E, nu = 210000.0, 0.3
    D = O.d_matrix(E, nu)
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    ref = np.zeros((6, 6))
    ref[:3, :3] = lam
    ref[np.arange(3), np.arange(3)] = lam + 2 * mu
    ref[np.arange(3, 6), np.arange(3, 6)] = mu
    assert np.allclose(D, ref, rtol=1e-14, atol=0)


def test_lut_reference_values():
    # material.rs:285-295
    t, v = [0.0, 100.0, 200.0], [100.0, 200.0, 400.0]
    for temp, exp in [(-10, 100), (0, 100), (50, 150), (100, 200), (150, 300), (200, 400), (250, 400)]:
        assert O.lut_interp(t, v, temp) == exp
    # material.rs:299-319
    assert O.lut_interp([0, 500], [210000.0, 150000.0], 250.0) == 180000.0
    assert abs(O.lut_interp([0, 500], [0.3, 0.35], 250.0) - 0.325) < 1e-12


# ---------------------------------------------------------------- elements
def _one_elem_mesh(nn):
    job = meshgen.c3d20_block(seed=11, nx=1, ny=1, nz=1) if nn == 20 else meshgen.c3d4_cantilever()
    m = job.mesh
    nodes = m.conn[m.el_ptr[0]:m.el_ptr[1]]
    return m, m.xyz[nodes]


def _rigid_modes(X):
    X = X - X.mean(axis=0)
    n = len(X)
    R = np.zeros((3 * n, 6))
    for i, (x, y, z) in enumerate(X):
        R[3 * i:3 * i + 3, :3] = np.eye(3)
        R[3 * i:3 * i + 3, 3:] = [[0, -z, y], [z, 0, -x], [-y, x, 0]]
    return R


@pytest.mark.parametrize("nn", [4, 20])
def test_element_symmetry_nullspace_rank(nn):
    m, X = _one_elem_mesh(nn)
    K, _ = O.element(m, X, np.zeros(nn))
    s = np.max(np.abs(K))
    assert np.max(np.abs(K - K.T)) < 1e-12 * s
    R = _rigid_modes(X)
    assert np.max(np.abs(K @ R)) < 1e-9 * s * np.max(np.abs(R))
    ev = np.linalg.eigvalsh((K + K.T) / 2)
    assert np.sum(ev > 1e-9 * ev.max()) == 3 * nn - 6
    assert ev.min() > -1e-9 * ev.max()


def test_c3d4_closed_form():
    """K = V B^T D B with barycentric gradients from the inverse of the
    4x4 coordinate matrix (independent of the J^-1 route)."""
    m, X = _one_elem_mesh(4)
    K, _ = O.element(m, X, np.zeros(4))
    A = np.hstack([np.ones((4, 1)), X])  # rows [1 x y z]
    Ainv = np.linalg.inv(A)
    grads = Ainv[1:, :]  # d lambda_i / d(x,y,z) as columns
    V = abs(np.linalg.det(A)) / 6.0
    B = np.zeros((6, 12))
    for i in range(4):
        dx, dy, dz = grads[:, i]
        B[0, 3 * i], B[1, 3 * i + 1], B[2, 3 * i + 2] = dx, dy, dz
        B[3, 3 * i], B[3, 3 * i + 1] = dy, dx
        B[4, 3 * i + 1], B[4, 3 * i + 2] = dz, dy
        B[5, 3 * i], B[5, 3 * i + 2] = dz, dx
    D = O.d_matrix(210000.0, 0.3)
    Kref = V * B.T @ D @ B
    assert np.allclose(K, Kref, rtol=1e-12, atol=1e-9 * np.abs(Kref).max())


@pytest.mark.parametrize("nn", [4, 20])
def test_internal_force_equals_k_u(nn):
    m, X = _one_elem_mesh(nn)
    u = np.random.default_rng(5).normal(size=3 * nn)
    K, f = O.element(m, X, np.zeros(nn), uel=u, want_f=True)
    assert np.allclose(f, K @ u, rtol=1e-11, atol=1e-11 * np.abs(K @ u).max())


def test_patch_test_c3d20_jittered():
    """u = eps x (linear field): internal nodes carry no force and the
    stress is D eps -> f_int vanishes away from the boundary."""
    job = meshgen.c3d20_block(seed=3, nx=3, ny=3, nz=3)
    m = job.mesh
    eps = np.array([[1e-3, 2e-4, -1e-4], [3e-4, -5e-4, 2e-4], [0.0, 1e-4, 7e-4]])
    u = (m.xyz @ eps.T).ravel()
    f = O.internal_force(m, job.node_temp, u)
    grid = meshgen.BlockGrid(3, 3, 3, 400.0 / 3, 100.0 / 3, 100.0 / 3, quadratic=True).coords()
    inner = np.all((grid > 1e-9) & (grid < np.array([400.0, 100.0, 100.0]) - 1e-9), axis=1)
    fi = f.reshape(-1, 3)[inner]
    assert inner.sum() > 10
    assert np.max(np.abs(fi)) < 1e-9 * np.max(np.abs(f))
    assert np.max(np.abs(f.reshape(-1, 3).sum(axis=0))) < 1e-9 * np.max(np.abs(f))


@pytest.mark.parametrize("nn", [4, 20])
def test_thermal_force(nn):
    """element.rs:453-520: exactly zero when the initial and current fields
    are equal (always so in the reference, state.rs:47-78); with a uniform
    rise dT and constant alpha the thermal strain is alpha*dT*I, i.e. the
    free-expansion field u = alpha*dT*x, so f_th = K (alpha dT X)."""
    m, X = _one_elem_mesh(nn)
    T = np.full(nn, 20.0)
    f0 = O.thermal_force(m, X, T, T, [0.0, 500.0], [1.2e-5, 1.5e-5], 20.0)
    assert np.all(f0 == 0.0)
    alpha, dT = 1.2e-5, 80.0
    f = O.thermal_force(m, X, T, T + dT, [0.0], [alpha], 20.0)
    K, _ = O.element(m, X, T)
    ref = K @ (alpha * dT * X).ravel()
    assert np.allclose(f, ref, rtol=1e-10, atol=1e-10 * np.abs(ref).max())


# ---------------------------------------------------------------- assembly
def test_triplets_to_csr_matches_scipy():
    import scipy.sparse as sp

    rng = np.random.default_rng(7)
    n, t = 50, 2000
    r = rng.integers(0, n, t).astype(np.int32)
    c = rng.integers(0, n, t).astype(np.int32)
    v = rng.integers(-100, 100, t).astype(np.float64)  # integer values: sums are exact
    csr = O.triplets_to_csr(n, r, c, v)
    ref = sp.coo_matrix((v, (r, c)), shape=(n, n)).tocsr()
    ref.sum_duplicates()
    ref.sort_indices()
    assert np.array_equal(csr.row_ptr, ref.indptr)
    assert np.array_equal(csr.col, ref.indices)
    assert np.array_equal(csr.val, ref.data)


def test_assembly_matches_dense_element_sum():
    """Global K equals the dense sum of element matrices scattered by DOF
    (drop rule only removes |v| <= 1e-12)."""
    job = meshgen.c3d20_block(seed=5, nx=2, ny=2, nz=1)
    m = job.mesh
    csr, md, _ = O.assemble_csr(m, job.node_temp)
    Kd = np.zeros((m.n_dof, m.n_dof))
    for e in range(m.n_el):
        nodes = m.conn[m.el_ptr[e]:m.el_ptr[e + 1]]
        Ke, _ = O.element(m, m.xyz[nodes], job.node_temp[nodes])
        dofs = (3 * nodes[:, None] + np.arange(3)).ravel()
        Kd[np.ix_(dofs, dofs)] += Ke
    assert np.allclose(csr.to_scipy().toarray(), Kd, rtol=1e-13, atol=1e-9)
    assert md == max(abs(v) for v in np.diag(Kd)) or md > 0


def test_max_diag_is_per_element_maximum():
    job = meshgen.c3d4_cantilever()
    m = job.mesh
    _, _, _, md = O.assemble_triplets(m, job.node_temp)
    best = 0.0
    for e in range(0, m.n_el, 7):
        nodes = m.conn[m.el_ptr[e]:m.el_ptr[e + 1]]
        Ke, _ = O.element(m, m.xyz[nodes], job.node_temp[nodes])
        best = max(best, np.max(np.abs(np.diag(Ke))))
    assert md >= best and md < 1.5 * best  # max over elements, not over the assembled diagonal


# ---------------------------------------------------------------- solvers
def test_pcg_against_dense_solve_cantilever():
    job = meshgen.c3d4_cantilever()
    st = job.steps[0]
    csr, md, pen = O.assemble_csr(job.mesh, job.node_temp, None, st.bc_dof)
    b = np.zeros(job.n_dof)
    np.add.at(b, st.load_dof, st.load_val)
    A = csr.to_scipy().toarray()
    xd = np.linalg.solve(A, b)
    x, it, crit, ok = O.pcg(csr, b, 1e-8, 1000, 0)
    assert ok and crit < 1e-8
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-8
    x2, it2, _, ok2 = O.pcg(csr, b, 1e-12, 100000, 1)
    assert ok2 and np.linalg.norm(x2 - xd) / np.linalg.norm(xd) < 1e-10
    xs = O.direct_solve(csr, b)
    assert np.linalg.norm(xs - xd) / np.linalg.norm(xd) < 1e-10  # two exact solvers, kappa ~ 1e6 penalty


def test_pcg_reports_not_converged():
    job = meshgen.c3d4_cantilever()
    st = job.steps[0]
    csr, _, _ = O.assemble_csr(job.mesh, job.node_temp, None, st.bc_dof)
    b = np.zeros(job.n_dof)
    np.add.at(b, st.load_dof, st.load_val)
    x, it, crit, ok = O.pcg(csr, b, 1e-8, 5, 0)
    assert not ok and it == 5


def test_pcg_par_is_the_same_algorithm():
    """The CPU baseline's PCG (parallel loops, per-iteration allocations as
    in iterative.rs) on one thread is bit for bit fxo_pcg (same sums in the
    same order); on several threads only the dot products' association
    changes; both stop rules and the not-converged outcome agree."""
    job = meshgen.c3d4_cantilever()
    st = job.steps[0]
    csr, _, _ = O.assemble_csr(job.mesh, job.node_temp, None, st.bc_dof)
    b = np.zeros(job.n_dof)
    np.add.at(b, st.load_dof, st.load_val)
    for mode, tol in ((0, 1e-8), (1, 1e-10)):
        x, it, crit, ok = O.pcg(csr, b, tol, 100000, mode, 1)
        x1, it1, crit1, ok1 = O.pcg_par(csr, b, tol, 100000, mode, 1)
        assert ok and ok1 and it == it1 and crit == crit1
        assert np.array_equal(x, x1)
        x4, it4, _, ok4 = O.pcg_par(csr, b, tol, 100000, mode, 4)
        assert ok4 and abs(it4 - it) <= 2
        assert np.linalg.norm(x4 - x) / np.linalg.norm(x) < 1e-9
    _, it5, _, ok5 = O.pcg_par(csr, b, 1e-8, 5, 0, 2)
    assert not ok5 and it5 == 5
    sec, its = O.pcg_par_time(csr, b, 1e-10, 1, 100000, 2)
    assert sec > 0 and abs(its - it) <= 2
    sec, its = O.pcg_par_time(csr, b, 1e-10, 1, 3, 2)
    assert sec < 0 and its == 3  # not converged within the limit
    sec, its = O.pcg_par_time(csr, b, 0.0, 0, 7, 2)
    assert sec > 0 and its == 7  # tol 0: a fixed iteration count


def test_jacobi_guard():
    csr = O.CSR(3, np.array([0, 1, 2, 2], np.int64), np.array([0, 1], np.int32), np.array([4.0, 1e-12]))
    assert np.array_equal(O.jacobi(csr), [0.25, 1.0, 1.0])


# ---------------------------------------------------------------- drivers / KAT
def test_amplitude_reference_cases():
    # amplitude.rs:108-141
    t, v = [0.0, 10.0, 20.0], [0.0, 100.0, 200.0]
    assert abs(O._interp(10.0, t, v, 0.0, 0.0) - 100.0) < 1e-15
    assert abs(O._interp(5.0, t, v, 0.0, 0.0) - 50.0) < 1e-15
    assert abs(O._interp(15.0, t, v, 5.0, 20.0) - 120.0) < 1e-15
    assert O._interp(-10.0, t, v, 0.0, 0.0) == 0.0
    assert O._interp(30.0, t, v, 0.0, 0.0) == 200.0


@pytest.mark.parametrize("name,node,expected,tol", [
    ("C3D4_staticstep", 17, [(9.356107e-06, -1.313901e-04, 1.719218e-03)], 1e-6),
    ("C3D20", 17, [(5.752112e-02, -2.497816e-04, -8.074951e-03)], 1e-6),
    ("C3D4_2staticstep", 17, [(9.356107e-06, -1.313901e-04, 1.719218e-03),
                             (2.603480e-04, -1.377070e-04, 1.001299e-03)], 1e-6),
])
def test_oracle_reference_decks(name, node, expected, tol):
    """calculix_parity.rs:23-180, 274-333 through the oracle (DIRECT)."""
    p = Project.from_jobname(deck(name))
    p.default_initial_temperature = 20.0
    job = Job.from_project(p)
    u = np.zeros(job.n_dof)
    timer = O.SolverTime()
    idx = p.get_index_for_node_id(node)
    for k in range(len(job.steps)):
        u = O.static_step(job, k, u, timer)
        assert np.all(np.abs(u[3 * idx:3 * idx + 3] - np.array(expected[k])) < tol)


def test_linear_step_counts():
    """Linear step: one solve, converged at NR iteration 1 (Appendix-A
    behaviour of static_step.rs: f_int, K, solve, f_int)."""
    job = meshgen.c3d4_cantilever()
    u, logs = O.run_job(job)
    inc = logs[0].increments
    assert len(inc) == 1 and inc[0]["converged"] and inc[0]["iters"] == [0] and len(inc[0]["rel"]) == 2
    assert inc[0]["rel"][0] == 1.0 and inc[0]["rel"][1] < 1e-3


def test_micro_increment_quirk():
    """0.1 steps sum to 0.9999999999999999 < 1: an 11th increment of ~1e-16
    runs and converges at iteration 0 (static_step.rs:71-84, 213)."""
    job = meshgen.c3d20_nlgeom_beam(nx=4, ny=1, nz=1, tip_load=10.0)
    job.steps[0].solver = SOLVER_DIRECT
    u, logs = O.run_job(job)
    incs = logs[0].increments
    assert len(incs) == 11
    assert incs[-1]["dt"] < 1e-15 and incs[-1]["iters"] == []


@pytest.mark.parametrize("name", ["cantilever", "block"])
def test_golden_fixture_metadata(name):
    meta = json.load(open(os.path.join(ROOT, "tests", "golden", f"{name}_oracle.json")))
    gen = {"cantilever": meshgen.c3d4_cantilever, "block": meshgen.c3d20_block}[name]()
    assert (meta["n_nodes"], meta["n_el"], meta["n_dof"]) == (gen.mesh.n_nodes, gen.mesh.n_el, gen.n_dof)
    if name == "cantilever":
        csr, md, _ = O.assemble_csr(gen.mesh, gen.node_temp)
        assert csr.nnz == meta["nnz"] and md == meta["max_diag"]
        u, _ = O.run_job(gen, solver_override=SOLVER_DIRECT)
        g = np.load(os.path.join(ROOT, "tests", "golden", "cantilever_oracle.npz"))["u"]
        assert np.max(np.abs(u - g)) <= 1e-14 * np.max(np.abs(g))


def _voigt_field(eps):
    """Displacement gradient of a linear field with engineering strains eps."""
    return np.array([[eps[0], 0.5 * eps[3], 0.5 * eps[5]],
                     [0.5 * eps[3], eps[1], 0.5 * eps[4]],
                     [0.5 * eps[5], 0.5 * eps[4], eps[2]]])


@pytest.mark.parametrize("nn", [4, 20])
def test_stress_recovery_patch(nn):
    """A linear displacement field gives every node the constant strain and
    sigma = D eps (C3D4 and jittered C3D20 reproduce linear fields exactly)."""
    job = meshgen.c3d4_cantilever() if nn == 4 else meshgen.c3d20_block(seed=9, nx=3, ny=2, nz=2)
    eps = np.array([1e-3, -2e-4, 3e-4, 5e-4, -1e-4, 2e-4])
    X = job.mesh.xyz.reshape(-1, 3)
    u = (X @ _voigt_field(eps).T).ravel()
    S, E = O.stress_strain(job.mesh, job.node_temp, u)
    E_mod, nu = float(job.mesh.lut_v_e[0, 0]), float(job.mesh.lut_v_nu[0, 0])
    assert np.max(np.abs(E - eps)) < 1e-15
    assert np.max(np.abs(S - O.d_matrix(E_mod, nu) @ eps)) < 1e-9 * np.max(np.abs(O.d_matrix(E_mod, nu) @ eps))


def test_stress_recovery_single_element_node_by_differences():
    """A block corner belongs to one element: its recovered strain is the
    symmetric gradient of the element's isoparametric displacement at that
    corner, computed here by central differences of the shape functions N
    (not the derivative tables) through the geometric map."""
    job = meshgen.c3d20_block(seed=4, nx=2, ny=2, nz=2)
    rng = np.random.default_rng(8)
    u = rng.normal(scale=1e-3, size=job.n_dof)
    S, E = O.stress_strain(job.mesh, job.node_temp, u)
    m = job.mesh
    e = 0  # the element holding node index 0 (block corner i=j=k=0)
    nodes = m.conn[m.el_ptr[e]:m.el_ptr[e + 1]]
    assert nodes[0] == 0 and np.sum(m.conn == 0) == 1
    X = m.xyz.reshape(-1, 3)[nodes]
    U = u.reshape(-1, 3)[nodes]
    h = 1e-6
    dX, dU = np.zeros((3, 3)), np.zeros((3, 3))
    for k in range(3):
        xi = np.array([-1.0, -1.0, -1.0])
        a, b = xi.copy(), xi.copy()
        a[k] += h
        b[k] -= h
        Na, Nb = O.shape_c3d20(*a)[0], O.shape_c3d20(*b)[0]
        dX[:, k] = (Na - Nb) @ X / (2 * h)
        dU[:, k] = (Na - Nb) @ U / (2 * h)
    grad = dU @ np.linalg.inv(dX)
    eps = np.array([grad[0, 0], grad[1, 1], grad[2, 2], grad[0, 1] + grad[1, 0], grad[1, 2] + grad[2, 1],
                    grad[2, 0] + grad[0, 2]])
    assert np.max(np.abs(E[0] - eps)) < 1e-8 * np.max(np.abs(eps))
    D = O.d_matrix(float(m.lut_v_e[0, 0]), float(m.lut_v_nu[0, 0]))
    assert np.max(np.abs(S[0] - D @ E[0])) < 1e-9 * np.max(np.abs(S[0]))


def test_stress_recovery_uses_node_temperature():
    """update_state is called with the node's own temperature
    (static_step.rs:345-346), so with a temperature-dependent E every node's
    averaged stress is D(T_node) times its averaged strain."""
    job = meshgen.c3d20_block(seed=6, nx=2, ny=2, nz=1)
    m = job.mesh
    m.lut_n_e[0] = 2
    m.lut_t_e[0, :2] = [0.0, 100.0]
    m.lut_v_e[0, :2] = [70000.0, 35000.0]
    rng = np.random.default_rng(3)
    T = rng.uniform(0.0, 100.0, size=len(m.xyz))
    u = rng.normal(scale=1e-3, size=job.n_dof)
    S, E = O.stress_strain(m, T, u)
    nu = float(m.lut_v_nu[0, 0])
    for n in range(len(T)):
        D = O.d_matrix(70000.0 + (35000.0 - 70000.0) * (T[n] / 100.0), nu)
        assert np.allclose(S[n], D @ E[n], rtol=1e-12, atol=1e-12 * np.abs(S).max())


def test_pattern_fingerprint_is_fnv1a():
    """The golden fixtures' pattern hash (oracle C helper) is byte-wise
    FNV-1a over row_ptr (int64 LE) then col (int32 LE): equal to a plain
    Python FNV-1a on a small pattern, and to the committed cantilever hash."""
    def fnv(data: bytes) -> str:
        h = 0xCBF29CE484222325
        for b in data:
            h ^= b
            h = (h * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
        return f"{h:016x}"

    rp = np.array([0, 2, 3, 5], np.int64)
    col = np.array([0, 2, 1, 0, 2], np.int32)
    assert O.pattern_fnv1a(rp, col) == fnv(rp.astype("<i8").tobytes() + col.astype("<i4").tobytes())
    assert O.pattern_fnv1a(rp, col) != O.pattern_fnv1a(rp, col[::-1].copy())
    job = meshgen.c3d4_cantilever()
    csr, _, _ = O.assemble_csr(job.mesh, job.node_temp)
    meta = json.load(open(os.path.join(ROOT, "tests", "golden", "cantilever_oracle.json")))
    assert O.pattern_fnv1a(csr.row_ptr, csr.col) == meta["pattern_fnv1a"]
