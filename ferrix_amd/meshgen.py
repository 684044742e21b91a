"""Seeded synthetic workloads: structured C3D4 (Kuhn split) and C3D20 blocks.

All random draws go through :class:`SplitMix64` so the sequences are fixed
across languages.  Node and element numbering rules:

* hex grid ``nx x ny x nz`` of cells; corner node (i, j, k) has
  ID ``1 + i + (nx+1)*(j + (ny+1)*k)``;
* C3D20 mid-edge nodes follow the corners: x-edges, then y-edges, then
  z-edges, each in (i fastest, then j, then k) order;
* C3D4: each hex (v0..v7 in CalculiX order) is cut into the 6 Kuhn tets
  ``[v0, va, vb, v6]`` for (a, b) in (1,2),(2,3),(3,7),(7,4),(4,5),(5,1);
* C3D20 element nodes in CalculiX order (8 corners, 4 bottom mid-edges,
  4 top mid-edges, 4 vertical mid-edges), matching the reference's
  ``node_local_coords`` (``src/solver/mesh_lib/elements/c3d20.rs:61-84``).

Clamped and loaded faces are picked by grid index, never by coordinate.
"""
from __future__ import annotations

import math
from typing import List, Tuple

import numpy as np

from .inp import LUT_MAX, Amplitude, IncrementData, MeshArrays
from .model import Job, StepSpec, SOLVER_ITERATIVE, STOP_REL_R2

_M64 = (1 << 64) - 1


class SplitMix64:
    """Vigna's SplitMix64; ``uniform()`` is in [0, 1) with 53 bits."""

    def __init__(self, seed: int):
        self.state = seed & _M64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _M64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)

    def uniform_array(self, n: int) -> np.ndarray:
        return np.array([self.uniform() for _ in range(n)], dtype=np.float64)


def _material_one(E: float, nu: float):
    ne = np.array([1], np.int32)
    te = np.zeros((1, LUT_MAX))
    ve = np.zeros((1, LUT_MAX))
    ve[0, 0] = E
    nn = np.array([1], np.int32)
    tn = np.zeros((1, LUT_MAX))
    vn = np.zeros((1, LUT_MAX))
    vn[0, 0] = nu
    return ne, te, ve, nn, tn, vn


class BlockGrid:
    """Structured grid numbering for an nx x ny x nz block of cells."""

    def __init__(self, nx, ny, nz, hx, hy, hz, quadratic: bool):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.h = (hx, hy, hz)
        self.quadratic = quadratic
        self.nc = (nx + 1) * (ny + 1) * (nz + 1)
        self.nxe = nx * (ny + 1) * (nz + 1) if quadratic else 0
        self.nye = (nx + 1) * ny * (nz + 1) if quadratic else 0
        self.nze = (nx + 1) * (ny + 1) * nz if quadratic else 0

    @property
    def n_nodes(self):
        return self.nc + self.nxe + self.nye + self.nze

    # IDs are 1-based; index = ID - 1 since IDs are dense and ascending.
    def corner(self, i, j, k):
        return 1 + i + (self.nx + 1) * (j + (self.ny + 1) * k)

    def xedge(self, i, j, k):
        return self.nc + 1 + i + self.nx * (j + (self.ny + 1) * k)

    def yedge(self, i, j, k):
        return self.nc + self.nxe + 1 + i + (self.nx + 1) * (j + self.ny * k)

    def zedge(self, i, j, k):
        return self.nc + self.nxe + self.nye + 1 + i + (self.nx + 1) * (j + (self.ny + 1) * k)

    def coords(self) -> np.ndarray:
        """Unjittered coordinates in ID order (vectorised)."""
        nx, ny, nz = self.nx, self.ny, self.nz
        hx, hy, hz = self.h
        blocks = []
        k, j, i = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
        blocks.append(np.stack([i * hx, j * hy, k * hz], -1).reshape(-1, 3))
        if self.quadratic:
            k, j, i = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx), indexing="ij")
            blocks.append(np.stack([(i + 0.5) * hx, j * hy, k * hz], -1).reshape(-1, 3))
            k, j, i = np.meshgrid(np.arange(nz + 1), np.arange(ny), np.arange(nx + 1), indexing="ij")
            blocks.append(np.stack([i * hx, (j + 0.5) * hy, k * hz], -1).reshape(-1, 3))
            k, j, i = np.meshgrid(np.arange(nz), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
            blocks.append(np.stack([i * hx, j * hy, (k + 0.5) * hz], -1).reshape(-1, 3))
        return np.concatenate(blocks).astype(np.float64)

    def hex_corners(self) -> np.ndarray:
        """(n_cells, 8) corner IDs in CalculiX order, cells i fastest."""
        nx, ny, nz = self.nx, self.ny, self.nz
        k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        i, j, k = i.ravel(), j.ravel(), k.ravel()
        c = self.corner
        return np.stack([c(i, j, k), c(i + 1, j, k), c(i + 1, j + 1, k), c(i, j + 1, k),
                         c(i, j, k + 1), c(i + 1, j, k + 1), c(i + 1, j + 1, k + 1),
                         c(i, j + 1, k + 1)], -1)

    def c3d20(self) -> np.ndarray:
        nx, ny, nz = self.nx, self.ny, self.nz
        k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        i, j, k = i.ravel(), j.ravel(), k.ravel()
        cor = self.hex_corners()
        X, Y, Z = self.xedge, self.yedge, self.zedge
        mids = np.stack([X(i, j, k), Y(i + 1, j, k), X(i, j + 1, k), Y(i, j, k),
                         X(i, j, k + 1), Y(i + 1, j, k + 1), X(i, j + 1, k + 1), Y(i, j, k + 1),
                         Z(i, j, k), Z(i + 1, j, k), Z(i + 1, j + 1, k), Z(i, j + 1, k)], -1)
        return np.concatenate([cor, mids], 1)

    def c3d4(self) -> np.ndarray:
        cor = self.hex_corners()
        pairs = [(1, 2), (2, 3), (3, 7), (7, 4), (4, 5), (5, 1)]
        tets = np.stack([np.stack([cor[:, 0], cor[:, a], cor[:, b], cor[:, 6]], -1) for a, b in pairs], 1)
        return tets.reshape(-1, 4)

    def face_nodes(self, axis: int, idx: int) -> np.ndarray:
        """IDs of all nodes whose grid coordinate along ``axis`` equals
        ``idx`` (an integer grid plane), ascending."""
        nx, ny, nz = self.nx, self.ny, self.nz
        out = []
        k, j, i = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
        g = (i, j, k)
        out.append(self.corner(i, j, k)[g[axis] == idx])
        if self.quadratic:
            # a mid-edge node lies on the plane when its edge lies in the plane
            k, j, i = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx), indexing="ij")
            if axis != 0:
                out.append(self.xedge(i, j, k)[(i, j, k)[axis] == idx])
            k, j, i = np.meshgrid(np.arange(nz + 1), np.arange(ny), np.arange(nx + 1), indexing="ij")
            if axis != 1:
                out.append(self.yedge(i, j, k)[(i, j, k)[axis] == idx])
            k, j, i = np.meshgrid(np.arange(nz), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
            if axis != 2:
                out.append(self.zedge(i, j, k)[(i, j, k)[axis] == idx])
        return np.sort(np.concatenate([o.ravel() for o in out]))


def _mesh(grid: BlockGrid, xyz: np.ndarray, conn_ids: np.ndarray, E: float, nu: float) -> MeshArrays:
    n_el, nn = conn_ids.shape
    return MeshArrays(
        node_ids=np.arange(1, grid.n_nodes + 1, dtype=np.int64),
        xyz=np.ascontiguousarray(xyz, dtype=np.float64),
        elem_ids=np.arange(1, n_el + 1, dtype=np.int64),
        el_ptr=np.arange(0, n_el * nn + 1, nn, dtype=np.int64),
        conn=np.ascontiguousarray((conn_ids - 1).ravel(), dtype=np.int32),
        mat=np.zeros(n_el, np.int32),
        **dict(zip(("lut_n_e", "lut_t_e", "lut_v_e", "lut_n_nu", "lut_t_nu", "lut_v_nu"),
                   _material_one(E, nu))),
    )


def _clamp_and_load(grid: BlockGrid, load_dir, total: float, amp=None):
    """x=0 face clamped in all 3 DOFs; ``total`` spread evenly over the
    x=max face nodes along unit direction ``load_dir``."""
    clamp = grid.face_nodes(0, 0) - 1
    tip = grid.face_nodes(0, grid.nx) - 1
    per = total / len(tip)
    amp = amp if amp is not None else Amplitude()
    loads = []
    for n in tip:
        for d in range(3):
            if load_dir[d] != 0.0:
                loads.append((3 * int(n) + d, per * load_dir[d], amp, 1))
    bcs = [(3 * int(n) + d, 0.0, Amplitude(), 1) for n in clamp for d in range(3)]
    return loads, bcs


def c3d4_cantilever() -> Job:
    """100x20x20 mm block, 20x4x4 hex cells of 5 mm cut into 1,920 C3D4
    (525 nodes, 1,575 DOF); E=210000, nu=0.3; x=0 clamped; -100 N in z over
    the 25 x=100 nodes."""
    g = BlockGrid(20, 4, 4, 5.0, 5.0, 5.0, quadratic=False)
    mesh = _mesh(g, g.coords(), g.c3d4(), 210000.0, 0.3)
    loads, bcs = _clamp_and_load(g, (0.0, 0.0, -1.0), 100.0)
    step = StepSpec.build(loads, bcs)
    return Job(mesh, np.zeros(g.n_nodes), [step], "c3d4_cantilever_20x4x4")


def c3d20_block(seed: int = 2, nx=40, ny=10, nz=10, jitter=True) -> Job:
    """400x100x100 mm block, 40x10x10 C3D20 (18,821 nodes, 56,463 DOF),
    coordinates jittered by U(-0.5, 0.5) mm per coordinate in node-ID order
    from SplitMix64(seed); E=70000, nu=0.33; x=0 clamped; -1000 N in z over
    the x=400 face nodes; Jacobi-PCG to ||r||/||b|| <= 1e-10."""
    g = BlockGrid(nx, ny, nz, 400.0 / nx, 100.0 / ny, 100.0 / nz, quadratic=True)
    xyz = g.coords()
    if jitter:
        rng = SplitMix64(seed)
        u = rng.uniform_array(3 * g.n_nodes).reshape(-1, 3)
        xyz = xyz + (u - 0.5) * (0.05 * 400.0 / nx)
    mesh = _mesh(g, xyz, g.c3d20(), 70000.0, 0.33)
    loads, bcs = _clamp_and_load(g, (0.0, 0.0, -1.0), 1000.0)
    step = StepSpec.build(loads, bcs, solver=SOLVER_ITERATIVE, stop_mode=STOP_REL_R2, tol=1e-10,
                          max_iter=100_000)
    return Job(mesh, np.zeros(g.n_nodes), [step], f"c3d20_block_{nx}x{ny}x{nz}")


def c3d20_nlgeom_beam(nx=80, ny=4, nz=4, tip_load=110000.0) -> Job:
    """1000x50x50 mm beam, 80x4x4 C3D20 (7,265 nodes, 21,795 DOF),
    E=210000, nu=0.3, NLGEOM; tip load in -z ramped by *AMPLITUDE (0,0,1,1)
    with *STATIC 0.1, 1., 1e-5, 0.1."""
    g = BlockGrid(nx, ny, nz, 1000.0 / nx, 50.0 / ny, 50.0 / nz, quadratic=True)
    mesh = _mesh(g, g.coords(), g.c3d20(), 210000.0, 0.3)
    amp = Amplitude(False, 0.0, 0.0, ([0.0, 1.0], [0.0, 1.0]))
    loads, bcs = _clamp_and_load(g, (0.0, 0.0, -1.0), tip_load, amp)
    inc = IncrementData(0.1, 1.0, 1e-5, 0.1, 10_000)
    step = StepSpec.build(loads, bcs, inc=inc, nlgeom=True, solver=SOLVER_ITERATIVE,
                          stop_mode=STOP_REL_R2, tol=1e-10, max_iter=100_000)
    return Job(mesh, np.zeros(g.n_nodes), [step], f"c3d20_nlgeom_beam_{nx}x{ny}x{nz}")


def c3d4_box(nx, ny, nz, h, E, nu, loads_dirs: List[Tuple[float, float, float]], total=100.0) -> List[Job]:
    """Structured Kuhn-tet box with one job per load case (same K)."""
    g = BlockGrid(nx, ny, nz, h[0], h[1], h[2], quadratic=False)
    mesh = _mesh(g, g.coords(), g.c3d4(), E, nu)
    jobs = []
    for d in loads_dirs:
        loads, bcs = _clamp_and_load(g, d, total)
        jobs.append(Job(mesh, np.zeros(g.n_nodes), [StepSpec.build(loads, bcs)], f"c3d4_box_{nx}x{ny}x{nz}"))
    return jobs


def _population_draw(rng: SplitMix64, index: int, dof_lo: float, dof_hi: float) -> dict:
    """One mesh of the batched-solve population (draw order: target DOF,
    aspect, E, nu, load-case count, directions)."""
    target = math.exp(math.log(dof_lo) + rng.uniform() * (math.log(dof_hi) - math.log(dof_lo)))
    aspect = 1.0 + 9.0 * rng.uniform()
    E = math.exp(math.log(1e3) + rng.uniform() * (math.log(3e5) - math.log(1e3)))
    nu = 0.2 + 0.25 * rng.uniform()
    ncase = 1 + int(rng.next_u64() % 4)
    dirs = []
    for _ in range(ncase):
        # uniform direction: z = U(-1,1), phi = U(0, 2pi)
        z = 2.0 * rng.uniform() - 1.0
        phi = 2.0 * math.pi * rng.uniform()
        r = math.sqrt(max(0.0, 1.0 - z * z))
        dirs.append((r * math.cos(phi), r * math.sin(phi), z))
    # cells: nx = aspect * n, ny = nz = n; 3 (nx+1)(n+1)^2 ~ target
    nodes = target / 3.0
    n = max(1, int(round((nodes / aspect) ** (1.0 / 3.0) - 1.0)))
    nx = max(1, int(round(aspect * (n + 1) - 1)))
    return dict(index=index, nx=nx, ny=n, nz=n, h=(10.0, 10.0, 10.0), E=E, nu=nu, dirs=dirs,
                n_dof=3 * (nx + 1) * (n + 1) * (n + 1))


def population_specs(n_meshes=512, seed=4, dof_lo=2e3, dof_hi=2e5):
    """Parameters of the batched-solve population: per mesh a target DOF
    (log-uniform), an aspect ratio U(1,10) of the x edge over the y/z edges,
    E log-uniform in [1e3, 3e5], nu U(0.2, 0.45) and 1-4 load cases, each a
    random unit direction times 100 N on the x=max face.  Returns a list of
    dicts; :func:`population_mesh` builds one."""
    rng = SplitMix64(seed)
    return [_population_draw(rng, m, dof_lo, dof_hi) for m in range(n_meshes)]


def population_mesh(spec) -> List[Job]:
    return c3d4_box(spec["nx"], spec["ny"], spec["nz"], spec["h"], spec["E"], spec["nu"], spec["dirs"])


def c3d4_box_nlgeom(spec) -> Job:
    """NLGEOM variant of a population mesh: its first load case, ramped by
    *AMPLITUDE (0,0, 1,1) over *STATIC 0.25, 1., 1e-5, 0.25 (Picard NLGEOM as
    in the reference), PCG to rel 1e-10."""
    g = BlockGrid(spec["nx"], spec["ny"], spec["nz"], *spec["h"], quadratic=False)
    mesh = _mesh(g, g.coords(), g.c3d4(), spec["E"], spec["nu"])
    amp = Amplitude(False, 0.0, 0.0, ([0.0, 1.0], [0.0, 1.0]))
    loads, bcs = _clamp_and_load(g, spec["dirs"][0], 100.0, amp)
    step = StepSpec.build(loads, bcs, inc=IncrementData(0.25, 1.0, 1e-5, 0.25, 1000), nlgeom=True,
                          solver=SOLVER_ITERATIVE, stop_mode=STOP_REL_R2, tol=1e-10, max_iter=100_000)
    return Job(mesh, np.zeros(g.n_nodes), [step], f"c3d4_box_nlgeom_{spec['nx']}x{spec['ny']}x{spec['nz']}")


def c3d20_block_cells(target_dof: float):
    """(nx, n) of a 4:1:1 C3D20 block with n_dof closest to target_dof."""
    best = None
    for n in range(1, 64):
        g = BlockGrid(4 * n, n, n, 1.0, 1.0, 1.0, quadratic=True)
        d = abs(3 * g.n_nodes - target_dof)
        if best is None or d < best[0]:
            best = (d, n)
    return 4 * best[1], best[1]


def request_specs(n_requests=1000, seed=5, mix=(0.7, 0.2, 0.1)):
    """The serving stream's requests (arrival times are drawn separately):
    kind 'linear' (a population mesh with its 1-4 load cases), 'nlgeom' (a
    population mesh, NLGEOM, first load case) or 'c3d20' (a jittered 4:1:1
    C3D20 block of log-uniform 1e4-5e5 DOF, one load case)."""
    rng = SplitMix64(seed)
    out = []
    for k in range(n_requests):
        u = rng.uniform()
        if u < mix[0]:
            out.append(dict(kind="linear", mesh=_population_draw(rng, k, 2e3, 2e5)))
        elif u < mix[0] + mix[1]:
            out.append(dict(kind="nlgeom", mesh=_population_draw(rng, k, 2e3, 2e5)))
        else:
            target = math.exp(math.log(1e4) + rng.uniform() * (math.log(5e5) - math.log(1e4)))
            nx, n = c3d20_block_cells(target)
            out.append(dict(kind="c3d20", nx=nx, n=n, seed=int(rng.next_u64() % (1 << 31))))
    return out


def request_jobs(req) -> List[Job]:
    """Jobs of one request: the load cases (sharing mesh and BCs) of a linear
    request, or the single job of an NLGEOM / C3D20 request."""
    if req["kind"] == "linear":
        return population_mesh(req["mesh"])
    if req["kind"] == "nlgeom":
        return [c3d4_box_nlgeom(req["mesh"])]
    return [c3d20_block(seed=req["seed"], nx=req["nx"], ny=req["n"], nz=req["n"])]


def to_inp(job: Job) -> str:
    """Write a job with one linear or NLGEOM step as a deck in the
    reference's keyword subset (used to round-trip the parser)."""
    m = job.mesh
    nn = int(m.el_ptr[1] - m.el_ptr[0])
    out = ["*HEADING", "synthetic", "*NODE"]
    for nid, (x, y, z) in zip(m.node_ids, m.xyz):
        out.append(f"{int(nid)}, {float(x)!r}, {float(y)!r}, {float(z)!r}")
    out.append(f"*ELEMENT, TYPE={'C3D4' if nn == 4 else 'C3D20'}, ELSET=EALL")
    for e in range(m.n_el):
        ids = [int(m.node_ids[c]) for c in m.conn[m.el_ptr[e]:m.el_ptr[e + 1]]]
        chunks = [ids[i:i + 10] for i in range(0, len(ids), 10)]
        line = f"{int(m.elem_ids[e])}, " + ", ".join(str(i) for i in chunks[0])
        for c in chunks[1:]:
            line += ",\n" + ", ".join(str(i) for i in c)
        out.append(line)
    out.append("*MATERIAL, NAME=MAT")
    out.append("*ELASTIC")
    out.append(f"{float(m.lut_v_e[0, 0])!r}, {float(m.lut_v_nu[0, 0])!r}")
    out.append("*SOLID SECTION, ELSET=EALL, MATERIAL=MAT")
    st = job.steps[0]
    out.append("*STEP, NLGEOM" if st.nlgeom else "*STEP")
    out.append("*STATIC")
    out.append(f"{float(st.initial_time_increment)!r}, {float(st.time_period)!r}, "
               f"{float(st.min_time_increment)!r}, {float(st.max_time_increment)!r}")
    # the reference applies only the first *BOUNDARY data line, so one
    # keyword per BC record
    for d, v in zip(st.bc_dof, st.bc_val):
        out.append("*BOUNDARY")
        out.append(f"{int(m.node_ids[d // 3])}, {d % 3 + 1}, {d % 3 + 1}, {float(v)!r}")
    out.append("*CLOAD")
    for d, v in zip(st.load_dof, st.load_val):
        out.append(f"{int(m.node_ids[d // 3])}, {d % 3 + 1}, {float(v)!r}")
    out.append("*END STEP")
    return "\n".join(out) + "\n"
