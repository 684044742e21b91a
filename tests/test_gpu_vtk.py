"""The VTK result writer driven by the GPU step driver (fx_static_step's
per-increment hook, the reference's writer.write_increment at
static_step.rs:211-250): one file per converged increment, the series
index with global times, and field values that read back to exactly what
the device holds."""
import json

import numpy as np
import pytest

from conftest import deck
from ferrix_amd.device import DeviceMesh, run_job
from ferrix_amd.inp import Project
from ferrix_amd.model import Job
from ferrix_amd.vtk import VtkWriter, mises

pytestmark = pytest.mark.gpu


def section(lines, header, n):
    i = lines.index(header)
    start = i + 2 if lines[i + 1] == "LOOKUP_TABLE default" else i + 1
    return np.array([[float(v) for v in lines[start + k].split()] for k in range(n)])


def test_vtk_writer_two_step_deck(ctx, tmp_path):
    p = Project.from_jobname(deck("C3D4_2staticstep"))
    p.default_initial_temperature = 20.0
    job = Job.from_project(p)
    n = job.mesh.n_nodes
    dm = DeviceMesh(ctx, job.mesh, job.node_temp)
    w = VtkWriter(str(tmp_path / "C3D4_2staticstep.inp"), job.mesh)
    seen = []
    write = w.write_increment

    def spy(step_id, inc_id, t, u, temps, S, E):
        seen.append((step_id, inc_id, t, u.copy(), S.copy(), E.copy()))
        return write(step_id, inc_id, t, u, temps, S, E)

    w.write_increment = spy
    u, reps = run_job(ctx, job, mesh=dm, writer=w)
    print("increments written", [(s, i, t) for s, i, t, *_ in seen], flush=True)
    assert len(seen) == sum(r.increments - r.cutbacks for r in reps) and len(seen) >= 2
    series = open(tmp_path / "C3D4_2staticstep" / "results.vtk.series").read()
    doc = json.loads(series.replace("},\n  ]", "}\n  ]"))
    assert [(f["name"], f["time"]) for f in doc["files"]] == [(f"step_{s}_{i}.vtk", t) for s, i, t, *_ in seen]
    # the last file holds the final state, exactly
    s_id, i_id, _, u_last, S_last, E_last = seen[-1]
    assert np.array_equal(u_last, u)
    S_now, E_now = dm.stress_strain()
    assert np.array_equal(S_last, S_now) and np.array_equal(E_last, E_now)
    lines = open(tmp_path / "C3D4_2staticstep" / f"step_{s_id}_{i_id}.vtk").read().split("\n")
    assert np.array_equal(section(lines, "POINTS %d float" % n, n), job.mesh.xyz)
    assert np.array_equal(section(lines, "VECTORS U float", n), u.reshape(n, 3))
    S = section(lines, "SCALARS S float 6", n)
    assert np.array_equal(S, S_now) and np.array_equal(section(lines, "SCALARS E float 6", n), E_now)
    assert np.array_equal(section(lines, "SCALARS NT float 1", n)[:, 0], job.node_temp)
    M = section(lines, "SCALARS MISES float 1", n)[:, 0]
    assert np.array_equal(M, [mises(S[k]) for k in range(n)])
    T = section(lines, "SCALARS TRESCA float 1", n)[:, 0]
    tol = 1e-9 * (np.abs(M) + 1.0)
    assert np.all(T >= M - tol) and np.all(T <= 2.0 / np.sqrt(3.0) * M + tol)  # Mises <= Tresca <= (2/sqrt 3) Mises


def test_vtk_writer_nlgeom_increments(ctx, tmp_path):
    """A C3D20 beam in ten ramped NLGEOM increments (plus the reference's
    micro-increment, 0.1 x 10 summing to 0.9999999999999999): one file per
    converged increment, named step_1_<n_inc>, global times rising to the
    end of the step, the last file holding the final displacement."""
    from ferrix_amd import meshgen

    job = meshgen.c3d20_nlgeom_beam(nx=8, ny=2, nz=2, tip_load=20000.0)
    w = VtkWriter(str(tmp_path / "beam.inp"), job.mesh)
    u, reps = run_job(ctx, job, writer=w)
    series = open(tmp_path / "beam" / "results.vtk.series").read()
    files = json.loads(series.replace("},\n  ]", "}\n  ]"))["files"]
    print("series", [(f["name"], f["time"]) for f in files], flush=True)
    assert len(files) == reps[0].increments - reps[0].cutbacks > 2
    assert [f["name"] for f in files] == [f"step_1_{k + 1}.vtk" for k in range(len(files))]
    times = [f["time"] for f in files]
    assert all(b > a for a, b in zip(times, times[1:]))
    assert times[-1] == pytest.approx(job.steps[0].time_period, rel=1e-12)
    n = job.mesh.n_nodes
    lines = open(tmp_path / "beam" / files[-1]["name"]).read().split("\n")
    assert np.array_equal(section(lines, "VECTORS U float", n), u.reshape(n, 3))
    assert lines.count("CELL_TYPES %d" % job.mesh.n_el) == 1 and "25" in lines


def test_writer_error_stops_the_step(ctx, tmp_path):
    """A writer that fails (the reference's GenericIo error out of
    write_increment) stops the step with that error; the series file is
    still closed, as main.rs:66 calls finish after a failed step."""
    p = Project.from_jobname(deck("C3D4_staticstep"))
    p.default_initial_temperature = 20.0
    job = Job.from_project(p)
    w = VtkWriter(str(tmp_path / "C3D4_staticstep.inp"), job.mesh)

    def broken(*a, **k):
        raise OSError("disk full")

    w.write_increment = broken
    with pytest.raises(OSError, match="disk full"):
        run_job(ctx, job, writer=w)
    assert open(tmp_path / "C3D4_staticstep" / "results.vtk.series").read().endswith("  ]\n}\n")
