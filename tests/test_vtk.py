"""Legacy VTK writer (src/solver/io/vtk.rs:21-333): number format, derived
fields and file layout, on the CPU.

The reference has no test of its writer, so the format is pinned by
vtk.rs's own text (restated in ferrix_amd/vtk.py), by Rust's float Display
(shortest round-trip digits, positional: the values below are its documented
outputs) and by closed forms for MISES and TRESCA.
"""
import json
import math
import os

import numpy as np
import pytest

from ferrix_amd.inp import MeshArrays
from ferrix_amd.vtk import VtkWriter, mises, rust_float, tresca


@pytest.mark.parametrize("x, text", [
    (1.0, "1"), (0.0, "0"), (-0.0, "-0"), (0.1, "0.1"), (-2.5, "-2.5"), (100.0, "100"),
    (0.1 + 0.2, "0.30000000000000004"), (1e-7, "0.0000001"), (1e21, "1000000000000000000000"),
    (123456.789, "123456.789"), (5e-324, "0." + "0" * 323 + "5"),
    (float("nan"), "NaN"), (float("inf"), "inf"), (float("-inf"), "-inf"),
])
def test_rust_float_display(x, text):
    assert rust_float(x) == text


def test_rust_float_round_trips_with_shortest_digits():
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.standard_normal(2000) * 10.0 ** rng.integers(-30, 30, 2000),
                         rng.integers(-10**6, 10**6, 200).astype(float)])
    for x in xs:
        t = rust_float(x)
        assert "e" not in t and float(t) == x
        significant = t.replace("-", "").replace(".", "").strip("0")
        assert len(significant) <= 17


def test_mises_and_tresca_closed_forms():
    s = 250.0
    assert mises([s, 0, 0, 0, 0, 0]) == s  # uniaxial
    assert tresca([s, 0, 0, 0, 0, 0]) == pytest.approx(s, rel=1e-15)
    t = 40.0
    assert mises([0, 0, 0, t, 0, 0]) == pytest.approx(math.sqrt(3.0) * t, rel=1e-15)  # pure shear
    assert tresca([0, 0, 0, t, 0, 0]) == pytest.approx(2.0 * t, rel=1e-15)
    assert mises([7.0, 7.0, 7.0, 0, 0, 0]) == 0.0  # hydrostatic
    assert tresca([3.0, -5.0, 1.0, 0, 0, 0]) == pytest.approx(8.0, rel=1e-15)  # principal axes
    # rotation invariance: the same principal stresses in a rotated frame
    c, sn = math.cos(0.3), math.sin(0.3)
    R = np.array([[c, -sn, 0], [sn, c, 0], [0, 0, 1.0]])
    T = R @ np.diag([3.0, -5.0, 1.0]) @ R.T
    v = [T[0, 0], T[1, 1], T[2, 2], T[0, 1], T[1, 2], T[0, 2]]
    assert tresca(v) == pytest.approx(8.0, rel=1e-13)
    assert mises(v) == pytest.approx(mises([3.0, -5.0, 1.0, 0, 0, 0]), rel=1e-13)


def _two_tets_and_a_loose_node():
    xyz = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1], [5, 5, 5.5]], float)
    conn = np.array([0, 1, 2, 3, 1, 2, 3, 4], np.int32)
    z = np.zeros((1, 16))
    return MeshArrays(node_ids=np.arange(1, 7), xyz=xyz, elem_ids=np.array([1, 2]),
                      el_ptr=np.array([0, 4, 8], np.int64), conn=conn, mat=np.zeros(2, np.int32),
                      lut_n_e=np.ones(1, np.int32), lut_t_e=z, lut_v_e=z + 1.0,
                      lut_n_nu=np.ones(1, np.int32), lut_t_nu=z, lut_v_nu=z + 0.3)


def test_writer_file_layout(tmp_path):
    mesh = _two_tets_and_a_loose_node()
    w = VtkWriter(str(tmp_path / "job.inp"), mesh)
    w.init()
    u = np.arange(18, dtype=float) * 0.25
    S = np.tile(np.array([1.0, 2.0, 3.0, 0.5, 0.0, -0.5]), (6, 1))
    E = S * 1e-6
    T = np.full(6, 20.0)
    p1 = w.write_increment(1, 1, 0.5, u, T, S, E)
    w.write_increment(1, 3, 1.0, u, T, S, E)
    w.finish()
    assert p1 == str(tmp_path / "job" / "step_1_1.vtk")
    lines = open(p1).read().split("\n")
    assert lines[:5] == ["# vtk DataFile Version 3.0", "CCX-RS Analysis Results", "ASCII",
                         "DATASET UNSTRUCTURED_GRID", "POINTS 6 float"]
    assert lines[5] == "0 0 0" and lines[10] == "5 5 5.5"
    i = lines.index("CELLS 2 10")
    assert lines[i + 1:i + 3] == ["4 0 1 2 3", "4 1 2 3 4"]
    assert lines[i + 3:i + 6] == ["CELL_TYPES 2", "10", "10"]
    j = lines.index("POINT_DATA 6")
    assert lines[j + 1] == "VECTORS U float" and lines[j + 2] == "0 0.25 0.5"
    k = lines.index("SCALARS S float 6")
    assert lines[k + 1] == "LOOKUP_TABLE default"
    assert lines[k + 2] == "1 2 3 0.5 0 -0.5"
    assert lines[k + 7] == "0.0 0.0 0.0 0.0 0.0 0.0"  # node 6: no element, no entry
    assert lines[lines.index("SCALARS E float 6") + 7] == "0.0 0.0 0.0 0.0 0.0 0.0"
    nt = lines.index("SCALARS NT float 1")
    assert lines[nt + 2:nt + 8] == ["20"] * 6
    m = lines.index("SCALARS MISES float 1")
    assert float(lines[m + 2]) == mises(S[0]) and lines[m + 7] == "0.0"
    t = lines.index("SCALARS TRESCA float 1")
    assert float(lines[t + 2]) == pytest.approx(tresca(S[0])) and lines[t + 7] == "0.0"
    assert lines[-1] == "" and lines[t + 7] == lines[-2]
    series = open(tmp_path / "job" / "results.vtk.series").read()
    assert series.startswith('{\n          "file-series-version" : "1.0",\n          "files" : [\n')
    assert '{ "name" : "step_1_1.vtk", "time" : 0.5 },\n' in series
    assert series.endswith('{ "name" : "step_1_3.vtk", "time" : 1 },\n  ]\n}\n')
    # the reference leaves a comma after the last entry; without it the index is JSON
    doc = json.loads(series.replace("},\n  ]", "}\n  ]"))
    assert [f["name"] for f in doc["files"]] == ["step_1_1.vtk", "step_1_3.vtk"]
    assert os.path.exists(tmp_path / "job" / "step_1_3.vtk")


def test_deck_requests_pick_their_path():
    """One linear *STATIC step joins the batch (C3D20 meshes as "c3d20");
    a deck with several steps or NLGEOM runs through the step driver."""
    from conftest import deck
    from ferrix_amd.inp import deck_stem
    from ferrix_amd.serving import request_from_deck

    kinds = {name: request_from_deck(0, 0.0, deck(name), 20.0).kind
             for name in ("C3D4_staticstep", "C3D20", "C3D4_2staticstep", "NLGEOM_test")}
    assert kinds == {"C3D4_staticstep": "linear", "C3D20": "c3d20", "C3D4_2staticstep": "nlgeom",
                     "NLGEOM_test": "nlgeom"}
    assert deck_stem("/a/b/C3D4.inp") == "C3D4" and deck_stem("x/C3D20.inp.gz") == "C3D20"
