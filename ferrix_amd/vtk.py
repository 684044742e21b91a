"""Legacy VTK result writer: the reference's ``VtkWriter``
(``src/solver/io/vtk.rs:21-333``) behind its ``ResultWriter`` trait
(``src/solver/io/writer.rs:7-34``).

For every converged increment one ``step_<step>_<inc>.vtk`` file (ASCII
unstructured grid: POINTS, CELLS, CELL_TYPES, then POINT_DATA U, S, E, NT,
MISES, TRESCA) goes to ``<deck dir>/<deck stem>/``, and ``results.vtk.series``
(ParaView's file-series index) lists it with the global time.  Text follows
``vtk.rs`` line for line:

* numbers print as Rust's ``{}`` prints an f64: the shortest digits that
  read back to the same double, positional (no exponent), ``-0`` for
  negative zero, ``NaN``/``inf`` (``rust_float``);
* nodes in index order (ascending node ID); a node that no element uses has
  no S/E entry in the reference (``static_step.rs:352-377``), so its S, E,
  MISES and TRESCA rows are the literal ``0.0`` texts of ``vtk.rs:186-300``;
* MISES = sqrt(((s11-s22)^2 + (s22-s33)^2 + (s33-s11)^2
  + 6 (s12^2 + s23^2 + s13^2)) / 2) in that association (``vtk.rs:247-256``);
  TRESCA = largest minus smallest eigenvalue of the stress tensor.

Cells are written in ascending element ID.  The reference walks its
element HashMap, whose order is unspecified (``vtk.rs:103``), so cell order
is not part of the format.  TRESCA's last bits come from the eigen solver
(nalgebra's SymmetricEigen there, LAPACK's here): equal to rounding, not bit
for bit.
"""
from __future__ import annotations

import math
import os
from typing import Optional

import numpy as np

from .inp import MeshArrays, deck_stem

SERIES_FILENAME = "results.vtk.series"
VTK_TETRA = 10
VTK_QUADRATIC_HEXAHEDRON = 25


def rust_float(x: float) -> str:
    """``format!("{}", x)`` for an f64."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "inf" if x > 0 else "-inf"
    return np.format_float_positional(x, unique=True, trim="-")


def mises(s: np.ndarray) -> float:
    s11, s22, s33, s12, s23, s13 = (float(v) for v in s)
    a, b, c = s11 - s22, s22 - s33, s33 - s11
    return math.sqrt((a * a + b * b + c * c + 6.0 * (s12 * s12 + s23 * s23 + s13 * s13)) / 2.0)


def tresca(s: np.ndarray) -> float:
    s11, s22, s33, s12, s23, s13 = (float(v) for v in s)
    ev = np.linalg.eigvalsh(np.array([[s11, s12, s13], [s12, s22, s23], [s13, s23, s33]]))
    return float(ev.max() - ev.min())


class VtkWriter:
    """``ResultWriter`` of the VTK format: ``init``, ``write_increment`` per
    converged increment, ``finish``.  ``deck_path`` names the job as the
    reference does: directory = the deck's directory joined with its stem
    (``Project::jobname``/``job_dir``, ``project.rs:133-154``)."""

    def __init__(self, deck_path: str, mesh: MeshArrays):
        stem = deck_stem(deck_path) or "Unknown"
        self.dirpath = os.path.join(os.path.dirname(deck_path) or ".", stem)
        self.mesh = mesh
        n = mesh.n_nodes
        counts = np.bincount(np.asarray(mesh.conn, np.int64), minlength=n) if mesh.conn.size else np.zeros(n, int)
        self.used = counts[:n] > 0  # nodes with an S/E entry
        self._cells = self._cell_text()

    def _series(self) -> str:
        return os.path.join(self.dirpath, SERIES_FILENAME)

    def init(self) -> None:
        os.makedirs(self.dirpath, exist_ok=True)
        with open(self._series(), "w") as f:
            f.write('{\n          "file-series-version" : "1.0",\n          "files" : [\n')

    def _cell_text(self) -> str:
        m = self.mesh
        ptr = np.asarray(m.el_ptr, np.int64)
        conn = np.asarray(m.conn, np.int64)
        lines, types, size = [], [], 0
        for e in range(m.n_el):
            nodes = conn[ptr[e]:ptr[e + 1]]
            if len(nodes) == 4:
                types.append(VTK_TETRA)
            elif len(nodes) == 20:
                types.append(VTK_QUADRATIC_HEXAHEDRON)
            else:
                continue
            lines.append(" ".join([str(len(nodes))] + [str(int(i)) for i in nodes]))
            size += len(nodes) + 1
        out = [f"CELLS {len(lines)} {size}"] + lines + [f"CELL_TYPES {len(types)}"] + [str(t) for t in types]
        return "\n".join(out) + "\n"

    def write_increment(self, step_id: int, inc_id: int, global_time: float, u: np.ndarray,
                        temperatures: np.ndarray, stress: Optional[np.ndarray], strain: Optional[np.ndarray]) -> str:
        """Writes step_<step_id>_<inc_id>.vtk and appends it to the series;
        u is (n_dof,), stress and strain (n_nodes, 6) nodal averages (or
        None when not computed).  Returns the file's path."""
        m = self.mesh
        n = m.n_nodes
        name = f"step_{step_id}_{inc_id}.vtk"
        path = os.path.join(self.dirpath, name)
        rf = rust_float
        out = ["# vtk DataFile Version 3.0", "CCX-RS Analysis Results", "ASCII", "DATASET UNSTRUCTURED_GRID",
               f"POINTS {n} float"]
        out += [f"{rf(x)} {rf(y)} {rf(z)}" for x, y, z in np.asarray(m.xyz, np.float64)]
        text = "\n".join(out) + "\n" + self._cells
        out = [f"POINT_DATA {n}", "VECTORS U float"]
        uu = np.asarray(u, np.float64).reshape(n, 3)
        out += [f"{rf(a)} {rf(b)} {rf(c)}" for a, b, c in uu]
        six_zero = "0.0 0.0 0.0 0.0 0.0 0.0"
        for label, field in (("S", stress), ("E", strain)):
            if field is None:
                continue
            out += [f"SCALARS {label} float 6", "LOOKUP_TABLE default"]
            out += [" ".join(rf(v) for v in field[k]) if self.used[k] else six_zero for k in range(n)]
        out += ["SCALARS NT float 1", "LOOKUP_TABLE default"]
        out += [rf(t) for t in np.asarray(temperatures, np.float64)]
        if stress is not None:
            out += ["SCALARS MISES float 1", "LOOKUP_TABLE default"]
            out += [rf(mises(stress[k])) if self.used[k] else "0.0" for k in range(n)]
            out += ["SCALARS TRESCA float 1", "LOOKUP_TABLE default"]
            out += [rf(tresca(stress[k])) if self.used[k] else "0.0" for k in range(n)]
        with open(path, "w") as f:
            f.write(text + "\n".join(out) + "\n")
        with open(self._series(), "a") as f:
            f.write(f'{{ "name" : "{name}", "time" : {rf(global_time)} }},\n')
        return path

    def finish(self) -> None:
        with open(self._series(), "a") as f:
            f.write("  ]\n}\n")
