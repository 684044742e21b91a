"""Abaqus/CalculiX ``.inp`` keyword surface, restated from FerriX's parser.

This is host-side model ingestion for the GPU path: it turns a deck into a
:class:`Project` whose flat arrays are what the C ABI in ``include/ferrix_hip.h``
takes.  It follows the reference behaviour line by line, quirks included:

* ``preprocess_inp`` -- trim, drop empty lines, upper-case, merge lines that
  end in ``,``, drop ``**`` comments, map CalculiX solver names
  (``src/solver/parser.rs:903-942``).
* keyword dispatch and data lines (``parser.rs:111-877``): ``*BOUNDARY``
  applies only its first data line (peeked, ``:442-446``), ``*AMPLITUDE``
  reads one line (``:394-426``), ``*CLOAD`` reads until the next keyword
  (``:462-470``), ``GENERATE`` sets are ignored (``:545-546``, ``:566-567``),
  step 1 inherits the pre-step loads and BCs (``:265-269``), ``*STATIC`` time
  values are read only with four or more fields (``:316``).
* node index = rank of the node ID in ascending order
  (``src/solver/mesh_lib/mesh.rs:35-49``).

Errors use :class:`FerrixError` subclasses named after the reference's
``FerrixError`` variants (``src/solver/error.rs:12-74``).
"""
from __future__ import annotations

import gzip
import os
import re
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np


# ---------------------------------------------------------------------------
# Errors (error.rs:12-74)
# ---------------------------------------------------------------------------
class FerrixError(Exception):
    """Base of all errors; mirrors ``FerrixError``."""


class ParseError(FerrixError):
    def __init__(self, line: int, message: str):
        super().__init__(f"Parsing error at line {line}: {message}")
        self.line = line


class UnsupportedKeyword(FerrixError):
    def __init__(self, line: int, keyword: str):
        super().__init__(f"Unsupported keyword at line {line}: {keyword}")
        self.keyword = keyword


class UnknownSolver(FerrixError):
    pass


class MaterialNotFound(FerrixError):
    pass


class ElsetNotFound(FerrixError):
    pass


class NumericalError(FerrixError):
    pass


class ConvergenceError(FerrixError):
    pass


class InvalidModelState(FerrixError):
    pass


# ---------------------------------------------------------------------------
# Model pieces
# ---------------------------------------------------------------------------
@dataclass
class Amplitude:
    """``amplitude.rs:10-98``: tabular (with shifts, step or total time) or
    the default ramp when ``data`` is None."""

    total_time: bool = False
    shift_x: float = 0.0
    shift_y: float = 0.0
    data: Optional[Tuple[List[float], List[float]]] = None


@dataclass
class Load:
    """``boundary_conds.rs:10-53`` (also used for BCs, ``:57-101``)."""

    node_id: int
    dof: int
    value: float
    amplitude: Amplitude
    origin_step: int


@dataclass
class IncrementData:
    """``increment.rs:8-31``."""

    initial_time_increment: float = 1.0
    time_period: float = 1.0
    min_time_increment: float = 1e-5
    max_time_increment: float = 1.0
    max_iterations: int = 10_000


@dataclass
class StaticStep:
    """``static_step.rs:23-36``; solver is DEFAULT, DIRECT or ITERATIVE."""

    solver: str
    increment_data: IncrementData
    loads: List[Load] = field(default_factory=list)
    bcs: List[Load] = field(default_factory=list)
    nlgeom: bool = False


@dataclass
class Material:
    """``material.rs:233-277``: LUTs are lists of (temperature, value)."""

    name: str
    density: Optional[List[Tuple[float, float]]] = None
    youngs_modulus: Optional[List[Tuple[float, float]]] = None
    poisson_ratio: Optional[List[Tuple[float, float]]] = None
    thermal_expansion: Optional[List[Tuple[float, float]]] = None
    reference_temperature: float = 0.0
    num_depvars: int = 0


def lut_interpolate(data: List[Tuple[float, float]], temp: float) -> float:
    """``material.rs:36-68`` clamped linear interpolation."""
    if len(data) == 1 or temp <= data[0][0]:
        return data[0][1]
    if temp >= data[-1][0]:
        return data[-1][1]
    lo, hi = 0, len(data) - 1
    while hi - lo > 1:
        mid = lo + (hi - lo) // 2
        if data[mid][0] <= temp:
            lo = mid
        else:
            hi = mid
    t0, v0 = data[lo]
    t1, v1 = data[hi]
    factor = (temp - t0) / (t1 - t0)
    return v0 + (v1 - v0) * factor


def _sorted_lut(data):
    # Vec::sort_by is stable (material.rs:24-26, parser.rs:621-633)
    return sorted(data, key=lambda p: p[0])


class Project:
    """``project.rs:32-59``: mesh, materials, steps, loads, BCs, amplitudes."""

    def __init__(self):
        self.nodes: Dict[int, Tuple[float, float, float]] = {}
        self.elements: Dict[int, Tuple[str, List[int]]] = {}
        self.node_sets: Dict[str, List[int]] = {}
        self.element_sets: Dict[str, List[int]] = {}
        self.materials: List[Material] = []
        self.element_materials: Dict[int, int] = {}
        self.steps: List[StaticStep] = []
        self.initial_loads: List[Load] = []
        self.initial_bcs: List[Load] = []
        self.initial_temperatures: Dict[int, float] = {}
        self.default_initial_temperature: float = 0.0
        self.amplitudes: Dict[str, Amplitude] = {}
        self.nodal_output: List[str] = []
        self.element_output: List[str] = []
        self.index_to_node_id: np.ndarray = np.zeros(0, np.int64)
        self._id_to_index: Dict[int, int] = {}

    # mesh.rs:35-55
    def build_node_mappings(self) -> None:
        ids = np.array(sorted(self.nodes.keys()), dtype=np.int64)
        self.index_to_node_id = ids
        self._id_to_index = {int(i): k for k, i in enumerate(ids)}

    def get_index_for_node_id(self, node_id: int) -> Optional[int]:
        return self._id_to_index.get(int(node_id))

    @property
    def n_nodes(self) -> int:
        return len(self.nodes)

    @property
    def n_dof(self) -> int:
        return 3 * len(self.nodes)

    @staticmethod
    def from_jobname(path: str) -> "Project":
        """``project.rs:93-127``: appends ``.inp`` when missing; ``.gz``
        files are read decompressed."""
        p = str(path)
        if not (p.lower().endswith(".inp") or p.lower().endswith(".inp.gz")):
            p = p + ".inp"
        if p.endswith(".gz"):
            with gzip.open(p, "rt") as fh:
                text = fh.read()
        else:
            with open(p) as fh:
                text = fh.read()
        return Parser(preprocess_inp(text)).parse()

    @staticmethod
    def from_string(text: str) -> "Project":
        return Parser(preprocess_inp(text)).parse()

    # ------------------------------------------------------------------
    def arrays(self) -> "MeshArrays":
        """Flat arrays in the reference's numbering: nodes by ascending ID,
        elements by ascending ID, connectivity as node indices."""
        return MeshArrays.from_project(self)


# ---------------------------------------------------------------------------
# Preprocessing (parser.rs:903-942)
# ---------------------------------------------------------------------------
_CCX = re.compile(r"ITERATIVE CHOLESKY|ITERATIVE SCALING|PASTIX|PARDISO|SPOOLES")
_CCX_MAP = {
    "ITERATIVE CHOLESKY": "ITERATIVE",
    "ITERATIVE SCALING": "ITERATIVE",
    "PASTIX": "DIRECT",
    "PARDISO": "DIRECT",
    "SPOOLES": "DIRECT",
}


def preprocess_inp(text: str) -> str:
    out = []
    for line in text.splitlines():
        line = line.strip()
        if not line:
            continue
        line = line.upper()
        line = line + " " if line.endswith(",") else line + "\n"
        if line.startswith("**"):
            continue
        out.append(line)
    return substitute_ccx_solvers("".join(out))


def substitute_ccx_solvers(text: str) -> str:
    return _CCX.sub(lambda m: _CCX_MAP[m.group(0)], text)


def get_keyword_arguments(line: str) -> Dict[str, Optional[str]]:
    """``parser.rs:879-890``."""
    out: Dict[str, Optional[str]] = {}
    for s in line.split(",")[1:]:
        if "=" in s:
            k, v = s.split("=", 1)
            out[k.strip()] = v.strip()
        else:
            out[s.strip()] = None
    return out


def get_positional_arguments(line: str) -> List[str]:
    return [s.strip() for s in line.split(",")]


_KEYWORDS = {
    "NODE", "ELEMENT", "NSET", "ELSET", "MATERIAL", "ELASTIC", "DENSITY", "SOLID_SECTION", "STEP",
    "STATIC", "BOUNDARY", "CLOAD", "END_STEP", "HEADING", "NODE_FILE", "EL_FILE", "AMPLITUDE",
    "SURFACE", "DLOAD", "DSLOAD", "NODE_PRINT", "EL_PRINT", "INCLUDE", "BEAM_SECTION", "EXPANSION",
    "SHELL_SECTION", "CONDUCTIVITY", "SPECIFIC_HEAT", "INITIAL_CONDITIONS", "DEPVAR",
}


def _f64(s: str) -> float:
    s = s.strip()
    if "_" in s:
        raise ValueError(s)
    return float(s)


def _usize(s: str) -> int:
    s = s.strip()
    if not re.fullmatch(r"\+?[0-9]+", s):
        raise ValueError(s)
    return int(s)


class Parser:
    """``parser.rs:62-147``."""

    def __init__(self, text: str):
        self.lines = text.split("\n")
        if self.lines and self.lines[-1] == "":
            self.lines.pop()
        self.project = Project()
        self.current_keyword: Optional[str] = None
        self.line_nr = 0
        self.element_type: Optional[str] = None
        self.elset_name: Optional[str] = None
        self.set_name: Optional[str] = None
        self.is_generate = False
        self.current_step: Optional[StaticStep] = None
        self.step_counter = 0
        self.step_loads: List[Load] = []
        self.step_bcs: List[Load] = []
        self.initial_condition_type: Optional[str] = None
        self.pos = 0

    # line cursor helpers (a peekable iterator)
    def _peek(self) -> Optional[str]:
        return self.lines[self.pos] if self.pos < len(self.lines) else None

    def _next(self) -> Optional[str]:
        if self.pos < len(self.lines):
            s = self.lines[self.pos]
            self.pos += 1
            return s
        return None

    def parse(self) -> Project:
        while self.pos < len(self.lines):
            self.line_nr = self.pos + 1
            line = self._next()
            if line.startswith("*"):
                self._keyword(line)
                continue
            line = line.strip()
            if not line:
                continue
            if self.current_keyword is not None:
                self._data(line)
        if self.current_step is not None:
            step = self.current_step
            step.loads = list(self.step_loads)
            step.bcs = list(self.step_bcs)
            self.project.steps.append(step)
            self.current_step = None
        self.project.build_node_mappings()
        return self.project

    def _push_step(self):
        if self.current_step is not None:
            step = self.current_step
            step.loads = list(self.step_loads)
            step.bcs = list(self.step_bcs)
            self.project.steps.append(step)
            self.current_step = None

    def _keyword(self, line: str) -> None:
        parts = [p.strip() for p in line.split(",")]
        kw = parts[0][1:].strip().upper() if parts[0].startswith("*") else ""
        key = kw.replace(" ", "_")
        if key not in _KEYWORDS:
            raise UnsupportedKeyword(self.line_nr, kw)
        self.current_keyword = key
        p = self.project
        if key == "ELEMENT":
            a = get_keyword_arguments(line)
            self.element_type = a.get("TYPE")
            self.elset_name = a.get("ELSET")
        elif key in ("NSET", "ELSET"):
            a = get_keyword_arguments(line)
            self.set_name = a.get(key)
            self.is_generate = "GENERATE" in a
        elif key == "MATERIAL":
            a = get_keyword_arguments(line)
            name = a.get("NAME")
            if name is None:
                raise ParseError(self.line_nr, "Material name not found")
            p.materials.append(Material(name=name))
        elif key == "EXPANSION":
            a = get_keyword_arguments(line)
            if p.materials and a.get("ZERO") is not None:
                try:
                    p.materials[-1].reference_temperature = _f64(a["ZERO"])
                except ValueError as e:
                    raise ParseError(self.line_nr, f"Invalid ZERO value in *EXPANSION: {e}")
        elif key == "INITIAL_CONDITIONS":
            a = get_keyword_arguments(line)
            t = a.get("TYPE")
            self.initial_condition_type = t.upper() if t is not None else None
        elif key == "SOLID_SECTION":
            a = get_keyword_arguments(line)
            elset = a.get("ELSET")
            if elset is None:
                raise ParseError(self.line_nr, "Elset not found in *SOLID SECTION")
            mname = a.get("MATERIAL")
            if mname is None:
                raise ParseError(self.line_nr, "Material not found in *SOLID SECTION")
            idx = next((i for i, m in enumerate(p.materials) if m.name == mname), None)
            if idx is None:
                raise MaterialNotFound(f"Material {mname} not found")
            if elset not in p.element_sets:
                raise ElsetNotFound(f"Elset {elset} not found")
            for eid in p.element_sets[elset]:
                p.element_materials[eid] = idx
        elif key == "STEP":
            if self.step_counter == 0:
                self.step_loads = list(p.initial_loads)
                self.step_bcs = list(p.initial_bcs)
            self.step_counter += 1
            self._push_step()
            a = get_keyword_arguments(line)
            nlgeom = "NLGEOM" in a
            try:
                max_inc = _usize(a["INC"]) if a.get("INC") is not None else 10_000
            except ValueError:
                max_inc = 10_000
            nxt = self._peek()
            if nxt is not None and nxt.startswith("*STATIC"):
                sa = get_keyword_arguments(nxt)
                sv = sa.get("SOLVER")
                if sv is None:
                    solver = "DEFAULT"
                elif sv in ("DIRECT", "ITERATIVE"):
                    solver = sv
                else:
                    raise UnknownSolver(f"Unknown solver type: {sv}")
                self._next()
                inc = IncrementData(max_iterations=max_inc)
                dl = self._peek()
                if dl is not None and not dl.startswith("*"):
                    args = get_positional_arguments(dl)
                    if len(args) >= 4:
                        names = ["initial time increment", "time period", "min time increment",
                                 "max time increment"]
                        vals = []
                        for s, nm in zip(args[:4], names):
                            try:
                                vals.append(_f64(s))
                            except ValueError as e:
                                raise ParseError(self.line_nr, f"Invalid {nm}: {e}")
                        (inc.initial_time_increment, inc.time_period, inc.min_time_increment,
                         inc.max_time_increment) = vals
                    self._next()
                self.current_step = StaticStep(solver=solver, increment_data=inc, nlgeom=nlgeom)
        elif key == "END_STEP":
            self._push_step()
        elif key == "AMPLITUDE":
            a = get_keyword_arguments(line)
            name = a.get("NAME")
            if name is None:
                raise ParseError(self.line_nr, "Amplitude card is missing a Name= argument")
            total = a.get("TIME") == "TOTAL TIME"

            def _opt(k):
                try:
                    return _f64(a[k]) if a.get(k) is not None else 0.0
                except ValueError:
                    return 0.0

            nxt = self._peek()
            if nxt is not None:
                fields = nxt.split(",")
                t = []
                for s in fields[0::2]:
                    try:
                        t.append(_f64(s))
                    except ValueError:
                        pass
                vals = []
                for s in fields[1::2]:
                    try:
                        vals.append(_f64(s))
                    except ValueError as e:
                        raise ParseError(self.line_nr, f"Invalid amplitude value: {e}")
                p.amplitudes[name] = Amplitude(total, _opt("SHIFTX"), _opt("SHIFTY"), (t, vals))
        elif key == "BOUNDARY":
            a = get_keyword_arguments(line)
            amp = a.get("AMPLITUDE")
            if a.get("OP") == "NEW":
                if self.current_step is not None:
                    self.step_bcs = []
                else:
                    p.initial_bcs = []
            nxt = self._peek()
            if nxt is not None and not nxt.startswith("*"):
                self._boundary(nxt, amp)
        elif key == "CLOAD":
            a = get_keyword_arguments(line)
            amp = a.get("AMPLITUDE")
            if a.get("OP") == "NEW":
                if self.current_step is not None:
                    self.step_loads = []
                else:
                    p.initial_loads = []
            while True:
                nxt = self._peek()
                if nxt is None or nxt.startswith("*"):
                    break
                self._cload(self._next(), amp)

    def _data(self, line: str) -> None:
        k = self.current_keyword
        p = self.project
        if k == "NODE":
            f = [s.strip() for s in line.split(",")]
            try:
                if len(f) != 4:
                    raise ValueError
                p.nodes[_usize(f[0])] = (_f64(f[1]), _f64(f[2]), _f64(f[3]))
            except ValueError:
                raise ParseError(self.line_nr, f"Invalid node definition: {line}")
        elif k == "ELEMENT":
            if self.element_type is None:
                return
            try:
                nums = [_usize(s) for s in line.split(",")]
            except ValueError:
                raise ParseError(0, "Integer conversion failed")
            if not nums:
                raise ParseError(0, "Line empty")
            eid, nodes = nums[0], nums[1:]
            t = self.element_type
            if t not in ("C3D4", "C3D20"):
                raise ParseError(0, f"Unknown element definition: {t}")
            need = 4 if t == "C3D4" else 20
            if len(nodes) != need:
                raise ParseError(0, f"Wrong node count for {t}")
            if self.elset_name:
                p.element_sets.setdefault(self.elset_name, []).append(eid)
            p.elements[eid] = (t, nodes)
        elif k == "NSET":
            if self.set_name is not None and not self.is_generate:
                ids = []
                for s in line.split(","):
                    s = s.strip()
                    if s:
                        try:
                            ids.append(_usize(s))
                        except ValueError:
                            pass
                p.node_sets.setdefault(self.set_name, []).extend(ids)
        elif k == "ELSET":
            if self.set_name is not None and not self.is_generate:
                parts = []
                for s in line.split(","):
                    s = s.strip()
                    if s:
                        try:
                            parts.append(_usize(s))
                        except ValueError:
                            pass
                if "," not in line and not parts:
                    other = p.element_sets.get(line.strip())
                    if other is not None:
                        p.element_sets.setdefault(self.set_name, []).extend(list(other))
                else:
                    p.element_sets.setdefault(self.set_name, []).extend(parts)
        elif k in ("ELASTIC", "DENSITY", "EXPANSION"):
            if not p.materials:
                return
            try:
                parts = [_f64(s) for s in line.split(",")]
            except ValueError as e:
                what = {"ELASTIC": "elastic constant", "DENSITY": "density value",
                        "EXPANSION": "expansion constant"}[k]
                raise ParseError(self.line_nr, f"Invalid {what}: {e}")
            m = p.materials[-1]
            if k == "ELASTIC":
                if len(parts) >= 2:
                    temp = parts[2] if len(parts) >= 3 else 0.0
                    m.youngs_modulus = _sorted_lut((m.youngs_modulus or []) + [(temp, parts[0])])
                    m.poisson_ratio = _sorted_lut((m.poisson_ratio or []) + [(temp, parts[1])])
            elif parts:
                temp = parts[1] if len(parts) >= 2 else 0.0
                if k == "DENSITY":
                    m.density = _sorted_lut((m.density or []) + [(temp, parts[0])])
                else:
                    m.thermal_expansion = _sorted_lut((m.thermal_expansion or []) + [(temp, parts[0])])
        elif k == "INITIAL_CONDITIONS":
            if self.initial_condition_type == "TEMPERATURE":
                f = [s.strip() for s in line.split(",")]
                if len(f) >= 2:
                    try:
                        temp = _f64(f[1])
                    except ValueError as e:
                        raise ParseError(self.line_nr, f"Invalid temperature value: {e}")
                    for nid in self._resolve(f[0]):
                        p.initial_temperatures[nid] = temp
        elif k == "DEPVAR":
            if p.materials:
                f = [s.strip() for s in line.split(",")]
                try:
                    p.materials[-1].num_depvars = _usize(f[0])
                except ValueError as e:
                    raise ParseError(self.line_nr, f"Invalid DEPVAR value: {e}")
        elif k == "NODE_FILE":
            p.nodal_output.extend(s.strip() for s in line.split(","))
        elif k == "EL_FILE":
            p.element_output.extend(s.strip() for s in line.split(","))

    def _resolve(self, target: str) -> List[int]:
        t = target.strip()
        if t in self.project.node_sets:
            return list(self.project.node_sets[t])
        try:
            return [_usize(t)]
        except ValueError:
            return []

    def _amp(self, name: Optional[str]) -> Amplitude:
        if name is None:
            return Amplitude()
        a = self.project.amplitudes.get(name)
        return Amplitude(a.total_time, a.shift_x, a.shift_y, a.data) if a is not None else Amplitude()

    def _origin(self) -> int:
        return self.step_counter if self.current_step is not None else 0

    def _cload(self, line: str, amp: Optional[str]) -> None:
        f = [s.strip() for s in line.split(",")]
        if len(f) < 3:
            return
        try:
            dof = _usize(f[1])
        except ValueError as e:
            raise ParseError(self.line_nr, f"Invalid DOF: {e}")
        try:
            val = _f64(f[2])
        except ValueError as e:
            raise ParseError(self.line_nr, f"Invalid load value: {e}")
        if 1 <= dof <= 3:
            for nid in self._resolve(f[0]):
                ld = Load(nid, dof - 1, val, self._amp(amp), self._origin())
                (self.step_loads if self.current_step is not None else self.project.initial_loads).append(ld)

    def _boundary(self, line: str, amp: Optional[str]) -> None:
        f = [s.strip() for s in line.split(",")]
        if len(f) < 2:
            return
        try:
            first = _usize(f[1])
        except ValueError as e:
            raise ParseError(self.line_nr, f"Invalid first DOF: {e}")
        if len(f) > 2 and f[2].strip():
            try:
                last = _usize(f[2])
            except ValueError as e:
                raise ParseError(self.line_nr, f"Invalid last DOF: {e}")
        else:
            last = first
        if len(f) > 3:
            try:
                val = _f64(f[3])
            except ValueError as e:
                raise ParseError(self.line_nr, f"Invalid boundary value: {e}")
        else:
            val = 0.0
        for nid in self._resolve(f[0]):
            for d in range(first, last + 1):
                if 1 <= d <= 3:
                    bc = Load(nid, d - 1, val, self._amp(amp), self._origin())
                    (self.step_bcs if self.current_step is not None else self.project.initial_bcs).append(bc)


# ---------------------------------------------------------------------------
# Flat arrays for the C ABI
# ---------------------------------------------------------------------------
LUT_MAX = 16


@dataclass
class MeshArrays:
    """Flat, index-ordered mesh for ``fx_mesh_upload``.

    node order = ascending node ID; element order = ascending element ID;
    ``conn[el_ptr[e]:el_ptr[e+1]]`` are node indices; ``mat`` is the material
    index per element; LUT tables are ``n_mat x LUT_MAX``."""

    node_ids: np.ndarray
    xyz: np.ndarray
    elem_ids: np.ndarray
    el_ptr: np.ndarray
    conn: np.ndarray
    mat: np.ndarray
    lut_n_e: np.ndarray
    lut_t_e: np.ndarray
    lut_v_e: np.ndarray
    lut_n_nu: np.ndarray
    lut_t_nu: np.ndarray
    lut_v_nu: np.ndarray

    @property
    def n_nodes(self) -> int:
        return int(self.xyz.shape[0])

    @property
    def n_el(self) -> int:
        return int(self.el_ptr.shape[0] - 1)

    @property
    def n_dof(self) -> int:
        return 3 * self.n_nodes

    @staticmethod
    def from_project(p: Project) -> "MeshArrays":
        ids = p.index_to_node_id
        xyz = np.array([p.nodes[int(i)] for i in ids], dtype=np.float64).reshape(-1, 3)
        eids = np.array(sorted(p.elements.keys()), dtype=np.int64)
        el_ptr = np.zeros(len(eids) + 1, np.int64)
        conn = []
        mat = np.zeros(len(eids), np.int32)
        for k, eid in enumerate(eids):
            _, nodes = p.elements[int(eid)]
            idx = []
            for nid in nodes:
                j = p.get_index_for_node_id(nid)
                if j is None:
                    raise FerrixError(f"Node {nid} not found in mesh")
                idx.append(j)
            conn.extend(idx)
            el_ptr[k + 1] = el_ptr[k] + len(nodes)
            m = p.element_materials.get(int(eid))
            if m is None:
                raise InvalidModelState(f"Element {int(eid)} has no material assigned.")
            mat[k] = m
        tabs = material_tables(p.materials)
        return MeshArrays(ids, xyz, eids, el_ptr, np.array(conn, np.int32), mat, *tabs)


def material_tables(materials: List[Material]):
    """Per-material E(T) and nu(T) LUTs padded to LUT_MAX points.  A missing
    E or nu is encoded as count 0 (the kernels report InvalidModelState)."""
    n = max(1, len(materials))
    ne = np.zeros(n, np.int32)
    te = np.zeros((n, LUT_MAX))
    ve = np.zeros((n, LUT_MAX))
    nn_ = np.zeros(n, np.int32)
    tn = np.zeros((n, LUT_MAX))
    vn = np.zeros((n, LUT_MAX))
    for k, m in enumerate(materials):
        for lut, cnt, tt, vv in ((m.youngs_modulus, ne, te, ve), (m.poisson_ratio, nn_, tn, vn)):
            if lut is None:
                continue
            if len(lut) > LUT_MAX:
                raise InvalidModelState(f"material {m.name}: more than {LUT_MAX} LUT points")
            cnt[k] = len(lut)
            for j, (t, v) in enumerate(lut):
                tt[k, j] = t
                vv[k, j] = v
    return ne, te, ve, nn_, tn, vn


def initial_node_temperatures(p: Project) -> np.ndarray:
    """``state.rs:47-64``: default temperature, then *INITIAL CONDITIONS."""
    t = np.full(p.n_nodes, p.default_initial_temperature, dtype=np.float64)
    for nid, temp in p.initial_temperatures.items():
        j = p.get_index_for_node_id(nid)
        if j is not None:
            t[j] = temp
    return t


def deck_stem(path: str) -> str:
    """The job name of a deck path: its file stem (``Project::jobname``,
    ``project.rs:133-147``), with a ``.gz`` of a compressed deck dropped
    first."""
    base = os.path.basename(str(path))
    if base.lower().endswith(".gz"):
        base = base[:-3]
    return os.path.splitext(base)[0]
