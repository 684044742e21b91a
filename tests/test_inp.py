"""Host-side model ingestion: the .inp keyword surface restated from
src/solver/parser.rs, checked against the reference's own parser and
material-parsing tests and its documented quirks."""
import numpy as np
import pytest

from ferrix_amd import meshgen
from ferrix_amd.inp import (Project, UnsupportedKeyword, get_keyword_arguments, get_positional_arguments,
                            preprocess_inp)
from ferrix_amd.model import Job

from conftest import deck


def test_preprocess_reference_case():
    # parser.rs:950-957
    inp = "**comment\n word  \n \t*keyword\n123.4\n4, 5, 6,\n7, 8, 9"
    assert preprocess_inp(inp) == "WORD\n*KEYWORD\n123.4\n4, 5, 6, 7, 8, 9\n"


def test_keyword_and_positional_arguments():
    # parser.rs:959-971
    assert get_keyword_arguments("*STEP , INC =  100")["INC"] == "100"
    assert get_positional_arguments("0.1, 1, 1E-05, 0.2") == ["0.1", "1", "1E-05", "0.2"]


def test_ccx_solver_names_map():
    assert "SOLVER=DIRECT" in preprocess_inp("*Static, Solver=Spooles")
    assert "SOLVER=ITERATIVE" in preprocess_inp("*static, solver=iterative cholesky")


def test_temperature_dependent_material():
    # tests/material_parsing.rs:5-37
    p = Project.from_string("""
*MATERIAL, NAME=TestMat
*DENSITY
1000.0, 0.0
900.0, 500.0
*ELASTIC
210000.0, 0.3, 0.0
150000.0, 0.35, 500.0
""")
    m = p.materials[0]
    assert m.name == "TESTMAT"
    from ferrix_amd.inp import lut_interpolate as li

    assert li(m.density, 250.0) == 950.0 and li(m.density, 600.0) == 900.0
    assert li(m.youngs_modulus, 250.0) == 180000.0
    assert abs(li(m.poisson_ratio, 250.0) - 0.325) < 1e-12


def test_expansion_and_initial_conditions():
    # tests/material_parsing.rs:39-60
    p = Project.from_string("""
*MATERIAL, NAME=ThermalMat
*EXPANSION, ZERO=20.0
1.2E-5, 0.0
1.5E-5, 500.0
*INITIAL CONDITIONS, TYPE=TEMPERATURE
1, 100.0
""")
    m = p.materials[0]
    assert m.reference_temperature == 20.0
    from ferrix_amd.inp import lut_interpolate as li

    assert li(m.thermal_expansion, 250.0) == 1.35e-5
    assert p.initial_temperatures[1] == 100.0


def test_unsupported_keyword():
    with pytest.raises(UnsupportedKeyword):
        Project.from_string("*FOO\n1\n")


def test_reference_deck_counts_and_numbering():
    p = Project.from_jobname(deck("C3D20"))
    assert (len(p.nodes), len(p.elements)) == (6330, 1116)
    ids = p.index_to_node_id
    assert np.all(np.diff(ids) > 0)
    assert p.get_index_for_node_id(int(ids[17])) == 17
    st = p.steps[0]
    assert st.solver == "DIRECT" and not st.nlgeom
    # *BOUNDARY op=New then a second *BOUNDARY whose first line is the fixed
    # set with DOFs 1..6 -> 3 records per node (DOFs 4-6 ignored)
    nset = p.node_sets["INTERNAL_SELECTION-1_FIXED-1"]
    assert len(st.bcs) == 3 * len(nset)


def test_cload_lines_after_node_print_are_dropped():
    """C3D4_2staticstep.inp: step 2 CLOAD lines after '*Node print' fall
    under that keyword and are ignored (parser.rs:462-470, 506)."""
    p = Project.from_jobname(deck("C3D4_2staticstep"))
    s2 = p.steps[1]
    # 145 lines before the interruption (first line of step 2's CLOAD block
    # through '436, 1, ...'), each one record
    assert len(s2.loads) == 145
    assert all(ld.origin_step == 2 for ld in s2.loads)


def test_nlgeom_deck_steps():
    p = Project.from_jobname(deck("NLGEOM_test"))
    assert [s.nlgeom for s in p.steps] == [True, True]
    # step 2 inherits the z loads of step 1 (OP=MOD) and adds x loads
    assert len(p.steps[0].loads) == 3 and len(p.steps[1].loads) == 6


def test_synthetic_deck_round_trip():
    job = meshgen.c3d4_cantilever()
    p = Project.from_string(meshgen.to_inp(job))
    j2 = Job.from_project(p)
    assert np.array_equal(j2.mesh.xyz, job.mesh.xyz)
    assert np.array_equal(j2.mesh.conn, job.mesh.conn)
    s1, s2 = job.steps[0], j2.steps[0]
    assert np.array_equal(s1.load_dof, s2.load_dof) and np.array_equal(s1.load_val, s2.load_val)
    assert np.array_equal(s1.bc_dof, s2.bc_dof)
