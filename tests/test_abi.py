"""The C-ABI library loads and exports every symbol include/ferrix_hip.h
declares; the ctypes layouts match the C structs; calls fail loudly (no CPU
fallback) when no device is present.  No compute calls: CPU only."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from ferrix_amd import _lib as L
from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "ferrix_hip.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:fx_status|void|const char \*)\s*(fx_\w+)\s*\(", text, re.M)))


def test_header_declares_expected_entry_points():
    names = declared_functions()
    for must in ("fx_ctx_create", "fx_mesh_upload", "fx_assemble", "fx_internal_force", "fx_apply_penalty",
                 "fx_solve", "fx_static_step", "fx_matrix_get", "fx_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = L.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # and the binding covers exactly the header
    assert set(L.SIGNATURES) == set(declared_functions())


def test_no_device_fails_loudly():
    lib = L.load()
    h = C.c_void_p()
    rc = lib.fx_ctx_create(0, C.byref(h))
    if rc == L.FX_OK:  # a GPU is present: nothing to check here
        lib.fx_ctx_destroy(h)
        pytest.skip("a device is present")
    assert rc == L.FX_ERR_NO_DEVICE
    assert len(lib.fx_last_error(None)) > 0


def test_null_handles_are_rejected():
    lib = L.load()
    assert lib.fx_assemble(None, 0, None, None) == L.FX_ERR_INVALID
    assert lib.fx_solve(None, None, None, None, None) == L.FX_ERR_INVALID
    assert lib.fx_static_step(None, None, None, None) == L.FX_ERR_INVALID
    lib.fx_mesh_destroy(None)
    lib.fx_ctx_destroy(None)


def test_solve_opts_default_is_reference():
    lib = L.load()
    o = L.fx_solve_opts()
    lib.fx_solve_opts_default(C.byref(o))
    assert (o.precision, o.stop_mode, o.tol, o.max_iter) == (L.FX_PREC_FP64, L.FX_STOP_ABS_SQRT_RZ, 1e-8, 1000)


STRUCTS = ["fx_material_table", "fx_solve_opts", "fx_solve_report", "fx_amplitudes", "fx_step_def", "fx_timer",
           "fx_step_report"]


def test_struct_layouts_match_the_header():
    """Compile a probe with gcc against the header and compare sizeof and
    field offsets with the ctypes mirrors."""
    lines = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{HEADER}"', "int main(void){"]
    for s in STRUCTS:
        lines.append(f'printf("{s} %zu\\n", sizeof({s}));')
        for f, _ in getattr(L, s)._fields_:
            lines.append(f'printf("{s}.{f} %zu\\n", offsetof({s}, {f}));')
    lines.append("return 0;}")
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "probe.c")
        exe = os.path.join(d, "probe")
        open(src, "w").write("\n".join(lines))
        subprocess.run(["gcc", "-std=c11", "-o", exe, src], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    got = dict(line.rsplit(" ", 1) for line in out.strip().splitlines())
    for s in STRUCTS:
        cs = getattr(L, s)
        assert int(got[s]) == C.sizeof(cs), s
        for f, _ in cs._fields_:
            assert int(got[f"{s}.{f}"]) == getattr(cs, f).offset, f"{s}.{f}"


def test_batch_null_handles_are_rejected():
    lib = L.load()
    out = C.c_void_p()
    assert lib.fx_batch_create(None, None, C.byref(out)) == L.FX_ERR_INVALID
    slot = C.c_int32()
    assert lib.fx_batch_add(None, None, 1, None, C.byref(slot)) == L.FX_ERR_INVALID
    assert lib.fx_batch_run(None, 0, None) == L.FX_ERR_INVALID
    assert lib.fx_batch_poll(None, 0, None, None) == L.FX_ERR_INVALID
    assert lib.fx_batch_get(None, 0, None) == L.FX_ERR_INVALID
    assert lib.fx_batch_remove(None, 0) == L.FX_ERR_INVALID
    lib.fx_batch_destroy(None)
