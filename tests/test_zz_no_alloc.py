"""GPU: fx_solve_csr_noalloc, the solve whose work vectors live in the
library's static per-device scratch array (no device allocation), against
fx_solve_csr on the same matrices through the C ABI.

Both run the same kernels in the same order on the same values, so the bar
is bit-for-bit equality of the solutions and equal iteration counts; each
is also within 1e-8 of the oracle's direct solve.  Two calls in a row, the
second on a smaller system, check that nothing of the first call's scratch
contents leaks into the second.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from ferrix_amd import _lib as L
from ferrix_amd import meshgen
from ferrix_amd.device import SolveOptions

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        try:
            _hip = C.CDLL("libamdhip64.so")
        except OSError:
            _hip = C.CDLL("/opt/rocm/lib/libamdhip64.so")
    return _hip


class DevArray:
    """A device copy of a host array (test-side allocation)."""

    def __init__(self, a):
        self.a = np.ascontiguousarray(a)
        self.p = C.c_void_p()
        nb = max(self.a.nbytes, 8)
        assert hip().hipMalloc(C.byref(self.p), C.c_size_t(nb)) == 0
        if self.a.nbytes:
            assert hip().hipMemcpy(self.p, self.a.ctypes.data_as(C.c_void_p), C.c_size_t(self.a.nbytes), 1) == 0

    def get(self):
        out = np.empty_like(self.a)
        assert hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        if self.p:
            hip().hipFree(self.p)
            self.p = C.c_void_p()


def system(job):
    st = job.steps[0]
    csr, md, pen = O.assemble_csr(job.mesh, job.node_temp, None, st.bc_dof)
    b = np.zeros(job.n_dof)
    np.add.at(b, st.load_dof, st.load_val)
    np.add.at(b, np.asarray(st.bc_dof), pen * np.asarray(st.bc_val, np.float64))
    return csr, b


def solve_both(ctx, csr, b, opts):
    lib = ctx.lib
    n = csr.n
    rp = np.ascontiguousarray(csr.row_ptr, np.int64)
    ci = np.ascontiguousarray(csr.col, np.int64)
    v = np.ascontiguousarray(csr.val, np.float64)
    o = opts.c()
    # the allocating entry point, host arrays
    x_h = np.zeros(n)
    rep_h = L.fx_solve_report()
    ctx.check(lib.fx_solve_csr(ctx.h, n, rp.ctypes.data_as(C.c_void_p), ci.ctypes.data_as(C.c_void_p),
                               v.ctypes.data_as(C.c_void_p), b.ctypes.data_as(C.c_void_p),
                               x_h.ctypes.data_as(C.c_void_p), C.byref(o), C.byref(rep_h)))
    # the no-allocation entry point, device arrays
    d = [DevArray(rp), DevArray(ci), DevArray(v), DevArray(b), DevArray(np.full(n, np.nan))]
    try:
        rep_d = L.fx_solve_report()
        ctx.check(lib.fx_solve_csr_noalloc(ctx.h, n, int(rp[-1]), d[0].p, d[1].p, d[2].p, d[3].p, d[4].p,
                                           C.byref(o), C.byref(rep_d)))
        x_d = d[4].get()
    finally:
        for a in d:
            a.free()
    return x_h, rep_h, x_d, rep_d


def test_noalloc_matches_allocating_solve_twice(ctx):
    opts = SolveOptions(stop_mode=L.FX_STOP_REL_R2, tol=1e-12, max_iter=100_000)
    big = system(meshgen.c3d20_block(seed=7, nx=10, ny=4, nz=3))  # 2,475 nodes, 7,425 DOF
    small = system(meshgen.c3d4_cantilever())                     # 525 nodes, 1,575 DOF
    assert small[0].n < big[0].n
    for name, (csr, b) in (("c3d20 10x4x3", big), ("cantilever c3d4", small)):
        print("case", name, "n", csr.n, "nnz", csr.nnz, flush=True)
        x_h, rep_h, x_d, rep_d = solve_both(ctx, csr, b, opts)
        assert rep_d.iterations == rep_h.iterations
        assert rep_d.final_criterion == rep_h.final_criterion
        assert np.array_equal(x_d, x_h), "no-allocation solve differs from fx_solve_csr"
        xd = O.direct_solve(csr, b)
        assert np.linalg.norm(x_d - xd) / np.linalg.norm(xd) < 1e-8


def test_noalloc_reference_rule(ctx):
    """The reference's stop rule (sqrt(r.z) < 1e-8, 1000 iterations)."""
    csr, b = system(meshgen.c3d4_cantilever())
    x_h, rep_h, x_d, rep_d = solve_both(ctx, csr, b, SolveOptions())
    assert rep_d.iterations == rep_h.iterations
    assert np.array_equal(x_d, x_h)


def test_noalloc_rejects_what_it_cannot_hold_or_check(ctx):
    lib = ctx.lib
    o = SolveOptions().c()
    rep = L.fx_solve_report()
    rp = DevArray(np.array([0, 2, 3], np.int64))
    ci = DevArray(np.array([0, 2, 1], np.int64))  # column 2 out of range for n = 2
    v = DevArray(np.ones(3))
    b = DevArray(np.ones(2))
    x = DevArray(np.zeros(2))
    try:
        # too large for the scratch array: rejected before anything runs
        rc = lib.fx_solve_csr_noalloc(ctx.h, 5_000_000, 0, rp.p, ci.p, v.p, b.p, x.p, C.byref(o), C.byref(rep))
        assert rc == L.FX_ERR_INVALID
        assert b"scratch" in lib.fx_last_error(ctx.h)
        # a malformed matrix: the device check rejects it before any gather
        rc = lib.fx_solve_csr_noalloc(ctx.h, 2, 3, rp.p, ci.p, v.p, b.p, x.p, C.byref(o), C.byref(rep))
        assert rc == L.FX_ERR_INVALID
        assert b"column" in lib.fx_last_error(ctx.h)
        # row_ptr[n] != nnz
        rc = lib.fx_solve_csr_noalloc(ctx.h, 2, 2, rp.p, ci.p, v.p, b.p, x.p, C.byref(o), C.byref(rep))
        assert rc == L.FX_ERR_INVALID
        assert b"row_ptr" in lib.fx_last_error(ctx.h)
    finally:
        for a in (rp, ci, v, b, x):
            a.free()
