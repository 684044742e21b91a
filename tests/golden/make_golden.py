"""Writes the golden fixtures from the CPU oracle (run from the repo root):

    python tests/golden/make_golden.py [cfg1 cfg2 cfg3 ...]

For every config: the oracle's filtered CSR pattern size and FNV-1a hash
(row_ptr as int64 LE, then col as int32 LE), max_diag, and the converged
displacement of the oracle's step driver with the DIRECT stand-in (SuperLU)
-- the FP64 CPU oracle the GPU results are compared against.  cfg3 is the
NLGEOM beam (10 increments + the 1e-16 micro-increment); its per-increment
NR iteration counts are stored too.
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402
from ferrix_amd import meshgen  # noqa: E402
from ferrix_amd.model import SOLVER_DIRECT  # noqa: E402


def fnv1a(*arrays) -> str:
    h = 0xCBF29CE484222325
    for a in arrays:
        for b in a.tobytes():
            h ^= b
            h = (h * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return f"{h:016x}"


def fnv1a_np(row_ptr, col) -> str:
    # byte-wise FNV-1a, vectorised per chunk (same result as fnv1a)
    h = np.uint64(0xCBF29CE484222325)
    prime = np.uint64(0x100000001B3)
    data = np.concatenate([np.ascontiguousarray(row_ptr, "<i8").view(np.uint8),
                           np.ascontiguousarray(col, "<i4").view(np.uint8)])
    with np.errstate(over="ignore"):
        for b in data:
            h = (h ^ np.uint64(b)) * prime
    return f"{int(h):016x}"


CONFIGS = {
    "cfg1": meshgen.cfg1_c3d4_cantilever,
    "cfg2": meshgen.cfg2_c3d20_block,
    "cfg3": meshgen.cfg3_c3d20_nlgeom_beam,
}


def main(names):
    for name in names:
        t0 = time.time()
        job = CONFIGS[name]()
        csr, md, _ = O.assemble_csr(job.mesh, job.node_temp)
        meta = dict(config=name, workload=job.name, n_nodes=job.mesh.n_nodes, n_el=job.mesh.n_el,
                    n_dof=job.n_dof, nnz=csr.nnz, max_diag=md,
                    pattern_fnv1a=fnv1a(np.ascontiguousarray(csr.row_ptr, "<i8"),
                                        np.ascontiguousarray(csr.col, "<i4")) if csr.nnz < 200_000 else None)
        u, logs = O.run_job(job, solver_override=SOLVER_DIRECT)
        meta["linear_solves"] = sum(lg.linear_solves for lg in logs)
        meta["increments"] = [[len(i["iters"]), i["converged"], i["dt"]] for lg in logs for i in lg.increments]
        meta["u_max_abs"] = float(np.max(np.abs(u)))
        meta["seconds"] = time.time() - t0
        np.savez_compressed(os.path.join(HERE, f"{name}_oracle.npz"), u=u, row_ptr_sum=np.int64(csr.row_ptr.sum()),
                            col_sum=np.int64(csr.col.astype(np.int64).sum()))
        with open(os.path.join(HERE, f"{name}_oracle.json"), "w") as fh:
            json.dump(meta, fh, indent=1)
        print(json.dumps(meta), flush=True)


if __name__ == "__main__":
    main(sys.argv[1:] or list(CONFIGS))
