"""Writes tests/golden/population_hard.npz (run from the repo root):

    python tests/golden/make_population_golden.py

The oracle's direct solves (SuperLU on the oracle's penalised CSR, the
stand-in for the reference's DIRECT solver) of the hardest members of the
batched-solve population (seed 4, 512 meshes): the two largest, the most
slender (aspect near 10, longest in cells), the ones with nu near 0.45, and
member 216 whose mixed solve once stalled.  The fixture keeps, per member
and load case, every 32nd displacement entry plus the full vector's L2 norm
and max |u|, enough for a relative-error check at small size.
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

MEMBERS = [160, 281, 80, 47, 365, 265, 67, 216]
STRIDE = 32


def rhs_of(jobs, pen):
    """Every load case's right-hand side after the penalty (all share the BC
    set of the first), as the batched driver builds them."""
    st0 = jobs[0].steps[0]
    rows = []
    for j in jobs:
        st = j.steps[0]
        b = np.zeros(j.n_dof)
        np.add.at(b, st.load_dof, st.load_val)
        np.add.at(b, np.asarray(st0.bc_dof), pen * np.asarray(st0.bc_val, np.float64))
        rows.append(b)
    return np.array(rows)


def main():
    specs = meshgen.population_specs()
    out, meta = {}, []
    for k in MEMBERS:
        t0 = time.time()
        jobs = meshgen.population_mesh(specs[k])
        csr, md, pen = O.assemble_csr(jobs[0].mesh, jobs[0].node_temp, None, jobs[0].steps[0].bc_dof)
        B = rhs_of(jobs, pen)
        U = np.array([O.direct_solve(csr, b) for b in B])
        out[f"m{k}_sub"] = U[:, ::STRIDE].copy()
        out[f"m{k}_norm"] = np.linalg.norm(U, axis=1)
        out[f"m{k}_maxabs"] = np.max(np.abs(U), axis=1)
        sp = specs[k]
        info = dict(member=k, n_dof=jobs[0].n_dof, n_rhs=len(jobs), cells=[sp["nx"], sp["ny"], sp["nz"]],
                    nu=sp["nu"], E=sp["E"], nnz=csr.nnz, seconds=round(time.time() - t0, 1))
        meta.append(info)
        print(json.dumps(info), flush=True)
    out["members"] = np.array(MEMBERS, np.int64)
    out["stride"] = np.int64(STRIDE)
    np.savez_compressed(os.path.join(HERE, "population_hard.npz"), **out)
    with open(os.path.join(HERE, "population_hard.json"), "w") as fh:
        json.dump(meta, fh, indent=1)


if __name__ == "__main__":
    main()
