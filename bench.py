"""Benchmark: DOF-solves/s and p50/p99 step latency of the static-step hot path
on one MI355X, mixed precision vs FP64 PCG.

Workload (default): linear *STATIC step on a 400x100x100 mm block meshed
with 40x10x10 C3D20 bricks (18,821 nodes, 56,463 DOF), coordinates jittered
by +/-0.5 mm (SplitMix64 seed 2), E=70000, nu=0.33, x=0 clamped, -1000 N tip
load; Jacobi-PCG to ||r||/||b|| <= 1e-10.  One step = element stiffness +
internal force + deterministic assembly + penalty BCs + PCG solve + the
convergence check at the new displacement (the reference's linear step:
two f_int passes, one K pass, one solve).  Mesh, loads and BCs are resident
in HBM before timing; every timed step starts from u = 0 on the device.

Arms: "fp64" (the reference's Jacobi-PCG, all FP64) and "mixed" (FP32-stored
K inside FP64 iterative refinement).  value = DOF-solves/s of the mixed arm
(n_dof x linear solves / wall seconds over the K timed steps); both arms'
numbers, the U error against the FP64 CPU oracle's solution and p50/p99
per-step latency are in the JSON line.

roofline: the dominant kernel is the PCG SpMV (k_spmv_p).  Its mean launch
duration is measured with HIP events on the library's stream
(fx_profile_spmv); achieved = algorithmic bytes per launch / duration.

cpu_baseline: the oracle (C restatement of the reference CPU path: serial
element loop + triplets + CSR, Jacobi-PCG with a row-parallel SpMV on 4
OpenMP threads, the reference's default rayon pool size), timed on a bounded
sample (full assembly + 200 PCG iterations) and extrapolated to the
iterations the FP64 solve needs.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0  # MI355X HBM3E


def device_sync():
    """hipDeviceSynchronize through the HIP runtime (no torch import)."""
    try:
        lib = ctypes.CDLL("libamdhip64.so")
    except OSError:
        lib = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
    rc = lib.hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed: {rc}")


def calib_copy(mib, reps=5):
    """hipMemcpyDtoD of `mib` MiB, `reps` times: a copy of known size whose
    FETCH_SIZE / WRITE_SIZE readings calibrate the counters."""
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(n)) == 0
    for _ in range(reps):
        assert hip.hipMemcpyDtoD(b, a, ctypes.c_size_t(n)) == 0
    device_sync()
    hip.hipFree(a)
    hip.hipFree(b)
    print(f"# calib copy {mib} MiB x {reps}", flush=True)


def cpu_baseline(job, fp64_iters, threads=4, sample_iters=200):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    st = job.steps[0]
    t0 = time.perf_counter()
    f_int = O.internal_force(job.mesh, job.node_temp, np.zeros(job.n_dof))
    t1 = time.perf_counter()
    r, c, v, md = O.assemble_triplets(job.mesh, job.node_temp)
    t2 = time.perf_counter()
    bd = np.asarray(st.bc_dof, np.int32)
    K = O.triplets_to_csr(job.n_dof, np.concatenate([r, bd]), np.concatenate([c, bd]),
                          np.concatenate([v, np.full(len(bd), md * 1e6)]))
    t3 = time.perf_counter()
    b = np.zeros(job.n_dof)
    np.add.at(b, st.load_dof, st.load_val)
    t_it = O.lib().fxo_pcg_time_iters(K.n, O._p(K.row_ptr), O._p(K.col), O._p(K.val), O._p(b), sample_iters,
                                      threads) / sample_iters
    t_fint = t1 - t0
    step_s = 2 * t_fint + (t2 - t1) + (t3 - t2) + fp64_iters * t_it
    return {
        "value": job.n_dof / step_s, "unit": "DOF-solves/s", "cores": threads, "kind": "port",
        "sample": (f"oracle C restatement on the host: full f_int pass ({t_fint:.2f} s), full serial element+"
                   f"triplet assembly ({t2 - t1:.2f} s), triplet->CSR ({t3 - t2:.2f} s), {sample_iters} Jacobi-PCG "
                   f"iterations on {threads} OpenMP threads ({t_it * 1e3:.2f} ms/iter) extrapolated to "
                   f"{fp64_iters} iterations; step = 2 f_int + assembly + CSR + solve = {step_s:.2f} s"),
        "seconds_per_step": step_s,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--arms", default="fp64,mixed")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--spmv-reps", type=int, default=100)
    ap.add_argument("--inner-tol", type=float, default=1e-3)
    ap.add_argument("--sweep-groups", action="store_true", help="print SpMV time per lane-group size")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="only run device-to-device copies of this many MiB (counter calibration) and exit")
    a = ap.parse_args()
    if a.calib_copy_mb > 0:
        calib_copy(a.calib_copy_mb)
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only; multi-GPU runs are not part of it "
              "(the batched and serving workloads shard by independent mesh, see the design notes).",
              file=sys.stderr)
        sys.exit(2)

    from ferrix_amd import _lib as L
    from ferrix_amd import meshgen
    from ferrix_amd.device import Context, DeviceMesh, SolveOptions, step_def

    job = meshgen.cfg2_c3d20_block()
    st = job.steps[0]
    golden = os.path.join(ROOT, "tests", "golden", "cfg2_oracle.npz")
    u_ref = np.load(golden)["u"] if os.path.exists(golden) else None
    print(f"# workload {job.name}: {job.mesh.n_nodes} nodes, {job.mesh.n_el} C3D20, {job.n_dof} DOF", flush=True)

    ctx = Context(0)
    dm = DeviceMesh(ctx, job.mesh, job.node_temp)
    arms = {}
    for arm in a.arms.split(","):
        prec = L.FX_PREC_MIXED if arm == "mixed" else L.FX_PREC_FP64
        so = SolveOptions(precision=prec, stop_mode=L.FX_STOP_REL_R2, tol=st.tol, max_iter=st.max_iter,
                          inner_tol=a.inner_tol, max_refine=100)
        sd, keep = step_def(st, 1, so, so)
        lat, reps = [], []

        def one():
            dm.zero_u()
            timer = L.fx_timer(0.0, 0.0, 1.0)
            rep = L.fx_step_report()
            ctx.check(ctx.lib.fx_static_step(dm.h, ctypes.byref(sd), ctypes.byref(timer), ctypes.byref(rep)))
            return rep

        for _ in range(a.warmup):
            one()
        device_sync()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            ts = time.perf_counter()
            reps.append(one())
            lat.append(time.perf_counter() - ts)
        device_sync()
        wall = time.perf_counter() - t0
        u = dm.get_u()
        solves = sum(r.linear_solves for r in reps)
        arm_res = {
            "dof_solves_per_s": job.n_dof * solves / wall,
            "ms_per_step": 1e3 * wall / a.steps,
            "p50_ms": 1e3 * float(np.percentile(lat, 50)),
            "p99_ms": 1e3 * float(np.percentile(lat, 99)),
            "pcg_iterations": int(reps[-1].pcg_iterations),
            "outer_fp64_spmv": int(reps[-1].outer_spmv),
            "ms_assembly": 1e3 * reps[-1].seconds_assembly,
            "ms_solve": 1e3 * reps[-1].seconds_solve,
        }
        if u_ref is not None:
            arm_res["u_rel_err_vs_oracle"] = float(np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref))
            arm_res["u_max_rel_err_vs_oracle"] = float(np.max(np.abs(u - u_ref)) / np.max(np.abs(u_ref)))
        arms[arm] = arm_res
        print(f"# {arm}: {json.dumps(arm_res)}", flush=True)

    head = "mixed" if "mixed" in arms else next(iter(arms))
    # roofline of the dominant kernel on the headline arm's operator
    dm.assemble(False)
    dm.apply_penalty(st.bc_dof, st.bc_val, np.zeros(job.n_dof))
    prec = L.FX_PREC_MIXED if head == "mixed" else L.FX_PREC_FP64
    if a.sweep_groups:
        for p_ in (L.FX_PREC_FP64, L.FX_PREC_MIXED):
            for g in (4, 8, 16, 32, 64):
                ms_, by_ = dm.profile_spmv(p_, a.spmv_reps, g)
                print(f"# sweep prec={p_} G={g}: {ms_ * 1e3:.2f} us  {by_ / (ms_ * 1e-3) / 1e9:.0f} GB/s", flush=True)
    ms, by = dm.profile_spmv(prec, a.spmv_reps)
    ms64, by64 = dm.profile_spmv(L.FX_PREC_FP64, a.spmv_reps)
    achieved = by / (ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None,
            "kernel": "k_spmv_p<G,float>" if prec == L.FX_PREC_MIXED else "k_spmv_p<G,double>",
            "bytes_per_launch": by, "us_per_launch": round(ms * 1e3, 2),
            "fp64_kernel": {"bytes_per_launch": by64, "us_per_launch": round(ms64 * 1e3, 2),
                            "achieved": round(by64 / (ms64 * 1e-3) / 1e9, 1)}}
    print(f"# roofline {json.dumps(roof)}", flush=True)

    cpu = None
    if not a.no_cpu_baseline:
        fp64_iters = arms.get("fp64", arms[head])["pcg_iterations"]
        cpu = cpu_baseline(job, fp64_iters)
        print(f"# cpu {json.dumps(cpu)}", flush=True)

    h = arms[head]
    out = {
        "metric": "DOF-solves/s + p50/p99 latency, mixed vs FP64 PCG at rel err <=1e-8",
        "value": round(h["dof_solves_per_s"], 1),
        "unit": "DOF-solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(h["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-stored K / fp64" if head == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": job.name + "_linear_static_jacobi_pcg_rel1e-10", "n_dof": job.n_dof,
                   "n_elements": job.mesh.n_el, "element": "C3D20", "nnz": int(dm.nnz)},
        "p50_ms": round(h["p50_ms"], 3),
        "p99_ms": round(h["p99_ms"], 3),
        "arms": arms,
        "mixed_over_fp64": (round(arms["mixed"]["dof_solves_per_s"] / arms["fp64"]["dof_solves_per_s"], 3)
                            if "mixed" in arms and "fp64" in arms else None),
        "roofline": roof,
        "cpu_baseline": {k: v for k, v in cpu.items() if k != "seconds_per_step"} if cpu else None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
