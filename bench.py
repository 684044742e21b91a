"""Benchmark: DOF-solves/s and p50/p99 latency of the static-step hot path on
one MI355X, mixed-precision batched PCG vs plain FP64 PCG, at relative
displacement error <= 1e-8.

Default workload ("batch"): the batched-solve population -- 512 independent
C3D4 Kuhn-tet box meshes (SplitMix64 seed 4; sizes log-uniform 2k-200k DOF,
aspect 1-10, E log-uniform 1e3-3e5 MPa, nu 0.2-0.45, x=0 clamped) with 1-4
load cases each (100 N in a random direction on the x=max face): 23.15 M DOF,
54.27 M DOF-solves per pass.  One step = the whole population: linear
assembly + penalty BCs of every mesh and a Jacobi-PCG solve of every load
case to ||b - K x||/||b|| <= 1e-10.  Meshes are resident in HBM before
timing; every step reassembles from u = 0.

Arms (identical work and stopping rule):
* "fp64_serial": the reference's pattern on the GPU -- one FP64 fx_solve per
  load case, system after system;
* "fp64_batched": continuous batching (fx_batch) of every load case in FP64
  Jacobi-PCG;
* "mixed_batched": the same batching with the FP32-stored, row-sum-corrected
  K and FP32 Krylov vectors, the residual replaced by the FP64 b - K x
  whenever the iterated one has fallen by inner_tol;
* "mixed_poly" (headline): the same batching with FP64 flexible CG whose
  preconditioner is a fixed number of Chebyshev steps on the Jacobi-scaled
  K stored in FP16 (FP64 iterates); outer SpMV, dots, residual and stop rule
  in FP64.
Load cases of a system share each pass over K; systems are admitted under an
operator working-set budget.  The headline arm times --steps after --warmup;
the comparison arms a fixed count (--compare-steps/--compare-warmup, one
each by default), recorded per arm.
value = DOF-solves/s of the headline arm (sum of n_dof x load cases / wall
seconds over the K timed steps).  p50/p99: per-system time from the start of
the step to its solution on the host.  Accuracy: every solution's relative
L2 error against the fp64_serial solution, and the hardest members'
solutions against the oracle's direct solves (tests/golden/population_hard).

roofline: the dominant kernel of the headline arm (kp_cheb_sell, the fused
FP16 SpMV + Chebyshev update; kb_spmv_sell<float, float> for mixed_batched);
every launch of it in the timed run is bracketed by HIP events on the
library's stream and the library sums each launch's algorithmic bytes on the
device; achieved = algorithmic bytes / measured duration.  traffic: the
committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary's ratio to the algorithmic
bytes, times this run's bytes per launch (when the summary was taken on the
same kernels and workload).

cpu_baseline: the oracle (C restatement of the reference CPU path: serial
element loop + triplets + CSR, Jacobi-PCG with the reference's rayon
structure on OpenMP threads) on 8 systems of the same population, every
phase run in full, on 4 threads (the reference's default pool) and on every
CPU the process may use; value is the faster leg, and the host's CPU model,
socket and core counts are recorded.

"--workload single": the C3D20 block (40x10x10, 56,463 DOF), one linear
static step per timed step, FP64 vs mixed, with the same roofline and
baseline figures for its single-system SpMV.  "--workload elements": the
block's C3D20 element stiffness, VALU kernel against FP64 MFMA.
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
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64, vector and matrix (datasheet)


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


def run_single(a):
    """Single-system workload: the C3D20 block, one linear static step."""
    from ferrix_amd import _lib as L
    from ferrix_amd import meshgen
    from ferrix_amd.device import Context, DeviceMesh, SolveOptions, step_def

    job = meshgen.c3d20_block()
    st = job.steps[0]
    golden = os.path.join(ROOT, "tests", "golden", "block_oracle.npz")
    u_ref = np.load(golden)["u"] if os.path.exists(golden) else None
    print(f"# workload {job.name}: {job.mesh.n_nodes} nodes, {job.mesh.n_el} C3D20, {job.n_dof} DOF", flush=True)

    ctx = Context(0)
    dm = DeviceMesh(ctx, job.mesh, job.node_temp)
    arms = {}
    for arm in a.arms.split(","):
        prec = {"mixed": L.FX_PREC_MIXED, "poly": L.FX_PREC_MIXED_POLY}.get(arm, L.FX_PREC_FP64)
        so = SolveOptions(precision=prec, stop_mode=L.FX_STOP_REL_R2, tol=st.tol, max_iter=st.max_iter,
                          inner_tol=a.inner_tol, max_refine=100, spmv_group=a.spmv_group)
        sd, keep = step_def(st, 1, so, so)
        lat, reps = [], []
        # the SpMV autotuner on this arm's operator, before the timed steps
        # (a step reassembles K on the same pattern; the choice stays)
        dm.assemble(False)
        # (the polynomial solve's FP16 inner operator takes the group tuned on
        # the FP32 one, its closest narrow operator: 16 lanes on this block,
        # where the FP64 operator's 32 ran the inner step 4% slower)
        tune_prec = L.FX_PREC_MIXED if prec == L.FX_PREC_MIXED_POLY else prec
        group, _ = dm.autotune_spmv(tune_prec) if a.spmv_group == 0 else (a.spmv_group, 0.0)

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
            "spmv_group": group,
        }
        if u_ref is not None:
            arm_res["u_rel_err_vs_oracle"] = float(np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref))
            arm_res["u_max_rel_err_vs_oracle"] = float(np.max(np.abs(u - u_ref)) / np.max(np.abs(u_ref)))
        arms[arm] = arm_res
        print(f"# {arm}: {json.dumps(arm_res)}", flush=True)

    head = next((h for h in ("poly", "mixed") if h in arms), next(iter(arms)))
    # roofline of the dominant kernel on the headline arm's operator
    dm.assemble(False)
    b = np.zeros(job.n_dof)
    np.add.at(b, st.load_dof, st.load_val)
    b, _ = dm.apply_penalty(st.bc_dof, st.bc_val, b)
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
    if head == "poly":
        # the polynomial solve's inner step (kp_cheb_lanes on the FP16
        # node-block operator): the same system once more as the one slot of
        # an instrumented batch, HIP events around every outer iteration's
        # inner launches, the batch's algorithmic byte count
        from ferrix_amd.device import Batch
        bt = Batch(ctx, SolveOptions(precision=L.FX_PREC_MIXED_POLY, stop_mode=L.FX_STOP_REL_R2, tol=st.tol,
                                     max_iter=st.max_iter, spmv_group=a.spmv_group))
        bt.time_spmv(True)
        bt.add(dm, b[None, :])
        bt.run()
        launches, nt, ms_t, by_t = bt.spmv_stats()
        bt.close()
        us = 1e3 * ms_t / max(nt, 1)
        ach = (by_t / max(launches, 1)) / (us * 1e-6) / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBPS, 4), "traffic": None, "kernel": "kp_cheb_lanes<G>",
                "bytes_per_launch": by_t / max(launches, 1), "us_per_launch": round(us, 2), "launches": launches,
                "fp64_kernel": roof["fp64_kernel"]}
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
        "dtype": DTYPE.get({"poly": "mixed_poly", "mixed": "mixed_batched"}.get(head, ""), "f64"),
        "data": "synthetic",
        "config": {"workload": job.name + "_linear_static_jacobi_pcg_rel1e-10", "n_dof": job.n_dof,
                   "n_elements": job.mesh.n_el, "element": "C3D20", "nnz": int(dm.nnz)},
        "p50_ms": round(h["p50_ms"], 3),
        "p99_ms": round(h["p99_ms"], 3),
        "arms": arms,
        "mixed_over_fp64": (round(arms["mixed"]["dof_solves_per_s"] / arms["fp64"]["dof_solves_per_s"], 3)
                            if "mixed" in arms and "fp64" in arms else None),
        "poly_over_fp64": (round(arms["poly"]["dof_solves_per_s"] / arms["fp64"]["dof_solves_per_s"], 3)
                           if "poly" in arms and "fp64" in arms else None),
        "roofline": roof,
        "cpu_baseline": {k: v for k, v in cpu.items() if k != "seconds_per_step"} if cpu else None,
    }
    print(json.dumps(out), flush=True)


def run_elements(a):
    """C3D20 element stiffness on the 40x10x10 C3D20 block (4,000 elements): the VALU
    kernel (the oracle's order, bit for bit) against the FP64-MFMA one
    (FX_C3D20_MFMA=1), each timed over --steps launches between HIP events,
    beside the device's measured FP64 ceilings.  Then the C3D4 kernel on the
    batch population's largest mesh at 64 / 128 / 256 / 512 elements per
    workgroup and the element autotuner's choice."""
    from ferrix_amd import meshgen
    from ferrix_amd.device import Context, DeviceMesh

    job = meshgen.c3d20_block()
    ctx = Context(0)
    dm = DeviceMesh(ctx, job.mesh, job.node_temp)
    valu_peak, mfma_peak = ctx.measure_fp64_peak(20)
    reps = a.steps or 50
    out = {}
    for name, mf in (("valu", False), ("mfma", True)):
        ms, fl = dm.profile_element(reps, mf)
        tf = fl / (ms * 1e-3) / 1e12
        ceiling = mfma_peak if mf else valu_peak
        out[name] = {"us_per_launch": round(ms * 1e3, 2), "flop_per_launch": fl, "tflops": round(tf, 2),
                     "frac_of_datasheet": round(tf / FP64_PEAK_TFLOPS, 4),
                     "frac_of_measured_ceiling": round(tf / ceiling, 4)}
        print(f"# {name}: {json.dumps(out[name])}", flush=True)
    # C3D4 element batch: elements per workgroup on the population's largest
    # mesh, and the element autotuner's choice
    j4 = meshgen.population_mesh(max(meshgen.population_specs(), key=lambda sp: sp["n_dof"]))[0]
    c3d4 = {"n_elements": j4.mesh.n_el, "us_per_launch": {}}
    old = os.environ.get("FX_C3D4_BLOCK")
    try:
        for bs in (64, 128, 256, 512):
            os.environ["FX_C3D4_BLOCK"] = str(bs)
            d4 = DeviceMesh(ctx, j4.mesh, j4.node_temp)
            c3d4["us_per_launch"][bs] = round(d4.profile_element(reps, False)[0] * 1e3, 2)
            d4.close()
    finally:
        if old is None:
            os.environ.pop("FX_C3D4_BLOCK", None)
        else:
            os.environ["FX_C3D4_BLOCK"] = old
    d4 = DeviceMesh(ctx, j4.mesh, j4.node_temp)
    bs, _, ms = d4.autotune_elements(reps)
    c3d4["tuned_block"], c3d4["tuned_us"] = bs, round(ms * 1e3, 2)
    print(f"# c3d4: {json.dumps(c3d4)}", flush=True)
    m = out["mfma"]
    print(json.dumps({
        "metric": "C3D20 element stiffness throughput (elements/s)", "value": round(job.mesh.n_el / (m["us_per_launch"] * 1e-6), 1),
        "unit": "elements/s", "n_gpus": 1, "steps": reps, "warmup": 1, "ms_per_step": m["us_per_launch"] * 1e-3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "c3d20_element_k_40x10x10_block", "n_elements": job.mesh.n_el},
        "arms": out, "c3d4_workgroup": c3d4, "fp64_ceilings_tflops": {"datasheet": FP64_PEAK_TFLOPS, "measured_valu": round(valu_peak, 2),
                                              "measured_mfma": round(mfma_peak, 2)},
        "roofline": {"bound": "mfma", "achieved": m["tflops"], "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": m["frac_of_datasheet"], "traffic": None, "kernel": "k_c3d20_mfma"}}), flush=True)


def population(n_meshes):
    from ferrix_amd import meshgen
    from ferrix_amd.batch import System

    specs = meshgen.population_specs(n_meshes=n_meshes)
    return specs, [System(k, meshgen.population_mesh(sp)) for k, sp in enumerate(specs)]


def host_info():
    """The host the CPU leg ran on: logical CPUs, the CPUs this process may
    use, the CPU model and the socket count (from /proc/cpuinfo)."""
    model, sockets = None, set()
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k = k.strip()
            if k == "model name" and model is None:
                model = v.strip()
            elif k == "physical id":
                sockets.add(v.strip())
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "model": model,
            "sockets": len(sockets) or None}


def host_threads():
    """The all-cores leg's thread count: the CPUs this process may use,
    capped by OMP_NUM_THREADS when that is set (the host's share)."""
    n = len(os.sched_getaffinity(0))
    cap = os.environ.get("OMP_NUM_THREADS", "")
    return min(n, int(cap)) if cap.isdigit() and int(cap) > 0 else n


def cpu_baseline_batch(specs, stride=64, thread_legs=(4, None)):
    """The reference CPU path restated by the oracle on a bounded sample of
    the population: every `stride`-th system (8 systems spanning the size
    range), each phase timed in full.  Per system: the serial element loop
    and triplet assembly (assembler.rs:37-151), triplets -> CSR with the
    penalty (static_step.rs:183-198), then every load case's Jacobi-PCG to
    ||r||/||b|| <= 1e-10 run to convergence with the reference's rayon
    structure (fxo_pcg_par: iterative.rs:26-136) on each thread count of
    `thread_legs` (None: every CPU this process may use, main.rs:20-22 with
    --num-threads).  Assembly and CSR are serial, as in the reference, and
    timed once."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from ferrix_amd import meshgen

    legs = [t if t is not None else host_threads() for t in thread_legs]
    legs = sorted(set(legs))
    t_all = time.perf_counter()
    t_asm = t_csr = 0.0
    t_solve = {t: 0.0 for t in legs}
    dofs, names, iters = 0, [], 0
    for k in range(0, len(specs), stride):
        sp = specs[k]
        jobs = meshgen.population_mesh(sp)
        st = jobs[0].steps[0]
        t0 = time.perf_counter()
        r, c, v, md = O.assemble_triplets(jobs[0].mesh, jobs[0].node_temp)
        t1 = time.perf_counter()
        bd = np.asarray(st.bc_dof, np.int32)
        K = O.triplets_to_csr(jobs[0].n_dof, np.concatenate([r, bd]), np.concatenate([c, bd]),
                              np.concatenate([v, np.full(len(bd), md * 1e6)]))
        t2 = time.perf_counter()
        t_asm += t1 - t0
        t_csr += t2 - t1
        for j in jobs:
            b = np.zeros(j.n_dof)
            np.add.at(b, j.steps[0].load_dof, j.steps[0].load_val)
            for t in legs:
                sec, it = O.pcg_par_time(K, b, 1e-10, 1, 100_000, t)
                if sec < 0:
                    raise RuntimeError(f"oracle PCG did not converge on system {k}")
                t_solve[t] += sec
            iters += it
            dofs += j.n_dof
        names.append(f"{sp['n_dof']}x{len(jobs)}")
    best = min(legs, key=lambda t: t_solve[t])
    per_leg = {str(t): {"threads": t, "dof_solves_per_s": dofs / (t_asm + t_csr + t_solve[t]),
                        "seconds_solve": round(t_solve[t], 2)} for t in legs}
    return {"value": dofs / (t_asm + t_csr + t_solve[best]), "unit": "DOF-solves/s", "cores": best,
            "kind": "port",
            "sample": (f"oracle C restatement of the reference CPU path on the host, {len(names)} systems of the "
                       f"population (every {stride}th: {', '.join(names)} DOF x load cases, {dofs} DOF-solves): "
                       f"serial element loop + triplets ({t_asm:.2f} s), triplets->CSR ({t_csr:.2f} s), then "
                       f"every load case's Jacobi-PCG to rel 1e-10 run in full ({iters} iterations) on "
                       + ", ".join(f"{t} threads ({t_solve[t]:.1f} s)" for t in legs)
                       + f"; value is the {best}-thread leg ({time.perf_counter() - t_all:.1f} s measured)"),
            "phases_s": {"assembly": round(t_asm, 3), "csr": round(t_csr, 3),
                         "solve": {str(t): round(t_solve[t], 3) for t in legs}},
            "legs": per_leg, "host": host_info()}


def oracle_hard_check(results, arms):
    """Part of the host leg: every arm's solutions of the population's
    hardest members (the largest, the most slender, nu near 0.45, and
    member 216) against the oracle's direct solves committed in
    tests/golden/population_hard.npz (every 32nd entry of each solution,
    plus its L2 norm); sets u_rel_err_vs_oracle_hard_max per arm."""
    base = os.path.join(ROOT, "tests", "golden", "population_hard")
    if not os.path.exists(base + ".npz"):
        return
    meta = json.load(open(base + ".json"))
    gold = np.load(base + ".npz")
    for arm, r in results.items():
        errs = []
        for m in meta:
            key = m["member"]
            if key not in r.x:
                continue
            sub = gold[f"m{key}_sub"]
            for q in range(sub.shape[0]):
                errs.append(float(np.linalg.norm(r.x[key][q][::32] - sub[q]) / np.linalg.norm(sub[q])))
        arms[arm]["u_rel_err_vs_oracle_hard_max"] = max(errs) if errs else None
        arms[arm]["oracle_hard_members"] = len(errs)


def counter_traffic(name, bytes_total, bytes_per_launch):
    """HBM bytes per launch from the committed rocprofv3 counter summary of
    this workload (FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel):
    its measured traffic/algorithmic ratio times this run's algorithmic
    bytes per launch; None when no summary is committed or it was taken on
    other kernels.  The match is on the step's total algorithmic bytes,
    which the workload fixes exactly (the launch count can differ by one
    trailing chunk of launches that find every system finished)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    ratio = ref = None
    for line in open(path):
        if line.startswith("traffic_over_algorithmic:"):
            ratio = float(line.split(":", 1)[1])
        if line.startswith("algorithmic_bytes_per_step:"):
            ref = float(line.split(":", 1)[1])
    if ratio is None or ref is None or abs(ref - bytes_total) > 1e-6 * ref:
        return None
    return round(ratio * bytes_per_launch)


# dominant kernel of each batched arm, and the committed rocprofv3 counter
# summary of it on the 512-mesh population
DOMINANT_KERNEL = {"mixed_poly": "kp_cheb_sell", "mixed_batched": "kb_spmv_sell<float, float>",
                   "fp64_batched": "kb_spmv_sell<double, double>"}
COUNTER_SUMMARY = {"mixed_poly": "kp_cheb_sell_batch512_counters.md",
                   "mixed_batched": "kb_spmv_sell_float_batch512_counters.md"}


# the arithmetic of each batched arm
DTYPE = {"mixed_poly": "f64 (inner Chebyshev steps on an f16-stored K, f64 iterates)",
         "mixed_batched": "fp32-stored K / fp64", "fp64_batched": "f64", "fp64_serial": "f64"}


def ratio(arms, x, y):
    if x in arms and y in arms:
        return round(arms[x]["dof_solves_per_s"] / arms[y]["dof_solves_per_s"], 3)
    return None


def run_batch(a):
    from ferrix_amd import _lib as L
    from ferrix_amd.batch import solve_batched, solve_serial, upload
    from ferrix_amd.device import Context, SolveOptions

    specs, systems = population(a.meshes)
    n_dof = sum(s.n_dof for s in systems)
    n_solves = sum(s.n_dof * s.n_rhs for s in systems)
    print(f"# workload batch: {len(systems)} C3D4 meshes, {n_dof} DOF, {n_solves} DOF-solves per step", flush=True)
    ctx = Context(0)
    # the meshes live on a context of their own (its own stream), so their
    # assembly overlaps the batch's iterations
    mctx = Context(0) if a.assemble_ahead else ctx
    t = time.perf_counter()
    upload(mctx, systems)
    print(f"# upload + symbolic {time.perf_counter() - t:.1f} s", flush=True)
    fp64 = SolveOptions(precision=L.FX_PREC_FP64, stop_mode=L.FX_STOP_REL_R2, tol=1e-10, max_iter=100_000)
    fp64_b = SolveOptions(precision=L.FX_PREC_FP64, stop_mode=L.FX_STOP_REL_R2, tol=1e-10, max_iter=100_000,
                          spmv_group=a.spmv_group)
    mixed = SolveOptions(precision=L.FX_PREC_MIXED, stop_mode=L.FX_STOP_REL_R2, tol=1e-10, max_iter=100_000,
                         inner_tol=a.inner_tol, max_refine=100, spmv_group=a.spmv_group)
    poly = SolveOptions(precision=L.FX_PREC_MIXED_POLY, stop_mode=L.FX_STOP_REL_R2, tol=1e-10, max_iter=400_000,
                        spmv_group=a.spmv_group, poly_degree=a.poly_degree)
    ws = int(a.ws_mb * (1 << 20))
    admit = (sorted(systems, key=lambda s: -s.n_dof * s.n_rhs) if a.admit_order == "largest" else systems)
    arms, results = {}, {}
    arm_list = a.arms.split(",")
    head = next((h for h in ("mixed_poly", "mixed_batched") if h in arm_list), arm_list[0])
    for arm in arm_list:
        if arm == "fp64_serial":
            step = lambda: solve_serial(systems, fp64)
        elif arm == "mixed_batched":
            step = lambda: solve_batched(ctx, admit, mixed, ws_bytes=ws, time_spmv=True,
                                         assemble_ahead=a.assemble_ahead)
        elif arm == "mixed_poly":
            shape = a.inner_shape if a.inner_shape == "tune" else int(a.inner_shape)
            step = lambda: solve_batched(ctx, admit, poly, ws_bytes=ws, time_spmv=True,
                                         assemble_ahead=a.assemble_ahead, inner_shape=shape)
        elif arm == "fp64_batched":
            step = lambda: solve_batched(ctx, admit, fp64_b, ws_bytes=ws, time_spmv=True,
                                         assemble_ahead=a.assemble_ahead)
        else:
            raise SystemExit(f"unknown arm {arm}")
        # the headline arm times --steps after --warmup; the comparison arms
        # a fixed, stated count (one of their steps is several times longer)
        n_steps, n_warm = (a.steps, a.warmup) if arm == head else (a.compare_steps, a.compare_warmup)
        for k in range(n_warm):
            ts = time.perf_counter()
            step()
            print(f"# {arm} warm-up step {k + 1}/{n_warm}: {time.perf_counter() - ts:.2f} s", flush=True)
        device_sync()
        t0 = time.perf_counter()
        runs = []
        for k in range(n_steps):
            ts = time.perf_counter()
            runs.append(step())
            print(f"# {arm} step {k + 1}/{n_steps}: {time.perf_counter() - ts:.2f} s", flush=True)
        device_sync()
        wall = time.perf_counter() - t0
        last = runs[-1]
        lat = np.concatenate([np.array(list(r.latency.values())) for r in runs])
        res = {
            "steps": n_steps,
            "warmup": n_warm,
            "dof_solves_per_s": n_solves * n_steps / wall,
            "ms_per_step": 1e3 * wall / n_steps,
            "p50_ms": 1e3 * float(np.percentile(lat, 50)),
            "p99_ms": 1e3 * float(np.percentile(lat, 99)),
            "failures": len(last.failures),
            "pcg_iterations_total": int(sum(last.iterations.values())),
        }
        if arm != "fp64_serial":
            res["first_run_at_ms"] = 1e3 * float(np.mean([r.first_run_at for r in runs]))
            res["collect_ms"] = 1e3 * float(np.mean([r.collect_s for r in runs]))
        if last.inner_tune:
            sh, t1, t2 = last.inner_tune
            res["inner_shape"] = {"chosen": "per_load_case_count" if sh else "one_launch",
                                  "us_one_launch": 1e3 * t1, "us_per_load_case_count": 1e3 * t2}
        elif arm == "mixed_poly":
            res["inner_shape"] = {"chosen": "per_load_case_count" if a.inner_shape == "1" else "one_launch"}
        if last.spmv_stats:
            launches, nt, ms, by = last.spmv_stats
            res["spmv"] = {"launches": launches, "timed": nt, "us_per_launch": 1e3 * ms / max(nt, 1),
                           "bytes_per_launch": by / max(launches, 1), "bytes_total": by,
                           "achieved_GBps": by / (ms * 1e-3) / 1e9 if ms > 0 and nt == launches else None}
        arms[arm] = res
        results[arm] = last
        print(f"# {arm}: {json.dumps(res)}", flush=True)

    # accuracy: every solution against the FP64 serial one; a sample against the oracle
    ref = results.get("fp64_serial")
    if ref is not None:
        for arm, r in results.items():
            if arm == "fp64_serial":
                continue
            errs = [float(np.max(np.linalg.norm(r.x[k] - ref.x[k], axis=1) / np.linalg.norm(ref.x[k], axis=1)))
                    for k in ref.x if k in r.x]
            arms[arm]["u_rel_err_vs_fp64_max"] = float(max(errs)) if errs else None
    roof = None
    st_ = arms[head].get("spmv")
    if st_ and st_["achieved_GBps"]:
        roof = {"bound": "hbm", "achieved": round(st_["achieved_GBps"], 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": round(st_["achieved_GBps"] / HBM_PEAK_GBPS, 4),
                "traffic": (counter_traffic(COUNTER_SUMMARY[head], st_["bytes_total"], st_["bytes_per_launch"])
                            if head in COUNTER_SUMMARY and a.meshes == 512 else None),
                "kernel": DOMINANT_KERNEL.get(head, "kb_spmv_sell<double, double>"),
                "us_per_launch": round(st_["us_per_launch"], 2), "bytes_per_launch": round(st_["bytes_per_launch"]),
                "launches": st_["launches"]}
    print(f"# roofline {json.dumps(roof)}", flush=True)
    cpu = None
    if not a.no_oracle_check:  # the checker: the oracle's direct solves of the hardest members
        oracle_hard_check(results, arms)
    if not a.no_cpu_baseline:  # the host leg: the oracle times the CPU path
        cpu = cpu_baseline_batch(specs)
    if cpu:
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
        "dtype": DTYPE.get(head, "f64"),
        "data": "synthetic",
        "config": {"workload": f"batched_linear_static_c3d4_{len(systems)}_meshes_2k-200k_dof_1-4_load_cases_rel1e-10",
                   "n_meshes": len(systems), "n_dof": n_dof, "dof_solves_per_step": n_solves,
                   "working_set_mb": a.ws_mb, "assembly": "overlapped" if a.assemble_ahead else "serial",
                   "admit_order": a.admit_order},
        "p50_ms": round(h["p50_ms"], 3),
        "p99_ms": round(h["p99_ms"], 3),
        "arms": arms,
        "mixed_arm": head,
        "mixed_over_fp64": ratio(arms, head, "fp64_serial"),
        "mixed_over_fp64_batched": ratio(arms, head, "fp64_batched"),
        "ratios": {f"{x}_over_{y}": ratio(arms, x, y) for x in arms for y in ("fp64_serial", "fp64_batched")
                   if x != y and y in arms},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def run_serve(a):
    """Serving stream: Poisson arrivals of linear C3D4, NLGEOM C3D4 and C3D20
    block requests (70/20/10), continuous batching vs the same stream served
    one request at a time with plain FP64 PCG."""
    from ferrix_amd import _lib as L
    from ferrix_amd.device import Context
    from ferrix_amd.serving import make_requests, serve_batched, serve_serial

    prec = {"mixed": L.FX_PREC_MIXED, "poly": L.FX_PREC_MIXED_POLY, "fp64": L.FX_PREC_FP64}[a.serve_precision]
    ctx = Context(0)
    warm = make_requests(4, rate=0.0, seed=a.seed + 1000)
    serve_batched(ctx, warm, prec)
    serve_serial(ctx, warm[:2])
    t = time.perf_counter()
    sat = make_requests(a.requests, rate=0.0, seed=a.seed)
    print(f"# serve: {len(sat)} requests generated in {time.perf_counter() - t:.1f} s", flush=True)
    r_sat = serve_batched(ctx, sat, prec, prep_context=not a.serve_shared_prep, log_every=100)
    capacity = len(sat) / r_sat.makespan
    rate = a.rate if a.rate > 0 else 0.7 * capacity
    print(f"# serve capacity (all at t=0): {capacity:.2f} req/s, {r_sat.dof_solves(sat) / r_sat.makespan:.0f} "
          f"DOF-solves/s; arrival rate {rate:.2f} req/s", flush=True)
    reqs = make_requests(a.requests, rate=rate, seed=a.seed)
    rb = serve_batched(ctx, reqs, prec, prep_context=not a.serve_shared_prep, log_every=100)
    print(f"# serve batched ({a.serve_precision}): makespan {rb.makespan:.1f} s, failures {len(rb.failures)}",
          flush=True)
    other = L.FX_PREC_FP64 if prec != L.FX_PREC_FP64 else L.FX_PREC_MIXED_POLY
    ro = serve_batched(ctx, reqs, other, prep_context=not a.serve_shared_prep, log_every=100)
    print(f"# serve batched (other precision): makespan {ro.makespan:.1f} s, failures {len(ro.failures)}",
          flush=True)
    rs = serve_serial(ctx, reqs, log_every=100)
    print(f"# serve serial fp64: makespan {rs.makespan:.1f} s, failures {len(rs.failures)}", flush=True)

    def max_err(r):
        err = 0.0
        for k, x in r.x.items():
            if k in rs.x:
                e = np.linalg.norm(x - rs.x[k], axis=1) / np.maximum(np.linalg.norm(rs.x[k], axis=1), 1e-300)
                err = max(err, float(e.max()))
        return err

    def summary(r):
        lat = np.array(list(r.latency.values()))
        return {"dof_solves_per_s": r.dof_solves(reqs) / r.makespan, "makespan_s": r.makespan,
                "p50_ms": 1e3 * float(np.percentile(lat, 50)), "p99_ms": 1e3 * float(np.percentile(lat, 99)),
                "failures": len(r.failures), "u_rel_err_vs_serial_fp64_max": max_err(r) if r is not rs else 0.0}

    sb, so_, ss = summary(rb), summary(ro), summary(rs)
    err = sb["u_rel_err_vs_serial_fp64_max"]
    kinds = {k: sum(1 for q in reqs if q.kind == k) for k in ("linear", "nlgeom", "c3d20")}
    out = {
        "metric": "DOF-solves/s + p50/p99 latency, mixed vs FP64 PCG at rel err <=1e-8",
        "value": round(sb["dof_solves_per_s"], 1),
        "unit": "DOF-solves/s",
        "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": round(1e3 * rb.makespan, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {L.FX_PREC_MIXED: "fp32-stored K / fp64", L.FX_PREC_MIXED_POLY: DTYPE["mixed_poly"]}.get(prec, "f64"),
        "data": "synthetic",
        "config": {"workload": f"serving_{len(reqs)}_requests_poisson_70lin_20nlgeom_10c3d20",
                   "requests": kinds, "rate_per_s": rate, "capacity_per_s": capacity,
                   "admission_context": "shared" if a.serve_shared_prep else "own"},
        "p50_ms": round(sb["p50_ms"], 3), "p99_ms": round(sb["p99_ms"], 3),
        "arms": {f"batched_{a.serve_precision}": sb,
                 "batched_" + ("fp64" if a.serve_precision != "fp64" else "poly"): so_,
                 "serial_fp64": ss},
        "admission": {"rule": "free device memory (hipMemGetInfo + reusable slot blocks) less 4 GiB",
                      "max_admitted": rb.max_admitted, "least_free_bytes": rb.min_free,
                      "nlgeom_solves_in_batch": rb.batched_nl_solves},
        "u_rel_err_vs_serial_fp64_max": err,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=["batch", "single", "serve", "elements"], default="batch")
    ap.add_argument("--requests", type=int, default=1000, help="serve: requests in the stream")
    ap.add_argument("--rate", type=float, default=0.0, help="serve: arrivals per second (0: 0.7 x capacity)")
    ap.add_argument("--seed", type=int, default=5, help="serve: request stream seed")
    ap.add_argument("--serve-precision", choices=["mixed", "poly", "fp64"], default="poly")
    ap.add_argument("--serve-shared-prep", action="store_true",
                    help="serve: admission thread on the batch's context (default: a context of its own)")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--arms", default=None)
    ap.add_argument("--compare-steps", type=int, default=1,
                    help="batch: timed steps of the comparison arms (fp64_serial, fp64_batched)")
    ap.add_argument("--compare-warmup", type=int, default=1,
                    help="batch: warm-up steps of the comparison arms")
    ap.add_argument("--meshes", type=int, default=512)
    ap.add_argument("--ws-mb", type=float, default=0.0, help="operator working-set budget of the batch (0: all)")
    ap.add_argument("--assemble-ahead", action="store_true",
                    help="batched arms: a host thread assembles the systems on a second context's stream while "
                         "the batch runs (default: each system is assembled just before its admission)")
    ap.add_argument("--admit-order", choices=["given", "largest"], default="given",
                    help="batched arms: admission order (largest: descending DOF x load cases)")
    ap.add_argument("--spmv-group", type=int, default=0, help="lanes per node row of the batched SpMV (0: auto)")
    ap.add_argument("--no-oracle-check", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--spmv-reps", type=int, default=100)
    ap.add_argument("--inner-tol", type=float, default=1e-3)
    ap.add_argument("--inner-shape", choices=["tune", "0", "1"], default="tune",
                    help="polynomial batch: Chebyshev step as one launch (0), one per load-case count (1), "
                         "or the faster of both timed on the running systems (tune)")
    ap.add_argument("--poly-degree", type=int, default=0,
                    help="mixed_poly: Chebyshev steps per outer iteration (0: the library default, 15)")
    ap.add_argument("--sweep-groups", action="store_true", help="print SpMV time per lane-group size")
    ap.add_argument("--fp64-peak", action="store_true",
                    help="only measure the device's FP64 VALU and MFMA throughput and exit")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="only run device-to-device copies of this many MiB (counter calibration) and exit")
    a = ap.parse_args()
    if a.calib_copy_mb > 0:
        calib_copy(a.calib_copy_mb)
        return
    if a.fp64_peak:
        from ferrix_amd.device import Context

        valu, mfma = Context(0).measure_fp64_peak(20)
        print(json.dumps({"fp64_peak_tflops": {"datasheet": FP64_PEAK_TFLOPS, "measured_valu_fma": round(valu, 2),
                                               "measured_mfma_16x16x4": round(mfma, 2)}}), flush=True)
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only; multi-GPU runs are not part of it "
              "(the batched and serving workloads shard by independent mesh, see the design notes).",
              file=sys.stderr)
        sys.exit(2)
    if a.workload == "serve":
        run_serve(a)
    elif a.workload == "elements":
        run_elements(a)
    elif a.workload == "single":
        a.steps = 10 if a.steps is None else a.steps
        a.warmup = 2 if a.warmup is None else a.warmup
        a.arms = a.arms or "fp64,mixed,poly"
        run_single(a)
    else:
        a.steps = 1 if a.steps is None else a.steps
        a.warmup = 1 if a.warmup is None else a.warmup
        a.arms = a.arms or "fp64_serial,fp64_batched,mixed_batched,mixed_poly"
        run_batch(a)


if __name__ == "__main__":
    main()
