/*
 * ferrix_hip.h -- C ABI of the MI355X (gfx950) static-step backend.
 *
 * The library sits behind FerriX's three seams for the static-step hot path
 * (paths relative to the FerriX source tree):
 *
 *   element seam   trait FiniteElement / Element::compute_stiffness_sdv,
 *                  compute_internal_force_sdv   (src/solver/mesh_lib/elements/element.rs:20-35, 217-447)
 *   assembly seam  Assembler::assemble / assemble_internal_force
 *                                               (src/solver/assembler.rs:37-44, 157-165)
 *   solver seam    trait Solver::solve + trait Preconditioner
 *                                               (src/solver/solvers/mod.rs:18-44,
 *                                                src/solver/preconditioner.rs:15-26)
 *   step driver    StaticStep::solve            (src/solver/step/static_step.rs:47-267)
 *
 * Conventions: plain C types only; every pointer argument is host memory
 * unless its name ends in _dev; arrays are copied, the caller keeps
 * ownership.  Node index = rank of the node ID in ascending order and
 * DOF = 3*index + d (src/solver/mesh_lib/mesh.rs:35-49).  Elements are
 * visited in the order given (the caller passes ascending element ID).
 * One host thread drives a batch; meshes may be uploaded, assembled and
 * destroyed from other threads on the same context (uploads overlap with a
 * running batch).  Errors never cross the
 * boundary as exceptions: every call returns an fx_status and
 * fx_last_error() holds the message.
 */
#ifndef FERRIX_HIP_H
#define FERRIX_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes; mapping onto FerrixError (src/solver/error.rs:12-74). */
typedef enum fx_status {
    FX_OK = 0,
    FX_ERR_INVALID = 1,           /* InvalidModelState: bad argument or model      */
    FX_ERR_OOM = 2,               /* InvalidModelState: device memory exhausted    */
    FX_ERR_HIP = 3,               /* HIP runtime failure (message has the detail)  */
    FX_ERR_SINGULAR_JACOBIAN = 4, /* NumericalError("Singular Jacobian"),
                                     element.rs:257-259                            */
    FX_ERR_NOT_CONVERGED = 5,     /* ConvergenceError: PCG (iterative.rs:117-120),
                                     NR / increments (static_step.rs:73-78,254-257) */
    FX_ERR_BREAKDOWN = 6,         /* NumericalError: p.Ap <= 0 or not finite        */
    FX_ERR_NO_DEVICE = 7          /* no usable gfx950 device                        */
} fx_status;

typedef struct fx_ctx fx_ctx;
typedef struct fx_mesh fx_mesh;

/* ---- context ---------------------------------------------------------- */
/* Binds HIP device `device`, creates one stream (default priority). */
fx_status fx_ctx_create(int device, fx_ctx **out);
/* Destroys the context and every batch and mesh still alive on it (their
 * handles become invalid). */
void fx_ctx_destroy(fx_ctx *ctx);
/* Message of the calling thread's last failing call ("" if none); ctx is
 * accepted for symmetry and may be NULL. */
const char *fx_last_error(const fx_ctx *ctx);
/* Library version string. */
const char *fx_version(void);

/* Device memory as an admission decision needs it (no counterpart in the
 * reference): *free_bytes and *total_bytes from hipMemGetInfo on the
 * context's device, *pool_bytes the released batch-slot blocks the context
 * keeps for reuse (allocated, but free for later slots).  Any pointer may be
 * NULL. */
fx_status fx_ctx_memory(fx_ctx *ctx, int64_t *free_bytes, int64_t *total_bytes, int64_t *pool_bytes);

/* ---- mesh ------------------------------------------------------------- */
#define FX_LUT_MAX 16

/* Temperature-dependent isotropic elasticity (material.rs:36-68, 202-230):
 * per material a clamped piecewise-linear LUT for E and one for nu, each with
 * n_* points (1..FX_LUT_MAX, temperatures ascending), rows of FX_LUT_MAX. */
typedef struct fx_material_table {
    int32_t n_mat;
    const int32_t *n_e;
    const double *t_e, *v_e;
    const int32_t *n_nu;
    const double *t_nu, *v_nu;
} fx_material_table;

/*
 * Uploads the mesh and builds the structural pattern (host symbolic phase):
 *   xyz        n_nodes x 3 coordinates in node-index order
 *   node_temp  n_nodes temperatures (initial == current; the reference never
 *              changes T after SolutionState::initialize, state.rs:47-78)
 *   el_ptr     n_el+1 offsets into conn; 4 nodes = C3D4, 20 nodes = C3D20
 *   conn       node indices per element (CalculiX node order)
 *   mat        material index per element
 * Replaces the Mesh/Project data the reference keeps in HashMaps
 * (src/solver/mesh_lib/mesh.rs:16-30, project.rs:32-59).
 */
fx_status fx_mesh_upload(fx_ctx *ctx, int64_t n_nodes, const double *xyz, const double *node_temp,
                         int64_t n_el, const int64_t *el_ptr, const int32_t *conn, const int32_t *mat,
                         const fx_material_table *mats, fx_mesh **out);
void fx_mesh_destroy(fx_mesh *mesh);

/* n_dof = 3 n_nodes; nnz_struct = structural (pre-filter) nonzeros;
 * device_bytes = HBM held by the mesh. */
fx_status fx_mesh_info(const fx_mesh *mesh, int64_t *n_dof, int64_t *nnz_struct, int64_t *device_bytes);

/* ---- displacement state (SolutionState::displacements, state.rs:16-27) -- */
/* u == NULL sets the resident displacement to zero without a host copy. */
fx_status fx_state_set_u(fx_mesh *mesh, const double *u);
fx_status fx_state_get_u(fx_mesh *mesh, double *u);

/* ---- assembly seam ---------------------------------------------------- */
/*
 * Assembler::assemble (assembler.rs:37-151): element stiffness for every
 * element with coordinates X (+ the resident u when nlgeom != 0), drop rule
 * |v| > 1e-12 per element contribution, contributions summed per CSR slot in
 * element order (deterministic), max |v| over kept element diagonal
 * contributions.  K stays resident; *max_diag_out and *nnz_out (filtered
 * nonzeros, i.e. what sprs to_csr would hold) are returned.
 */
fx_status fx_assemble(fx_mesh *mesh, int32_t nlgeom, double *max_diag_out, int64_t *nnz_out);

/*
 * Assembler::assemble_internal_force (assembler.rs:157-241) at the resident u;
 * coordinates move with u when nlgeom != 0 (u_conf, static_step.rs:121-126).
 * f_int_out (n_dof, host) may be NULL to keep the result on the device.
 */
fx_status fx_internal_force(fx_mesh *mesh, int32_t nlgeom, double *f_int_out);

/*
 * Nodal stress and strain at the resident u (StaticStep::calculate_stress_strain,
 * static_step.rs:271-379, with Element::calculate_stress_strain_at_local_coords,
 * element.rs:526-564): per element and node, B at the node's local coordinates
 * on the undeformed geometry, strain = B u_el, stress = D(T_node) strain;
 * summed over the elements of each node in element order and divided by
 * their number.  stress_out, strain_out: n_nodes x 6 (host), Voigt
 * [xx, yy, zz, xy, yz, zx], engineering shear; nodes of no element get 0.
 */
fx_status fx_stress_strain(fx_mesh *mesh, double *stress_out, double *strain_out);

/*
 * Filtered CSR of the resident K (the sprs CsMat the reference hands to the
 * solver, static_step.rs:198): row_ptr (n_dof+1), col_idx (nnz), values
 * (nnz, may be NULL).  Columns ascending per row.
 */
fx_status fx_matrix_get(fx_mesh *mesh, int64_t *row_ptr, int32_t *col_idx, double *values);

/*
 * Penalty boundary conditions (static_step.rs:183-195): when the last
 * assembly's max_diag > 0, penalty = 1e6 * max_diag is added to K[d,d] once
 * per record (duplicates add twice) and rhs[d] += penalty * (target - u[d])
 * per record in order.  rhs_inout (n_dof) is host memory; *penalty_out
 * receives the penalty (0 when not applied).
 */
fx_status fx_apply_penalty(fx_mesh *mesh, int64_t n_bc, const int64_t *dofs, const double *targets,
                           double *rhs_inout, double *penalty_out);

/* ---- solver seam ------------------------------------------------------ */
typedef enum fx_precision {
    FX_PREC_FP64 = 0,  /* Jacobi-PCG in FP64 (the reference's IterativeSolver)   */
    FX_PREC_MIXED = 1, /* FP32-stored K, FP64 residuals and norms; PCG with
                          reliable updates: the residual is replaced by the
                          FP64 b - K x whenever the iterated one has fallen
                          by inner_tol, keeping the search direction if the
                          two agree within 10%.  fx_solve keeps FP64 Krylov
                          vectors, a batch FP32 ones (dots in FP64) */
    FX_PREC_MIXED_POLY = 2 /* flexible CG in FP64 (x, r, the directions,
                          K64 z, every dot product, the stop rule on the
                          FP64 residual) preconditioned by poly_degree
                          Chebyshev steps on the Jacobi-scaled K stored in
                          FP16, with FP64 iterates; the Chebyshev interval
                          comes from a power iteration per system when it
                          joins.  A batch precision; fx_solve and
                          fx_static_step run the system as the one slot of a
                          batch kept with the mesh (same bits as alone in a
                          batch).  iterations in the report count the inner
                          steps (max_iter bounds them), outer_spmv the FP64
                          outer iterations */
} fx_precision;

typedef enum fx_stop {
    FX_STOP_ABS_SQRT_RZ = 0, /* sqrt(r.z) < tol (iterative.rs:63; tol 1e-8)     */
    FX_STOP_REL_R2 = 1       /* ||b - K x||_2 / ||b||_2 <= tol                 */
} fx_stop;

typedef struct fx_solve_opts {
    int32_t precision;  /* fx_precision */
    int32_t stop_mode;  /* fx_stop */
    double tol;         /* reference: 1e-8 (static_step.rs:204) */
    int64_t max_iter;   /* reference: 1000; iterations with the working operator */
    double inner_tol;   /* MIXED: reduction of the iterated residual between
                           residual replacements; 1e-3 */
    int32_t max_refine; /* MIXED: residual replacements before NOT_CONVERGED */
    int32_t spmv_group; /* lanes per node row in the SpMV (0: the mesh's tuned
                           group, fx_mesh_autotune_spmv, else a row-length
                           heuristic) */
    int32_t poly_degree; /* MIXED_POLY: Chebyshev steps per outer iteration,
                            1..32 (0: 15); max_iter bounds the inner steps
                            of a load case */
} fx_solve_opts;

typedef struct fx_solve_report {
    int64_t iterations;     /* SpMVs with the working operator (PCG iterations) */
    int64_t outer_spmv;     /* MIXED: FP64 residual evaluations */
    int32_t refinements;    /* MIXED: residual replacements (segments) */
    int32_t spmv_group;     /* lanes per node row used */
    double final_criterion; /* last stopping-rule value */
    double seconds;         /* wall time of the solve on the context stream */
} fx_solve_report;

/* Fills opts with the reference's values: FP64, sqrt(r.z) < 1e-8, 1000. */
void fx_solve_opts_default(fx_solve_opts *opts);

/*
 * Solver::solve with the Jacobi preconditioner (solvers/mod.rs:18-44,
 * iterative.rs:26-136, preconditioner.rs:39-58) on the resident K
 * (after fx_assemble / fx_apply_penalty): x0 = 0, returns x (n_dof).
 * FX_ERR_NOT_CONVERGED after max_iter iterations, like the reference.
 */
fx_status fx_solve(fx_mesh *mesh, const double *rhs, double *x_out, const fx_solve_opts *opts,
                   fx_solve_report *report);

/*
 * Solver::solve on the caller's matrix: the seam's exact contract
 * (src/solver/solvers/mod.rs:18-44), called by the step driver with the CSR
 * it has just built (static_step.rs:198-204), here with the Jacobi
 * preconditioner built from that matrix's diagonal (preconditioner.rs:39-50:
 * 1/K_ii if |K_ii| > 1e-9, else 1; no stored diagonal: 1) and
 * IterativeSolver's recurrences (iterative.rs:26-121), x0 = 0.
 *   n        rows (= columns)
 *   row_ptr  n+1 offsets, row_ptr[0] = 0, non-decreasing (sprs indptr, usize)
 *   col_idx  row_ptr[n] column indices, strictly ascending per row (sprs
 *            indices, usize); both triangles stored
 *   values   row_ptr[n] values
 *   rhs, x_out  n each (host)
 * All arrays are host memory and are copied; the call returns after the
 * solution is in x_out.  opts: FX_PREC_FP64 only (the mixed solve needs the
 * node-block operator of a mesh, fx_solve); NULL = the reference's values.
 * A malformed matrix gives FX_ERR_INVALID before anything runs;
 * FX_ERR_NOT_CONVERGED after max_iter iterations and FX_ERR_BREAKDOWN as for
 * fx_solve, with *report filled either way.
 */
fx_status fx_solve_csr(fx_ctx *ctx, int64_t n, const int64_t *row_ptr, const int64_t *col_idx,
                       const double *values, const double *rhs, double *x_out, const fx_solve_opts *opts,
                       fx_solve_report *report);

/* ---- step driver (StaticStep::solve, static_step.rs:47-267) ----------- */
/* Amplitudes (amplitude.rs:10-98); entry 0 must be the default ramp. */
typedef struct fx_amplitudes {
    int32_t n;
    const int32_t *total_time; /* n: 1 = TIME=TOTAL TIME */
    const double *shift_x;     /* n */
    const double *shift_y;     /* n */
    const int32_t *has_data;   /* n: 0 = default ramp */
    const int64_t *ptr;        /* n+1 offsets into times/values */
    const double *times;
    const double *values;
} fx_amplitudes;

typedef enum fx_solver_kind {
    FX_SOLVER_DEFAULT = 0,  /* the reference maps Default to DIRECT          */
    FX_SOLVER_DIRECT = 1,   /* served by PCG with direct_opts (no GPU Cholesky) */
    FX_SOLVER_ITERATIVE = 2 /* PCG with solve_opts                            */
} fx_solver_kind;

/* Result-writer hook (ResultWriter::write_increment, io/writer.rs:21-25,
 * called at static_step.rs:211-250): after every converged increment, with
 * its displacement committed as the resident u, fx_static_step calls
 * on_increment(user, step_id, inc_id, global_time).  inc_id counts the
 * step's increments as the reference's n_inc does (attempts that were cut
 * back included); global_time is SolverTime::global_time.  The callee may
 * read the state (fx_state_get_u, fx_stress_strain) but must not change it
 * or run a step on the mesh.  A nonzero return stops the step with
 * FX_ERR_INVALID (the reference's GenericIo error).  NULL: no call. */
typedef int32_t (*fx_increment_fn)(void *user, int32_t step_id, int32_t inc_id, double global_time);

/* A caller-supplied linear solver for the step's Newton iterations: the
 * Solver seam the step driver calls with the matrix it has just built
 * (static_step.rs:198-204, solvers/mod.rs:18-44).  Called with the mesh's K
 * assembled and penalised on the device (fx_matrix_get reads it), rhs the
 * penalised residual (host, n values) and x the solution buffer (host, n
 * values).  Returns FX_OK with x filled and *iterations set, or an
 * fx_status that stops the step with that status.  A continuous-batching
 * server passes one that solves K in its running batch. */
typedef int32_t (*fx_solve_fn)(void *user, fx_mesh *mesh, int64_t n, const double *rhs, double *x,
                               int64_t *iterations);

typedef struct fx_step_def {
    int32_t step_id; /* 1-based, as passed to Step::solve */
    int32_t nlgeom;
    int32_t solver;  /* fx_solver_kind */
    double initial_time_increment, time_period, min_time_increment, max_time_increment;
    int64_t max_increments;
    int64_t n_loads; /* *CLOAD records in deck order (boundary_conds.rs:10-53) */
    const int64_t *load_dof;
    const double *load_val;
    const int32_t *load_amp;
    const int32_t *load_origin;
    int64_t n_bcs; /* *BOUNDARY records in deck order (boundary_conds.rs:57-101) */
    const int64_t *bc_dof;
    const double *bc_val;
    const int32_t *bc_amp;
    const int32_t *bc_origin;
    fx_amplitudes amps;
    fx_solve_opts solve_opts;  /* ITERATIVE */
    fx_solve_opts direct_opts; /* DEFAULT / DIRECT */
    fx_increment_fn on_increment; /* NULL, or the result writer's hook */
    void *on_increment_user;
    fx_solve_fn solve;             /* NULL (the step's own PCG), or a caller solver */
    void *solve_user;
} fx_step_def;

/* SolverTime (time.rs:8-63). */
typedef struct fx_timer {
    double global_time, local_time, local_max;
} fx_timer;

typedef struct fx_step_report {
    int32_t increments;      /* attempted increments */
    int32_t cutbacks;        /* increments that failed NR and halved dt */
    int32_t nr_iterations;   /* NR iterations over all increments */
    int32_t linear_solves;
    int64_t pcg_iterations;  /* summed over linear solves */
    int64_t outer_spmv;      /* MIXED residual SpMVs */
    double last_rel_residual;
    double seconds;
    double seconds_assembly; /* element + assembly + penalty */
    double seconds_solve;
    int64_t host_syncs;      /* times the host waited on the device (stream
                                synchronisations) during the step */
} fx_step_report;

/*
 * Runs one *STATIC step on the resident displacement (set it with
 * fx_state_set_u first; read it back with fx_state_get_u).  Calls
 * timer->new_step itself (Step::solve, steps.rs:28-42).
 */
fx_status fx_static_step(fx_mesh *mesh, const fx_step_def *step, fx_timer *timer, fx_step_report *report);

/* ---- instrumentation (no counterpart in the reference) -------------------
 * Times `reps` back-to-back launches of the PCG SpMV kernel on the resident
 * K (FP64 values, or FP32 values for FX_PREC_MIXED) between two HIP events on
 * the context stream.  *ms_per_launch: mean launch duration; *bytes_per_launch:
 * algorithmic bytes of one launch (values + one column index per 3x3 block +
 * node-row pointers + two gathered and two written FP64 vectors).  Used by
 * the benchmark's roofline figure. */
fx_status fx_profile_spmv(fx_mesh *mesh, int32_t precision, int32_t reps, int32_t spmv_group,
                          double *ms_per_launch, double *bytes_per_launch);

/* FP64 throughput ceilings of the device (instrumentation, no counterpart
 * in the reference): register-only kernels that keep every SIMD busy with
 * v_fma_f64 chains and with v_mfma_f64_16x16x4_f64 accumulators, `reps`
 * timed launches each.  The measured TFLOP/s that the element kernels' rates
 * are stated against, beside the datasheet figure. */
fx_status fx_measure_fp64_peak(fx_ctx *ctx, int32_t reps, double *valu_tflops, double *mfma_tflops);

/* Element-K timing (instrumentation, no counterpart in the reference): times
 * `reps` back-to-back launches of the mesh's element stiffness kernels (K
 * only, reference coordinates) between two HIP events on the context
 * stream, after one warm-up launch.  mfma: 0 the VALU kernels (bit for bit
 * the oracle's order), 1 the C3D20 kernel on FP64 MFMA (the path
 * FX_C3D20_MFMA=1 selects).  *ms_per_launch: mean launch duration;
 * *flop_per_launch: the launch's FLOP count (VALU: 1,283,040 per C3D20
 * element, the dense B^T D B contraction; MFMA: 860,160 per C3D20 element,
 * 2 x 16 x 16 x 4 per v_mfma_f64_16x16x4_f64 issued; 2,592 per C3D4 element
 * in either). */
fx_status fx_profile_element(fx_mesh *mesh, int32_t reps, int32_t mfma, double *ms_per_launch,
                             double *flop_per_launch);

/* SpMV autotuner (no counterpart in the reference): times the PCG SpMV
 * kernel as fx_profile_spmv does for every lane-group size 2, 4, ..., 64 on
 * the mesh's assembled K in the given precision and keeps the fastest as the
 * mesh's group; fx_solve and fx_batch_add use it whenever their options
 * leave spmv_group at 0 (until then a row-length heuristic picks it).
 * *spmv_group_out, *ms_per_launch (either may be NULL): the choice and its
 * time.  The choice is stored on the context for the mesh's bucket
 * (precision, block columns per node row, floor(log2(n_dof))): a later mesh
 * of the same bucket takes it with one timed run of that group instead of
 * the sweep.  Overwrites the mesh's PCG work vectors; call between solves. */
fx_status fx_mesh_autotune_spmv(fx_mesh *mesh, int32_t precision, int32_t reps, int32_t *spmv_group_out,
                                double *ms_per_launch);

/* Element-kernel autotuner (no counterpart in the reference): times `reps`
 * element-K launches (reference coordinates) of the C3D4 kernel at 64, 128,
 * 256 and 512 elements per workgroup and keeps the fastest for the mesh
 * (default 128); with allow_mfma != 0 it also times the C3D20 kernel on
 * VALU and on FP64 MFMA (the FX_C3D20_MFMA=1 kernel, within 1e-13 of the
 * oracle rather than bit for bit) and keeps the faster, else the VALU
 * kernel.  Choices are stored on the context per bucket (element type,
 * floor(log2(element count)), and allow_mfma for C3D20): a later mesh of a
 * timed bucket takes them untimed.  The workgroup size does not change any
 * value.  *c3d4_block_out,
 * *c3d20_mfma_out (-1 without C3D20 elements) and *ms_per_launch (the
 * chosen kernels over all elements) may be NULL.  Overwrites the element
 * buffers: call between steps or before assembling. */
fx_status fx_mesh_autotune_elements(fx_mesh *mesh, int32_t reps, int32_t allow_mfma, int32_t *c3d4_block_out,
                                    int32_t *c3d20_mfma_out, double *ms_per_launch);

/* ---- batched solves ------------------------------------------------------
 * The reference runs one job at a time: one Solver::solve per step on one K
 * (solvers/mod.rs:18-44, static_step.rs:200-228).  A batch replaces N such
 * calls on N resident meshes -- each with 1..4 load cases sharing its K --
 * by one set of kernel launches over all of them.  Per load case the
 * recurrences and stopping rule are those of fx_solve with the same options.
 * Meshes with short node rows (C3D4) run a sliced-ELL SpMV in a batch, whose
 * row sums associate differently from fx_solve's lane-group SpMV: results
 * agree to rounding.  Meshes on the lane-group SpMV (C3D20, or uploaded with
 * FX_SELL=0) give FP64 solutions bitwise equal to fx_solve's.  Either way a
 * system's result does not depend on what else shares the batch.  Systems
 * may join while others run (continuous batching); a mesh must not be
 * reassembled while it is in a batch, and destroying a mesh drops its slots.
 * All calls run on the context stream.  A mesh may belong to another context
 * on the same device: its assembly can then run on that context's stream
 * (from another host thread) while the batch runs, and fx_batch_add orders
 * the batch after the work queued on the mesh's stream so far.  Threads: the
 * calls on one batch are serialised by the batch, and destroying a mesh (or
 * its context) from another host thread while the mesh has slots in a batch
 * of another context waits for that batch's calls and drops the slots; a
 * batch itself must not be destroyed while another thread calls it.
 */
typedef struct fx_batch fx_batch;

/* opts: precision FX_PREC_FP64 or FX_PREC_MIXED (NULL = reference values). */
fx_status fx_batch_create(fx_ctx *ctx, const fx_solve_opts *opts, fx_batch **out);
void fx_batch_destroy(fx_batch *batch);

/* Admits one system: the mesh's resident K (after fx_assemble and
 * fx_apply_penalty) with n_rhs (1..4) right-hand sides, rhs[r*n_dof + i].
 * Its solve starts from x = 0 on the next fx_batch_run.  *slot identifies it
 * until fx_batch_remove. */
fx_status fx_batch_add(fx_batch *batch, fx_mesh *mesh, int32_t n_rhs, const double *rhs, int32_t *slot);

/* Assembles the mesh's linear K (Assembler::assemble, assembler.rs:37-151),
 * builds the n_rhs load vectors from n_loads load records (load case
 * load_case[k] in 0..n_rhs-1, DOF load_dof[k], value load_val[k]; per case
 * and DOF the values added in record order from 0.0, as static_step.rs:
 * 102-108 accumulates f_ext), applies the penalty BCs (static_step.rs:183-
 * 195: 1e6 x max_diag per record, duplicates adding again) to K and to each
 * load vector (u = the resident displacement), and admits the system like
 * fx_batch_add -- with no host wait: max_diag, the penalty and the
 * right-hand sides stay on the device (only the records travel), and an
 * element failure (singular Jacobian, missing material values) finishes
 * the slot with that status (fx_batch_poll).  The same arithmetic as
 * fx_assemble + fx_apply_penalty + fx_batch_add on the accumulated vectors,
 * bit for bit.  No counterpart in the reference (its loop assembles and
 * solves one system at a time). */
fx_status fx_batch_add_system(fx_batch *batch, fx_mesh *mesh, int32_t n_rhs, int64_t n_loads,
                              const int32_t *load_case, const int64_t *load_dof, const double *load_val,
                              int64_t n_bc, const int64_t *bc_dofs, const double *bc_targets, int32_t *slot);

/* Advances every running system by up to max_iters PCG iterations
 * (max_iters <= 0: until none is running).  *n_running: systems still
 * running on return, as far as the host has seen (with overlap, see
 * fx_batch_set_overlap, some may have finished on the GPU meanwhile). */
fx_status fx_batch_run(fx_batch *batch, int64_t max_iters, int32_t *n_running);

/* Overlap of the GPU's iterations with the caller's work between
 * fx_batch_run calls (no counterpart in the reference, whose solver has no
 * batch).  on != 0: a budgeted fx_batch_run returns once the budget is
 * launched and the oldest chunk of it has completed, with up to three chunks
 * (16 iterations each, or the outer iterations holding about 16 inner
 * steps) still running on the GPU, so it keeps iterating while the caller
 * collects and admits systems; the next fx_batch_run takes them up.
 * fx_batch_poll, fx_batch_finished, and fx_batch_get and fx_batch_remove
 * of finished slots do not wait for them; the calls that depend on them
 * (removing a running slot, admitting past the slot capacity, the
 * autotuner, the statistics) do.  Device memory freed meanwhile by another
 * thread (hipFree waits for the whole device) then waits for those chunks.
 * on = 0 (default): fx_batch_run returns with nothing in flight. */
fx_status fx_batch_set_overlap(fx_batch *batch, int32_t on);

/* *state: 0 running, 1 converged, -fx_status when it stopped on an error
 * (FX_ERR_NOT_CONVERGED, FX_ERR_BREAKDOWN).  report as for fx_solve
 * (final_criterion: the largest over the load cases; seconds: from
 * fx_batch_add to completion). */
fx_status fx_batch_poll(fx_batch *batch, int32_t slot, int32_t *state, fx_solve_report *report);

/* The slots whose solves have finished (converged or stopped on an error)
 * and are still in the batch, in slot order: the first `cap` of them into
 * slots[], their number into *n (which may exceed cap).  Lets a driver
 * collect results without polling every running slot. */
fx_status fx_batch_finished(fx_batch *batch, int32_t *slots, int32_t cap, int32_t *n);

/* Copies the solutions, x[r*n_dof + i].  A finished slot of a
 * polynomial batch: from the pinned host array the device filled before it
 * reported the finish (no GPU call; a batch serving fx_solve copies on the
 * device instead); another finished slot: copied on the context's copy
 * stream, past any chunks in flight; a running slot: its current iterate,
 * after the work queued so far. */
fx_status fx_batch_get(fx_batch *batch, int32_t slot, double *x);

/* Frees the slot for reuse (a running system is dropped). */
fx_status fx_batch_remove(fx_batch *batch, int32_t slot);

/* Inner-step autotuner of the polynomial batch (FX_PREC_MIXED_POLY; no
 * counterpart in the reference, whose solver has no batch).  The batch runs
 * its sliced-ELL Chebyshev step in one of two shapes: 0, one launch over
 * every running system, with the registers and LDS that four load cases
 * need; 1, one launch per load-case count, each sized for its own count
 * (more wavefronts in flight for systems with one or two load cases, at up
 * to three more launches per step).  Both compute the same values bit for
 * bit.  fx_batch_autotune times `reps` launches of each shape on the
 * batch's running systems (call it between fx_batch_run calls while systems
 * run; the timed steps do not change any solve) and keeps the faster;
 * *shape_out and the two times in ms per inner step (0 with no sliced-ELL
 * system running) may be NULL.  fx_batch_set_inner_shape sets the shape
 * directly (default 0). */
fx_status fx_batch_autotune(fx_batch *batch, int32_t reps, int32_t *shape_out, double *ms_one,
                            double *ms_split);
fx_status fx_batch_set_inner_shape(fx_batch *batch, int32_t shape);

/* Instrumentation (no counterpart in the reference): on != 0 brackets every
 * SpMV launch of later fx_batch_run calls with two HIP events on the
 * context stream. */
fx_status fx_batch_time_spmv(fx_batch *batch, int32_t on);

/* SpMV launches issued so far, launches timed and their summed duration
 * (ms), and the summed algorithmic bytes of all launches: per system whose
 * load cases are running, the operator values in the working precision, one
 * 4-byte column index per 3x3 block, node-row pointers (and the FP32
 * operator's 3x3 correction per node), plus per DOF and running load case
 * p gathered once and Ap written (16 bytes in FP64, 8 in the mixed solve,
 * whose Krylov vectors are FP32; p itself is formed by a separate streaming
 * kernel).  A mixed system on the sliced-ELL operator (C3D4 meshes) uses the
 * symmetric FP32 operator: its upper 3x3 blocks (diagonal included) count
 * once with their column index (40 bytes), each lower block, read transposed
 * from the upper record of its column's row, as its record offset and
 * column (8 bytes); set FX_SSYM=0 in the environment before fx_batch_add to
 * keep the full FP32 operator.  A mixed system's
 * residual replacement runs in the same launch: FP64 values, column indices,
 * row pointers and the Jacobi scaling, plus per DOF and load case xacc and
 * x gathered once, b read and the new residual written (28 bytes). */
fx_status fx_batch_spmv_stats(fx_batch *batch, int64_t *launches, int64_t *timed, double *ms_timed,
                              double *bytes_total);

/* ---- solve without device allocation --------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (serving loops that must not call hipMalloc, harnesses that forbid
 * it): fx_solve_csr on arrays already in device memory, with every work
 * vector (r, z, p, Ap, the Jacobi scaling, block partials, PCG scalars) in
 * one static device array of 64 MiB per GPU, resolved once per device by
 * fx_ctx_create and held until the process ends.  Same kernels, same
 * results as fx_solve_csr, bit for bit.
 *   n, nnz                 rows and stored entries; the work vectors must
 *                          fit the 64 MiB array (about 1.3 M rows), else
 *                          FX_ERR_INVALID and nothing runs
 *   row_ptr_dev, col_dev   int64, as for fx_solve_csr, in device memory;
 *                          checked on the device (row_ptr[0] = 0,
 *                          row_ptr[n] = nnz, non-decreasing, columns in
 *                          range and strictly ascending) before any gather
 *   val_dev, rhs_dev, x_dev  device memory; x_dev receives the solution
 * Runs on the context's stream and returns after the solve has finished.
 * Every call on one GPU uses that GPU's one array and overwrites what the
 * last call left there: calls that overlap on one GPU (from several host
 * threads or contexts) race and must be serialised by the caller.
 */
fx_status fx_solve_csr_noalloc(fx_ctx *ctx, int64_t n, int64_t nnz, const int64_t *row_ptr_dev,
                               const int64_t *col_dev, const double *val_dev, const double *rhs_dev, double *x_dev,
                               const fx_solve_opts *opts, fx_solve_report *report);

#ifdef __cplusplus
}
#endif
#endif /* FERRIX_HIP_H */
