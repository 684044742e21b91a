// fx_internal.h -- shared declarations of the gfx950 backend (not part of the ABI).
#pragma once

#include <cfloat>

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <array>
#include <map>
#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ferrix_hip.h"

struct fx_mesh;

namespace fx {
constexpr int kSliceRows = 64;  // node rows per slice of the sliced-ELL operator (one wavefront)
constexpr int kPolyMax = 32;    // most Chebyshev inner steps per outer iteration (polynomial batch)
// Host waits on the device (stream synchronisations) made by this thread;
// fx_static_step reports the count of a step
extern thread_local int64_t t_host_syncs;
}

namespace fx {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
struct Error {
    fx_status code;
    std::string msg;
};

#define FX_HIP(call)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            throw ::fx::Error{e_ == hipErrorOutOfMemory ? FX_ERR_OOM : FX_ERR_HIP,                     \
                              std::string(#call) + ": " + hipGetErrorString(e_)};                      \
    } while (0)

// ---------------------------------------------------------------------------
// device buffer (RAII)
// ---------------------------------------------------------------------------
template <class T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p)
            (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        if (count == n && p)
            return;
        release();
        if (count == 0)
            return;
        FX_HIP(hipMalloc(&p, count * sizeof(T)));
        n = count;
    }
    void upload(const T *h, size_t count, hipStream_t s) {
        alloc(count);
        if (count)
            FX_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T> &v, hipStream_t s) { upload(v.data(), v.size(), s); }
    size_t bytes() const { return n * sizeof(T); }
};

// ---------------------------------------------------------------------------
// constants shared by host and device
// ---------------------------------------------------------------------------
constexpr int kLutMax = FX_LUT_MAX;
constexpr double kDropTol = 1e-12;        // assembler.rs:135
constexpr double kPenaltyScale = 1.0e6;   // static_step.rs:185
constexpr double kJacobiGuard = 1e-9;     // preconditioner.rs:44

// material LUT row on the device
struct MatLut {
    int32_t n_e, n_nu;
    double t_e[kLutMax], v_e[kLutMax], t_nu[kLutMax], v_nu[kLutMax];
};

// Gauss table of one element type: dN_local (3 x nn) and N (nn) per point.
struct GaussTable {
    int ngp, nn;
    std::vector<double> dN;  // ngp * 3 * nn
    std::vector<double> N;   // ngp * nn
    std::vector<double> w;   // ngp
};
GaussTable gauss_table(int nn);  // fx_shape.cpp
std::vector<double> c3d20_node_dn();  // fx_shape.cpp: dN at the 20 node positions

// device-side status word written by kernels
struct DevStatus {
    int32_t code;        // fx_status of the first failure
    int32_t pad;
    int64_t bad_element; // smallest failing element position (atomicMin)
};

// element block of one type inside a mesh
struct ElemBlock {
    int nn = 0;
    int64_t n_el = 0;
    DBuf<int32_t> conn;     // n_el * nn node indices
    DBuf<int32_t> mat;      // n_el
    DBuf<int64_t> gpos;     // n_el global element positions (error reporting)
    int64_t ke_offset = 0;  // first value of this block in the K contribution buffer
    int64_t fe_offset = 0;  // first value in the f contribution buffer
};

// PCG state in device memory.  Each field is written by block 0 of one
// kernel kind only and read by the other kind (rz is double-buffered by
// iteration parity), so no kernel reads a field it may be writing.
struct PcgScalars {
    double tol;
    int64_t max_iter;
    int32_t stop_mode;
    int32_t status;  // 0 running, 1 converged, 2 breakdown, 3 max_iter
    int64_t it;      // next iteration index (written by k_update)
    int64_t it_k1;   // current iteration index (written by k_spmv_p)
    double rz[2];    // r.z at iteration parity (k_spmv_p)
    double rr, bb, crit;  // last r.r, b.b, stop-rule value (k_spmv_p)
    double pap, alpha;    // last p.Ap and alpha (k_update)
    // mixed precision: the first iteration after a residual replacement
    // (set by k_resume_scalars, cleared by k_update) decides whether the
    // direction is kept, comparing the replaced r.r with rr_prev, the
    // iterated r.r the previous segment ended on
    int32_t resume;
    int32_t pad;
    double rr_prev;
};

// One captured batch of PCG iterations.
struct PcgGraph {
    int vt, group, chunk;
    const void *val;
    hipGraphExec_t exec;
};

// Gauss tables in device memory, one set per context (uploaded at creation).
struct DevTables {
    DBuf<double> n4, dn4, w4, n20, dn20, w20;
    DBuf<double> dn20n;  // dN at the C3D20 node positions (stress recovery)
};

}  // namespace fx

struct fx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // result copies that must not wait behind the work queued on `stream`
    // (fx_batch_get of finished slots); made on first use (copy_once), so
    // contexts that never copy results (admission, Newton workers) hold
    // no second stream
    hipStream_t copy_stream = nullptr;
    std::once_flag copy_once;
    int n_cu = 0;
    fx::DevTables tables;
    std::mutex reg;                   // guards the two registries below
    std::vector<fx_mesh *> meshes;    // live meshes, destroyed with the context
    std::vector<fx_batch *> batches;  // live batches, destroyed with the context
    // device blocks of released batch slots, reused by later slots (bytes ->
    // block); freed with the context
    std::mutex pool_mx;
    std::multimap<size_t, void *> pool;
    size_t pool_bytes = 0;
    // the same for pinned host blocks (the polynomial batch's result
    // exports), freed with the context
    std::mutex pin_mx;
    std::multimap<size_t, void *> pin_pool;
    size_t pin_bytes = 0;
    // autotuner choices by size bucket (fx::tune_key): a mesh whose bucket
    // was already timed on this context takes the stored choice untimed
    std::mutex tune_mx;
    std::map<std::array<int64_t, 3>, int> tune;
};

namespace fx {
// Autotuner bucket: (what was tuned, a shape class, floor(log2(size))).
inline std::array<int64_t, 3> tune_key(int64_t what, int64_t cls, int64_t size) {
    int64_t lg = 0;
    while (size > 1) {
        size >>= 1;
        ++lg;
    }
    return {what, cls, lg};
}
// The stored choice of a bucket, or -1.
inline int tune_get(fx_ctx *ctx, const std::array<int64_t, 3> &k) {
    std::lock_guard<std::mutex> lk(ctx->tune_mx);
    auto it = ctx->tune.find(k);
    return it == ctx->tune.end() ? -1 : it->second;
}
inline void tune_put(fx_ctx *ctx, const std::array<int64_t, 3> &k, int v) {
    std::lock_guard<std::mutex> lk(ctx->tune_mx);
    ctx->tune[k] = v;
}

// A device block of at least `bytes` from the context's pool (the smallest
// that fits, if it is at most twice the request), else a new hipMalloc;
// *cap: the block's size.
void *pool_take(fx_ctx *ctx, size_t bytes, size_t *cap);
// Returns a block to the pool (freed instead once the pool holds 16 GiB);
// work that used it must have completed on the context stream.
void pool_give(fx_ctx *ctx, void *p, size_t cap);
void pool_clear(fx_ctx *ctx);
// Pinned host blocks (hipHostMalloc), pooled the same way (at most 4 GiB
// kept); the device may write such a block only while its owner holds it.
void *pin_take(fx_ctx *ctx, size_t bytes, size_t *cap);
void pin_give(fx_ctx *ctx, void *p, size_t cap);
void pin_clear(fx_ctx *ctx);

// A device array owned through the context's pool.
template <typename T>
struct PoolBuf {
    fx_ctx *ctx = nullptr;
    T *p = nullptr;
    size_t n = 0, cap = 0;
    PoolBuf() = default;
    PoolBuf(const PoolBuf &) = delete;
    PoolBuf &operator=(const PoolBuf &) = delete;
    ~PoolBuf() { release(); }
    void release() {
        if (p)
            pool_give(ctx, p, cap);
        p = nullptr;
        n = cap = 0;
    }
    void alloc(fx_ctx *c, size_t count) {
        release();
        ctx = c;
        if (count == 0)
            return;
        p = (T *)pool_take(c, count * sizeof(T), &cap);
        n = count;
    }
    void upload(fx_ctx *c, const std::vector<T> &v, hipStream_t s) {
        alloc(c, v.size());
        if (n)
            FX_HIP(hipMemcpyAsync(p, v.data(), n * sizeof(T), hipMemcpyHostToDevice, s));
    }
};

// A pinned host array from the context's pinned pool (returned on release).
template <typename T>
struct PinBuf {
    fx_ctx *ctx = nullptr;
    T *p = nullptr;
    size_t n = 0, cap = 0;
    PinBuf() = default;
    PinBuf(const PinBuf &) = delete;
    PinBuf &operator=(const PinBuf &) = delete;
    ~PinBuf() { release(); }
    void release() {
        if (p)
            pin_give(ctx, p, cap);
        p = nullptr;
        n = cap = 0;
    }
    void alloc(fx_ctx *c, size_t count) {
        release();
        ctx = c;
        if (count == 0)
            return;
        p = (T *)pin_take(c, count * sizeof(T), &cap);
        n = count;
    }
};
}  // namespace fx

struct fx_mesh {
    fx_ctx *ctx = nullptr;
    // slots of this mesh in batches of other contexts (a slot decrements it
    // when it goes): while 0, destroying the mesh touches no such batch
    std::atomic<int> foreign_slots{0};
    int64_t n_nodes = 0, n_dof = 0, n_el = 0;
    // node-block structural pattern (3x3 blocks)
    int64_t nblk = 0;           // structural node-pair blocks
    int64_t nnz_struct = 0;     // 9 * nblk
    std::vector<int32_t> h_nb_ptr, h_nb_col;
    fx::DBuf<double> xyz, temp;
    fx::DBuf<fx::MatLut> mats;
    int n_mat = 0;
    fx::ElemBlock blk4, blk20;
    int64_t ke_total = 0, fe_total = 0;
    fx::DBuf<double> ke, fe;             // element contributions
    fx::DBuf<int32_t> nb_ptr, nb_col;    // n_nodes+1, nblk
    fx::DBuf<int32_t> blk_row;              // per 3x3 block: its row node
    fx::DBuf<int32_t> blk_cptr, blk_cbase;  // per block: contribution bases into ke
    fx::DBuf<int32_t> nod_cptr, nod_cbase;  // per node: bases into fe
    fx::DBuf<int32_t> diag_slot;         // per dof: slot of K[d,d]
    fx::DBuf<double> kval;               // nnz_struct values (super-pattern)
    fx::DBuf<float> kval32;              // FP32 operator of the mixed solve ...
    fx::DBuf<float> kcorr32;             // ... and its per-node 3x3 row-sum corrections
    fx::DBuf<_Float16> kval16, sval16;   // Jacobi-scaled FP16 operator (node blocks / sliced ELL)
    // sliced-ELL copy of the operator for the solve (meshes with short node
    // rows, see fx_sell.h): slice pointers, column nodes, values
    bool sell = false;
    fx::DBuf<int32_t> sl_ptr, scol;
    fx::DBuf<double> sval64;
    fx::DBuf<float> sval32;
    // symmetric sliced ELL (FP32, the mixed batch's operator on sliced-ELL
    // meshes): per row its upper blocks (column >= row) stored like sval
    // (uval, su_ptr, scol_u; padding: zero values, own column), and per
    // lower block the index of its transpose's record in the owner row
    // (lrec: kk * 576 + lane) with its column (lcol); padding points at the
    // zero record after the last block column.  lblk: the transpose's
    // node-block index (conversion only).  kcorr32s: the row-sum correction
    // of this operator.
    bool ssym = false;  // built (int32 record offsets fit, pattern symmetric)
    fx::DBuf<int32_t> su_ptr, scol_u, sll_ptr, lrec, lcol, lblk, ndiag;
    fx::DBuf<float> uval32, kcorr32s;
    int64_t nblk_u = 0, nblk_l = 0;       // upper (diagonal included) and lower blocks
    int64_t su_cols = 0;                  // block columns of the upper part (the zero record follows)
    fx::DBuf<uint8_t> keep;              // 1 if the slot survives the drop rule
    fx::DBuf<double> inv_diag;
    fx::DBuf<float> inv_diag32;
    fx::DBuf<unsigned long long> max_diag_bits;
    fx::DBuf<fx::DevStatus> status;
    fx::DBuf<fx::PcgScalars> pcg;
    fx::DBuf<double> part;               // reduction partials
    fx::DBuf<uint32_t> counters;
    // vectors
    fx::DBuf<double> u, f_int, f_ext, rhs, du, x, r, z, p0, p1, ap, tmp;
    fx::DBuf<double> se;                 // stress recovery: 12 values per (element, node)
    fx::DBuf<double> red2;               // 2 doubles: norm results
    double last_max_diag = 0.0;
    // device flag that, when set, makes the K assembly (element K, max_diag
    // reset, gather, device-side penalty) return at once: the converged
    // Newton iteration's decision (fx_newton.hip); nullptr outside it
    const int32_t *k_skip = nullptr;
    // fx_batch_add_system: the penalty records of the last admission (kept
    // while they repeat), the right-hand sides it penalised, and the event
    // after which another context's batch has read them
    fx::DBuf<int32_t> adm_ud_dof, adm_ud_ptr, adm_rec_order;
    fx::DBuf<double> adm_target, adm_rhs;
    fx::DBuf<int32_t> adm_lkey;          // the load records' (case, DOF) targets, rhs-major
    fx::DBuf<double> adm_lval;           // and their sums
    std::vector<int64_t> adm_key_dof;
    std::vector<double> adm_key_val;
    int64_t adm_n_unique = 0;
    hipEvent_t adm_rhs_ev = nullptr;
    bool k_valid = false;
    int spmv_group = 0;                  // lanes per node row from fx_mesh_autotune_spmv (0: heuristic)
    int c3d4_block = 128;                // threads per workgroup of k_c3d4 (fx_mesh_autotune_elements)
    int c20_mfma = -1;                   // C3D20 element K on MFMA: 1 / 0, -1: FX_C3D20_MFMA decides
    int64_t device_bytes = 0;
    std::vector<fx::PcgGraph> graphs;    // captured iteration batches
    // fx_solve / fx_static_step in FX_PREC_MIXED_POLY: a polynomial batch on
    // the mesh's context with this mesh as its only system (fx::poly_solve),
    // destroyed with the mesh
    fx_batch *poly_batch = nullptr;
    ~fx_mesh() {
        for (auto &g : graphs)
            (void)hipGraphExecDestroy(g.exec);
        if (adm_rhs_ev)
            (void)hipEventDestroy(adm_rhs_ev);
    }
};

namespace fx {
// launchers (defined in the .hip translation units)
void launch_element(fx_mesh *m, bool want_k, bool want_f, bool nlgeom);
double profile_element(fx_mesh *m, int reps, bool mfma, double *flop, int types = 3);  // ms per launch
double autotune_elements(fx_mesh *m, int reps, bool allow_mfma);                      // ms per launch
void launch_recover(fx_mesh *m);         // per (element, node) strain and stress at the resident u
void launch_gather_recover(fx_mesh *m, double *S, double *E);  // nodal averages (device arrays)
void launch_assemble(fx_mesh *m);
void launch_gather_f(fx_mesh *m);
// s: the stream to launch on (nullptr: the mesh's context stream)
void launch_jacobi(fx_mesh *m, bool want32, hipStream_t s = nullptr);
void launch_convert32(fx_mesh *m, hipStream_t s = nullptr);
void launch_sell(fx_mesh *m, bool fp32, hipStream_t s = nullptr);  // sliced-ELL copy of kval (FP64) or kval32 (FP32)
void launch_ssym(fx_mesh *m, hipStream_t s = nullptr);  // symmetric FP32 sliced-ELL operator + its row-sum correction
void launch_scaled16(fx_mesh *m, hipStream_t s = nullptr);  // Jacobi-scaled FP16 operator (after launch_jacobi)
void check_status(fx_mesh *m, const char *what);
void reset_status(fx_mesh *m);
double device_max_diag(fx_mesh *m);
// PCG (fx_pcg.hip)
fx_status pcg_solve(fx_mesh *m, const double *b_dev, double *x_dev, const fx_solve_opts &o, fx_solve_report *rep);
// pcg_solve in FX_PREC_MIXED_POLY (fx_batch.hip)
fx_status poly_solve(fx_mesh *m, const double *b_dev, double *x_dev, const fx_solve_opts &o, fx_solve_report *rep);
// unlinks and frees the mesh's polynomial batch (the caller holds the context's registry lock)
void drop_poly_batch(fx_mesh *m);
// the FP64 solve in pieces (device-resident Newton iteration, fx_newton.hip):
// Jacobi + start scalars written on the device (*skip != 0: stopped before
// the first iteration) + x = 0, r = b; `iters` iterations queued with no
// status read; the adaptive chunked finish; SpMVs done of a final state
int pcg_group(fx_mesh *m, const fx_solve_opts &o);
void pcg_prepare(fx_mesh *m);  // the solve's vectors (allocates; not under graph capture)
void pcg_start_dev(fx_mesh *m, const double *b_dev, const fx_solve_opts &o, const int32_t *skip);
void pcg_queue_fp64(fx_mesh *m, int group, int64_t iters);
int pcg_continue_fp64(fx_mesh *m, double tol, int64_t max_iter, int group, PcgScalars *out);
int64_t pcg_iterations_done(const PcgScalars &h);
void spmv64(fx_mesh *m, const double *x, double *y, int group, double *abs_y = nullptr);
double profile_spmv(fx_mesh *m, int precision, int reps, int group, double *bytes);
int default_group(fx_mesh *m);  // lanes per node row for the SpMV (row-length heuristic)
int autotune_spmv(fx_mesh *m, int precision, int reps, double *best_ms);  // sets m->spmv_group
void measure_fp64_peak(fx_ctx *ctx, int reps, double *valu_tflops, double *mfma_tflops);  // fx_peak.hip
void set_error(const std::string &msg);  // message of the calling thread's last failure
// Solver::solve on a caller's CSR (fx_csr.hip): host arrays (work vectors
// from the context pool), or device arrays with the work vectors in the
// per-device scratch array (no device allocation)
fx_status solve_csr_host(fx_ctx *ctx, int64_t n, const int64_t *row_ptr, const int64_t *col_idx,
                         const double *values, const double *rhs, double *x_out, const fx_solve_opts &o,
                         fx_solve_report *rep);
fx_status solve_csr_noalloc(fx_ctx *ctx, int64_t n, int64_t nnz, const int64_t *row_ptr_dev,
                            const int64_t *col_dev, const double *val_dev, const double *rhs_dev, double *x_dev,
                            const fx_solve_opts &o, fx_solve_report *rep);
void csr_scratch_setup(int device);  // resolves the device's scratch address (device current)
size_t csr_scratch_bytes();
// Mixed-precision stop on the true FP64 residual b - K x, which has a
// rounding floor: each row sum cancels terms of total magnitude (|K| |x|)_i,
// so the computed residual cannot fall much below eps * || |K| |x| ||.  (The
// FP64 PCG's recursively updated residual keeps falling below that floor,
// which is why its stop test can ask for more than the true residual can
// show -- for a small right-hand side, a Newton correction, by orders of
// magnitude.)  A residual evaluation that no longer halves the criterion
// while ||r|| is within 64 eps of || |K| |x| || has reached the floor: no
// further correction can improve x.
__host__ __device__ inline bool at_rounding_floor(double crit, double crit_prev, double rnorm, double absnorm) {
    return crit > 0.5 * crit_prev && rnorm <= 64.0 * DBL_EPSILON * absnorm;
}
void batch_free(fx_batch *b);                 // deletes a batch (its context's stream is idle)
fx_status batch_add_impl(fx_batch *B, fx_mesh *m, int32_t n_rhs, const double *rhs, bool rhs_dev,
                         const DevStatus *elem_status, int32_t *slot_out);
void batch_drop_mesh(fx_batch *b, fx_mesh *m);  // removes the slots that use m
void batch_drop_mesh_foreign(fx_mesh *m);        // the same in live batches of other contexts
void axpy_dev(fx_mesh *m, int64_t n, double a, const double *x, double *y);
// step helpers (fx_assemble.hip)
void launch_residual(fx_mesh *m, const double *f_ext, const double *f_int, double *r, int64_t n);
void launch_masked_norms(fx_mesh *m, const double *r, const double *f, const uint8_t *mask, double *out2);
void launch_penalty(fx_mesh *m, int64_t n_unique, const int32_t *ud_dof, const int32_t *ud_ptr,
                    const double *rec_target, const int32_t *rec_order, double *rhs, const double *u);
// the same with the factor 1e6 x max_diag read on the device from the last
// assembly (no penalty when max_diag is not positive)
void launch_penalty_dev(fx_mesh *m, int64_t n_unique, const int32_t *ud_dof, const int32_t *ud_ptr,
                        const double *rec_target, const int32_t *rec_order, double *rhs, const double *u,
                        bool rhs_only = false);  // rhs_only: another right-hand side, K untouched
void launch_reset_status(fx_mesh *m);  // reset_status by a kernel (graph capture)
void launch_scatter(fx_mesh *m, int64_t n, const int32_t *idx, const double *val, double *dst);
void launch_fill(fx_mesh *m, double *dst, int64_t n, double v, hipStream_t s = nullptr);
int64_t count_kept(fx_mesh *m);
}  // namespace fx
