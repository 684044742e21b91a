// fx_pcg.hip -- Jacobi-preconditioned CG on gfx950.
//
// Reference: src/solver/solvers/iterative.rs:26-136 (PCG, CSR SpMV) and
// src/solver/preconditioner.rs:39-58 (Jacobi).  Same recurrences, same
// per-entry arithmetic (x += a p, r -= a Ap, z = r * inv_diag,
// p = z + beta p, beta = rz_new / rz_old), the stopping rule evaluated at the
// top of every iteration i < max_iter, x0 = 0.
//
// One iteration i = two kernels and no host round trip:
//   k_spmv_p(i) : reduces the r.z / r.r partials of the previous update,
//                 evaluates the stop rule, forms p_i = z + beta p_{i-1} on the
//                 fly for every gathered column (ping-pong p buffers),
//                 Ap = K p_i and the p_i . Ap block partials;
//   k_update(i) : reduces the p.Ap partials, alpha = rz / pAp, updates x, r,
//                 z and writes the r.z / r.r block partials.
// No block waits for another inside a kernel: each kernel writes per-block
// partial sums and EVERY block of the next kernel reduces all of them in the
// same fixed order (so all blocks agree bitwise and results repeat run to
// run).  Kernel boundaries give the cross-XCD coherence, so there are no
// device-scope fences or atomics in the iteration.  Block 0 records the
// scalars the host polls; each field is written by one kernel kind and read
// only by the other one (rz is double-buffered by iteration parity).
//
// SpMV layout: the structural pattern is made of 3x3 node blocks; values of
// node row a are stored per DOF row (scalar CSR order): row 3a+p, column
// slot j (block j/3, component j%3) at val[9*nb_ptr[a] + p*3*deg + j].
// A group of G lanes owns node row a and walks the 3*deg column slots: lane
// j loads the three rows' values at j (coalesced across lanes), one column
// index per 3 slots and one entry of z and of p_{i-1}.  Workgroups are
// renumbered so that each XCD (round-robin dispatch) sweeps a contiguous
// range of node rows and its L2 keeps the gathered columns.
#include "fx_pcg_core.h"

#include <cfloat>
#include <chrono>
#include <cstdlib>

namespace fx {
namespace {

template <int G, typename VT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8))) void k_spmv_p(int64_t n_nodes, const int32_t *__restrict__ nb_ptr,
                                                   const int32_t *__restrict__ nb_col, const VT *__restrict__ val,
                                                   const double *__restrict__ z, double *__restrict__ pa,
                                                   double *__restrict__ pb, double *__restrict__ ap,
                                                   PcgScalars *__restrict__ sc, const double *__restrict__ part2,
                                                   int nparts2, double *__restrict__ part1,
                                                   const float *__restrict__ kc) {
    int64_t it;
    double beta, gamma;
    if (!pcg_head(sc, part2, nparts2, it, beta, gamma))
        return;
    // the FP64 solve never replaces its residual: keep its arithmetic as is
    const double gz = sizeof(VT) == 4 ? gamma : 1.0;
    const double *pold = (it & 1) ? pb : pa;
    double *pnew = (it & 1) ? pa : pb;

    const unsigned bid = xcd_swizzle(blockIdx.x, gridDim.x);
    const int64_t a = (int64_t)bid * (kBlock / G) + threadIdx.x / G;
    const int lane = threadIdx.x % G;
    double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
    if (a < n_nodes) {
        const int b0 = nb_ptr[a], ncol = 3 * (nb_ptr[a + 1] - b0);
        const VT *v0 = val + 9 * (int64_t)b0;
        // kUnroll slots per lane in flight; the accumulation order is slot
        // j, j+G, j+2G, ... exactly as in a plain strided loop
        for (int j0 = lane; j0 < ncol; j0 += kUnroll * G) {
            int c[kUnroll];  // DOF index (n_dof < 2^31)
            double w0[kUnroll], w1[kUnroll], w2[kUnroll], zz[kUnroll], pp[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                const int j = j0 + u * G;
                const int kb = j / 3;
                c[u] = j < ncol ? 3 * nb_col[b0 + kb] + (j - 3 * kb) : -1;
            }
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                const int j = j0 + u * G;
                if (c[u] >= 0) {
                    w0[u] = (double)v0[j];
                    w1[u] = (double)v0[ncol + j];
                    w2[u] = (double)v0[2 * ncol + j];
                    zz[u] = z[c[u]];
                    pp[u] = pold[c[u]];
                }
            }
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                if (c[u] >= 0) {
                    const double xj = (sizeof(VT) == 4 ? gz * zz[u] : zz[u]) + beta * pp[u];
                    acc0 += w0[u] * xj;
                    acc1 += w1[u] * xj;
                    acc2 += w2[u] * xj;
                }
            }
        }
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
        acc0 += __shfl_xor(acc0, off, G);
        acc1 += __shfl_xor(acc1, off, G);
        acc2 += __shfl_xor(acc2, off, G);
    }
    double loc[1] = {0.0};
    if (lane == 0 && a < n_nodes) {
        const int64_t d = 3 * a;
        const double q0 = (sizeof(VT) == 4 ? gz * z[d] : z[d]) + beta * pold[d];
        const double q1 = (sizeof(VT) == 4 ? gz * z[d + 1] : z[d + 1]) + beta * pold[d + 1];
        const double q2 = (sizeof(VT) == 4 ? gz * z[d + 2] : z[d + 2]) + beta * pold[d + 2];
        if constexpr (sizeof(VT) == 4) {
            // FP32 operator: add the node's 3x3 row-sum correction block
            const float *c = kc + 9 * a;
            acc0 += ((double)c[0] * q0 + (double)c[1] * q1) + (double)c[2] * q2;
            acc1 += ((double)c[3] * q0 + (double)c[4] * q1) + (double)c[5] * q2;
            acc2 += ((double)c[6] * q0 + (double)c[7] * q1) + (double)c[8] * q2;
        }
        pnew[d] = q0;
        pnew[d + 1] = q1;
        pnew[d + 2] = q2;
        ap[d] = acc0;
        ap[d + 1] = acc1;
        ap[d + 2] = acc2;
        loc[0] = ((0.0 + q0 * acc0) + q1 * acc1) + q2 * acc2;
    }
    block_sum<1>(loc);
    if (threadIdx.x == 0)
        part1[bid] = loc[0];  // indexed by row group, independent of dispatch order
}

// y = K x (FP64 values), same column-sliced layout; used for true residuals.
template <int G>
__global__ __launch_bounds__(kBlock) void k_spmv(int64_t n_nodes, const int32_t *__restrict__ nb_ptr,
                                                 const int32_t *__restrict__ nb_col, const double *__restrict__ val,
                                                 const double *__restrict__ x, double *__restrict__ y,
                                                 double *__restrict__ ya) {
    const unsigned bid = xcd_swizzle(blockIdx.x, gridDim.x);
    const int64_t a = (int64_t)bid * (kBlock / G) + threadIdx.x / G;
    const int lane = threadIdx.x % G;
    double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
    double abs0 = 0.0, abs1 = 0.0, abs2 = 0.0;  // |K| |x|: the magnitude the row sums cancel from
    if (a < n_nodes) {
        const int b0 = nb_ptr[a], ncol = 3 * (nb_ptr[a + 1] - b0);
        const double *v0 = val + 9 * (int64_t)b0;
        for (int j = lane; j < ncol; j += G) {
            const int kb = j / 3;
            const double xj = x[3 * (int64_t)nb_col[b0 + kb] + (j - 3 * kb)];
            acc0 += v0[j] * xj;
            acc1 += v0[ncol + j] * xj;
            acc2 += v0[2 * ncol + j] * xj;
            abs0 += fabs(v0[j] * xj);
            abs1 += fabs(v0[ncol + j] * xj);
            abs2 += fabs(v0[2 * ncol + j] * xj);
        }
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
        acc0 += __shfl_xor(acc0, off, G);
        acc1 += __shfl_xor(acc1, off, G);
        acc2 += __shfl_xor(acc2, off, G);
        abs0 += __shfl_xor(abs0, off, G);
        abs1 += __shfl_xor(abs1, off, G);
        abs2 += __shfl_xor(abs2, off, G);
    }
    if (lane == 0 && a < n_nodes) {
        y[3 * a] = acc0;
        y[3 * a + 1] = acc1;
        y[3 * a + 2] = acc2;
        if (ya) {
            ya[3 * a] = abs0;
            ya[3 * a + 1] = abs1;
            ya[3 * a + 2] = abs2;
        }
    }
}

// r = b - K x ; block partials of (r.r, b.b, r.(inv r), (|K||x|).(|K||x|)) -> part.
__global__ __launch_bounds__(kBlock) void k_true_residual(int64_t n, const double *__restrict__ b,
                                                          const double *__restrict__ kx,
                                                          const double *__restrict__ kxa,
                                                          const double *__restrict__ inv,
                                                          double *__restrict__ r, double *__restrict__ part) {
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        const double ri = b[i] - kx[i];
        r[i] = ri;
        v[0] = ri * ri;
        v[1] = b[i] * b[i];
        v[2] = ri * (ri * inv[i]);
        v[3] = kxa[i] * kxa[i];
    }
    block_sum<4>(v);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < 4; ++k)
            part[4 * blockIdx.x + k] = v[k];
}

// One block: fixed-order sum of NV-wide partials into out (device).
template <int NV>
__global__ __launch_bounds__(kBlock) void k_sum_partials(const double *__restrict__ part, int nparts,
                                                         double *__restrict__ out) {
    double red[NV];
    reduce_partials<NV>(part, nparts, red);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k)
            out[k] = red[k];
}

// Residual replacement in the mixed solve (own DOFs): x = 0, r = r64,
// z = r inv; partials (r.z, r.r, r.p) with p the kept direction (what the
// next SpMV reads as p_old) -> part2.
__global__ __launch_bounds__(kBlock) void k_resume(int64_t n, const double *__restrict__ r64,
                                                   double *__restrict__ x, double *__restrict__ r,
                                                   double *__restrict__ z, const double *__restrict__ p,
                                                   const double *__restrict__ inv, double *__restrict__ part2) {
    double v[3] = {0.0, 0.0, 0.0};
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        const double ri = r64[i];
        const double zi = ri * inv[i];
        x[i] = 0.0;
        r[i] = ri;
        z[i] = zi;
        v[0] = ri * zi;
        v[1] = ri * ri;
        v[2] = ri * p[i];
    }
    block_sum<3>(v);
    if (threadIdx.x == 0) {
        part2[3 * blockIdx.x] = v[0];
        part2[3 * blockIdx.x + 1] = v[1];
        part2[3 * blockIdx.x + 2] = v[2];
    }
}

// The next segment of the mixed solve: running again, relative stop at
// `tol` of the replaced residual (bb = its r.r), the resume flag for the
// first SpMV head, and the iterated r.r the last segment ended on.
__global__ void k_resume_scalars(PcgScalars *__restrict__ sc, double bb, double tol) {
    if (blockIdx.x != 0 || threadIdx.x != 0)
        return;
    sc->rr_prev = sc->rr;
    sc->bb = bb;
    sc->tol = tol;
    sc->stop_mode = FX_STOP_REL_R2;
    sc->status = 0;
    sc->resume = 1;
}

__global__ void k_add(int64_t n, const double *__restrict__ d, double *__restrict__ x) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        x[i] += d[i];
}

// partial buffers: part1 (one per SpMV block) then part2 (3 per update block)
inline double *part1_of(fx_mesh *m) { return m->part.p; }
inline double *part2_of(fx_mesh *m) { return m->part.p + (m->n_nodes + 64); }

template <int G, typename VT>
void launch_spmv_p(fx_mesh *m, const VT *val, hipStream_t s) {
    const unsigned nb = nblocks(m->n_nodes, kBlock / G);
    hipLaunchKernelGGL((k_spmv_p<G, VT>), dim3(nb), dim3(kBlock), 0, s, m->n_nodes, m->nb_ptr.p, m->nb_col.p, val,
                       m->z.p, m->p0.p, m->p1.p, m->ap.p, m->pcg.p, part2_of(m), (int)nblocks(m->n_dof, kBlock),
                       part1_of(m), sizeof(VT) == 4 ? m->kcorr32.p : nullptr);
}

// row groups of the SpMV (= its partial count)
inline int64_t spmv_groups(const fx_mesh *m, int G) { return nblocks(m->n_nodes, kBlock / G); }

template <typename VT>
void launch_spmv_any(fx_mesh *m, const VT *val, int G, hipStream_t s) {
    switch (G) {
    case 2: launch_spmv_p<2, VT>(m, val, s); break;
    case 4: launch_spmv_p<4, VT>(m, val, s); break;
    case 8: launch_spmv_p<8, VT>(m, val, s); break;
    case 16: launch_spmv_p<16, VT>(m, val, s); break;
    case 32: launch_spmv_p<32, VT>(m, val, s); break;
    default: launch_spmv_p<64, VT>(m, val, s); break;
    }
}

template <typename VT>
void launch_iteration(fx_mesh *m, const VT *val, int G, hipStream_t s) {
    launch_spmv_any<VT>(m, val, G, s);
    hipLaunchKernelGGL(k_update, dim3(nblocks(m->n_dof, kBlock)), dim3(kBlock), 0, s, m->n_dof, m->x.p, m->r.p,
                       m->z.p, m->p0.p, m->p1.p, m->ap.p, m->inv_diag.p, m->pcg.p, part1_of(m),
                       (int)spmv_groups(m, G), part2_of(m));
}

template <int G>
void launch_spmv_g(fx_mesh *m, const double *x, double *y, double *ya, hipStream_t s) {
    hipLaunchKernelGGL((k_spmv<G>), dim3(nblocks(m->n_nodes, kBlock / G)), dim3(kBlock), 0, s, m->n_nodes,
                       m->nb_ptr.p, m->nb_col.p, m->kval.p, x, y, ya);
}

}  // namespace

// Average columns per DOF row -> lanes per node row (about 4 slots a lane).
int default_group(fx_mesh *m) {
    const double ncol = 3.0 * (double)m->nblk / (double)std::max<int64_t>(1, m->n_nodes);
    int G = 2;
    while (G < 64 && 2 * G <= ncol / 4.0)
        G *= 2;
    return G;
}

void spmv64(fx_mesh *m, const double *x, double *y, int G, double *ya) {
    hipStream_t s = m->ctx->stream;
    switch (G) {
    case 2: launch_spmv_g<2>(m, x, y, ya, s); break;
    case 4: launch_spmv_g<4>(m, x, y, ya, s); break;
    case 8: launch_spmv_g<8>(m, x, y, ya, s); break;
    case 16: launch_spmv_g<16>(m, x, y, ya, s); break;
    case 32: launch_spmv_g<32>(m, x, y, ya, s); break;
    default: launch_spmv_g<64>(m, x, y, ya, s); break;
    }
    FX_HIP(hipGetLastError());
}

static void ensure_vectors(fx_mesh *m) {
    for (auto *b : {&m->x, &m->r, &m->z, &m->p0, &m->p1, &m->ap, &m->tmp})
        b->alloc(m->n_dof);
}

static void pcg_start(fx_mesh *m, const double *b_dev, double tol, int64_t max_iter, int stop_mode) {
    hipStream_t s = m->ctx->stream;
    PcgScalars init{};
    init.tol = tol;
    init.max_iter = max_iter;
    init.stop_mode = stop_mode;
    FX_HIP(hipMemcpyAsync(m->pcg.p, &init, sizeof init, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_init, dim3(nblocks(m->n_dof, kBlock)), dim3(kBlock), 0, s, m->n_dof, b_dev, m->x.p, m->r.p,
                       m->z.p, m->p0.p, m->inv_diag.p, part2_of(m));
    FX_HIP(hipGetLastError());
}

// Times `reps` back-to-back launches of the PCG SpMV kernel (the same
// kernel the solve runs, on the context stream) between two HIP events.
// The PCG state comes from an init pass on b = 1 with tol 0, so every launch
// does the full work (the stop rule never fires).
double profile_spmv(fx_mesh *m, int precision, int reps, int G, double *bytes) {
    hipStream_t s = m->ctx->stream;
    const int64_t n = m->n_dof;
    if (G <= 0)
        G = m->spmv_group > 0 ? m->spmv_group : default_group(m);
    ensure_vectors(m);
    launch_jacobi(m, false);
    if (precision == FX_PREC_MIXED)
        launch_convert32(m);
    launch_fill(m, m->tmp.p, n, 1.0);
    pcg_start(m, m->tmp.p, 0.0, INT64_MAX, FX_STOP_ABS_SQRT_RZ);
    hipEvent_t e0, e1;
    FX_HIP(hipEventCreate(&e0));
    FX_HIP(hipEventCreate(&e1));
    auto one = [&]() {
        if (precision == FX_PREC_MIXED)
            launch_spmv_any<float>(m, m->kval32.p, G, s);
        else
            launch_spmv_any<double>(m, m->kval.p, G, s);
    };
    for (int k = 0; k < 3; ++k)
        one();
    FX_HIP(hipEventRecord(e0, s));
    for (int k = 0; k < reps; ++k)
        one();
    FX_HIP(hipEventRecord(e1, s));
    FX_HIP(hipGetLastError());
    FX_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    FX_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const double vs = precision == FX_PREC_MIXED ? 4.0 : 8.0;
    // values + one column index per 3x3 block + node-row pointers
    // + z and p gathered (each at least once) + p_new and Ap written
    // (+ the FP32 operator's 3x3 correction block per node)
    *bytes = vs * (double)m->nnz_struct + 4.0 * (double)m->nblk + 4.0 * (double)(m->n_nodes + 1) +
             32.0 * (double)n + (precision == FX_PREC_MIXED ? 36.0 * (double)m->n_nodes : 0.0);
    return (double)ms / reps;
}

// The SpMV autotuner: times the PCG SpMV kernel (profile_spmv, `reps`
// launches after 3 warm-up launches) for every lane-group size and keeps
// the fastest as the mesh's group, which solves and batches then use when
// their options leave spmv_group at 0.
int autotune_spmv(fx_mesh *m, int precision, int reps, double *best_ms) {
    // bucket: precision, block columns per node row, DOF count
    const auto key = tune_key(1, 256 * precision + (m->nblk + m->n_nodes / 2) / std::max<int64_t>(1, m->n_nodes),
                              m->n_dof);
    const int known = tune_get(m->ctx, key);
    if (known > 0) {
        double bytes = 0.0;
        const double t = profile_spmv(m, precision, reps, known, &bytes);
        m->spmv_group = known;
        if (best_ms)
            *best_ms = t;
        return known;
    }
    int best = 0;
    double best_t = HUGE_VAL;
    for (int G = 2; G <= 64; G *= 2) {
        double bytes = 0.0;
        const double t = profile_spmv(m, precision, reps, G, &bytes);
        if (t < best_t) {
            best_t = t;
            best = G;
        }
    }
    m->spmv_group = best;
    tune_put(m->ctx, key, best);
    if (best_ms)
        *best_ms = best_t;
    return best;
}

// Each chunk of iterations (iterate_chunks) is one captured hipGraph per
// chunk size; FX_PCG_GRAPH=0 issues plain launches instead of graphs.
static bool use_graphs() {
    const char *e = std::getenv("FX_PCG_GRAPH");
    return !(e && e[0] == '0');
}

template <typename VT>
static hipGraphExec_t batch_graph(fx_mesh *m, const VT *val, int G, int chunk) {
    const int vt = sizeof(VT) == 4 ? 1 : 0;
    for (auto &g : m->graphs)
        if (g.vt == vt && g.group == G && g.chunk == chunk && g.val == (const void *)val)
            return g.exec;
    hipStream_t s = m->ctx->stream;
    hipGraph_t graph;
    FX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < chunk; ++k)
        launch_iteration<VT>(m, val, G, s);
    FX_HIP(hipStreamEndCapture(s, &graph));
    hipGraphExec_t exec;
    FX_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    (void)hipGraphDestroy(graph);
    m->graphs.push_back({vt, G, chunk, (const void *)val, exec});
    return exec;
}

// Iterates the PCG state in m (started by pcg_start, or resumed) until its
// stop rule fires; returns the final status (1 converged, 2 breakdown,
// 3 max_iter).
template <typename VT>
static int pcg_iterate(fx_mesh *m, const VT *val, double tol, int64_t max_iter, int G, PcgScalars *out,
                       int first_chunk = 8) {
    hipStream_t s = m->ctx->stream;
    const bool graphs = use_graphs();
    return iterate_chunks(
        s, m->pcg.p, tol, max_iter,
        [&](int chunk) {
            if (graphs) {
                FX_HIP(hipGraphLaunch(batch_graph<VT>(m, val, G, chunk), s));
            } else {
                for (int k = 0; k < chunk; ++k)
                    launch_iteration<VT>(m, val, G, s);
                FX_HIP(hipGetLastError());
            }
        },
        out, first_chunk);
}

// One PCG solve on the resident operator `val` with rhs b_dev; x in m->x.
template <typename VT>
static int run_pcg(fx_mesh *m, const VT *val, const double *b_dev, double tol, int64_t max_iter, int stop_mode,
                   int G, PcgScalars *out) {
    pcg_start(m, b_dev, tol, max_iter, stop_mode);
    return pcg_iterate<VT>(m, val, tol, max_iter, G, out);
}

// Continues a mixed solve after its residual was replaced by r64 (r.r =
// rr64): x = 0, r = r64, z = r inv, the kept direction's r.p, and a new
// segment that stops at a relative reduction `tol` of r64.
static void pcg_resume(fx_mesh *m, const double *r64, double rr64, double tol, const PcgScalars &h) {
    hipStream_t s = m->ctx->stream;
    const double *p = (h.it & 1) ? m->p1.p : m->p0.p;  // what the next SpMV reads as p_old
    hipLaunchKernelGGL(k_resume, dim3(nblocks(m->n_dof, kBlock)), dim3(kBlock), 0, s, m->n_dof, r64, m->x.p,
                       m->r.p, m->z.p, p, m->inv_diag.p, part2_of(m));
    hipLaunchKernelGGL(k_resume_scalars, dim3(1), dim3(64), 0, s, m->pcg.p, rr64, tol);
    FX_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// The FP64 solve in pieces for the device-resident Newton iteration
// (fx_newton.hip): a start that a graph can hold (the scalars written by a
// kernel, not copied from the host), iterations queued without a status
// read, and the adaptive chunked finish of pcg_solve.
// ---------------------------------------------------------------------------
namespace {
// The start scalars of a solve; *skip != 0 (the Newton iteration has already
// converged) marks the solve stopped before its first iteration, so every
// queued SpMV and update returns at once.
__global__ void k_pcg_set(PcgScalars *__restrict__ sc, double tol, int64_t max_iter, int stop_mode,
                          const int32_t *__restrict__ skip) {
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    PcgScalars init{};
    init.tol = tol;
    init.max_iter = max_iter;
    init.stop_mode = stop_mode;
    init.status = skip && *skip ? 1 : 0;
    *sc = init;
}
}  // namespace

int pcg_group(fx_mesh *m, const fx_solve_opts &o) {
    return o.spmv_group > 0 ? o.spmv_group : (m->spmv_group > 0 ? m->spmv_group : default_group(m));
}

void pcg_prepare(fx_mesh *m) { ensure_vectors(m); }

void pcg_start_dev(fx_mesh *m, const double *b_dev, const fx_solve_opts &o, const int32_t *skip) {
    hipStream_t s = m->ctx->stream;
    launch_jacobi(m, false);
    hipLaunchKernelGGL(k_pcg_set, dim3(1), dim3(64), 0, s, m->pcg.p, o.tol, o.max_iter, o.stop_mode, skip);
    hipLaunchKernelGGL(k_init, dim3(nblocks(m->n_dof, kBlock)), dim3(kBlock), 0, s, m->n_dof, b_dev, m->x.p, m->r.p,
                       m->z.p, m->p0.p, m->inv_diag.p, part2_of(m));
    FX_HIP(hipGetLastError());
}

void pcg_queue_fp64(fx_mesh *m, int G, int64_t iters) {
    hipStream_t s = m->ctx->stream;
    const bool graphs = use_graphs();
    while (iters > 0) {
        int chunk = 64;
        while (chunk > iters)
            chunk /= 2;
        if (graphs) {
            FX_HIP(hipGraphLaunch(batch_graph<double>(m, m->kval.p, G, chunk), s));
        } else {
            for (int k = 0; k < chunk; ++k)
                launch_iteration<double>(m, m->kval.p, G, s);
            FX_HIP(hipGetLastError());
        }
        iters -= chunk;
    }
}

int pcg_continue_fp64(fx_mesh *m, double tol, int64_t max_iter, int G, PcgScalars *out) {
    // a solve that ran past its queued iterations: chunks of 64 from the start
    return pcg_iterate<double>(m, m->kval.p, tol, max_iter, G, out, 64);
}

int64_t pcg_iterations_done(const PcgScalars &h) { return iterations_done(h); }

fx_status pcg_solve(fx_mesh *m, const double *b_dev, double *x_dev, const fx_solve_opts &o, fx_solve_report *rep) {
    if (o.precision == FX_PREC_MIXED_POLY)
        return poly_solve(m, b_dev, x_dev, o, rep);
    hipStream_t s = m->ctx->stream;
    auto t0 = std::chrono::steady_clock::now();
    const int64_t n = m->n_dof;
    const int G = o.spmv_group > 0 ? o.spmv_group : (m->spmv_group > 0 ? m->spmv_group : default_group(m));
    ensure_vectors(m);
    const bool mixed = o.precision == FX_PREC_MIXED;
    launch_jacobi(m, false);
    fx_solve_report R{};
    R.spmv_group = G;
    fx_status st = FX_OK;
    if (!mixed) {
        PcgScalars h;
        const int code = run_pcg<double>(m, m->kval.p, b_dev, o.tol, o.max_iter, o.stop_mode, G, &h);
        R.iterations = iterations_done(h);
        R.final_criterion = h.crit;
        if (code == 2)
            st = FX_ERR_BREAKDOWN;
        else if (code == 3)
            st = FX_ERR_NOT_CONVERGED;
        if (x_dev != m->x.p)
            FX_HIP(hipMemcpyAsync(x_dev, m->x.p, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    } else {
        // The FP32-stored operator with FP64 vectors, and reliable updates
        // of the residual: x = 0 (in xacc); repeat r64 = b - K64 xacc
        // (FP64), stop on the outer rule; the first time start PCG on r64,
        // afterwards replace PCG's residual by r64 keeping its direction
        // when the two residuals agree within 10% (k_resume, pcg_head); run
        // until the iterated residual has fallen by inner_tol; xacc += x.
        launch_convert32(m);
        DBuf<double> xacc, rtrue, kx, kxa, out4;
        xacc.alloc(n);
        rtrue.alloc(n);
        kx.alloc(n);
        kxa.alloc(n);
        out4.alloc(4);
        launch_fill(m, xacc.p, n, 0.0);
        double bb = -1.0, crit_prev = HUGE_VAL;
        int passes = 0;
        PcgScalars h{};
        const unsigned nb = nblocks(n, kBlock);
        for (;;) {
            spmv64(m, xacc.p, kx.p, G, kxa.p);
            hipLaunchKernelGGL(k_true_residual, dim3(nb), dim3(kBlock), 0, s, n, b_dev, kx.p, kxa.p, m->inv_diag.p,
                               rtrue.p, m->tmp.p);
            hipLaunchKernelGGL(k_sum_partials<4>, dim3(1), dim3(kBlock), 0, s, m->tmp.p, (int)nb, out4.p);
            FX_HIP(hipGetLastError());
            double h3[4];
            FX_HIP(hipMemcpyAsync(h3, out4.p, sizeof h3, hipMemcpyDeviceToHost, s));
            FX_HIP(hipStreamSynchronize(s));
            ++t_host_syncs;
            R.outer_spmv++;
            if (bb < 0.0)
                bb = h3[1];
            double crit;
            bool conv;
            if (o.stop_mode == FX_STOP_ABS_SQRT_RZ) {
                crit = std::sqrt(h3[2]);
                conv = crit < o.tol;
            } else {
                crit = bb > 0.0 ? std::sqrt(h3[0]) / std::sqrt(bb) : 0.0;
                conv = crit <= o.tol;
            }
            conv = conv || at_rounding_floor(crit, crit_prev, std::sqrt(h3[0]), std::sqrt(h3[3]));
            crit_prev = crit;
            R.final_criterion = crit;
            if (conv)
                break;
            if (passes >= o.max_refine || (passes > 0 && h.it >= o.max_iter) || !std::isfinite(crit)) {
                st = FX_ERR_NOT_CONVERGED;
                break;
            }
            int code;
            if (passes == 0) {
                code = run_pcg<float>(m, m->kval32.p, rtrue.p, o.inner_tol, o.max_iter, FX_STOP_REL_R2, G, &h);
            } else {
                pcg_resume(m, rtrue.p, h3[0], o.inner_tol, h);
                code = pcg_iterate<float>(m, m->kval32.p, o.inner_tol, o.max_iter, G, &h);
            }
            R.iterations = iterations_done(h);  // the iteration index runs on across segments
            ++passes;
            if (code == 2) {
                st = FX_ERR_BREAKDOWN;
                break;
            }
            hipLaunchKernelGGL(k_add, dim3(nblocks(n, 256)), dim3(256), 0, s, n, m->x.p, xacc.p);
            FX_HIP(hipGetLastError());
        }
        R.refinements = passes;
        FX_HIP(hipMemcpyAsync(x_dev, xacc.p, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    FX_HIP(hipStreamSynchronize(s));
    ++t_host_syncs;
    R.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rep)
        *rep = R;
    return st;
}

}  // namespace fx
