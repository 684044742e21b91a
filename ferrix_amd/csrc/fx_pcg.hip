// fx_pcg.hip -- Jacobi-preconditioned CG on gfx950.
//
// Reference: src/solver/solvers/iterative.rs:26-136 (PCG, CSR SpMV) and
// src/solver/preconditioner.rs:39-58 (Jacobi).  Same recurrences, same
// per-entry arithmetic (x += a p, r -= a Ap, z = r * inv_diag,
// p = z + beta p, beta = rz_new / rz_old), same stopping rule evaluated at
// the top of every iteration, x0 = 0.  Dot products and row sums are
// reduced in a fixed order (lane groups, then waves, then blocks in index
// order), so a solve is bitwise repeatable.
//
// One iteration = two kernels, no host round trip:
//   k_spmv_p : p_i = z + beta p_{i-1} formed on the fly for every gathered
//              column (ping-pong p buffers), Ap = K p_i, p_i . Ap;
//              the last block computes alpha.
//   k_update : x, r, z updates, r.z and r.r; the last block computes beta,
//              advances the iteration count and evaluates the stop rule.
// Kernels read the iteration state from device memory, so the host issues
// batches of iterations (captured once in a hipGraph) and polls the status
// word between batches; iterations after convergence return at once.
//
// SpMV layout: the structural pattern is made of 3x3 node blocks.  Values of
// node row a are stored row-major per DOF row (the scalar CSR order), so
// block k of row a holds rows p=0..2 at val[9*nb_ptr[a] + p*3*deg + 3k + q].
// A group of G lanes owns one node row; a lane reads one column index per
// 9 values and gathers 3 consecutive entries of z and p.  Workgroups are
// renumbered so that each XCD (round-robin dispatch) sweeps a contiguous
// range of node rows, which keeps gathered columns in its own L2.
#include "fx_internal.h"

#include <chrono>

namespace fx {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ unsigned xcd_swizzle(unsigned id, unsigned nb) {
    const unsigned q = nb / 8, r = nb % 8, x = id % 8, i = id / 8;
    return x * q + (x < r ? x : r) + i;
}

template <int NV>
__device__ __forceinline__ void block_reduce_to_part(double (&v)[NV], double *part) {
    __shared__ double s[kBlock / 64][NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double t = v[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
            t += __shfl_xor(t, off, 64);
        v[k] = t;
    }
    const int w = threadIdx.x / 64;
    if (threadIdx.x % 64 == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k)
            s[w][k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double t = 0.0;
#pragma unroll
            for (int j = 0; j < kBlock / 64; ++j)
                t += s[j][k];
            part[NV * blockIdx.x + k] = t;
        }
    }
}

// Returns true in thread 0 of the last block to arrive; that thread then
// sums the partials in block order into out[].
template <int NV>
__device__ __forceinline__ bool last_block_sum(const double *part, uint32_t *cnt, double (&out)[NV]) {
    __shared__ bool last;
    if (threadIdx.x == 0) {
        __threadfence();
        last = atomicAdd(cnt, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last || threadIdx.x != 0)
        return false;
    __threadfence();
#pragma unroll
    for (int k = 0; k < NV; ++k)
        out[k] = 0.0;
    for (unsigned b = 0; b < gridDim.x; ++b)
#pragma unroll
        for (int k = 0; k < NV; ++k)
            out[k] += __hip_atomic_load(&part[NV * b + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    *cnt = 0;
    return true;
}

__device__ __forceinline__ void stop_rule(PcgScalars *sc) {
    double crit;
    bool conv;
    if (sc->stop_mode == FX_STOP_ABS_SQRT_RZ) {
        crit = sqrt(sc->rz);
        conv = crit < sc->tol;
    } else {
        crit = sc->bb > 0.0 ? sqrt(sc->rr) / sqrt(sc->bb) : 0.0;
        conv = crit <= sc->tol;
    }
    sc->crit = crit;
    if (conv)
        sc->status = 1;
    else if (sc->iter >= sc->max_iter)
        sc->status = 3;
}

__global__ __launch_bounds__(kBlock) void k_init(int64_t n, const double *__restrict__ b, double *__restrict__ x,
                                                 double *__restrict__ r, double *__restrict__ z,
                                                 double *__restrict__ p0, const double *__restrict__ inv,
                                                 PcgScalars *__restrict__ sc, double *__restrict__ part) {
    double v[3] = {0.0, 0.0, 0.0};
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        double bi = b[i];
        double zi = bi * inv[i];
        x[i] = 0.0;
        r[i] = bi;
        z[i] = zi;
        p0[i] = zi;
        v[0] = bi * zi;
        v[1] = bi * bi;
        v[2] = bi * bi;
    }
    block_reduce_to_part<3>(v, part);
    double s[3];
    if (last_block_sum<3>(part, &sc->cnt_b, s)) {
        sc->rz = s[0];
        sc->rz0 = s[0];
        sc->rr = s[1];
        sc->bb = s[2];
        sc->beta = 0.0;
        sc->iter = 0;
        sc->status = 0;
        stop_rule(sc);
    }
}

template <int G, typename VT>
__global__ __launch_bounds__(kBlock) void k_spmv_p(int64_t n_nodes, const int32_t *__restrict__ nb_ptr,
                                                   const int32_t *__restrict__ nb_col, const VT *__restrict__ val,
                                                   const double *__restrict__ z, double *__restrict__ pa,
                                                   double *__restrict__ pb, double *__restrict__ ap,
                                                   PcgScalars *__restrict__ sc, double *__restrict__ part) {
    if (sc->status != 0)
        return;
    const double beta = sc->beta;
    const int64_t it = sc->iter;
    const double *pold = (it & 1) ? pb : pa;
    double *pnew = (it & 1) ? pa : pb;
    const unsigned bid = xcd_swizzle(blockIdx.x, gridDim.x);
    const int64_t a = (int64_t)bid * (kBlock / G) + threadIdx.x / G;
    const int lane = threadIdx.x % G;
    double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
    if (a < n_nodes) {
        const int b0 = nb_ptr[a], deg = nb_ptr[a + 1] - b0;
        const VT *v = val + 9 * (int64_t)b0;
        const int stride = 3 * deg;
        for (int k = lane; k < deg; k += G) {
            const int64_t c = 3 * (int64_t)nb_col[b0 + k];
            const double x0 = z[c] + beta * pold[c];
            const double x1 = z[c + 1] + beta * pold[c + 1];
            const double x2 = z[c + 2] + beta * pold[c + 2];
            const VT *w = v + 3 * k;
            acc0 += (double)w[0] * x0;
            acc0 += (double)w[1] * x1;
            acc0 += (double)w[2] * x2;
            acc1 += (double)w[stride] * x0;
            acc1 += (double)w[stride + 1] * x1;
            acc1 += (double)w[stride + 2] * x2;
            acc2 += (double)w[2 * stride] * x0;
            acc2 += (double)w[2 * stride + 1] * x1;
            acc2 += (double)w[2 * stride + 2] * x2;
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
        const double q0 = z[d] + beta * pold[d];
        const double q1 = z[d + 1] + beta * pold[d + 1];
        const double q2 = z[d + 2] + beta * pold[d + 2];
        pnew[d] = q0;
        pnew[d + 1] = q1;
        pnew[d + 2] = q2;
        ap[d] = acc0;
        ap[d + 1] = acc1;
        ap[d + 2] = acc2;
        loc[0] = ((0.0 + q0 * acc0) + q1 * acc1) + q2 * acc2;
    }
    block_reduce_to_part<1>(loc, part);
    double s[1];
    if (last_block_sum<1>(part, &sc->cnt_a, s)) {
        const double pap = s[0];
        const double alpha = sc->rz / pap;
        sc->pap = pap;
        sc->alpha = alpha;
        if (!(pap > 0.0) || !isfinite(alpha))
            sc->status = 2;
    }
}

__global__ __launch_bounds__(kBlock) void k_update(int64_t n, double *__restrict__ x, double *__restrict__ r,
                                                   double *__restrict__ z, const double *__restrict__ pa,
                                                   const double *__restrict__ pb, const double *__restrict__ ap,
                                                   const double *__restrict__ inv, PcgScalars *__restrict__ sc,
                                                   double *__restrict__ part) {
    if (sc->status != 0)
        return;
    const double alpha = sc->alpha;
    const int64_t it = sc->iter;
    const double *p = (it & 1) ? pa : pb;  // p_i was written to the "new" buffer
    double v[2] = {0.0, 0.0};
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        const double xi = x[i] + alpha * p[i];
        const double ri = r[i] - alpha * ap[i];
        const double zi = ri * inv[i];
        x[i] = xi;
        r[i] = ri;
        z[i] = zi;
        v[0] = ri * zi;
        v[1] = ri * ri;
    }
    block_reduce_to_part<2>(v, part);
    double s[2];
    if (last_block_sum<2>(part, &sc->cnt_b, s)) {
        sc->beta = s[0] / sc->rz;
        sc->rz = s[0];
        sc->rr = s[1];
        sc->iter = it + 1;
        stop_rule(sc);
    }
}

// y = K x (plain SpMV, FP64 values), used for true residuals.
template <int G>
__global__ __launch_bounds__(kBlock) void k_spmv(int64_t n_nodes, const int32_t *__restrict__ nb_ptr,
                                                 const int32_t *__restrict__ nb_col, const double *__restrict__ val,
                                                 const double *__restrict__ x, double *__restrict__ y) {
    const unsigned bid = xcd_swizzle(blockIdx.x, gridDim.x);
    const int64_t a = (int64_t)bid * (kBlock / G) + threadIdx.x / G;
    const int lane = threadIdx.x % G;
    double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
    if (a < n_nodes) {
        const int b0 = nb_ptr[a], deg = nb_ptr[a + 1] - b0;
        const double *v = val + 9 * (int64_t)b0;
        const int stride = 3 * deg;
        for (int k = lane; k < deg; k += G) {
            const int64_t c = 3 * (int64_t)nb_col[b0 + k];
            const double x0 = x[c], x1 = x[c + 1], x2 = x[c + 2];
            const double *w = v + 3 * k;
            acc0 += w[0] * x0;
            acc0 += w[1] * x1;
            acc0 += w[2] * x2;
            acc1 += w[stride] * x0;
            acc1 += w[stride + 1] * x1;
            acc1 += w[stride + 2] * x2;
            acc2 += w[2 * stride] * x0;
            acc2 += w[2 * stride + 1] * x1;
            acc2 += w[2 * stride + 2] * x2;
        }
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
        acc0 += __shfl_xor(acc0, off, G);
        acc1 += __shfl_xor(acc1, off, G);
        acc2 += __shfl_xor(acc2, off, G);
    }
    if (lane == 0 && a < n_nodes) {
        y[3 * a] = acc0;
        y[3 * a + 1] = acc1;
        y[3 * a + 2] = acc2;
    }
}

// r = b - K x ; returns via part/last block: r.r, b.b, r.(inv r)
__global__ __launch_bounds__(kBlock) void k_true_residual(int64_t n, const double *__restrict__ b,
                                                          const double *__restrict__ kx, const double *__restrict__ inv,
                                                          double *__restrict__ r, double *__restrict__ part,
                                                          uint32_t *__restrict__ cnt, double *__restrict__ out) {
    double v[3] = {0.0, 0.0, 0.0};
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        const double ri = b[i] - kx[i];
        r[i] = ri;
        v[0] = ri * ri;
        v[1] = b[i] * b[i];
        v[2] = ri * (ri * inv[i]);
    }
    block_reduce_to_part<3>(v, part);
    double s[3];
    if (last_block_sum<3>(part, cnt, s)) {
        out[0] = s[0];
        out[1] = s[1];
        out[2] = s[2];
    }
}

__global__ void k_add(int64_t n, const double *__restrict__ d, double *__restrict__ x) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        x[i] += d[i];
}

inline unsigned nblocks(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

template <int G, typename VT>
void launch_spmv_p(fx_mesh *m, const VT *val, hipStream_t s) {
    unsigned nb = nblocks(m->n_nodes, kBlock / G);
    hipLaunchKernelGGL((k_spmv_p<G, VT>), dim3(nb), dim3(kBlock), 0, s, m->n_nodes, m->nb_ptr.p, m->nb_col.p, val,
                       m->z.p, m->p0.p, m->p1.p, m->ap.p, m->pcg.p, m->part.p);
}

template <typename VT>
void launch_iteration(fx_mesh *m, const VT *val, int G, hipStream_t s) {
    switch (G) {
    case 2: launch_spmv_p<2, VT>(m, val, s); break;
    case 4: launch_spmv_p<4, VT>(m, val, s); break;
    case 8: launch_spmv_p<8, VT>(m, val, s); break;
    case 16: launch_spmv_p<16, VT>(m, val, s); break;
    case 32: launch_spmv_p<32, VT>(m, val, s); break;
    default: launch_spmv_p<64, VT>(m, val, s); break;
    }
    hipLaunchKernelGGL(k_update, dim3(nblocks(m->n_dof, kBlock)), dim3(kBlock), 0, s, m->n_dof, m->x.p, m->r.p,
                       m->z.p, m->p0.p, m->p1.p, m->ap.p, m->inv_diag.p, m->pcg.p, m->part.p);
}

template <int G>
void launch_spmv_g(fx_mesh *m, const double *x, double *y, hipStream_t s) {
    hipLaunchKernelGGL((k_spmv<G>), dim3(nblocks(m->n_nodes, kBlock / G)), dim3(kBlock), 0, s, m->n_nodes,
                       m->nb_ptr.p, m->nb_col.p, m->kval.p, x, y);
}

// Average node-row degree -> lanes per node row.
int default_group(fx_mesh *m) {
    double deg = (double)m->nblk / (double)std::max<int64_t>(1, m->n_nodes);
    if (deg <= 6)
        return 2;
    if (deg <= 14)
        return 4;
    if (deg <= 28)
        return 8;
    if (deg <= 56)
        return 16;
    return 32;
}

struct PcgGraph {
    hipGraphExec_t exec = nullptr;
    int G = 0, chunk = 0;
    const void *val = nullptr;
};

}  // namespace

void spmv64(fx_mesh *m, const double *x, double *y, int G) {
    hipStream_t s = m->ctx->stream;
    switch (G) {
    case 2: launch_spmv_g<2>(m, x, y, s); break;
    case 4: launch_spmv_g<4>(m, x, y, s); break;
    case 8: launch_spmv_g<8>(m, x, y, s); break;
    case 16: launch_spmv_g<16>(m, x, y, s); break;
    case 32: launch_spmv_g<32>(m, x, y, s); break;
    default: launch_spmv_g<64>(m, x, y, s); break;
    }
    FX_HIP(hipGetLastError());
}

// Host loop of one PCG solve on the resident operator `val` (FP64 or FP32
// values) with rhs b_dev; result in m->x.  Returns the device status.
template <typename VT>
static int run_pcg(fx_mesh *m, const VT *val, const double *b_dev, double tol, int64_t max_iter, int stop_mode,
                   int G, PcgScalars *out) {
    hipStream_t s = m->ctx->stream;
    PcgScalars init{};
    init.tol = tol;
    init.max_iter = max_iter;
    init.stop_mode = stop_mode;
    FX_HIP(hipMemcpyAsync(m->pcg.p, &init, sizeof init, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_init, dim3(nblocks(m->n_dof, kBlock)), dim3(kBlock), 0, s, m->n_dof, b_dev, m->x.p, m->r.p,
                       m->z.p, m->p0.p, m->inv_diag.p, m->pcg.p, m->part.p);
    FX_HIP(hipGetLastError());
    // batches of iterations; the batch grows so that long solves poll rarely
    int chunk = 8;
    PcgScalars h{};
    for (;;) {
        for (int k = 0; k < chunk; ++k)
            launch_iteration<VT>(m, val, G, s);
        FX_HIP(hipGetLastError());
        FX_HIP(hipMemcpyAsync(&h, m->pcg.p, sizeof h, hipMemcpyDeviceToHost, s));
        FX_HIP(hipStreamSynchronize(s));
        if (h.status != 0)
            break;
        if (chunk < 64)
            chunk *= 2;
    }
    *out = h;
    return h.status;
}

fx_status pcg_solve(fx_mesh *m, const double *b_dev, double *x_dev, const fx_solve_opts &o, fx_solve_report *rep) {
    hipStream_t s = m->ctx->stream;
    auto t0 = std::chrono::steady_clock::now();
    const int64_t n = m->n_dof;
    int G = o.spmv_group > 0 ? o.spmv_group : (m->spmv_group > 0 ? m->spmv_group : default_group(m));
    for (auto *b : {&m->x, &m->r, &m->z, &m->p0, &m->p1, &m->ap, &m->tmp})
        b->alloc(n);
    const bool mixed = o.precision == FX_PREC_MIXED;
    launch_jacobi(m, false);
    fx_solve_report R{};
    R.spmv_group = G;
    fx_status st = FX_OK;
    if (!mixed) {
        PcgScalars h;
        int code = run_pcg<double>(m, m->kval.p, b_dev, o.tol, o.max_iter, o.stop_mode, G, &h);
        R.iterations = h.iter;
        R.final_criterion = h.crit;
        if (code == 2)
            st = FX_ERR_BREAKDOWN;
        else if (code == 3)
            st = FX_ERR_NOT_CONVERGED;
        if (x_dev != m->x.p)
            FX_HIP(hipMemcpyAsync(x_dev, m->x.p, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    } else {
        // FP64 iterative refinement around an FP32-stored operator:
        //   x = 0, r = b;  repeat: solve K32 d = r to inner_tol (PCG with FP64
        //   vectors), x += d, r = b - K64 x (FP64), stop on the outer rule.
        launch_convert32(m);
        DBuf<double> xacc, rtrue, kx, out3;
        xacc.alloc(n);
        rtrue.alloc(n);
        kx.alloc(n);
        out3.alloc(3);
        FX_HIP(hipMemsetAsync(xacc.p, 0, n * sizeof(double), s));
        FX_HIP(hipMemcpyAsync(rtrue.p, b_dev, n * sizeof(double), hipMemcpyDeviceToDevice, s));
        double bb = -1.0;
        int64_t budget = o.max_iter;
        int passes = 0;
        for (;;) {
            // outer criterion on the true residual
            spmv64(m, xacc.p, kx.p, G);
            hipLaunchKernelGGL(k_true_residual, dim3(nblocks(n, kBlock)), dim3(kBlock), 0, s, n, b_dev, kx.p,
                               m->inv_diag.p, rtrue.p, m->part.p, m->counters.p, out3.p);
            FX_HIP(hipGetLastError());
            double h3[3];
            FX_HIP(hipMemcpyAsync(h3, out3.p, sizeof h3, hipMemcpyDeviceToHost, s));
            FX_HIP(hipStreamSynchronize(s));
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
            R.final_criterion = crit;
            if (conv)
                break;
            if (passes >= o.max_refine || budget <= 0 || !std::isfinite(crit)) {
                st = FX_ERR_NOT_CONVERGED;
                break;
            }
            // inner solve to a relative reduction inner_tol of this residual
            PcgScalars h;
            int code = run_pcg<float>(m, m->kval32.p, rtrue.p, o.inner_tol, budget, FX_STOP_REL_R2, G, &h);
            R.iterations += h.iter;
            budget -= h.iter;
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
        FX_HIP(hipStreamSynchronize(s));
    }
    FX_HIP(hipStreamSynchronize(s));
    R.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rep)
        *rep = R;
    return st;
}

}  // namespace fx
