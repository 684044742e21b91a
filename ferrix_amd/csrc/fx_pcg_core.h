// fx_pcg_core.h -- the PCG pieces shared by the mesh solve (fx_pcg.hip) and
// the caller-CSR solve (fx_csr.hip): XCD-contiguous workgroup numbering, the
// fixed-order reduction of block partials, the start of a solve, the head of
// every SpMV launch (stop rule, beta) and the x/r/z update.
//
// Reference: src/solver/solvers/iterative.rs:26-121 (recurrences, stop rule
// at the top of each iteration, error after max_iter) and
// src/solver/preconditioner.rs:52-58 (z = r * inv_diag).
#pragma once

#include "fx_reduce.h"

namespace fx {
namespace {
constexpr int kBlock = 256;
constexpr int kUnroll = 2;  // column slots per lane with loads in flight

__device__ __forceinline__ unsigned xcd_swizzle(unsigned id, unsigned nb) {
    const unsigned q = nb / 8, r = nb % 8, x = id % 8, i = id / 8;
    return x * q + (x < r ? x : r) + i;
}

// Sum of part[NV*b + k] over b < nparts in a fixed order (each thread its
// rows b = tid, tid + kBlock, ... in turn, then the block sum); the result
// is returned to every thread of the block.  A thread loads four of its
// rows before adding them: a row past the end reloads the last one, and
// its add is not taken.
template <int NV>
__device__ __forceinline__ void reduce_partials(const double *__restrict__ part, int nparts, double (&out)[NV]) {
    __shared__ double bcast[NV];
    constexpr int U = 4;
#pragma unroll
    for (int k = 0; k < NV; ++k)
        out[k] = 0.0;
    for (int b0 = threadIdx.x; b0 < nparts; b0 += U * kBlock) {
        double t[U][NV];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int b = b0 + u * kBlock < nparts ? b0 + u * kBlock : nparts - 1;
#pragma unroll
            for (int k = 0; k < NV; ++k)
                t[u][k] = part[NV * b + k];
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int k = 0; k < NV; ++k)
                out[k] = b0 + u * kBlock < nparts ? out[k] + t[u][k] : out[k];
    }
    block_sum<NV>(out);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k)
            bcast[k] = out[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k)
        out[k] = bcast[k];
}

// x = 0, r = b, z = r inv, p_{-1} = z; partials (r.z, r.r, b.b) -> part2.
__global__ __launch_bounds__(kBlock) void k_init(int64_t n, const double *__restrict__ b, double *__restrict__ x,
                                                 double *__restrict__ r, double *__restrict__ z,
                                                 double *__restrict__ p0, const double *__restrict__ inv,
                                                 double *__restrict__ part2) {
    double v[3] = {0.0, 0.0, 0.0};
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        const double bi = b[i];
        const double zi = bi * inv[i];
        x[i] = 0.0;
        r[i] = bi;
        z[i] = zi;
        p0[i] = zi;
        v[0] = bi * zi;
        v[1] = bi * bi;
        v[2] = bi * bi;
    }
    block_sum<3>(v);
    if (threadIdx.x == 0) {
        part2[3 * blockIdx.x] = v[0];
        part2[3 * blockIdx.x + 1] = v[1];
        part2[3 * blockIdx.x + 2] = v[2];
    }
}

// Head of an SpMV launch: the block-uniform PCG status, r.z and r.r (and
// b.b at iteration 0) from the previous update's partials, the stop rule
// (block 0 records it), beta and gamma (p_new = gamma z + beta p_old).
// Right after a residual replacement (sc->resume; the third partial is then
// r.p of the kept direction) the direction is kept only when the replaced
// residual is within 10% of the iterated one the segment ended on, with
// gamma = 1 - beta (r.p)/(r.z) restoring r.p = 0; otherwise beta = 0.
// Returns false when the solve has stopped.
__device__ __forceinline__ bool pcg_head(PcgScalars *__restrict__ sc, const double *__restrict__ part2, int nparts2,
                                         int64_t &it_out, double &beta_out, double &gamma_out) {
    __shared__ int32_t s_status;
    if (threadIdx.x == 0)
        s_status = sc->status;
    __syncthreads();  // one block-wide view (block 0 may update it below)
    if (s_status != 0)
        return false;
    const int64_t it = sc->it;
    // r.z, r.r (and b.b at i = 0) after the previous update
    double red[3];
    reduce_partials<3>(part2, nparts2, red);
    const double rz = red[0], rr = red[1];
    const double bb = it == 0 ? red[2] : sc->bb;
    double crit;
    bool conv;
    if (sc->stop_mode == FX_STOP_ABS_SQRT_RZ) {
        crit = sqrt(rz);
        conv = crit < sc->tol;
    } else {
        crit = bb > 0.0 ? sqrt(rr) / sqrt(bb) : 0.0;
        conv = crit <= sc->tol;
    }
    const bool stop_max = it >= sc->max_iter;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        sc->rz[it & 1] = rz;
        sc->rr = rr;
        sc->crit = crit;
        sc->it_k1 = it;
        if (it == 0)
            sc->bb = bb;
        if (stop_max)
            sc->status = 3;
        else if (conv)
            sc->status = 1;
    }
    if (stop_max || conv)
        return false;
    double beta = it == 0 ? 0.0 : rz / sc->rz[(it - 1) & 1];
    double gamma = 1.0;
    if (sc->resume) {
        const bool keep = it > 0 && rr <= 1.21 * sc->rr_prev;
        beta = keep ? beta : 0.0;
        gamma = keep ? 1.0 - beta * (rz > 0.0 ? red[2] / rz : 0.0) : 1.0;
    }
    it_out = it;
    beta_out = beta;
    gamma_out = gamma;
    return true;
}

__global__ __launch_bounds__(kBlock) void k_update(int64_t n, double *__restrict__ x, double *__restrict__ r,
                                                   double *__restrict__ z, const double *__restrict__ pa,
                                                   const double *__restrict__ pb, const double *__restrict__ ap,
                                                   const double *__restrict__ inv, PcgScalars *__restrict__ sc,
                                                   const double *__restrict__ part1, int nparts1,
                                                   double *__restrict__ part2) {
    __shared__ int32_t s_status;
    if (threadIdx.x == 0)
        s_status = sc->status;
    __syncthreads();
    if (s_status != 0)
        return;
    const int64_t it = sc->it_k1;
    double red[1];
    reduce_partials<1>(part1, nparts1, red);
    const double pap = red[0];
    const double alpha = sc->rz[it & 1] / pap;
    if (!(pap > 0.0) || !isfinite(alpha)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            sc->status = 2;
            sc->pap = pap;
        }
        return;
    }
    const double *p = (it & 1) ? pa : pb;  // p_i is in the "new" buffer of iteration i
    double v[3] = {0.0, 0.0, 0.0};
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
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
    block_sum<3>(v);
    if (threadIdx.x == 0) {
        part2[3 * blockIdx.x] = v[0];
        part2[3 * blockIdx.x + 1] = v[1];
        part2[3 * blockIdx.x + 2] = 0.0;
        if (blockIdx.x == 0) {
            sc->it = it + 1;
            sc->alpha = alpha;
            sc->pap = pap;
            sc->resume = 0;  // read by the SpMV head that ran before this kernel
        }
    }
}

inline unsigned nblocks(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

// Iterations are issued in chunks and the status block `sc_dev` is read
// after each chunk (one host round trip per chunk).  The next chunk is the
// largest power of two (at most 64) not above the iterations still
// predicted from the observed convergence rate, so few launches run past
// convergence (they return at once).  launch(chunk) issues that many
// iterations on stream s.  Returns the final status (1 converged,
// 2 breakdown, 3 max_iter) with the scalars in *out.
template <class LaunchChunk>
int iterate_chunks(hipStream_t s, PcgScalars *sc_dev, double tol, int64_t max_iter, LaunchChunk &&launch,
                   PcgScalars *out, int first_chunk = 8) {
    PcgScalars h{};
    int chunk = first_chunk;
    double crit_prev = -1.0;
    int64_t it_prev = 0;
    for (;;) {
        launch(chunk);
        FX_HIP(hipMemcpyAsync(&h, sc_dev, sizeof h, hipMemcpyDeviceToHost, s));
        FX_HIP(hipStreamSynchronize(s));
        ++t_host_syncs;
        if (h.status != 0)
            break;
        // predicted iterations left from the mean log-reduction per iteration
        double left = 64.0;
        const int64_t did = h.it_k1 - it_prev;
        if (crit_prev > 0.0 && did > 0 && h.crit > 0.0 && h.crit < crit_prev) {
            const double rate = (std::log(crit_prev) - std::log(h.crit)) / (double)did;
            left = (std::log(h.crit) - std::log(tol > 0.0 ? tol : 1e-300)) / rate;
        }
        left = std::min<double>(left, (double)(max_iter - h.it + 1));
        chunk = 1;
        while (chunk * 2 <= left && chunk < 64)
            chunk *= 2;
        crit_prev = h.crit;
        it_prev = h.it_k1;
    }
    *out = h;
    return h.status;
}

// SpMVs done when the solve stopped: the stop rule fires in the SpMV head
// of iteration it_k1 (it_k1 SpMVs done); a breakdown fires in the update
// after that iteration's SpMV.
inline int64_t iterations_done(const PcgScalars &h) { return h.status == 2 ? h.it_k1 + 1 : h.it_k1; }

}  // namespace
}  // namespace fx
