// fx_batch.hip -- many independent systems (meshes x load cases) advanced by
// one set of kernel launches: the engine of the batched workload and of the
// serving loop.
//
// A system ("slot") is one resident mesh's K (after assembly and penalty)
// with R <= 4 right-hand sides, the load cases that share that K.  Vectors of
// a slot are interleaved by right-hand side (v[dof*R + r]), so one pass over
// K serves all its load cases.  Every workgroup of a launch works for exactly
// one slot (the block tables hold (slot, local block) pairs) and block
// partial sums are kept per slot, so a slot's arithmetic -- and its result --
// is the same whichever other slots share the launch.  Per right-hand side the
// recurrences and the stop rule are those of fx_pcg.hip (the reference's
// Jacobi-PCG, src/solver/solvers/iterative.rs:26-121).
//
// Mixed precision: PCG on the FP32-stored K (with its row-sum correction
// blocks) and FP32 Krylov vectors, with reliable updates of the residual
// (Sleijpen & van der Vorst, "Reliable updated residuals in hybrid Bi-CG
// methods", Computing 56, 1996; the mixed-precision CG form of Clark et al.,
// "Solving lattice QCD systems of equations using mixed precision solvers on
// GPUs", Comput. Phys. Commun. 181, 2010).  When the iterated residual of a
// slot's load case has fallen below inner_tol times its value at the last
// replacement (or meets the stop rule), the slot spends one iteration on a
// replacement instead of a Krylov step:
//   SpMV launch:  r64 = b - K64 (xacc + x), in FP64, with the true stop
//                 criterion's partial sums;
//   kb_red1:      stop rule on r64 (and the rounding-floor test);
//   kb_update:    xacc += x, x = 0, r = r64, z = r inv;
//   kb_red2:      beta = (r.z)/(r.z)_old and the correction that makes the
//                 kept direction p orthogonal to the new r again,
//                 p <- p - ((r.p)/(r.z)) z, folded into the next p update as
//                 p_new = gamma z + beta p_old.
// The Krylov direction survives the replacement when the new residual is
// within 10% of the iterated one, so a well-conditioned system keeps its
// convergence rate; otherwise the solve restarts from p = z, as a restarted
// refinement would.  Slots run independently,
// and new slots join between chunks of iterations (continuous batching).
#include "fx_batch.h"

#include <chrono>
#include <cstdlib>
#include <cstring>
#include <memory>

namespace fx {

namespace {

// Per-iteration kernel sequence of a batch: kb_pupdate (p = gamma z +
// beta p_old, streamed, so the SpMV gathers one vector instead of two),
// kb_spmv_*, kb_red1, kb_update, kb_red2.  The vector kernels write block
// partial sums; one block per
// system (kb_red1 / kb_red2) sums them in a fixed order and takes the
// scalar decisions (alpha, beta, stop rule), so every block of the next
// vector kernel reads a handful of scalars instead of re-summing all
// partials.  The sums have exactly the association of fx_pcg.hip's
// reduce_partials (thread-strided, then block_sum), so FP64 results are
// bitwise those of fx_solve.

// Solve start for the right-hand sides with status 0:
// x = 0, r = b, z = r inv, p0 = z (stored as XT); partials (r.z, r.r of the
// stored values, b.b of the FP64 right-hand side) -> part2.
template <typename XT>
__global__ __launch_bounds__(kBB) void kb_init(BSys *__restrict__ sys, const int2 *__restrict__ tab) {
    const int2 t = tab[blockIdx.x];
    BSys &S = sys[t.x];
    const int lb = t.y, R = S.R;
    double v[3 * kRmax];
#pragma unroll
    for (int k = 0; k < 3 * kRmax; ++k)
        v[k] = 0.0;
    for (int tile = 0; tile < kTiles<XT>; ++tile) {
        const int64_t i = ((int64_t)lb * kTiles<XT> + tile) * kBB + threadIdx.x;
        if (i >= S.n_dof)
            break;
#pragma unroll
        for (int r = 0; r < kRmax; ++r) {
            if (r >= R || S.rhs[r].status != 0)  // set by the host before this launch
                continue;
            const int64_t q = i * R + r;
            const double bt = S.btrue[q];
            const double bi = (double)(XT)bt;
            const double zi = (double)(XT)(bi * S.inv[i]);
            ((XT *)S.x)[q] = (XT)0;
            ((XT *)S.r)[q] = (XT)bi;
            ((XT *)S.z)[q] = (XT)zi;
            ((XT *)S.p0)[q] = (XT)zi;
            v[3 * r] += bi * zi;
            v[3 * r + 1] += bi * bi;
            v[3 * r + 2] += bt * bt;
        }
    }
    block_sum<3 * kRmax>(v);
    if (threadIdx.x == 0)
        for (int r = 0; r < R; ++r)
            for (int k = 0; k < 3; ++k)
                S.part2[((size_t)lb * R + r) * 3 + k] = v[3 * r + k];
}

// One block per listed system, after kb_init or kb_update: r.z, r.r (and
// b.b at iteration 0) per right-hand side, the stop rule of fx_pcg.hip's
// k_spmv_p, beta, and the mask of right-hand sides the next SpMV works on.
// Mixed slots: the stop rule on the iterated residual, or its fall below
// delta * rmax, starts a residual replacement of all the slot's running
// right-hand sides (status 5, S.upd) instead of a Krylov step.  After a
// replacement (S.upd set, partials r.z, r.r, r.p of the new residual):
// beta and gamma of the next direction.
__global__ __launch_bounds__(kBB) void kb_red2(BSys *__restrict__ sys, const int32_t *__restrict__ list,
                                               BFin *__restrict__ fin) {
    const int slot = list[blockIdx.x];
    BSys &S = sys[slot];
    __shared__ BRhs H[kRmax];
    stage_rhs(S, H);
    const int R = S.R;
    const int upd = S.upd;
    bool any = false;
    for (int r = 0; r < R; ++r)
        any |= H[r].status == 0;
    if (!any) {  // uniform: only this block's thread 0 writes these states, in earlier launches
        if (threadIdx.x == 0) {
            if (upd)
                S.upd = 0;  // every replaced right-hand side stopped
            if (!S.reported) {  // the slot has just finished: tell the host
                BFin &F = fin[slot];
                int64_t it = 0, nupd = 0;
                for (int r = 0; r < kRmax; ++r) {
                    const BRhs &h = H[r];
                    F.crit[r] = h.crit;
                    F.status[r] = h.status;
                    if (r < R) {
                        it = max(it, h.status == 2 ? h.it_k1 + 1 : h.it_k1);
                        nupd = max(nupd, (int64_t)h.nupd);
                    }
                }
                F.it = it;
                F.nupd = nupd;
                __threadfence_system();
                F.done = 1;
                S.reported = 1;
            }
        }
        return;
    }
    double red[3 * kRmax];
    bsum_partials<3 * kRmax>(S.part2, S.nparts2, 3 * R, red);
    if (threadIdx.x != 0)
        return;
    int go = 0;
    // loops over kRmax with a guard (not to R) keep red[] in registers
    if (upd) {
#pragma unroll
        for (int r = 0; r < kRmax; ++r) {
            const BRhs &c = H[r];
            BRhs &h = S.rhs[r];
            if (r >= R || c.status != 0)
                continue;
            const double rz = red[3 * r], rr = red[3 * r + 1], rp = red[3 * r + 2];
            const int64_t it = c.it;
            // keep the direction only while the replaced residual is within
            // 10% of the iterated one (c.rr): a larger gap -- the FP32
            // operator's error on this x, of order eps32 * cond(K) -- means
            // the iterated Krylov space belongs to the FP32 system, and
            // continuing it can stall the solve for thousands of iterations;
            // restart from p = z instead
            const double beta = it == 0 || rr > 1.21 * c.rr ? 0.0 : rz / c.rz[(it - 1) & 1];
            h.beta = beta;
            h.gamma = 1.0 - beta * (rz > 0.0 ? rp / rz : 0.0);
            h.rz[it & 1] = rz;
            h.rr = rr;
            h.it_k1 = it;
            go |= 1 << r;
        }
        S.upd = 0;
        S.go = go;
        if (go)
            atomicAdd(S.bytes_acc, S.spmv_bytes0 + (double)__popc(go) * S.spmv_bytes_rhs);
        return;
    }
    int trig = 0;
    int64_t itv = S.it;
#pragma unroll
    for (int r = 0; r < kRmax; ++r) {
        const BRhs &c = H[r];
        BRhs &h = S.rhs[r];
        if (r >= R || c.status != 0)
            continue;
        const double rz = red[3 * r], rr = red[3 * r + 1];
        const int64_t it = c.it;
        const double bb = it == 0 ? red[3 * r + 2] : c.bb;
        double crit;
        bool conv;
        if (c.stop_mode == FX_STOP_ABS_SQRT_RZ) {
            crit = sqrt(rz);
            conv = crit < c.tol;
        } else {
            crit = bb > 0.0 ? sqrt(rr) / sqrt(bb) : 0.0;
            conv = crit <= c.tol;
        }
        const bool stop_max = it >= c.max_iter;
        const double rmax = it == 0 ? sqrt(bb) : c.rmax;
        h.beta = it == 0 ? 0.0 : rz / c.rz[(it - 1) & 1];
        h.gamma = 1.0;  // a plain CG step (gamma != 1 only right after a replacement)
        h.rz[it & 1] = rz;
        h.rr = rr;
        h.crit = crit;
        h.it_k1 = it;
        if (it == 0) {
            h.bb = bb;
            h.rmax = rmax;
        }
        if (stop_max)
            h.status = 3;
        else if (c.delta > 0.0 && (conv || sqrt(rr) < c.delta * rmax))
            trig |= 1 << r;
        else if (conv)
            h.status = 1;
        else
            go |= 1 << r;
        itv = it;  // the running right-hand sides of a system share it
    }
    S.it = itv;
    if (trig) {  // replace the residual of every running right-hand side of the slot
        const int all = trig | go;
        for (int r = 0; r < R; ++r)
            if ((all >> r) & 1)
                S.rhs[r].status = 5;
        S.upd = all;
        S.trig = trig;
        S.go = 0;
        atomicAdd(S.bytes_acc, S.upd_bytes0 + (double)__popc(all) * S.upd_bytes_rhs);
        return;
    }
    S.go = go;
    if (go)  // the next SpMV launch works on this system (integer-valued: exact in any order)
        atomicAdd(S.bytes_acc, S.spmv_bytes0 + (double)__popc(go) * S.spmv_bytes_rhs);
}

// Tail of a batched SpMV row group: the FP32 operator's correction block
// on the node's own p, Ap, and p.Ap block partials (right-hand sides
// outside `go` are computed but not stored).
template <typename VT, typename XT, int RR>
__device__ __forceinline__ void bspmv_tail(const BSys &S, int64_t a, int go, gptr<const XT> pv,
                                           double (&acc)[3][RR], double (&loc)[kRmax]) {
    gptr<XT> ap = gp((XT *)S.ap);
    double qv[3][RR];
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int r = 0; r < RR; ++r)
            qv[p][r] = (double)pv[(3 * a + p) * RR + r];
    if constexpr (sizeof(VT) == 4) {
        gptr<const float> c = gp(S.kc) + 9 * a;
        float cc[9];
#pragma unroll
        for (int e = 0; e < 9; ++e)
            cc[e] = c[e];
#pragma unroll
        for (int r = 0; r < RR; ++r)
#pragma unroll
            for (int p = 0; p < 3; ++p)
                acc[p][r] += ((double)cc[3 * p] * qv[0][r] + (double)cc[3 * p + 1] * qv[1][r]) +
                             (double)cc[3 * p + 2] * qv[2][r];
    }
#pragma unroll
    for (int r = 0; r < RR; ++r) {
        if (!((go >> r) & 1))
            continue;
        double sum = 0.0;
#pragma unroll
        for (int p = 0; p < 3; ++p) {
            const int64_t q = (3 * a + p) * RR + r;
            const double aps = (double)(XT)acc[p][r];  // Ap as stored
            ap[q] = (XT)aps;
            sum += qv[p][r] * aps;
        }
        loc[r] = sum;
    }
}

template <int RR>
__device__ __forceinline__ void load_coefs(const BSys &S, int go, double (&gam)[RR], double (&beta)[RR]) {
#pragma unroll
    for (int r = 0; r < RR; ++r) {
        const bool on = (go >> r) & 1;
        beta[r] = on ? S.rhs[r].beta : 0.0;
        gam[r] = on ? S.rhs[r].gamma : 1.0;
    }
}

template <int G, typename VT, typename XT, int RR>
__device__ __forceinline__ void bspmv_lanes(const BSys &S, int lb, int go, double (&loc)[kRmax]) {
    // p of this iteration, written by kb_pupdate into the "new" buffer
    gptr<const XT> pv = gp((const XT *)((S.it & 1) ? S.p0 : S.p1));
    gptr<const VT> val = gp((const VT *)S.val);
    gptr<const int32_t> nb_ptr = gp(S.nb_ptr), nb_col = gp(S.nb_col);
    const int64_t a = (int64_t)lb * (kBB / G) + threadIdx.x / G;
    const int lane = threadIdx.x % G;
    double acc[3][RR];
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int r = 0; r < RR; ++r)
            acc[p][r] = 0.0;
    if (a < S.n_nodes) {
        const int b0 = nb_ptr[a], ncol = 3 * (nb_ptr[a + 1] - b0);
        gptr<const VT> v0 = val + 9 * (int64_t)b0;
        // two slots per lane in flight; accumulation order j, j+G, j+2G, ...
        // as in a plain strided loop
        for (int j0 = lane; j0 < ncol; j0 += 2 * G) {
            int c[2];
            double w[2][3];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int j = j0 + u * G;
                const int kb = j / 3;
                c[u] = j < ncol ? (3 * nb_col[b0 + kb] + (j - 3 * kb)) * RR : -1;
            }
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int j = j0 + u * G;
                if (c[u] >= 0) {
                    w[u][0] = (double)v0[j];
                    w[u][1] = (double)v0[ncol + j];
                    w[u][2] = (double)v0[2 * ncol + j];
                }
            }
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                if (c[u] < 0)
                    continue;
                double xj[RR];
#pragma unroll
                for (int r = 0; r < RR; ++r)
                    xj[r] = (double)pv[c[u] + r];
#pragma unroll
                for (int r = 0; r < RR; ++r) {
                    acc[0][r] += w[u][0] * xj[r];
                    acc[1][r] += w[u][1] * xj[r];
                    acc[2][r] += w[u][2] * xj[r];
                }
            }
        }
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1)
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int r = 0; r < RR; ++r)
                acc[p][r] += __shfl_xor(acc[p][r], off, G);
    if (lane == 0 && a < S.n_nodes)
        bspmv_tail<VT, XT, RR>(S, a, go, pv, acc, loc);
}

template <typename VT, typename XT, int RR>
__device__ __forceinline__ void bspmv_sell(const BSys &S, int lb, int go, double (&loc)[kRmax]) {
    gptr<const XT> pv = gp((const XT *)((S.it & 1) ? S.p0 : S.p1));  // written by kb_pupdate
    const int64_t a = (int64_t)lb * kBB + threadIdx.x;
    double acc[3][RR];
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int r = 0; r < RR; ++r)
            acc[p][r] = 0.0;
    if (a < S.n_nodes) {
        if (sizeof(VT) == 4 && S.ssym) {
            // lower blocks from their owners' records, then the row's own upper blocks
            ssym_lower_rows<XT, RR>(gp(S.sll_ptr), gp(S.lrec), gp(S.lcol), gp(S.uval), a, pv, acc);
            sell_node_rows<float, XT, RR>(gp(S.su_ptr), gp(S.scol_u), gp(S.uval), a, pv, acc);
        } else {
            sell_node_rows<VT, XT, RR>(gp(S.sl_ptr), gp(S.scol), gp((const VT *)S.sval), a, pv, acc);
        }
        bspmv_tail<VT, XT, RR>(S, a, go, pv, acc, loc);
    }
}

// Residual replacement, SpMV part, for node row a (G lanes): r64 = b -
// K64 (xacc + x) at the right-hand sides in `upd` -> S.b, and the partial
// sums r64.r64, r64.(inv r64) and (|K||y|).(|K||y|), y = xacc + x, -> v
// (lane 0).  xacc + x is the FP64 solution the update kernel stores next.
// One right-hand side at a time: the row's FP64 values are read once per
// replaced load case (from cache after the first), and only 6 running sums
// are live, so the pass does not raise the register count of the SpMV
// kernel it shares.  Every sum keeps its order: columns in lane-strided
// order, the lane reduction, then rows into v in p order.
template <int G>
__device__ __forceinline__ void bresid_row(const BSys &S, int64_t a, int lane, int upd, double (&v)[3 * kRmax]) {
    const int R = S.R;
    gptr<const double> xacc = gp((const double *)S.xacc), val = gp(S.val64);
    gptr<const float> x = gp((const float *)S.x);
    gptr<const int32_t> nb_ptr = gp(S.nb_ptr), nb_col = gp(S.nb_col);
    gptr<const double> bt = gp(S.btrue), inv = gp(S.inv);
    gptr<double> rb = gp(S.b);
    const bool own = a < S.n_nodes;
    const int b0 = own ? nb_ptr[a] : 0, ncol = own ? 3 * (nb_ptr[a + 1] - b0) : 0;
    gptr<const double> v0 = val + 9 * (int64_t)b0;
#pragma unroll 1
    for (int r = 0; r < R; ++r) {
        if (!((upd >> r) & 1))
            continue;
        double acc[3] = {0.0, 0.0, 0.0}, mag[3] = {0.0, 0.0, 0.0};
        for (int j = lane; j < ncol; j += G) {
            const int kb = j / 3;
            const int64_t c = (3 * (int64_t)nb_col[b0 + kb] + (j - 3 * kb)) * R + r;
            const double w0 = v0[j], w1 = v0[ncol + j], w2 = v0[2 * ncol + j];
            const double yj = xacc[c] + (double)x[c];
            acc[0] += w0 * yj;
            acc[1] += w1 * yj;
            acc[2] += w2 * yj;
            mag[0] += fabs(w0 * yj);
            mag[1] += fabs(w1 * yj);
            mag[2] += fabs(w2 * yj);
        }
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1)
#pragma unroll
            for (int p = 0; p < 3; ++p) {
                acc[p] += __shfl_xor(acc[p], off, G);
                mag[p] += __shfl_xor(mag[p], off, G);
            }
        if (lane != 0 || !own)
            continue;
#pragma unroll
        for (int p = 0; p < 3; ++p) {
            const int64_t d = 3 * a + p, q = d * R + r;
            const double ri = bt[q] - acc[p];
            rb[q] = ri;
#pragma unroll
            for (int k = 0; k < kRmax; ++k)  // v[3 r + .] with r in a register
                if (k == r) {
                    v[3 * k] += ri * ri;
                    v[3 * k + 1] += ri * (ri * inv[d]);
                    v[3 * k + 2] += mag[p] * mag[p];
                }
        }
    }
}

// one block of a replacement pass: `sub` groups of kBB / G node rows from a0;
// block partials -> part1 (3 per right-hand side)
template <int G>
__device__ __forceinline__ void bresid_block(const BSys &S, int lb, int64_t a0, int sub) {
    double v[3 * kRmax];
#pragma unroll
    for (int k = 0; k < 3 * kRmax; ++k)
        v[k] = 0.0;
    const int upd = S.upd;
    for (int k = 0; k < sub; ++k)
        bresid_row<G>(S, a0 + (int64_t)k * (kBB / G) + threadIdx.x / G, threadIdx.x % G, upd, v);
    block_sum<3 * kRmax>(v);
    if (threadIdx.x == 0)
        for (int r = 0; r < S.R; ++r)
            for (int k = 0; k < 3; ++k)
                S.part1[((size_t)lb * S.R + r) * 3 + k] = v[3 * r + k];
}

// p_new = gamma z + beta p_old (as stored) for the running right-hand
// sides, into the buffer the next SpMV gathers: the "new" buffer of
// iteration S.it.  Own DOFs only; kTiles<XT> DOF tiles per block.  gamma is
// 1 except right after a residual replacement (the re-orthogonalisation of
// the kept direction); the FP64 solve never replaces its residual and
// forms z + beta p_old exactly as fx_pcg.hip does.
// Every tile's loads are issued before the first store: a lane past the
// system's last DOF loads that DOF again (in bounds, value unused), so no
// load sits behind a per-lane branch, where the compiler would wait for it
// before the branch joins.
template <typename XT, int RR>
__device__ __forceinline__ void bpupdate(const BSys &S, int lb, int go) {
    double gam[RR], beta[RR];
    load_coefs<RR>(S, go, gam, beta);
    gptr<const XT> z = gp((const XT *)S.z);
    gptr<const XT> pold = gp((const XT *)((S.it & 1) ? S.p1 : S.p0));
    gptr<XT> pnew = gp((XT *)((S.it & 1) ? S.p0 : S.p1));
    constexpr int T = kTiles<XT>;
    const int64_t n = S.n_dof;  // >= 1: a system with no DOF has no blocks
    XT zz[T][RR], pp[T][RR];
#pragma unroll
    for (int tile = 0; tile < T; ++tile) {
        const int64_t i = ((int64_t)lb * T + tile) * kBB + threadIdx.x;
        const int64_t ic = i < n ? i : n - 1;
#pragma unroll
        for (int r = 0; r < RR; ++r) {
            zz[tile][r] = z[ic * RR + r];
            pp[tile][r] = pold[ic * RR + r];
        }
    }
#pragma unroll
    for (int tile = 0; tile < T; ++tile) {
        const int64_t i = ((int64_t)lb * T + tile) * kBB + threadIdx.x;
        if (i >= n)
            break;
#pragma unroll
        for (int r = 0; r < RR; ++r)
            if ((go >> r) & 1) {
                const double zv = (double)zz[tile][r], pv = (double)pp[tile][r];
                pnew[i * RR + r] = sizeof(XT) == 4 ? (XT)(gam[r] * zv + beta[r] * pv) : (XT)(zv + beta[r] * pv);
            }
    }
}

template <typename XT>
__global__ __launch_bounds__(kBB) void kb_pupdate(BSys *__restrict__ sys, const int2 *__restrict__ tab) {
    const int2 t = tab[blockIdx.x];
    const BSys &S = sys[t.x];
    const int go = S.go;
    if (!go)
        return;
    switch (S.R) {
    case 1: bpupdate<XT, 1>(S, t.y, go); break;
    case 2: bpupdate<XT, 2>(S, t.y, go); break;
    case 3: bpupdate<XT, 3>(S, t.y, go); break;
    default: bpupdate<XT, 4>(S, t.y, go); break;
    }
}

// Batched SpMV for every (system, row group) of the table: lane
// groups on the node-block layout (kb_spmv_p) or one thread per node row on
// the sliced-ELL operator (kb_spmv_sell); the right-hand side count of the
// block's system selects a straight-line instance.  The operand is the p
// that kb_pupdate stored: one gathered vector per load case.
template <int G, typename VT, typename XT>
__global__ __launch_bounds__(kBB) void kb_spmv_p(BSys *__restrict__ sys, const int2 *__restrict__ tab) {
    const int2 t = tab[blockIdx.x];
    const BSys &S = sys[t.x];
    if constexpr (sizeof(VT) == 4) {
        if (S.upd) {  // this system replaces its residuals in this iteration
            bresid_block<G>(S, t.y, (int64_t)t.y * (kBB / G), 1);
            return;
        }
    }
    const int go = S.go;
    if (!go)
        return;
    double loc[kRmax] = {0.0, 0.0, 0.0, 0.0};
    switch (S.R) {
    case 1: bspmv_lanes<G, VT, XT, 1>(S, t.y, go, loc); break;
    case 2: bspmv_lanes<G, VT, XT, 2>(S, t.y, go, loc); break;
    case 3: bspmv_lanes<G, VT, XT, 3>(S, t.y, go, loc); break;
    default: bspmv_lanes<G, VT, XT, 4>(S, t.y, go, loc); break;
    }
    block_sum<kRmax>(loc);
    if (threadIdx.x == 0)
        for (int r = 0; r < S.R; ++r)
            S.part1[(size_t)t.y * S.R + r] = loc[r];
}

template <typename VT, typename XT>
__global__ __launch_bounds__(kBB) void kb_spmv_sell(BSys *__restrict__ sys, const int2 *__restrict__ tab) {
    const int2 t = tab[blockIdx.x];
    const BSys &S = sys[t.x];
    if constexpr (sizeof(VT) == 4) {
        if (S.upd) {  // the block's kBB node rows, 8 lanes per row on the FP64 node-block values
            bresid_block<8>(S, t.y, (int64_t)t.y * kBB, 8);
            return;
        }
    }
    const int go = S.go;
    if (!go)
        return;
    double loc[kRmax] = {0.0, 0.0, 0.0, 0.0};
    switch (S.R) {
    case 1: bspmv_sell<VT, XT, 1>(S, t.y, go, loc); break;
    case 2: bspmv_sell<VT, XT, 2>(S, t.y, go, loc); break;
    case 3: bspmv_sell<VT, XT, 3>(S, t.y, go, loc); break;
    default: bspmv_sell<VT, XT, 4>(S, t.y, go, loc); break;
    }
    block_sum<kRmax>(loc);
    if (threadIdx.x == 0)
        for (int r = 0; r < S.R; ++r)
            S.part1[(size_t)t.y * S.R + r] = loc[r];
}

// One block per listed system, after kb_spmv_p: p.Ap, alpha, breakdown
// (k_update's rule in fx_pcg.hip) and the mask the update works on.  After
// a replacement pass: the stop rule on the FP64 residual (with the
// rounding-floor test) per replaced right-hand side; S.go = those that go on.
__global__ __launch_bounds__(kBB) void kb_red1(BSys *__restrict__ sys, const int32_t *__restrict__ list) {
    BSys &S = sys[list[blockIdx.x]];
    __shared__ BRhs H[kRmax];
    stage_rhs(S, H);
    const int upd = S.upd;
    if (upd) {
        const int R = S.R;
        double red[3 * kRmax];
        bsum_partials<3 * kRmax>(S.part1, S.nparts1, 3 * R, red);
        if (threadIdx.x != 0)
            return;
        int go2 = 0;
#pragma unroll
        for (int r = 0; r < kRmax; ++r) {
            if (r >= R || !((upd >> r) & 1))
                continue;
            const BRhs &c = H[r];
            BRhs &h = S.rhs[r];
            const double rr = red[3 * r], rzt = red[3 * r + 1], aa = red[3 * r + 2];
            double crit;
            bool conv;
            if (c.stop_mode == FX_STOP_ABS_SQRT_RZ) {
                crit = sqrt(rzt);
                conv = crit < c.tol;
            } else {
                crit = c.bb > 0.0 ? sqrt(rr) / sqrt(c.bb) : 0.0;
                conv = crit <= c.tol;
            }
            // the floor test compares the ends of two full cycles: it runs,
            // and prev moves, only for a right-hand side whose own residual
            // fell by delta (or met the stop rule), not for one replaced
            // early along with the others of its slot
            if ((S.trig >> r) & 1) {
                conv = conv || at_rounding_floor(crit, c.prev, sqrt(rr), sqrt(aa));
                h.prev = crit;
            }
            h.crit = crit;
            const int32_t nupd = c.nupd + 1;
            h.nupd = nupd;
            if (conv) {
                h.status = 1;
            } else if (!isfinite(crit) || nupd >= c.max_upd) {
                h.status = 3;
            } else {
                h.status = 0;
                h.rmax = sqrt(rr);
                go2 |= 1 << r;
            }
        }
        S.go = go2;
        return;
    }
    const int go = S.go;
    if (!go)
        return;
    const int R = S.R;
    double red[kRmax];
    bsum_partials<kRmax>(S.part1, S.nparts1, R, red);
    if (threadIdx.x != 0)
        return;
    int go2 = 0;
#pragma unroll
    for (int r = 0; r < kRmax; ++r) {
        if (r >= R || !((go >> r) & 1))
            continue;
        const BRhs &c = H[r];
        BRhs &h = S.rhs[r];
        const double pap = red[r];
        const double al = c.rz[c.it_k1 & 1] / pap;
        h.pap = pap;
        h.alpha = al;
        if (!(pap > 0.0) || !isfinite(al)) {
            h.status = 2;
        } else {
            h.it = c.it_k1 + 1;
            go2 |= 1 << r;
        }
    }
    S.go = go2;
}

template <typename XT, int RR>
__device__ __forceinline__ void bupdate(const BSys &S, int lb, int go, double (&v)[2 * kRmax]) {
    double alpha[RR];
#pragma unroll
    for (int r = 0; r < RR; ++r)
        alpha[r] = ((go >> r) & 1) ? S.rhs[r].alpha : 0.0;
    gptr<const XT> p = gp((const XT *)((S.it & 1) ? S.p0 : S.p1));  // p_i: the "new" buffer of iteration i
    gptr<XT> x = gp((XT *)S.x), rv = gp((XT *)S.r), z = gp((XT *)S.z);
    gptr<const XT> ap = gp((const XT *)S.ap);
    // every tile's loads before the first store, as in bpupdate (lanes past
    // the last DOF reload it and store nothing)
    constexpr int T = kTiles<XT>;
    const int64_t n = S.n_dof;
    XT xo[T][RR], ro[T][RR], po[T][RR], apo[T][RR];
    double inv[T];
#pragma unroll
    for (int tile = 0; tile < T; ++tile) {
        const int64_t i = ((int64_t)lb * T + tile) * kBB + threadIdx.x;
        const int64_t ic = i < n ? i : n - 1;
        inv[tile] = gp(S.inv)[ic];
#pragma unroll
        for (int r = 0; r < RR; ++r) {
            const int64_t q = ic * RR + r;
            xo[tile][r] = x[q];
            ro[tile][r] = rv[q];
            po[tile][r] = p[q];
            apo[tile][r] = ap[q];
        }
    }
#pragma unroll
    for (int tile = 0; tile < T; ++tile) {
        const int64_t i = ((int64_t)lb * T + tile) * kBB + threadIdx.x;
        if (i >= n)
            break;
#pragma unroll
        for (int r = 0; r < RR; ++r) {
            if (!((go >> r) & 1))
                continue;
            const int64_t q = i * RR + r;
            const double xi = (double)xo[tile][r] + alpha[r] * (double)po[tile][r];
            const double ri = (double)(XT)((double)ro[tile][r] - alpha[r] * (double)apo[tile][r]);  // values as stored
            const double zi = (double)(XT)(ri * inv[tile]);
            x[q] = (XT)xi;
            rv[q] = (XT)ri;
            z[q] = (XT)zi;
            v[2 * r] += ri * zi;
            v[2 * r + 1] += ri * ri;
        }
    }
}

// Residual replacement, update part (own DOFs only): xacc += x, x = 0 for
// the replaced right-hand sides; r = r64, z = r inv for those that go on,
// with the partials r.z, r.r, r.p (p: the kept direction).
__device__ __forceinline__ void breplace(const BSys &S, int lb, int upd, int go, double (&v)[3 * kRmax]) {
    const int R = S.R;
    gptr<float> x = gp((float *)S.x), rv = gp((float *)S.r), z = gp((float *)S.z);
    gptr<const float> p = gp((const float *)((S.it & 1) ? S.p1 : S.p0));  // what the next SpMV reads as p_old
    gptr<double> xacc = gp(S.xacc);
    gptr<const double> rb = gp((const double *)S.b);
    for (int tile = 0; tile < kTiles<float>; ++tile) {
        const int64_t i = ((int64_t)lb * kTiles<float> + tile) * kBB + threadIdx.x;
        if (i >= S.n_dof)
            break;
        const double inv = gp(S.inv)[i];
#pragma unroll
        for (int r = 0; r < kRmax; ++r) {
            if (r >= R || !((upd >> r) & 1))
                continue;
            const int64_t q = i * R + r;
            xacc[q] += (double)x[q];
            x[q] = 0.f;
            if ((go >> r) & 1) {
                const double ri = (double)(float)rb[q];
                const double zi = (double)(float)(ri * inv);
                rv[q] = (float)ri;
                z[q] = (float)zi;
                v[3 * r] += ri * zi;
                v[3 * r + 1] += ri * ri;
                v[3 * r + 2] += ri * (double)p[q];
            }
        }
    }
}

template <typename XT>
__global__ __launch_bounds__(kBB) void kb_update(BSys *__restrict__ sys, const int2 *__restrict__ tab) {
    const int2 t = tab[blockIdx.x];
    const BSys &S = sys[t.x];
    const int go = S.go;
    if constexpr (sizeof(XT) == 4) {
        const int upd = S.upd;
        if (upd) {
            double v[3 * kRmax];
#pragma unroll
            for (int k = 0; k < 3 * kRmax; ++k)
                v[k] = 0.0;
            breplace(S, t.y, upd, go, v);
            block_sum<3 * kRmax>(v);
            if (threadIdx.x == 0)
                for (int r = 0; r < S.R; ++r)
                    for (int k = 0; k < 3; ++k)
                        S.part2[((size_t)t.y * S.R + r) * 3 + k] = v[3 * r + k];
            return;
        }
    }
    if (!go)
        return;
    double v[2 * kRmax];
#pragma unroll
    for (int k = 0; k < 2 * kRmax; ++k)
        v[k] = 0.0;
    switch (S.R) {
    case 1: bupdate<XT, 1>(S, t.y, go, v); break;
    case 2: bupdate<XT, 2>(S, t.y, go, v); break;
    case 3: bupdate<XT, 3>(S, t.y, go, v); break;
    default: bupdate<XT, 4>(S, t.y, go, v); break;
    }
    block_sum<2 * kRmax>(v);
    if (threadIdx.x == 0)
        for (int r = 0; r < S.R; ++r) {
            S.part2[((size_t)t.y * S.R + r) * 3] = v[2 * r];
            S.part2[((size_t)t.y * S.R + r) * 3 + 1] = v[2 * r + 1];
            S.part2[((size_t)t.y * S.R + r) * 3 + 2] = 0.0;
        }
}

template <typename VT>
void launch_bspmv(int G, unsigned nb, hipStream_t s, BSys *sys, const int2 *tab) {
    switch (G) {
    case 2: hipLaunchKernelGGL((kb_spmv_p<2, VT, VT>), dim3(nb), dim3(kBB), 0, s, sys, tab); break;
    case 4: hipLaunchKernelGGL((kb_spmv_p<4, VT, VT>), dim3(nb), dim3(kBB), 0, s, sys, tab); break;
    case 8: hipLaunchKernelGGL((kb_spmv_p<8, VT, VT>), dim3(nb), dim3(kBB), 0, s, sys, tab); break;
    case 16: hipLaunchKernelGGL((kb_spmv_p<16, VT, VT>), dim3(nb), dim3(kBB), 0, s, sys, tab); break;
    case 32: hipLaunchKernelGGL((kb_spmv_p<32, VT, VT>), dim3(nb), dim3(kBB), 0, s, sys, tab); break;
    default: hipLaunchKernelGGL((kb_spmv_p<64, VT, VT>), dim3(nb), dim3(kBB), 0, s, sys, tab); break;
    }
}

}  // namespace
}  // namespace fx

// ===========================================================================
// host side
// ===========================================================================
struct fx_batch {
    fx_ctx *ctx = nullptr;
    fx_solve_opts opts{};
    int G = 0;  // lane group forced by the options (0: each system's own)
    int poly_deg = 0;  // polynomial batch: FP32 Chebyshev steps per outer iteration
    struct Slot {
        bool used = false;
        fx_mesh *mesh = nullptr;
        bool foreign = false;  // the mesh belongs to another context (counted in its foreign_slots)
        ~Slot() {
            if (foreign)
                mesh->foreign_slots--;
        }
        int R = 0;
        int G = 8;      // lanes per node row (lane-group SpMV, replacement pass)
        int phase = 0;  // 0 running, 2 finished
        fx_status err = FX_OK;
        int64_t iters = 0, outer = 0;  // Krylov iterations (slowest load case), residual replacements
        double crit[fx::kRmax] = {0, 0, 0, 0};
        // device vectors from the context's pool: a later slot reuses them
        fx::PoolBuf<double> b, rin, x, r, z, p0, p1, ap, xacc, part1, part2;
        fx::PoolBuf<double> zq0, zq1, zq2, rs;  // polynomial batch: Chebyshev iterates, scaled residual
        fx::PinBuf<double> hx;                 // polynomial batch: the solutions exported at the finish
        bool pending = false;                  // polynomial batch: admitted, its start sequence not yet run
        std::chrono::steady_clock::time_point t0;
        double seconds = 0.0;
    };
    std::vector<std::unique_ptr<Slot>> slots;
    std::vector<fx::BSys> h_sys;
    fx::DBuf<fx::BSys> d_sys;
    static constexpr int kGroups = 6;  // lane groups 2, 4, ..., 64
    // Iteration chunks are pipelined: while chunk k runs, the host processes
    // the state of chunk k-1.  Each of the kSets chunks that may be in
    // flight has its own block tables (device, and their pinned host
    // staging) and its own events.  Three: a budgeted fx_batch_run with
    // overlap leaves up to three queued, which at full load (8 ms each on
    // the 512-mesh population) cover the caller's collection and admission
    // between two calls.
    static constexpr int kSets = 3;
    struct ChunkSet {
        // tab3s / tab3: the polynomial batch's FP64 outer SpMV (sliced-ELL and node-block operators)
        fx::DBuf<int2> tab1g[kGroups], tab1s, tab2, tab3s, tab3;
        fx::DBuf<int32_t> run;
        int n1g[kGroups] = {0, 0, 0, 0, 0, 0};
        int n1s = 0, n2 = 0, n3s = 0, n3 = 0, nrun = 0;
        int n1s_r[fx::kRmax] = {0, 0, 0, 0};  // polynomial batch: tab1s segments by load-case count
        int inner_rows = fx::kBB;               // polynomial batch: node rows per tab1s entry
        int64_t version = -1;          // of the running-slot set these tables were built for
        int2 *h_tab = nullptr;         // pinned staging of the tables
        size_t h_tab_cap = 0;
        hipEvent_t done = nullptr;
        std::vector<hipEvent_t> ev;    // SpMV timing: 2 per iteration
        int ev_used = 0;
    };
    ChunkSet cs[kSets];
    // chunks launched and not yet processed (the oldest in cs[head]); with
    // `overlap` (fx_batch_set_overlap) a budgeted fx_batch_run returns with
    // up to three still running, so the GPU keeps iterating while the caller
    // collects and admits systems
    int inflight = 0, head = 0;
    bool overlap = false;
    // polynomial batch: the sliced-ELL inner step as one launch (false) or
    // one launch per load-case count (fx_batch_autotune)
    bool inner_split = false;
    int64_t run_version = 0;           // bumped when the tables must be rebuilt
    int64_t tab_blocks = 0, dead_blocks = 0;  // in the current tables, and of slots finished since
    fx::BFin *fin = nullptr;           // pinned host memory, one per slot capacity
    size_t fin_cap = 0;
    fx::DBuf<int2> d_tab4;
    fx::DBuf<int32_t> d_list2;
    // fx_batch_get: pinned host staging of a slot's solutions, rhs-major,
    // written by the device and copied out by the host
    double *h_get = nullptr;
    size_t h_get_cap = 0;
    std::vector<int2> h_tab_start;
    // SpMV timing (instrumentation): when on, every SpMV launch is bracketed
    // by two HIP events on the stream; the device counter sums the launches'
    // algorithmic bytes (kb_red2)
    bool timing = false;
    int64_t spmv_launches = 0, timed = 0;
    double timed_ms = 0.0;
    fx::DBuf<double> d_bytes;
    // orders the batch's stream after work queued on the stream of an
    // admitted mesh's own context (re-recorded at each such admission)
    hipEvent_t adm_ev = nullptr;
    // the mesh whose single solves this batch serves (fx::poly_solve), or null
    fx_mesh *owner = nullptr;
    // serialises the batch's calls with a foreign mesh's removal from
    // another host thread (batch_drop_mesh_foreign)
    std::mutex mx;
    fx_batch();
    ~fx_batch();
};

namespace {
// every live batch of the process, whatever its context: a mesh admitted
// to a batch of another context is dropped from it when the mesh goes
std::mutex g_batches_mx;
std::vector<fx_batch *> g_batches;
}  // namespace

fx_batch::fx_batch() {
    std::lock_guard<std::mutex> lk(g_batches_mx);
    g_batches.push_back(this);
}

fx_batch::~fx_batch() {
    if (owner)
        owner->poly_batch = nullptr;
    {
        // the slots go first, under the registry's lock: a foreign mesh being
        // destroyed meanwhile either drops its slot here before, or finds
        // its count of foreign slots already down and no batch to visit
        std::lock_guard<std::mutex> lk(g_batches_mx);
        slots.clear();
        g_batches.erase(std::find(g_batches.begin(), g_batches.end(), this));
    }
    {
        if (adm_ev)
            (void)hipEventDestroy(adm_ev);
        if (fin)
            (void)hipHostFree(fin);
        if (h_get)
            (void)hipHostFree(h_get);
        for (auto &c : cs) {
            for (auto &e : c.ev)
                if (e)
                    (void)hipEventDestroy(e);
            if (c.done)
                (void)hipEventDestroy(c.done);
            if (c.h_tab)
                (void)hipHostFree(c.h_tab);
        }
    }
}

namespace {

using fx::BRhs;
using fx::BSys;
using fx::kBB;
using fx::kRmax;

bool is_mixed(const fx_batch *B) { return B->opts.precision == FX_PREC_MIXED; }
bool is_poly(const fx_batch *B) { return B->opts.precision == FX_PREC_MIXED_POLY; }

// mixed batches on sliced-ELL meshes use the symmetric FP32 operator
// (FX_SSYM=0: the full one)
bool use_ssym() {
    const char *e = std::getenv("FX_SSYM");
    return !(e && e[0] == '0');
}

void start_solve(fx_batch *B, fx_batch::Slot &sl, BSys &S) {
    for (int r = 0; r < kRmax; ++r) {
        BRhs &h = S.rhs[r];
        h = BRhs{};
        if (r >= sl.R) {
            h.status = 4;
            continue;
        }
        h.status = 0;
        h.tol = B->opts.tol;
        h.stop_mode = B->opts.stop_mode;
        h.max_iter = B->opts.max_iter;
        h.gamma = 1.0;
        if (is_mixed(B)) {
            h.delta = B->opts.inner_tol;
            h.max_upd = B->opts.max_refine;
            h.prev = HUGE_VAL;
        }
    }
}

void fill_sys(fx_batch *B, fx_batch::Slot &sl, BSys &S) {
    fx_mesh *m = sl.mesh;
    S.nb_ptr = m->nb_ptr.p;
    S.nb_col = m->nb_col.p;
    S.val = is_mixed(B) ? (const void *)m->kval32.p : (const void *)m->kval.p;
    S.val64 = m->kval.p;
    S.kc = is_mixed(B) ? m->kcorr32.p : nullptr;
    S.sell = m->sell ? 1 : 0;
    S.sl_ptr = m->sl_ptr.p;
    S.scol = m->scol.p;
    S.sval = is_mixed(B) ? (const void *)m->sval32.p : (const void *)m->sval64.p;
    S.ssym = is_mixed(B) && m->sell && m->ssym && use_ssym() ? 1 : 0;
    if (S.ssym) {
        S.su_ptr = m->su_ptr.p;
        S.scol_u = m->scol_u.p;
        S.sll_ptr = m->sll_ptr.p;
        S.lrec = m->lrec.p;
        S.lcol = m->lcol.p;
        S.uval = m->uval32.p;
        S.kc = m->kcorr32s.p;  // the correction of the symmetric operator
    }
    S.inv = m->inv_diag.p;
    S.b = is_mixed(B) ? sl.rin.p : sl.b.p;
    S.btrue = sl.b.p;
    S.x = sl.x.p;
    S.r = sl.r.p;
    S.z = sl.z.p;
    S.p0 = sl.p0.p;
    S.p1 = sl.p1.p;
    S.ap = sl.ap.p;
    S.xacc = sl.xacc.p;
    S.part1 = sl.part1.p;
    S.part2 = sl.part2.p;
    S.n_nodes = m->n_nodes;
    S.n_dof = m->n_dof;
    S.R = sl.R;
    S.G = sl.G;
    S.nparts1 = m->sell ? fx::nblk(m->n_nodes, kBB) : fx::nblk(m->n_nodes, kBB / sl.G);
    // SpMV algorithmic bytes: values in the working precision, one column
    // index per 3x3 block, node-row pointers (+ the FP32 operator's 3x3
    // correction per node); per right-hand side p gathered once and Ap
    // written, in the vector precision (FP32 in the mixed solve)
    S.bytes_acc = B->d_bytes.p;
    S.spmv_bytes0 = (is_mixed(B) ? 4.0 : 8.0) * (double)m->nnz_struct + 4.0 * (double)m->nblk +
                    4.0 * (double)(m->n_nodes + 1) + (is_mixed(B) ? 36.0 * (double)m->n_nodes : 0.0);
    if (S.ssym)  // upper blocks once (values + column), lower blocks as (record, column), the correction
        S.spmv_bytes0 = 40.0 * (double)m->nblk_u + 8.0 * (double)m->nblk_l + 36.0 * (double)m->n_nodes +
                        16.0 * (double)((m->n_nodes + fx::kSliceRows - 1) / fx::kSliceRows + 1);
    S.spmv_bytes_rhs = (is_mixed(B) ? 8.0 : 16.0) * (double)m->n_dof;
    // replacement pass: FP64 values, block columns, row pointers, the
    // Jacobi scaling; per right-hand side xacc and x gathered once, b read
    // and the new residual written
    S.upd_bytes0 = 8.0 * (double)m->nnz_struct + 4.0 * (double)m->nblk + 4.0 * (double)(m->n_nodes + 1) +
                   8.0 * (double)m->n_dof;
    S.upd_bytes_rhs = 28.0 * (double)m->n_dof;
    S.nparts2 = fx::nblk(m->n_dof, kBB * (is_mixed(B) ? fx::kTiles<float> : fx::kTiles<double>));
    if (is_poly(B)) {
        // FP64 x, r, direction pp, its image ps, w = K z; FP32 Chebyshev
        // iterates and residual copy
        S.poly = 1;
        S.deg = B->poly_deg;
        S.b = sl.b.p;
        S.pp = sl.p0.p;
        S.ps = sl.p1.p;
        S.pw = sl.ap.p;
        S.zq[0] = sl.zq0.p;
        S.zq[1] = sl.zq1.p;
        S.zq[2] = sl.zq2.p;
        S.rs = sl.rs.p;
        S.hx = sl.hx.p;
        S.val16 = m->kval16.p;
        S.sval16 = m->sval16.p;
        S.sval16b = m->sval16.p ? m->sval16.p + (size_t)m->scol.n * 8 : nullptr;
        S.sval64 = m->sval64.p;
        S.nparts3 = m->sell ? S.nparts1 : fx::nblk(m->n_nodes, kBB / fx::kOuterG);
        // one inner step: the FP16 scaled operator (values, one column index
        // per 3x3 block, node-row pointers), and per load case the FP64 z_j
        // gathered once, z_{j-1} and rs read, z_{j+1} written
        S.cheb_bytes0 = 2.0 * (double)m->nnz_struct + 4.0 * (double)m->nblk + 4.0 * (double)(m->n_nodes + 1);
        S.cheb_bytes_rhs = 32.0 * (double)m->n_dof;
    }
}

// workgroups are dispatched round-robin over the 8 XCDs: permute an SpMV
// table so that each XCD sweeps a contiguous range of row groups
void xcd_permute(std::vector<int2> &t) {
    const unsigned nb = (unsigned)t.size(), q = nb / 8, rr = nb % 8;
    std::vector<int2> perm(nb);
    for (unsigned id = 0; id < nb; ++id) {
        const unsigned x = id % 8, i = id / 8;
        perm[id] = t[x * q + (x < rr ? x : rr) + i];
    }
    t.swap(perm);
}

inline int group_index(int G) {  // 2 -> 0, 4 -> 1, ..., 64 -> 5
    int i = 0;
    while ((2 << i) < G)
        ++i;
    return i;
}

void ensure_pinned(int2 *&p, size_t &cap, size_t n) {
    if (cap >= n)
        return;
    if (p)
        FX_HIP(hipHostFree(p));
    p = nullptr;
    cap = 0;
    const size_t c = std::max<size_t>(n, 2 * cap + 1024);
    FX_HIP(hipHostMalloc((void **)&p, c * sizeof(int2), hipHostMallocDefault));
    cap = c;
}

// Block tables of the running slots for chunk set `set`: tab1g / tab1s SpMV
// blocks (lane-group and sliced-ELL operators), tab2 vector blocks, run the
// running slots.  The tables are staged in the set's pinned buffer and
// uploaded on the stream without waiting: the set's previous chunk has been
// processed (its event waited), so neither buffer is still being read.
//
// The SpMV-shaped tables are ordered by the slots' load-case count, four
// first: a block's work grows with the count, so the longest blocks lead a
// launch and the shortest fill its tail, and neighbouring workgroups on a CU
// carry equal work.  Each count's segment is permuted on its own
// (xcd_permute; over the whole table the XCD ranges are then only rotated,
// and a launch over one segment starts them at block 0).  The vector table
// keeps the slots' order, unpermuted: the polynomial batch's update ran
// 653 us per launch so, against 752 us ordered by count and 705 us ordered
// and permuted.  The order changes no value: a block's partial sums go to
// its own slot.
struct ByCount {
    std::vector<int2> r[kRmax];
    void add(int R, int k, int b) { r[R - 1].push_back(make_int2(k, b)); }
    std::vector<int2> flat(bool permute, int *counts = nullptr) {
        std::vector<int2> t;
        for (int q = kRmax - 1; q >= 0; --q) {
            if (counts)
                counts[q] = (int)r[q].size();
            if (permute)
                xcd_permute(r[q]);
            t.insert(t.end(), r[q].begin(), r[q].end());
        }
        return t;
    }
};

void build_tables(fx_batch *B, int set) {
    fx_batch::ChunkSet &C = B->cs[set];
    ByCount b1g[fx_batch::kGroups], b1s, b3s, b3;
    std::vector<int2> t2;
    std::vector<int> run;
    // the polynomial batch's inner step: kInnerRowsLarge rows per workgroup
    // when the running sliced-ELL systems hold enough rows to fill the GPU
    // many times over, kBB otherwise
    int64_t poly_rows = 0;
    for (size_t k = 0; k < B->slots.size(); ++k) {
        auto &sl = *B->slots[k];
        if (sl.used && sl.phase == 0 && !sl.pending && B->h_sys[k].poly && B->h_sys[k].sell)
            poly_rows += B->h_sys[k].n_nodes;
    }
    C.inner_rows = poly_rows >= fx::kInnerRowsLargeMin ? fx::kInnerRowsLarge : kBB;
    for (size_t k = 0; k < B->slots.size(); ++k) {
        auto &sl = *B->slots[k];
        if (!sl.used || sl.phase != 0 || sl.pending)
            continue;
        run.push_back((int)k);
        const BSys &S = B->h_sys[k];
        // the polynomial batch's sliced-ELL inner step: C.inner_rows rows a block
        const int np1 = S.poly && S.sell ? fx::nblk(S.n_nodes, C.inner_rows) : S.nparts1;
        for (int b = 0; b < np1; ++b)
            (S.sell ? b1s : b1g[group_index(S.G)]).add(S.R, (int)k, b);
        for (int b = 0; b < S.nparts2; ++b)
            t2.push_back(make_int2((int)k, b));
        if (S.poly)
            for (int b = 0; b < S.nparts3; ++b)
                (S.sell ? b3s : b3).add(S.R, (int)k, b);
    }
    std::vector<int2> t1g[fx_batch::kGroups];
    for (int g = 0; g < fx_batch::kGroups; ++g)
        t1g[g] = b1g[g].flat(true);
    std::vector<int2> t1s = b1s.flat(true, C.n1s_r), t3s = b3s.flat(true), t3 = b3.flat(true);
    // staged: every table below and the running-slot list
    size_t total = t1s.size() + t2.size() + t3s.size() + t3.size() + (run.size() + 1) / 2 + 1;
    for (auto &t : t1g)
        total += t.size();
    ensure_pinned(C.h_tab, C.h_tab_cap, total);
    hipStream_t s = B->ctx->stream;
    size_t off = 0;
    auto room = [&](size_t n) {
        if (off + n > C.h_tab_cap)
            throw fx::Error{FX_ERR_INVALID, "batch tables: staging buffer too small (internal error)"};
    };
    auto stage = [&](fx::DBuf<int2> &d, std::vector<int2> &t) {
        if (t.empty())
            return;
        room(t.size());
        std::copy(t.begin(), t.end(), C.h_tab + off);
        d.alloc(std::max<size_t>(d.n, t.size()));
        FX_HIP(hipMemcpyAsync(d.p, C.h_tab + off, t.size() * sizeof(int2), hipMemcpyHostToDevice, s));
        off += t.size();
    };
    for (int g = 0; g < fx_batch::kGroups; ++g) {
        C.n1g[g] = (int)t1g[g].size();
        stage(C.tab1g[g], t1g[g]);
    }
    C.n1s = (int)t1s.size();
    stage(C.tab1s, t1s);
    C.n3s = (int)t3s.size();
    stage(C.tab3s, t3s);
    C.n3 = (int)t3.size();
    stage(C.tab3, t3);
    C.n2 = (int)t2.size();
    room(t2.size() + (run.size() + 1) / 2);
    std::copy(t2.begin(), t2.end(), C.h_tab + off);  // vector blocks: no XCD permutation
    if (!t2.empty()) {
        C.tab2.alloc(std::max<size_t>(C.tab2.n, t2.size()));
        FX_HIP(hipMemcpyAsync(C.tab2.p, C.h_tab + off, t2.size() * sizeof(int2), hipMemcpyHostToDevice, s));
        off += t2.size();
    }
    C.nrun = (int)run.size();
    if (C.nrun) {
        int32_t *h = (int32_t *)(C.h_tab + off);
        std::copy(run.begin(), run.end(), h);
        C.run.alloc(std::max<size_t>(C.run.n, run.size()));
        FX_HIP(hipMemcpyAsync(C.run.p, h, run.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
    }
    C.version = B->run_version;
    int64_t blocks = (int64_t)t1s.size() + (int64_t)t2.size() + (int64_t)t3s.size() + (int64_t)t3.size();
    for (auto &t : t1g)
        blocks += (int64_t)t.size();
    B->tab_blocks = blocks;
    B->dead_blocks = 0;
}

void drain(fx_batch *B);

void reserve_sys(fx_batch *B, size_t n) {
    if (B->d_sys.p && B->d_sys.n >= n)
        return;
    const size_t cap = std::max<size_t>({n, 2 * B->d_sys.n, 16});
    // the finish summaries: pinned host memory the kernels write, replaced
    // once no chunk is in flight
    drain(B);
    fx::BFin *f = nullptr;
    FX_HIP(hipHostMalloc((void **)&f, cap * sizeof(fx::BFin), hipHostMallocDefault));
    std::memset(f, 0, cap * sizeof(fx::BFin));
    if (B->fin) {
        std::memcpy(f, B->fin, B->fin_cap * sizeof(fx::BFin));
        FX_HIP(hipHostFree(B->fin));
    }
    B->fin = f;
    B->fin_cap = cap;
    BSys *p = nullptr;
    FX_HIP(hipMalloc(&p, cap * sizeof(BSys)));
    if (B->d_sys.p) {
        FX_HIP(hipMemcpyAsync(p, B->d_sys.p, B->d_sys.n * sizeof(BSys), hipMemcpyDeviceToDevice, B->ctx->stream));
        FX_HIP(hipStreamSynchronize(B->ctx->stream));
        B->d_sys.release();
    }
    B->d_sys.p = p;
    B->d_sys.n = cap;
}

// an admission that threw: the slot goes back to the free list, and the
// running slots' tables are unchanged
void release_failed_slot(fx_batch *B, int k) {
    if (k < 0 || k >= (int)B->slots.size())
        return;
    B->slots[k].reset(new fx_batch::Slot());
    if (k < (int)B->h_sys.size())
        B->h_sys[k] = BSys{};
}

// The start sequence of the polynomial batch's pending slots, all together:
// the power iteration (lambda_max of D^-1 K32), the Chebyshev coefficients,
// x = 0, r = b and the first stop test; their solves run from the next
// chunk on.
void start_pending_poly(fx_batch *B) {
    std::vector<int> list;
    for (size_t k = 0; k < B->slots.size(); ++k)
        if (B->slots[k] && B->slots[k]->used && B->slots[k]->pending)
            list.push_back((int)k);
    if (list.empty())
        return;
    hipStream_t s = B->ctx->stream;
    std::vector<int2> tv, ts, tg[fx_batch::kGroups];
    for (int k : list) {
        const BSys &S = B->h_sys[k];
        for (int b = 0; b < S.nparts2; ++b)
            tv.push_back(make_int2(k, b));
        for (int b = 0; b < S.nparts1; ++b)
            (S.sell ? ts : tg[group_index(S.G)]).push_back(make_int2(k, b));
    }
    xcd_permute(ts);
    for (auto &t : tg)
        xcd_permute(t);
    // one staging array: vector table, sliced-ELL table, lane-group tables, slot list
    std::vector<int2> all(tv);
    all.insert(all.end(), ts.begin(), ts.end());
    for (auto &t : tg)
        all.insert(all.end(), t.begin(), t.end());
    const size_t n_list2 = (list.size() + 1) / 2;
    all.resize(all.size() + n_list2);
    std::memcpy(all.data() + (all.size() - n_list2), list.data(), list.size() * sizeof(int));
    B->d_tab4.alloc(std::max<size_t>(B->d_tab4.n, all.size()));
    FX_HIP(hipMemcpyAsync(B->d_tab4.p, all.data(), all.size() * sizeof(int2), hipMemcpyHostToDevice, s));
    const int2 *base = B->d_tab4.p;
    const int2 *p_tv = base, *p_ts = base + tv.size();
    const int2 *p_tg[fx_batch::kGroups];
    int n_tg[fx_batch::kGroups];
    size_t off = tv.size() + ts.size();
    for (int g = 0; g < fx_batch::kGroups; ++g) {
        p_tg[g] = base + off;
        n_tg[g] = (int)tg[g].size();
        off += tg[g].size();
    }
    const int32_t *p_list = (const int32_t *)(base + off);
    fx::launch_poly_power_start(s, B->d_sys.p, p_tv, (int)tv.size(), p_ts, (int)ts.size(), p_tg, n_tg, p_list,
                                (int)list.size(), B->fin);
    FX_HIP(hipGetLastError());
    FX_HIP(hipStreamSynchronize(s));  // the staging array is reused
    for (int k : list)
        B->slots[k]->pending = false;
    B->run_version++;
}

// kb_init over the vector blocks of `slots`, then kb_red2 on them: their
// solves start with the next iteration chunk.
void start_slots(fx_batch *B, const std::vector<int> &slots) {
    if (slots.empty())
        return;
    hipStream_t s = B->ctx->stream;
    std::vector<int2> &t = B->h_tab_start;
    t.clear();
    for (int k : slots)
        for (int b = 0; b < B->h_sys[k].nparts2; ++b)
            t.push_back(make_int2(k, b));
    B->d_tab4.alloc(std::max<size_t>(B->d_tab4.n, t.size()));
    B->d_list2.alloc(std::max<size_t>(B->d_list2.n, slots.size()));
    FX_HIP(hipMemcpyAsync(B->d_tab4.p, t.data(), t.size() * sizeof(int2), hipMemcpyHostToDevice, s));
    FX_HIP(hipMemcpyAsync(B->d_list2.p, slots.data(), slots.size() * sizeof(int), hipMemcpyHostToDevice, s));
    if (is_mixed(B))
        hipLaunchKernelGGL(fx::kb_init<float>, dim3((unsigned)t.size()), dim3(kBB), 0, s, B->d_sys.p, B->d_tab4.p);
    else
        hipLaunchKernelGGL(fx::kb_init<double>, dim3((unsigned)t.size()), dim3(kBB), 0, s, B->d_sys.p,
                           B->d_tab4.p);
    hipLaunchKernelGGL(fx::kb_red2, dim3((unsigned)slots.size()), dim3(kBB), 0, s, B->d_sys.p, B->d_list2.p,
                       B->fin);
    FX_HIP(hipGetLastError());
    FX_HIP(hipStreamSynchronize(s));  // the host tables are reused
}

void launch_iterations(fx_batch *B, int n, int set) {
    fx_batch::ChunkSet &C = B->cs[set];
    hipStream_t s = B->ctx->stream;
    if (B->timing && (int)C.ev.size() < 2 * n) {
        const size_t old = C.ev.size();
        C.ev.resize(2 * n, nullptr);
        for (size_t k = old; k < C.ev.size(); ++k)
            FX_HIP(hipEventCreate(&C.ev[k]));
    }
    if (is_poly(B)) {  // n outer iterations, each poly_deg FP32 inner steps
        const int2 *tg[fx_batch::kGroups];
        for (int g = 0; g < fx_batch::kGroups; ++g)
            tg[g] = C.tab1g[g].p;
        for (int k = 0; k < n; ++k)
            fx::launch_poly_outer(s, B->d_sys.p, B->poly_deg, C.tab2.p, C.n2, C.tab1s.p, C.n1s, C.n1s_r,
                                  B->inner_split, C.inner_rows, tg, C.n1g, C.tab3s.p, C.n3s, C.tab3.p, C.n3,
                                  C.run.p, C.nrun, B->fin,
                                  B->timing ? C.ev[2 * k] : nullptr, B->timing ? C.ev[2 * k + 1] : nullptr);
        C.ev_used = B->timing ? 2 * n : 0;
        B->spmv_launches += (int64_t)n * B->poly_deg;
        FX_HIP(hipGetLastError());
        if (!C.done)
            FX_HIP(hipEventCreateWithFlags(&C.done, hipEventDisableTiming));
        FX_HIP(hipEventRecord(C.done, s));
        return;
    }
    for (int k = 0; k < n; ++k) {
        if (is_mixed(B))
            hipLaunchKernelGGL(fx::kb_pupdate<float>, dim3(C.n2), dim3(kBB), 0, s, B->d_sys.p, C.tab2.p);
        else
            hipLaunchKernelGGL(fx::kb_pupdate<double>, dim3(C.n2), dim3(kBB), 0, s, B->d_sys.p, C.tab2.p);
        if (B->timing)
            FX_HIP(hipEventRecord(C.ev[2 * k], s));
        for (int g = 0; g < fx_batch::kGroups; ++g)
            if (C.n1g[g]) {
                if (is_mixed(B))
                    fx::launch_bspmv<float>(2 << g, C.n1g[g], s, B->d_sys.p, C.tab1g[g].p);
                else
                    fx::launch_bspmv<double>(2 << g, C.n1g[g], s, B->d_sys.p, C.tab1g[g].p);
            }
        if (C.n1s) {
            if (is_mixed(B))
                hipLaunchKernelGGL((fx::kb_spmv_sell<float, float>), dim3(C.n1s), dim3(kBB), 0, s, B->d_sys.p,
                                   C.tab1s.p);
            else
                hipLaunchKernelGGL((fx::kb_spmv_sell<double, double>), dim3(C.n1s), dim3(kBB), 0, s, B->d_sys.p,
                                   C.tab1s.p);
        }
        if (B->timing)
            FX_HIP(hipEventRecord(C.ev[2 * k + 1], s));
        hipLaunchKernelGGL(fx::kb_red1, dim3(C.nrun), dim3(kBB), 0, s, B->d_sys.p, C.run.p);
        if (is_mixed(B))
            hipLaunchKernelGGL(fx::kb_update<float>, dim3(C.n2), dim3(kBB), 0, s, B->d_sys.p, C.tab2.p);
        else
            hipLaunchKernelGGL(fx::kb_update<double>, dim3(C.n2), dim3(kBB), 0, s, B->d_sys.p, C.tab2.p);
        hipLaunchKernelGGL(fx::kb_red2, dim3(C.nrun), dim3(kBB), 0, s, B->d_sys.p, C.run.p, B->fin);
    }
    C.ev_used = B->timing ? 2 * n : 0;
    B->spmv_launches += n;
    FX_HIP(hipGetLastError());
    if (!C.done)
        FX_HIP(hipEventCreateWithFlags(&C.done, hipEventDisableTiming));
    FX_HIP(hipEventRecord(C.done, s));
}

// after the set's chunk has completed
void collect_sample(fx_batch *B, int set) {
    fx_batch::ChunkSet &C = B->cs[set];
    for (int k = 0; k + 1 < C.ev_used; k += 2) {
        float ms = 0.f;
        FX_HIP(hipEventElapsedTime(&ms, C.ev[k], C.ev[k + 1]));
        B->timed_ms += ms;
        B->timed += is_poly(B) ? B->poly_deg : 1;  // a polynomial pair brackets the inner steps
    }
    C.ev_used = 0;
}

// After a chunk has completed: the slots whose summary kb_red2 has written
// (none of their load cases runs) finish.  A finished slot may stay in the
// block tables for a while; its blocks see no running load case and return
// at once.  The tables are rebuilt when a quarter of their blocks belong to
// finished slots.
void advance_phases(fx_batch *B, int set) {
    collect_sample(B, set);
    for (size_t k = 0; k < B->slots.size(); ++k) {
        auto &sl = *B->slots[k];
        if (!sl.used || sl.phase != 0)
            continue;
        const fx::BFin &F = B->fin[k];
        if (!((volatile const int32_t &)F.done))
            continue;
        sl.iters = F.it;
        sl.outer = F.nupd;
        for (int r = 0; r < sl.R; ++r) {
            sl.crit[r] = F.crit[r];
            if (F.status[r] == 2)
                sl.err = FX_ERR_BREAKDOWN;
            else if (F.status[r] == 3 && sl.err == FX_OK)
                sl.err = FX_ERR_NOT_CONVERGED;
            else if (F.status[r] >= 16 && sl.err == FX_OK)
                sl.err = (fx_status)(F.status[r] - 16);  // the mesh's element kernels failed
        }
        sl.phase = 2;
        sl.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - sl.t0).count();
        const BSys &S = B->h_sys[k];
        B->dead_blocks += S.nparts1 + S.nparts2 + (S.poly ? S.nparts3 : 0);
    }
    if (4 * B->dead_blocks > B->tab_blocks)
        B->run_version++;
}

// Waits for and processes the chunks a budgeted fx_batch_run left in flight.
// The calls that change what those chunks read or write (the finish
// summaries' buffer, a running slot's vectors, the tables they were built
// with) drain first.  A finished slot needs no drain: every kernel returns
// at once for it (no load case runs, its summary is written), so its
// vectors are final and nothing in flight reads or writes them.
void drain(fx_batch *B) {
    while (B->inflight > 0) {
        FX_HIP(hipEventSynchronize(B->cs[B->head].done));
        advance_phases(B, B->head);
        B->head = (B->head + 1) % fx_batch::kSets;
        --B->inflight;
    }
}

}  // namespace

namespace fx {
void batch_free(fx_batch *b) { delete b; }

void batch_drop_mesh_foreign(fx_mesh *m) {
    if (m->foreign_slots.load() == 0)
        return;
    std::lock_guard<std::mutex> lk(g_batches_mx);
    for (fx_batch *B : g_batches) {
        if (B->ctx == m->ctx)
            continue;
        bool uses = false;
        for (auto &sl : B->slots)
            uses |= sl && sl->used && sl->mesh == m;
        if (!uses)
            continue;
        std::lock_guard<std::mutex> bl(B->mx);  // no call of the batch runs meanwhile
        (void)hipSetDevice(B->ctx->device);
        (void)hipStreamSynchronize(B->ctx->stream);  // the batch's stream may still read the mesh
        batch_drop_mesh(B, m);
    }
}

void batch_drop_mesh(fx_batch *B, fx_mesh *m) {
    for (size_t k = 0; k < B->slots.size(); ++k)
        if (B->slots[k] && B->slots[k]->used && B->slots[k]->mesh == m) {
            if (B->slots[k]->phase == 0)
                B->run_version++;
            B->h_sys[k] = BSys{};
            B->slots[k].reset(new fx_batch::Slot());
        }
}
}  // namespace fx

extern "C" {

fx_status fx_batch_create(fx_ctx *ctx, const fx_solve_opts *opts, fx_batch **out) {
    if (!ctx || !out)
        return FX_ERR_INVALID;
    fx_batch *B = nullptr;
    try {
        FX_HIP(hipSetDevice(ctx->device));
        B = new fx_batch();
        B->ctx = ctx;
        if (opts)
            B->opts = *opts;
        else
            fx_solve_opts_default(&B->opts);
        if (B->opts.precision != FX_PREC_FP64 && B->opts.precision != FX_PREC_MIXED &&
            B->opts.precision != FX_PREC_MIXED_POLY) {
            delete B;
            fx::set_error("fx_batch_create: unknown precision");
            return FX_ERR_INVALID;
        }
        if (B->opts.precision == FX_PREC_MIXED_POLY) {
            B->poly_deg = B->opts.poly_degree > 0 ? B->opts.poly_degree : 15;
            if (B->poly_deg > fx::kPolyMax) {
                delete B;
                fx::set_error("fx_batch_create: poly_degree must be 0 (default 15) or 1.." +
                              std::to_string(fx::kPolyMax));
                return FX_ERR_INVALID;
            }
        }
        if (B->opts.precision == FX_PREC_MIXED &&
            !(B->opts.inner_tol > 0.0 && B->opts.inner_tol < 1.0 && B->opts.max_refine >= 1)) {
            delete B;
            fx::set_error("fx_batch_create: MIXED needs 0 < inner_tol < 1 and max_refine >= 1");
            return FX_ERR_INVALID;
        }
        const int g = B->opts.spmv_group;
        if (!(g == 0 || g == 2 || g == 4 || g == 8 || g == 16 || g == 32 || g == 64)) {
            delete B;
            fx::set_error("fx_batch_create: spmv_group must be 0 or a power of two in 2..64");
            return FX_ERR_INVALID;
        }
        // 0: chosen from the first system admitted (as fx_solve chooses it)
        B->G = g;
        B->d_bytes.alloc(1);
        FX_HIP(hipMemsetAsync(B->d_bytes.p, 0, sizeof(double), ctx->stream));
        FX_HIP(hipStreamSynchronize(ctx->stream));
        {
            std::lock_guard<std::mutex> lk(ctx->reg);
            ctx->batches.push_back(B);
        }
        *out = B;
    } catch (const fx::Error &e) {
        delete B;
        fx::set_error(e.msg);
        return e.code;
    } catch (const std::exception &e) {
        delete B;
        fx::set_error(e.what());
        return FX_ERR_OOM;
    }
    return FX_OK;
}

void fx_batch_destroy(fx_batch *B) {
    if (!B)
        return;
    fx_ctx *ctx = B->ctx;
    {
        std::lock_guard<std::mutex> lk(ctx->reg);
        auto it = std::find(ctx->batches.begin(), ctx->batches.end(), B);
        if (it == ctx->batches.end())
            return;  // not a live batch of its context
        ctx->batches.erase(it);
    }
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete B;
}

}  // extern "C"

namespace fx {
namespace {
// interleaved right-hand sides b[i R + r] from rhs-major ones (device)
__global__ void k_interleave(int64_t n, int R, const double *__restrict__ src, double *__restrict__ dst) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n * R)
        dst[q] = src[(q % R) * n + q / R];
}

// A slot whose mesh's element kernels failed (singular Jacobian, no
// material values) stops before its first iteration with that failure,
// reported as status 16 + fx_status.
__global__ void k_slot_elem_check(BSys *__restrict__ S, const DevStatus *__restrict__ st) {
    if (threadIdx.x != 0 || blockIdx.x != 0 || st->code == 0)
        return;
    for (int r = 0; r < S->R; ++r)
        S->rhs[r].status = 16 + st->code;
}
}  // namespace

// fx_batch_add with the right-hand sides on the host (rhs_dev false) or in
// device memory written on the mesh's stream (rhs_dev true), and optionally
// the mesh's element status to check on the device.
fx_status batch_add_impl(fx_batch *B, fx_mesh *m, int32_t n_rhs, const double *rhs, bool rhs_dev,
                         const DevStatus *elem_status, int32_t *slot_out) {
    if (!B || !m || !rhs || !slot_out)
        return FX_ERR_INVALID;
    std::lock_guard<std::mutex> lk(B->mx);
    fx_ctx *ctx = B->ctx;
    int k = -1;  // the slot taken, released again if admission fails
    try {
        if (m->ctx->device != ctx->device)
            throw fx::Error{FX_ERR_INVALID, "fx_batch_add: mesh on another device"};
        if (n_rhs < 1 || n_rhs > kRmax)
            throw fx::Error{FX_ERR_INVALID, "fx_batch_add: 1 to 4 right-hand sides per system"};
        if (!m->k_valid)
            throw fx::Error{FX_ERR_INVALID, "fx_batch_add: assemble the mesh first"};
        if (m->n_dof * n_rhs >= ((int64_t)1 << 31))
            throw fx::Error{FX_ERR_INVALID, "fx_batch_add: n_dof x n_rhs must stay below 2^31"};
        FX_HIP(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        if (m->ctx != ctx) {
            // a mesh of another context on this device, e.g. assembled by
            // another host thread on its own stream while this batch runs:
            // the batch's stream waits for what that stream holds so far,
            // and the mesh-level conversions below run on the batch's stream
            if (!B->adm_ev)
                FX_HIP(hipEventCreateWithFlags(&B->adm_ev, hipEventDisableTiming));
            FX_HIP(hipEventRecord(B->adm_ev, m->ctx->stream));
            FX_HIP(hipStreamWaitEvent(s, B->adm_ev, 0));
        }
        const int G = B->G > 0 ? B->G : (m->spmv_group > 0 ? m->spmv_group : fx::default_group(m));
        k = 0;
        while (k < (int)B->slots.size() && B->slots[k] && B->slots[k]->used)
            ++k;
        if (k == (int)B->slots.size())
            B->slots.emplace_back(new fx_batch::Slot());
        else
            B->slots[k].reset(new fx_batch::Slot());
        auto &sl = *B->slots[k];
        sl.mesh = m;
        sl.R = n_rhs;
        const int64_t n = m->n_dof, nv = n * n_rhs;
        // interleave the right-hand sides: the caller's layout is rhs-major
        if (rhs_dev) {
            sl.b.alloc(ctx, nv);
            hipLaunchKernelGGL(k_interleave, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, s, n, n_rhs, rhs,
                               sl.b.p);
            FX_HIP(hipGetLastError());
            if (m->ctx != ctx) {  // the mesh's stream may overwrite rhs only after this read
                if (!m->adm_rhs_ev)
                    FX_HIP(hipEventCreateWithFlags(&m->adm_rhs_ev, hipEventDisableTiming));
                FX_HIP(hipEventRecord(m->adm_rhs_ev, s));
            }
        } else {
            std::vector<double> inter(nv);
            for (int r = 0; r < n_rhs; ++r)
                for (int64_t i = 0; i < n; ++i)
                    inter[i * n_rhs + r] = rhs[(int64_t)r * n + i];
            sl.b.upload(ctx, inter, s);
        }
        for (auto *v : {&sl.x, &sl.r, &sl.z, &sl.p0, &sl.p1, &sl.ap})
            if (v != &sl.z || !is_poly(B))
                v->alloc(ctx, is_mixed(B) ? (nv + 1) / 2 : nv);  // the mixed solve stores FP32 vectors
        if (is_poly(B)) {
            for (auto *v : {&sl.zq0, &sl.zq1, &sl.zq2, &sl.rs})
                v->alloc(ctx, nv);
            // the solutions go to pinned host memory as the slot finishes
            // (a single solve's batch copies them on the device instead)
            if (!B->owner)
                sl.hx.alloc(ctx, nv);
        }
        if (is_mixed(B)) {
            sl.rin.alloc(ctx, nv);  // replaced residuals: written by each replacement pass before use
            sl.xacc.alloc(ctx, nv);
            fx::launch_fill(m, sl.xacc.p, nv, 0.0, s);
            fx::launch_convert32(m, s);  // K may have been reassembled since the last conversion
        }
        if (is_poly(B)) {
            // the FP64 outer SpMV's sliced-ELL operator, the Jacobi scaling
            // and the inner steps' scaled FP16 operator built from it
            fx::launch_sell(m, false, s);
            fx::launch_jacobi(m, false, s);
            fx::launch_scaled16(m, s);
        } else {
            if (is_mixed(B) && m->sell && m->ssym && use_ssym())
                fx::launch_ssym(m, s);
            else
                fx::launch_sell(m, is_mixed(B), s);
            fx::launch_jacobi(m, false, s);
        }
        sl.G = G;
        // >= the SpMV's row groups (1 partial per load case) and the
        // replacement pass's (3 per load case); polynomial batch: the power
        // iteration's (2 per row group) and the outer SpMV's (5 per load case)
        size_t np1 = (size_t)std::max(fx::nblk(m->n_nodes, kBB / G), 1) * n_rhs * 4;
        if (is_poly(B))
            np1 = std::max<size_t>(2 * (size_t)std::max(fx::nblk(m->n_nodes, m->sell ? kBB : kBB / G), 1),
                                   5 * (size_t)n_rhs * (size_t)std::max(fx::nblk(m->n_nodes, kBB / fx::kOuterG), 1));
        sl.part1.alloc(ctx, np1);
        sl.part2.alloc(ctx, (size_t)std::max(fx::nblk(n, kBB), 1) * n_rhs * 3);
        sl.used = true;
        sl.phase = 0;
        sl.t0 = std::chrono::steady_clock::now();
        if ((int)B->h_sys.size() < (int)B->slots.size())
            B->h_sys.resize(B->slots.size());
        BSys &S = B->h_sys[k];
        S = BSys{};
        fill_sys(B, sl, S);
        start_solve(B, sl, S);
        // running slots' device state is newer than the host copy: upload
        // this entry only
        reserve_sys(B, B->h_sys.size());
        std::memset(&B->fin[k], 0, sizeof(fx::BFin));
        FX_HIP(hipMemcpyAsync(B->d_sys.p + k, &S, sizeof(BSys), hipMemcpyHostToDevice, s));
        if (elem_status) {
            hipLaunchKernelGGL(k_slot_elem_check, dim3(1), dim3(64), 0, s, B->d_sys.p + k, elem_status);
            FX_HIP(hipGetLastError());
        }
        if (is_poly(B))
            sl.pending = true;  // started with the other pending slots by the next fx_batch_run
        else
            start_slots(B, std::vector<int>{k});  // start this slot's solve now
        B->run_version++;
        if (m->ctx != ctx) {
            sl.foreign = true;
            m->foreign_slots++;
        }
        *slot_out = k;
    } catch (const fx::Error &e) {
        release_failed_slot(B, k);
        fx::set_error(e.msg);
        return e.code;
    } catch (const std::exception &e) {
        release_failed_slot(B, k);
        fx::set_error(e.what());
        return FX_ERR_OOM;
    }
    return FX_OK;
}

// fx_solve and fx_static_step in FX_PREC_MIXED_POLY: the system as the one
// slot of a polynomial batch kept with the mesh (created on the mesh's
// context on first use, and again when the options change), b_dev in and
// x_dev out on the mesh's stream; the report is fx_batch_poll's (iterations:
// inner steps; outer_spmv: outer iterations).  The arithmetic is that of
// the same system alone in any polynomial batch with these options.
fx_status poly_solve(fx_mesh *m, const double *b_dev, double *x_dev, const fx_solve_opts &o, fx_solve_report *rep) {
    fx_batch *B = m->poly_batch;
    if (B) {
        const fx_solve_opts &k = B->opts;
        if (k.stop_mode != o.stop_mode || k.tol != o.tol || k.max_iter != o.max_iter ||
            k.poly_degree != o.poly_degree || k.spmv_group != o.spmv_group) {
            fx_batch_destroy(B);  // its destructor clears m->poly_batch
            B = nullptr;
        }
    }
    if (!B) {
        const fx_status st = fx_batch_create(m->ctx, &o, &B);
        if (st != FX_OK)
            return st;  // the message is set
        B->owner = m;
        m->poly_batch = B;
    }
    int32_t slot = -1;
    fx_status st = batch_add_impl(B, m, 1, b_dev, true, nullptr, &slot);
    if (st != FX_OK)
        return st;
    int32_t running = 0;
    st = fx_batch_run(B, 0, &running);
    if (st == FX_OK) {
        int32_t state = 0;
        fx_solve_report r{};
        st = fx_batch_poll(B, slot, &state, &r);
        if (st == FX_OK) {
            if (rep)
                *rep = r;
            const fx_batch::Slot &sl = *B->slots[slot];
            FX_HIP(hipMemcpyAsync(x_dev, sl.x.p, (size_t)m->n_dof * sizeof(double), hipMemcpyDeviceToDevice,
                                  B->ctx->stream));
            st = state == 1 ? FX_OK : (fx_status)(-state);
        }
    }
    (void)fx_batch_remove(B, slot);  // waits for the stream: the copy has completed
    return st;
}

void drop_poly_batch(fx_mesh *m) {
    fx_batch *B = m->poly_batch;
    if (!B)
        return;
    m->poly_batch = nullptr;
    B->owner = nullptr;
    auto &v = m->ctx->batches;
    auto it = std::find(v.begin(), v.end(), B);
    if (it != v.end())
        v.erase(it);
    delete B;
}

}  // namespace fx

namespace {
// x[r*n + i] = h[i*R + r]
void deinterleave(const double *h, int64_t n, int R, double *x) {
    if (R == 1) {
        std::memcpy(x, h, (size_t)n * sizeof(double));
        return;
    }
    for (int64_t i = 0; i < n; ++i)
        for (int r = 0; r < R; ++r)
            x[(int64_t)r * n + i] = h[i * R + r];
}
}  // namespace

extern "C" {

fx_status fx_batch_add(fx_batch *B, fx_mesh *m, int32_t n_rhs, const double *rhs, int32_t *slot_out) {
    return fx::batch_add_impl(B, m, n_rhs, rhs, false, nullptr, slot_out);
}

fx_status fx_batch_run(fx_batch *B, int64_t max_iters, int32_t *n_running_out) {
    if (!B)
        return FX_ERR_INVALID;
    std::lock_guard<std::mutex> lk(B->mx);
    fx_ctx *ctx = B->ctx;
    try {
        FX_HIP(hipSetDevice(ctx->device));
        if (is_poly(B))
            start_pending_poly(B);
        int64_t done_iters = 0;
        int processed = 0;
        for (;;) {
            // keep kSets chunks in flight while slots run and the budget lasts
            int running = 0;
            for (auto &sl : B->slots)
                running += sl && sl->used && sl->phase == 0;
            const bool budget = max_iters <= 0 || done_iters < max_iters;
            if (B->inflight < fx_batch::kSets && running > 0 && budget) {
                const int set = (B->head + B->inflight) % fx_batch::kSets;
                if (B->cs[set].version != B->run_version)
                    build_tables(B, set);
                // a chunk: 16 iterations, or (polynomial batch) the outer
                // iterations holding about 16 FP32 inner steps, at least one
                const int per = is_poly(B) ? B->poly_deg : 1;
                int chunk = std::max(1, 16 / per);
                if (max_iters > 0)
                    chunk = (int)std::min<int64_t>(chunk, std::max<int64_t>(1, (max_iters - done_iters) / per));
                launch_iterations(B, chunk, set);
                done_iters += (int64_t)chunk * per;
                ++B->inflight;
                continue;
            }
            // overlap: a budgeted call returns once the budget is launched
            // and at least one chunk has been processed; the rest stays in
            // flight
            if (B->inflight == 0 || (B->overlap && !budget && processed > 0))
                break;
            FX_HIP(hipEventSynchronize(B->cs[B->head].done));
            advance_phases(B, B->head);
            B->head = (B->head + 1) % fx_batch::kSets;
            --B->inflight;
            ++processed;
        }
        if (n_running_out) {
            int c = 0;
            for (auto &sl : B->slots)
                c += sl && sl->used && sl->phase == 0;
            *n_running_out = c;
        }
    } catch (const fx::Error &e) {
        fx::set_error(e.msg);
        return e.code;
    } catch (const std::exception &e) {
        fx::set_error(e.what());
        return FX_ERR_OOM;
    }
    return FX_OK;
}

fx_status fx_batch_poll(fx_batch *B, int32_t slot, int32_t *state, fx_solve_report *rep) {
    if (!B)
        return FX_ERR_INVALID;
    std::lock_guard<std::mutex> lk(B->mx);  // the slot table may change under a foreign mesh's removal
    if (!B || slot < 0 || slot >= (int)B->slots.size() || !B->slots[slot] || !B->slots[slot]->used)
        return FX_ERR_INVALID;
    auto &sl = *B->slots[slot];
    if (state)
        *state = sl.phase == 2 ? (sl.err == FX_OK ? 1 : -(int)sl.err) : 0;
    if (rep) {
        *rep = fx_solve_report{};
        rep->iterations = sl.iters;
        rep->outer_spmv = sl.outer;
        rep->refinements = (int32_t)sl.outer;
        rep->spmv_group = sl.G;
        double c = 0.0;
        for (int r = 0; r < sl.R; ++r)
            c = std::max(c, sl.crit[r]);
        rep->final_criterion = c;
        rep->seconds = sl.seconds;
    }
    return FX_OK;
}

fx_status fx_batch_finished(fx_batch *B, int32_t *slots, int32_t cap, int32_t *n) {
    if (!B || !n || cap < 0 || (cap > 0 && !slots))
        return FX_ERR_INVALID;
    std::lock_guard<std::mutex> lk(B->mx);
    int32_t c = 0;
    for (size_t k = 0; k < B->slots.size(); ++k) {
        const auto &sl = B->slots[k];
        if (sl && sl->used && sl->phase == 2) {
            if (c < cap)
                slots[c] = (int32_t)k;
            ++c;
        }
    }
    *n = c;
    return FX_OK;
}

fx_status fx_batch_get(fx_batch *B, int32_t slot, double *x) {
    if (!B)
        return FX_ERR_INVALID;
    std::lock_guard<std::mutex> lk(B->mx);  // the slot table may change under a foreign mesh's removal
    if (!B || !x || slot < 0 || slot >= (int)B->slots.size() || !B->slots[slot] || !B->slots[slot]->used)
        return FX_ERR_INVALID;
    auto &sl = *B->slots[slot];
    try {
        const int64_t n = sl.mesh->n_dof, nv = n * sl.R;
        const double *src = is_mixed(B) ? sl.xacc.p : sl.x.p;
        // a finished slot's solutions are final (the kernel that wrote its
        // summary ran after the last update): copied on the context's copy
        // stream, past the chunks in flight; a running slot's current
        // iterate: after the work queued so far
        const int R = sl.R;
        if (sl.phase == 2 && sl.hx.p) {
            // exported by the device before it wrote the finish summary
            deinterleave(sl.hx.p, n, R, x);
            return FX_OK;
        }
        fx_ctx *ctx = B->ctx;
        FX_HIP(hipSetDevice(ctx->device));
        std::call_once(ctx->copy_once, [ctx] {
            if (hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking) != hipSuccess)
                ctx->copy_stream = nullptr;
        });
        hipStream_t s = sl.phase == 2 && ctx->copy_stream ? ctx->copy_stream : ctx->stream;
        // one copy command into pinned host memory (the DMA engine: no
        // workgroups queued behind the chunks in flight; a pageable copy
        // would be staged in pieces), then the host writes x rhs-major
        if (B->h_get_cap < (size_t)nv) {
            if (B->h_get)
                FX_HIP(hipHostFree(B->h_get));
            B->h_get = nullptr;
            B->h_get_cap = 0;
            FX_HIP(hipHostMalloc((void **)&B->h_get, (size_t)nv * sizeof(double), hipHostMallocDefault));
            B->h_get_cap = (size_t)nv;
        }
        FX_HIP(hipMemcpyAsync(B->h_get, src, (size_t)nv * sizeof(double), hipMemcpyDeviceToHost, s));
        FX_HIP(hipStreamSynchronize(s));
        deinterleave(B->h_get, n, R, x);
    } catch (const fx::Error &e) {
        fx::set_error(e.msg);
        return e.code;
    }
    return FX_OK;
}

fx_status fx_batch_set_overlap(fx_batch *B, int32_t on) {
    if (!B)
        return FX_ERR_INVALID;
    std::lock_guard<std::mutex> lk(B->mx);
    if (!on) {
        try {
            FX_HIP(hipSetDevice(B->ctx->device));
            drain(B);
        } catch (const fx::Error &e) {
            fx::set_error(e.msg);
            return e.code;
        }
    }
    B->overlap = on != 0;
    return FX_OK;
}

fx_status fx_batch_set_inner_shape(fx_batch *B, int32_t shape) {
    if (!B)
        return FX_ERR_INVALID;
    std::lock_guard<std::mutex> lk(B->mx);
    if (shape != 0 && shape != 1) {
        fx::set_error("fx_batch_set_inner_shape: shape must be 0 or 1");
        return FX_ERR_INVALID;
    }
    B->inner_split = shape == 1;
    return FX_OK;
}

fx_status fx_batch_autotune(fx_batch *B, int32_t reps, int32_t *shape_out, double *ms_one, double *ms_split) {
    if (!B)
        return FX_ERR_INVALID;
    std::lock_guard<std::mutex> lk(B->mx);
    try {
        if (!is_poly(B) || reps <= 0)
            throw fx::Error{FX_ERR_INVALID, "fx_batch_autotune: needs a polynomial batch (FX_PREC_MIXED_POLY) and reps > 0"};
        FX_HIP(hipSetDevice(B->ctx->device));
        hipStream_t s = B->ctx->stream;
        drain(B);
        FX_HIP(hipStreamSynchronize(s));
        // the running systems' tables (no chunk is in flight now)
        fx_batch::ChunkSet &C = B->cs[0];
        if (C.version != B->run_version)
            build_tables(B, 0);
        double t1 = 0.0, t2 = 0.0;
        if (C.n1s > 0) {
            t1 = fx::time_poly_inner(s, B->d_sys.p, C.tab1s.p, C.n1s, C.n1s_r, false, C.inner_rows, reps);
            t2 = fx::time_poly_inner(s, B->d_sys.p, C.tab1s.p, C.n1s, C.n1s_r, true, C.inner_rows, reps);
            B->inner_split = t2 < t1;
        }
        if (shape_out)
            *shape_out = B->inner_split ? 1 : 0;
        if (ms_one)
            *ms_one = t1;
        if (ms_split)
            *ms_split = t2;
    } catch (const fx::Error &e) {
        fx::set_error(e.msg);
        return e.code;
    }
    return FX_OK;
}

fx_status fx_batch_time_spmv(fx_batch *B, int32_t on) {
    if (!B)
        return FX_ERR_INVALID;
    std::lock_guard<std::mutex> lk(B->mx);
    B->timing = on != 0;
    return FX_OK;
}

fx_status fx_batch_spmv_stats(fx_batch *B, int64_t *launches, int64_t *timed, double *ms_timed,
                              double *bytes_total) {
    if (!B)
        return FX_ERR_INVALID;
    std::lock_guard<std::mutex> lk(B->mx);
    double by = 0.0;
    try {
        drain(B);  // the timing samples of the chunks in flight
        FX_HIP(hipMemcpyAsync(&by, B->d_bytes.p, sizeof(double), hipMemcpyDeviceToHost, B->ctx->stream));
        FX_HIP(hipStreamSynchronize(B->ctx->stream));
    } catch (const fx::Error &e) {
        fx::set_error(e.msg);
        return e.code;
    }
    if (launches)
        *launches = B->spmv_launches;
    if (timed)
        *timed = B->timed;
    if (ms_timed)
        *ms_timed = B->timed_ms;
    if (bytes_total)
        *bytes_total = by;
    return FX_OK;
}

fx_status fx_batch_remove(fx_batch *B, int32_t slot) {
    if (!B)
        return FX_ERR_INVALID;
    std::lock_guard<std::mutex> lk(B->mx);  // the slot table may change under a foreign mesh's removal
    if (!B || slot < 0 || slot >= (int)B->slots.size() || !B->slots[slot] || !B->slots[slot]->used)
        return FX_ERR_INVALID;
    // a running slot: the chunks in flight still iterate it (and may write
    // its finish summary, which a later slot here starts from zero); a
    // finished one: they do not touch it
    if (B->slots[slot]->phase == 0) {
        try {
            drain(B);
        } catch (const fx::Error &e) {
            fx::set_error(e.msg);
            return e.code;
        }
    }
    if (B->inflight == 0)
        (void)hipStreamSynchronize(B->ctx->stream);
    const bool was_running = B->slots[slot]->phase == 0;
    if (slot < (int)B->h_sys.size())
        B->h_sys[slot] = BSys{};
    if (was_running)
        B->run_version++;
    B->slots[slot].reset(new fx_batch::Slot());
    return FX_OK;
}

}  // extern "C"
