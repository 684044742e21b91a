// fx_batch.h -- what the batch kernels share: the per-slot system record the
// device reads (BSys), the per-right-hand-side solver state (BRhs), the
// finish summary the host polls (BFin), and the fixed-order partial-sum and
// state-staging helpers.  Used by fx_batch.hip (Jacobi-PCG batches and the
// host engine) and fx_batch_poly.hip (the polynomial-preconditioned mixed
// batch).
#pragma once

#include "fx_reduce.h"
#include "fx_sell.h"

namespace fx {

constexpr int kRmax = 4;
constexpr int kBB = 256;  // threads per block in the batch kernels
constexpr int kPowerSteps = 20;   // power steps of the lambda_max estimate (polynomial batch)
constexpr int kOuterG = 8;        // lanes per node row of the polynomial batch's FP64 outer SpMV
// node rows per workgroup of the polynomial batch's sliced-ELL Chebyshev
// step: eight 64-row slices (their overlapping column ranges share L1) when
// the running systems hold at least kInnerRowsLargeMin node rows, else kBB
// (four slices: twice the workgroups for a small batch).  Measured on the
// 512-mesh population (7.7 M node rows): 256 / 512 / 1024 rows ran 546 /
// 538 / 554 us per launch; on the serving stream (up to about 0.6 M node
// rows running) 256 rows kept the lower latencies.
constexpr int kInnerRowsLarge = 512;
constexpr int64_t kInnerRowsLargeMin = (int64_t)2 << 20;
// DOF tiles of kBB per block of the vector kernels (kb_init, kb_update):
// four with FP32 vectors, so a block's fixed costs (descriptor reads, the
// block reduction) are spread over enough bytes; one with FP64 vectors,
// whose partial sums then associate exactly as fx_pcg.hip's.
template <typename XT>
constexpr int kTiles = sizeof(XT) == 4 ? 4 : 1;

struct BRhs {
    double rz[2], rr, bb, crit, pap, alpha, beta, tol;
    double gamma;  // p_new = gamma z + beta p_old (1 except right after a replacement)
    double rmax;   // mixed: true residual norm at the last replacement
    double delta;  // mixed: replace when the iterated residual norm < delta * rmax (0: FP64 slot)
    double prev;   // mixed: true criterion at the last replacement its own residual asked for (floor test)
    int64_t it, it_k1, max_iter;
    int32_t status;  // 0 running, 1 converged, 2 breakdown, 3 max_iter / replacement limit, 4 idle,
                     // 5 residual replacement in this iteration
    int32_t stop_mode;
    int32_t nupd, max_upd;  // mixed: replacements done, allowed
};

// What the host needs of a slot once none of its load cases runs, written
// by the GPU into pinned host memory (no copy command at chunk ends).
struct BFin {
    int64_t it, nupd;
    double crit[kRmax];
    int32_t status[kRmax];
    int32_t done;  // written last
    int32_t pad;
};

struct BSys {
    const int32_t *nb_ptr, *nb_col;
    const void *val;                    // double (FP64 slots) or float (mixed inner)
    const double *val64, *inv;
    const float *kc;                    // FP32 operator: per-node 3x3 row-sum corrections
    const int32_t *sl_ptr, *scol;       // sliced-ELL operator (fx_sell.h) when sell != 0
    const void *sval;
    int32_t sell;
    int32_t ssym;                       // mixed sliced-ELL slot on the symmetric FP32 operator
    const int32_t *su_ptr, *scol_u, *sll_ptr, *lrec, *lcol;
    const float *uval;
    double *b;                          // FP64: right-hand sides; mixed: the last replaced residual b - K64 xacc
    const double *btrue;                // right-hand sides
    void *x, *r, *z, *p0, *p1, *ap;    // PCG vectors, XT (double, or float in the mixed solve)
    double *xacc;                       // mixed: FP64 solution; x holds the increment since the last replacement
    double *part1, *part2;
    int64_t n_nodes, n_dof;
    int64_t it;                         // iteration index of the running right-hand sides
    int32_t R, nparts1, nparts2;
    int32_t G;                          // lanes per node row of the lane-group SpMV of this system
    double *bytes_acc;                  // instrumentation: algorithmic bytes of the SpMV launches issued
    double spmv_bytes0, spmv_bytes_rhs; // per launch: operator part, per active right-hand side
    double upd_bytes0, upd_bytes_rhs;   // the same for a residual replacement pass
    int32_t go;                         // right-hand sides the next SpMV / update work on (bit mask)
    int32_t upd;                        // mixed: right-hand sides whose residual is replaced in this iteration
    int32_t trig;                       // mixed: those of them whose own residual asked for it
    int32_t reported;                   // the slot's BFin has been written
    // polynomial-preconditioned mixed batch (FX_PREC_MIXED_POLY): flexible
    // CG in FP64 on x, r (FP64 in x and r), the direction pp, ps = K pp and
    // pw = K z; z from `deg` Chebyshev steps on the Jacobi-scaled FP16
    // operator (val16: node blocks, sval16: sliced ELL) with FP64 iterates:
    // three rotating iterates zq and the scaled residual rs
    int32_t poly, deg;
    double *pp, *ps, *pw;
    double *zq[3], *rs;
    const _Float16 *val16, *sval16, *sval16b;  // sliced ELL: 8-entry records, then the 9th entries
    const double *sval64;               // sliced-ELL FP64 operator (sell slots): the outer SpMV's
    int32_t nparts3;                    // row groups of the FP64 outer SpMV (sliced ELL: kBB node rows;
                                        // node blocks: kBB / kOuterG)
    int32_t pad3;
    // result export (polynomial batch, when hx is set): the outer iteration
    // in which the slot stops sets exp; the next one's kp_upd copies x into
    // the pinned host array hx and its kp_stop writes the finish summary, so
    // the host reads the solutions straight from hx once it sees the summary
    double *hx;
    int32_t exp, pad4;
    double lam;                         // Rayleigh-quotient estimate of lambda_max(D^-1 K32)
    double cheb[2 * kPolyMax + 2];      // [0] = 1/theta; per inner step j >= 1: c1_j at 2j, c2_j at 2j+1
    double cheb_bytes0, cheb_bytes_rhs; // algorithmic bytes of one inner step: operator part, per load case
    BRhs rhs[kRmax];
};

// polynomial batch launchers (fx_batch_poly.hip); tables as the Jacobi-PCG
// batch's: (slot, row group) pairs for the SpMV-shaped kernels, (slot, DOF
// tile) pairs for the vector kernels, slot lists for the one-block kernels
void launch_poly_power_start(hipStream_t s, BSys *sys, const int2 *tab_vec, int n_vec, const int2 *tab_in_sell,
                             int n_in_sell, const int2 *const *tab_in_g, const int *n_in_g, const int32_t *list,
                             int n_list, BFin *fin);
void launch_poly_outer(hipStream_t s, BSys *sys, int deg, const int2 *tab_vec, int n_vec, const int2 *tab_in_sell,
                       int n_in_sell, const int *n_in_sell_r, bool split, int inner_rows, const int2 *const *tab_in_g,
                       const int *n_in_g, const int2 *tab_out_sell, int n_out_sell, const int2 *tab_out, int n_out,
                       const int32_t *list, int n_list, BFin *fin, hipEvent_t ev0, hipEvent_t ev1);
double time_poly_inner(hipStream_t s, BSys *sys, const int2 *tab, int n, const int *n_r, bool split, int rows,
                       int reps);

inline int nblk(int64_t n, int per) { return (int)((n + per - 1) / per); }

// sums nv <= NV values over nparts blocks (layout part[b*nv + k]); valid in
// thread 0; fixed order.  A row's NV loads are issued together: out[k] for
// k >= nv (never read) sums the row's first value again, an in-bounds load
// with no branch around it.
template <int NV>
__device__ __forceinline__ void bsum_partials(const double *part, int nparts, int nv, double (&out)[NV]) {
#pragma unroll
    for (int k = 0; k < NV; ++k)
        out[k] = 0.0;
    gptr<const double> pp = gp(part);
    for (int b = threadIdx.x; b < nparts; b += kBB) {
        double v[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k)
            v[k] = pp[(size_t)b * nv + (k < nv ? k : 0)];
#pragma unroll
        for (int k = 0; k < NV; ++k)
            out[k] += v[k];
    }
    block_sum<NV>(out);
}

// The launch's view of a system's right-hand side states, copied to LDS by
// one load per thread: the one-block reduction kernels take their decisions'
// inputs from it instead of waiting on a load per right-hand side, and write
// the new states to the global record (a field written there is not read
// back in the same launch).
__device__ __forceinline__ void stage_rhs(const BSys &S, BRhs (&H)[kRmax]) {
    constexpr int kWords = kRmax * (int)sizeof(BRhs) / 4;
    static_assert(sizeof(BRhs) % 4 == 0 && kWords <= kBB, "one word per thread");
    if (threadIdx.x < kWords)
        ((int32_t *)H)[threadIdx.x] = gp((const int32_t *)S.rhs)[threadIdx.x];
    __syncthreads();
}

}  // namespace fx
