// fx_sell.h -- sliced-ELL node-block operator for meshes with short node rows.
//
// Node rows are grouped in slices of 64 (one wavefront).  Within a slice the
// 3x3 blocks of block column k of all 64 rows are stored together, entry
// e = 3p+q of the block at sval[((kk*9 + e) * 64) + lane] with kk =
// sl_ptr[slice] + k, and the block's column node at scol[kk*64 + lane].  One
// thread owns one node row: every load of a block column is one coalesced
// 64-lane access, and the gathered operand entries of a block are 3
// consecutive DOFs.  Rows shorter than their slice's width are padded; the
// padding is never read (each thread stops at its own row length).
//
// The accumulation order of a DOF row is block column k ascending (the
// node-block pattern's sorted columns), q = 0, 1, 2 within a block, one
// multiply-add at a time: the single-system and the batched solves share
// this function and so produce bitwise equal results.
#pragma once

#include "fx_internal.h"

namespace fx {

constexpr int kSlice = kSliceRows;

// Pointers the compiler cannot prove global (loaded from a descriptor in
// memory) are cast to the global address space so that loads and stores are
// global_* instructions with their own wait counters rather than flat_*.
template <typename T>
using gptr = __attribute__((address_space(1))) T *;
template <typename T>
__device__ __forceinline__ gptr<T> gp(T *p) {
    return (gptr<T>)p;
}

// p_new = z + beta p_old at one vector entry, rounded to the stored type;
// the mixed-precision solve scales z by gamma (1 except in the iteration
// after a residual replacement, where it folds in the re-orthogonalisation
// of the kept direction).  FP64 vectors: the plain update.
template <typename XT>
__device__ __forceinline__ double pnew_at(gptr<const XT> z, gptr<const XT> pold, int64_t at, double gam,
                                          double beta) {
    if constexpr (sizeof(XT) == 4)
        return (double)(XT)(gam * (double)z[at] + beta * (double)pold[at]);
    else
        return (double)(XT)((double)z[at] + beta * (double)pold[at]);
}

// acc[p][r] += sum over the row's blocks of K(a,p; c,q) * p_new(c,q,r)
// for r < RR right-hand sides (vectors interleaved: v[dof*RR + r], stored as
// XT: double, or float in the mixed-precision inner solve).
template <typename VT, typename XT, int RR>
__device__ __forceinline__ void sell_node_rows(gptr<const int32_t> nb_ptr, gptr<const int32_t> sl_ptr,
                                               gptr<const int32_t> scol, gptr<const VT> sval, int64_t a,
                                               gptr<const XT> z, gptr<const XT> pold, const double (&gam)[RR],
                                               const double (&beta)[RR], double (&acc)[3][RR]) {
    const int64_t s = a / kSlice;
    const int lane = (int)(a % kSlice);
    const int k0 = sl_ptr[s];
    const int k1 = k0 + (nb_ptr[a + 1] - nb_ptr[a]);
    for (int k = k0; k < k1; k += 2) {
        int c[2];
        double v[2][9];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int kk = k + u;
            c[u] = kk < k1 ? scol[(size_t)kk * kSlice + lane] : -1;
            if (kk < k1)
#pragma unroll
                for (int e = 0; e < 9; ++e)
                    v[u][e] = (double)sval[((size_t)kk * 9 + e) * kSlice + lane];
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            if (c[u] < 0)
                continue;
            double x[3][RR];
#pragma unroll
            for (int q = 0; q < 3; ++q)
#pragma unroll
                for (int r = 0; r < RR; ++r) {
                    const int64_t at = (3 * (int64_t)c[u] + q) * RR + r;
                    x[q][r] = pnew_at<XT>(z, pold, at, gam[r], beta[r]);  // as its owner stores it
                }
#pragma unroll
            for (int r = 0; r < RR; ++r)
#pragma unroll
                for (int p = 0; p < 3; ++p) {
                    acc[p][r] += v[u][3 * p] * x[0][r];
                    acc[p][r] += v[u][3 * p + 1] * x[1][r];
                    acc[p][r] += v[u][3 * p + 2] * x[2][r];
                }
        }
    }
}

}  // namespace fx
