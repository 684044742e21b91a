// fx_sell.h -- sliced-ELL node-block operator for meshes with short node rows.
//
// Node rows are grouped in slices of 64 (one wavefront).  Within a slice the
// 3x3 blocks of block column k of all 64 rows are stored together (kk =
// sl_ptr[slice] + k), 576 values per block column, and the block's column
// node at scol[kk*64 + lane].  FP32 values: entry e = 3p+q of the block at
// sval[(kk*9 + e) * 64 + lane].  FP64 values: entries 0..7 as four 16-byte
// pairs per lane, pair m at sval[kk*576 + m*128 + lane*2], entry 8 at
// sval[kk*576 + 512 + lane] (sell_index): five loads per block instead of
// nine.  One thread owns one node row: every load of a block column is one
// coalesced 64-lane access, and the gathered operand entries of a block are
// 3 consecutive DOFs.  Rows shorter than their slice's width are padded with
// zero blocks whose column is the row's own node (node 0 for the lanes past
// the last row of the mesh).
//
// The accumulation order of a DOF row is block column k ascending (the
// node-block pattern's sorted columns), q = 0, 1, 2 within a block, one
// multiply-add at a time.
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

// acc[p][r] += sum over the row's blocks of K(a,p; c,q) * p(c,q,r)
// for r < RR right-hand sides (vectors interleaved: v[dof*RR + r], stored as
// XT: double, or float in the mixed-precision solve).
//
// The loop runs over the slice's width, which is the same for the 64 rows
// of a slice, so one wavefront has one trip count (scalar loop control, no
// per-lane branches).  A row's padding blocks past its own length have zero
// values and point at the row's own node, so they add exact zeros.  Column
// indices are loaded one step ahead: a step issues the operand gathers and
// the value loads of two blocks together, behind one memory latency instead
// of two.
template <typename VT, typename XT, int RR>
__device__ __forceinline__ void gather_block(gptr<const XT> pv, int c, double (&x)[3][RR]) {
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int r = 0; r < RR; ++r)
            x[q][r] = (double)pv[(3 * (int64_t)c + q) * RR + r];
}

template <typename VT, int RR>
__device__ __forceinline__ void madd_block(const VT (&v)[9], const double (&x)[3][RR], double (&acc)[3][RR]) {
#pragma unroll
    for (int r = 0; r < RR; ++r)
#pragma unroll
        for (int p = 0; p < 3; ++p) {
            acc[p][r] += (double)v[3 * p] * x[0][r];
            acc[p][r] += (double)v[3 * p + 1] * x[1][r];
            acc[p][r] += (double)v[3 * p + 2] * x[2][r];
        }
}

// Position of entry e of block column kk for row `lane` of its slice.
template <typename VT>
__host__ __device__ __forceinline__ size_t sell_index(size_t kk, int e, int lane) {
    if constexpr (sizeof(VT) == 8)
        return e < 8 ? kk * 576 + (size_t)(e / 2) * 128 + (size_t)lane * 2 + (e & 1) : kk * 576 + 512 + lane;
    else
        return (kk * 9 + e) * kSlice + lane;
}

typedef double fx_d2 __attribute__((ext_vector_type(2)));

// The nine values of block column kk of row `lane`.
template <typename VT>
__device__ __forceinline__ void load_sell_block(gptr<const VT> sval, size_t kk, int lane, VT (&v)[9]) {
    if constexpr (sizeof(VT) == 8) {
        gptr<const fx_d2> rec = (gptr<const fx_d2>)(sval + kk * 576);
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const fx_d2 w = rec[m * kSlice + lane];
            v[2 * m] = w.x;
            v[2 * m + 1] = w.y;
        }
        v[8] = sval[kk * 576 + 512 + lane];
    } else {
#pragma unroll
        for (int e = 0; e < 9; ++e)
            v[e] = sval[(kk * 9 + e) * kSlice + lane];
    }
}

template <typename VT, typename XT, int RR>
__device__ __forceinline__ void sell_node_rows(gptr<const int32_t> sl_ptr, gptr<const int32_t> scol,
                                               gptr<const VT> sval, int64_t a, gptr<const XT> pv,
                                               double (&acc)[3][RR]) {
    const int64_t s = a / kSlice;
    const int lane = (int)(a % kSlice);
    // every active lane of the wavefront belongs to slice s
    const int k0 = __builtin_amdgcn_readfirstlane(sl_ptr[s]);
    const int kw = __builtin_amdgcn_readfirstlane(sl_ptr[s + 1]);
    int cn0 = scol[(size_t)k0 * kSlice + lane];
    int cn1 = k0 + 1 < kw ? scol[(size_t)(k0 + 1) * kSlice + lane] : 0;
    for (int k = k0; k < kw; k += 2) {
        const bool two = k + 1 < kw;
        const int c0 = cn0, c1 = cn1;
        double x0[3][RR], x1[3][RR];
        VT v0[9], v1[9];
        gather_block<VT, XT, RR>(pv, c0, x0);
        load_sell_block<VT>(sval, (size_t)k, lane, v0);
        if (two) {
            gather_block<VT, XT, RR>(pv, c1, x1);
            load_sell_block<VT>(sval, (size_t)k + 1, lane, v1);
        }
        if (k + 2 < kw)
            cn0 = scol[(size_t)(k + 2) * kSlice + lane];
        if (k + 3 < kw)
            cn1 = scol[(size_t)(k + 3) * kSlice + lane];
        madd_block<VT, RR>(v0, x0, acc);
        if (two)
            madd_block<VT, RR>(v1, x1, acc);
    }
}

// acc[p][r] += K(a,p; c,q) p(c,q,r) over the row's lower blocks (c < a) of
// the symmetric operator: K(a, c) = K(c, a)^T, read from the record the
// owner row c stores (lrec: its offset in the sval layout, entry e at
// +e*64), in ascending c.  Padding points at the zero record.
template <typename XT, int RR>
__device__ __forceinline__ void ssym_lower_rows(gptr<const int32_t> sll_ptr, gptr<const int32_t> lrec,
                                                gptr<const int32_t> lcol, gptr<const float> uval, int64_t a,
                                                gptr<const XT> pv, double (&acc)[3][RR]) {
    const int64_t s = a / kSlice;
    const int lane = (int)(a % kSlice);
    const int k0 = __builtin_amdgcn_readfirstlane(sll_ptr[s]);
    const int kw = __builtin_amdgcn_readfirstlane(sll_ptr[s + 1]);
    if (k0 >= kw)
        return;  // no lower blocks in this slice
    // record offsets and columns one step ahead, as in sell_node_rows; an
    // index past the slice's width reloads its last entry (unused)
    auto at = [&](int k) { return (size_t)(k < kw ? k : kw - 1) * kSlice + lane; };
    int rn0 = lrec[at(k0)], cn0 = lcol[at(k0)];
    int rn1 = lrec[at(k0 + 1)], cn1 = lcol[at(k0 + 1)];
    for (int k = k0; k < kw; k += 2) {
        const bool two = k + 1 < kw;
        const int r0 = rn0, c0 = cn0, r1 = rn1, c1 = cn1;
        double x0[3][RR], x1[3][RR];
        float v0[9], v1[9];
        gather_block<float, XT, RR>(pv, c0, x0);
#pragma unroll
        for (int e = 0; e < 9; ++e)
            v0[e] = uval[r0 + e * kSlice];
        if (two) {
            gather_block<float, XT, RR>(pv, c1, x1);
#pragma unroll
            for (int e = 0; e < 9; ++e)
                v1[e] = uval[r1 + e * kSlice];
        }
        rn0 = lrec[at(k + 2)];
        cn0 = lcol[at(k + 2)];
        rn1 = lrec[at(k + 3)];
        cn1 = lcol[at(k + 3)];
        // K(a,c)(p,q) = K(c,a)(q,p) = v[3q + p]
#pragma unroll
        for (int r = 0; r < RR; ++r)
#pragma unroll
            for (int p = 0; p < 3; ++p) {
                acc[p][r] += (double)v0[p] * x0[0][r];
                acc[p][r] += (double)v0[3 + p] * x0[1][r];
                acc[p][r] += (double)v0[6 + p] * x0[2][r];
            }
        if (two)
#pragma unroll
            for (int r = 0; r < RR; ++r)
#pragma unroll
                for (int p = 0; p < 3; ++p) {
                    acc[p][r] += (double)v1[p] * x1[0][r];
                    acc[p][r] += (double)v1[3 + p] * x1[1][r];
                    acc[p][r] += (double)v1[6 + p] * x1[2][r];
                }
    }
}

}  // namespace fx
