// fx_reduce.h -- deterministic block and grid reductions (device helpers).
//
// Every sum has a fixed association independent of scheduling: lanes ->
// wave (butterfly shuffles) -> waves of a block in index order -> block
// partials, and the last block to finish (atomic ticket) sums the partials
// with the same fixed tree.  No floating-point atomics, so results are
// bitwise repeatable run to run.
#pragma once

#include "fx_internal.h"

namespace fx {

constexpr int kRedBlock = 256;

// Sum of v over the block; valid in thread 0.  All threads must call it.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
    __shared__ double s[kRedBlock / 64][NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double t = v[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
            t += __shfl_xor(t, off, 64);
        v[k] = t;
    }
    const int w = threadIdx.x / 64;
    __syncthreads();  // s[] may still be read by a previous call
    if (threadIdx.x % 64 == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k)
            s[w][k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double t = 0.0;
#pragma unroll
            for (int j = 0; j < kRedBlock / 64; ++j)
                t += s[j][k];
            v[k] = t;
        }
}

// Block partial of v -> part[NV*blockIdx + k]; then the last block to arrive
// sums all partials (all its threads take part) into out (valid in thread 0
// of the last block, the only thread for which it returns true) and resets
// the ticket counter.  All threads must call it.
template <int NV>
__device__ __forceinline__ bool grid_sum(double (&v)[NV], double *part, uint32_t *cnt, double (&out)[NV]) {
    __shared__ bool last;
    block_sum<NV>(v);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k)
            part[NV * blockIdx.x + k] = v[k];
        __threadfence();
        last = atomicAdd(cnt, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last)
        return false;
    __threadfence();
#pragma unroll
    for (int k = 0; k < NV; ++k)
        out[k] = 0.0;
    for (unsigned b = threadIdx.x; b < gridDim.x; b += blockDim.x)
#pragma unroll
        for (int k = 0; k < NV; ++k)
            out[k] += __hip_atomic_load(&part[NV * b + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    block_sum<NV>(out);
    if (threadIdx.x == 0)
        *cnt = 0;
    return threadIdx.x == 0;
}

}  // namespace fx
