// fx_peak.hip -- FP64 throughput microbenchmarks (instrumentation, no
// counterpart in the reference): the measured ceilings that the element
// kernels' FLOP rates are stated against, beside the datasheet figure.
//
// k_peak_valu: every lane runs 8 independent chains x = fma(x, a, b)
// (v_fma_f64); 2 FLOP per lane and step.
// k_peak_mfma: every wavefront runs 8 independent accumulators through
// v_mfma_f64_16x16x4_f64 (C[16x16] += A[16x4] B[4x16]); 2*16*16*4 FLOP per
// instruction.  Operands come from registers, so both measure issue rate,
// not memory.  Each thread stores one sum so that nothing is dead code.
#include "fx_internal.h"

namespace fx {
namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int kPeakBlock = 256;
constexpr int kPeakSteps = 2048;
constexpr int kChains = 8;

__global__ __launch_bounds__(kPeakBlock) void k_peak_valu(double *__restrict__ out, double a, double b) {
    double x[kChains];
#pragma unroll
    for (int i = 0; i < kChains; ++i)
        x[i] = 1e-3 * (double)(threadIdx.x + i);
    for (int it = 0; it < kPeakSteps; ++it)
#pragma unroll
        for (int i = 0; i < kChains; ++i)
            x[i] = fma(x[i], a, b);
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < kChains; ++i)
        s += x[i];
    out[(size_t)blockIdx.x * kPeakBlock + threadIdx.x] = s;
}

__global__ __launch_bounds__(kPeakBlock) void k_peak_mfma(double *__restrict__ out, double a, double b) {
    v4d c[kChains];
#pragma unroll
    for (int i = 0; i < kChains; ++i)
        c[i] = v4d{0.0, 0.0, 0.0, 0.0};
    const double av = a + 1e-12 * (double)threadIdx.x;
    for (int it = 0; it < kPeakSteps; ++it)
#pragma unroll
        for (int i = 0; i < kChains; ++i)
            c[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, b, c[i], 0, 0, 0);
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < kChains; ++i)
        s += c[i][0] + c[i][1] + c[i][2] + c[i][3];
    out[(size_t)blockIdx.x * kPeakBlock + threadIdx.x] = s;
}

}  // namespace

// TFLOP/s of `reps` timed launches (after one warm-up) of each kernel over
// 8 blocks per compute unit.
void measure_fp64_peak(fx_ctx *ctx, int reps, double *valu_tflops, double *mfma_tflops) {
    hipStream_t s = ctx->stream;
    const unsigned nb = (unsigned)std::max(1, ctx->n_cu) * 8u;
    DBuf<double> out;
    out.alloc((size_t)nb * kPeakBlock);
    hipEvent_t e0, e1;
    FX_HIP(hipEventCreate(&e0));
    FX_HIP(hipEventCreate(&e1));
    auto time_kernel = [&](bool mfma) {
        for (int k = 0; k <= reps; ++k) {
            if (k == 1)
                FX_HIP(hipEventRecord(e0, s));
            if (mfma)
                hipLaunchKernelGGL(k_peak_mfma, dim3(nb), dim3(kPeakBlock), 0, s, out.p, 0.999999, 1e-6);
            else
                hipLaunchKernelGGL(k_peak_valu, dim3(nb), dim3(kPeakBlock), 0, s, out.p, 0.999999, 1e-6);
        }
        FX_HIP(hipGetLastError());
        FX_HIP(hipEventRecord(e1, s));
        FX_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        FX_HIP(hipEventElapsedTime(&ms, e0, e1));
        return (double)ms / reps;
    };
    const double ms_valu = time_kernel(false), ms_mfma = time_kernel(true);
    FX_HIP(hipEventDestroy(e0));
    FX_HIP(hipEventDestroy(e1));
    const double threads = (double)nb * kPeakBlock, waves = threads / 64.0;
    const double flop_valu = threads * kPeakSteps * kChains * 2.0;
    const double flop_mfma = waves * kPeakSteps * kChains * (2.0 * 16 * 16 * 4);
    *valu_tflops = flop_valu / (ms_valu * 1e-3) / 1e12;
    *mfma_tflops = flop_mfma / (ms_mfma * 1e-3) / 1e12;
}

}  // namespace fx
