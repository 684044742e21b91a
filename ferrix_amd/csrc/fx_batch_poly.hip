// fx_batch_poly.hip -- the polynomial-preconditioned mixed-precision batch
// (FX_PREC_MIXED_POLY).
//
// Per load case: flexible CG in FP64 (FCG(1): each new direction is made
// K-orthogonal to the previous one; Notay, "Flexible conjugate gradients",
// SIAM J. Sci. Comput. 22, 2000) whose preconditioner is `deg` steps of the
// Chebyshev iteration (Saad, "Iterative Methods for Sparse Linear Systems",
// 2nd ed., Alg. 12.1, in its three-term form) on the Jacobi-scaled operator
// A = S K S, S = diag(sqrt(inv_diag)), stored in FP16: z = S p(A) S r.  The
// outer iteration -- w = K64 z, every dot product, x, r and the stop rule --
// is FP64; r is the FP64 recursively updated residual, as in the FP64 PCG.
//
// Precision split, measured on the batched population's hardest members: the
// rounding of the Chebyshev *iterates* decides the preconditioner's quality
// (FP32 iterates cost 1.6x the inner steps of exact ones), the operator's
// does not (FP16 and FP64 operators take the same steps).  So the operator
// the inner steps stream is FP16 (2 bytes a value, a quarter of FP64) and
// the iterates FP64.  A Chebyshev polynomial needs no inner products: each
// inner step is one launch with the SpMV and the update fused, and its
// rounding only perturbs the preconditioner, which the flexible outer
// iteration absorbs.  The outer FP64 iterations are about deg times fewer
// than FP64 PCG iterations.
//
// Chebyshev interval [lo, hi] of A (the spectrum of D^-1 K): hi = 1.1 x the
// Rayleigh quotient after kPowerSteps power steps from a fixed start vector
// (2-4% below lambda_max on the population's meshes, so hi stays above it),
// lo = 1e-3 hi.  Eigenvalues below lo are damped less; the outer iteration
// resolves them.
//
// Sequences (every slot of a launch at the same stage, each slot's
// arithmetic independent of the others):
//   start:  kp_pinit, kPowerSteps+1 x kp_power, kp_lambda, kp_init, kp_stop
//   outer:  deg x kp_cheb, kp_outer, kp_coef, kp_upd, kp_stop
// kp_init and kp_upd also write the next outer iteration's inner start (the
// scaled residual rs = S r and z_1 = rs / theta; z_0 = 0 is not stored), and
// the last inner step stores its iterate unscaled (z = S z_deg+1).
#include "fx_batch.h"

namespace fx {
namespace {

constexpr double kPolyMargin = 1.1;     // hi = margin x Rayleigh quotient
constexpr double kPolyLoRatio = 1e-3;   // lo = ratio x hi
constexpr double kPowerScale = 0.25;    // keeps the unnormalised power iterates in range

// start vector of the power iteration: a fixed hash of the DOF index in [-1, 1)
__device__ __forceinline__ double power_v0(int64_t i) {
    uint64_t h = (uint64_t)i * 0x9E3779B97F4A7C15ull;
    h ^= (uint64_t)i >> 7;
    h = (h * 0xBF58476D1CE4E5B9ull) & 0xFFFFFFFFull;
    return (double)h / 2147483648.0 - 1.0;
}

// the Jacobi scaling of DOF i
__device__ __forceinline__ double jscale(const BSys &S, int64_t i) { return sqrt(gp(S.inv)[i]); }

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// One block column of the FP16 sliced ELL: its 8-entry record and 9th entry.
__device__ __forceinline__ void load_block16(const BSys &S, size_t kk, int lane, _Float16 (&v)[9]) {
    const u32x4 w = gp((const u32x4 *)S.sval16)[kk * kSlice + lane];
    __builtin_memcpy(v, &w, sizeof w);
    v[8] = gp(S.sval16b)[kk * kSlice + lane];
}

// acc[p][r] += (A q)(3a+p, r) over the node row's blocks, sliced-ELL FP16
// operator, one thread per row: the loop of fx_sell.h's sell_node_rows (the
// slice's width as the trip count, two blocks a step, column indices one step
// ahead) with two loads per block's values.
template <int RR>
__device__ __forceinline__ void sell_apply16(const BSys &S, int64_t a, gptr<const double> pv, double (&acc)[3][RR]) {
    gptr<const int32_t> sl_ptr = gp(S.sl_ptr), scol = gp(S.scol);
    const int64_t s = a / kSlice;
    const int lane = (int)(a % kSlice);
    const int k0 = __builtin_amdgcn_readfirstlane(sl_ptr[s]);
    const int kw = __builtin_amdgcn_readfirstlane(sl_ptr[s + 1]);
    int cn0 = scol[(size_t)k0 * kSlice + lane];
    int cn1 = k0 + 1 < kw ? scol[(size_t)(k0 + 1) * kSlice + lane] : 0;
    for (int k = k0; k < kw; k += 2) {
        const bool two = k + 1 < kw;
        const int c0 = cn0, c1 = cn1;
        double x0[3][RR], x1[3][RR];
        _Float16 v0[9], v1[9];
        gather_block<_Float16, double, RR>(pv, c0, x0);
        load_block16(S, (size_t)k, lane, v0);
        if (two) {
            gather_block<_Float16, double, RR>(pv, c1, x1);
            load_block16(S, (size_t)k + 1, lane, v1);
        }
        if (k + 2 < kw)
            cn0 = scol[(size_t)(k + 2) * kSlice + lane];
        if (k + 3 < kw)
            cn1 = scol[(size_t)(k + 3) * kSlice + lane];
        madd_block<_Float16, RR>(v0, x0, acc);
        if (two)
            madd_block<_Float16, RR>(v1, x1, acc);
    }
}

// The same on the node-block FP16 values, G lanes per row (column slots
// j = lane, lane + G, ...), reduced over the lanes.
template <int G, int RR>
__device__ __forceinline__ void lanes_apply16(const BSys &S, int64_t a, int lane, gptr<const double> pv,
                                              double (&acc)[3][RR]) {
    if (a < S.n_nodes) {
        gptr<const int32_t> nb_ptr = gp(S.nb_ptr), nb_col = gp(S.nb_col);
        const int b0 = nb_ptr[a], ncol = 3 * (nb_ptr[a + 1] - b0);
        gptr<const _Float16> v0 = gp(S.val16) + 9 * (int64_t)b0;
        for (int j = lane; j < ncol; j += G) {
            const int kb = j / 3;
            const int64_t c = (3 * (int64_t)nb_col[b0 + kb] + (j - 3 * kb)) * RR;
            const double w0 = (double)v0[j], w1 = (double)v0[ncol + j], w2 = (double)v0[2 * ncol + j];
#pragma unroll
            for (int r = 0; r < RR; ++r) {
                const double xj = pv[c + r];
                acc[0][r] += w0 * xj;
                acc[1][r] += w1 * xj;
                acc[2][r] += w2 * xj;
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
}

template <int RR>
__device__ __forceinline__ void zero_acc(double (&acc)[3][RR]) {
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int r = 0; r < RR; ++r)
            acc[p][r] = 0.0;
}

// ---------------------------------------------------------------------------
// power iteration on A (RR = 1 layout in zq[0], zq[1]): v_{j+1} = 0.25 A v_j;
// after the last step the partials v.(A v) and v.v -> part1 (2 per row group)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void power_row_tail(const BSys &S, int64_t a, int step, bool last, gptr<const double> pv,
                                               const double (&acc)[3][1], double (&loc)[2]) {
    if (last) {
#pragma unroll
        for (int p = 0; p < 3; ++p) {
            const double q = pv[3 * a + p];
            loc[0] += q * acc[p][0];
            loc[1] += q * q;
        }
    } else {
        gptr<double> vn = gp(S.zq[(step + 1) & 1]);
#pragma unroll
        for (int p = 0; p < 3; ++p)
            vn[3 * a + p] = kPowerScale * acc[p][0];
    }
}

__device__ __forceinline__ void write_power_partials(const BSys &S, int lb, bool last, double (&loc)[2]) {
    if (!last)
        return;
    block_sum<2>(loc);
    if (threadIdx.x == 0) {
        S.part1[2 * (size_t)lb] = loc[0];
        S.part1[2 * (size_t)lb + 1] = loc[1];
    }
}

__global__ __launch_bounds__(kBB) void kp_power_sell(BSys *__restrict__ sys, const int2 *__restrict__ tab, int step,
                                                     int last) {
    const int2 t = tab[blockIdx.x];
    const BSys &S = sys[t.x];
    gptr<const double> pv = gp((const double *)S.zq[step & 1]);
    const int64_t a = (int64_t)t.y * kBB + threadIdx.x;
    double loc[2] = {0.0, 0.0};
    if (a < S.n_nodes) {
        double acc[3][1];
        zero_acc<1>(acc);
        sell_apply16<1>(S, a, pv, acc);
        power_row_tail(S, a, step, last, pv, acc, loc);
    }
    write_power_partials(S, t.y, last, loc);
}

template <int G>
__global__ __launch_bounds__(kBB) void kp_power_lanes(BSys *__restrict__ sys, const int2 *__restrict__ tab, int step,
                                                      int last) {
    const int2 t = tab[blockIdx.x];
    const BSys &S = sys[t.x];
    gptr<const double> pv = gp((const double *)S.zq[step & 1]);
    const int64_t a = (int64_t)t.y * (kBB / G) + threadIdx.x / G;
    const int lane = threadIdx.x % G;
    double loc[2] = {0.0, 0.0};
    double acc[3][1];
    zero_acc<1>(acc);
    lanes_apply16<G, 1>(S, a, lane, pv, acc);
    if (lane == 0 && a < S.n_nodes)
        power_row_tail(S, a, step, last, pv, acc, loc);
    write_power_partials(S, t.y, last, loc);
}

// start vector of the power iteration (own DOFs, vector tiles)
__global__ __launch_bounds__(kBB) void kp_pinit(BSys *__restrict__ sys, const int2 *__restrict__ tab) {
    const int2 t = tab[blockIdx.x];
    const BSys &S = sys[t.x];
    const int64_t i = (int64_t)t.y * kBB + threadIdx.x;
    if (i < S.n_dof)
        gp(S.zq[0])[i] = power_v0(i);
}

// One block per slot: lambda_max estimate, the Chebyshev coefficients of the
// three-term recurrence z_{j+1} = z_j + c1_j (z_j - z_{j-1}) + c2_j (rs - A z_j),
// z_0 = 0, z_1 = rs / theta (Saad Alg. 12.1: rho_0 = 1/sigma,
// rho_j = 1/(2 sigma - rho_{j-1}), c1_j = rho_j rho_{j-1}, c2_j = 2 rho_j / delta).
__global__ __launch_bounds__(kBB) void kp_lambda(BSys *__restrict__ sys, const int32_t *__restrict__ list) {
    BSys &S = sys[list[blockIdx.x]];
    double red[2];
    bsum_partials<2>(S.part1, S.nparts1, 2, red);
    if (threadIdx.x != 0)
        return;
    const double lam = red[1] > 0.0 ? red[0] / red[1] : 1.0;
    const double hi = kPolyMargin * (lam > 0.0 && isfinite(lam) ? lam : 1.0), lo = kPolyLoRatio * hi;
    const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
    S.lam = lam;
    S.cheb[0] = 1.0 / theta;
    double rho = 1.0 / sigma;
    for (int j = 1; j <= S.deg; ++j) {
        const double rn = 1.0 / (2.0 * sigma - rho);
        S.cheb[2 * j] = rn * rho;
        S.cheb[2 * j + 1] = 2.0 * rn / delta;
        rho = rn;
    }
}

// ---------------------------------------------------------------------------
// outer iteration
// ---------------------------------------------------------------------------

// x = 0, r = b, pp = ps = 0 for the load cases with status 0, and the first
// inner start (rs = S b, z_1 = rs / theta); partials (r.r, r.(D^-1 r), b.b)
// -> part2 (3 per load case)
__global__ __launch_bounds__(kBB) void kp_init(BSys *__restrict__ sys, const int2 *__restrict__ tab) {
    const int2 t = tab[blockIdx.x];
    const BSys &S = sys[t.x];
    const int R = S.R;
    double v[3 * kRmax];
#pragma unroll
    for (int k = 0; k < 3 * kRmax; ++k)
        v[k] = 0.0;
    const int64_t i = (int64_t)t.y * kBB + threadIdx.x;
    if (i < S.n_dof) {
        const double inv = gp(S.inv)[i], si = jscale(S, i), ith = S.cheb[0];
#pragma unroll
        for (int r = 0; r < kRmax; ++r) {
            if (r >= R || S.rhs[r].status != 0)
                continue;
            const int64_t q = i * R + r;
            const double bi = gp(S.btrue)[q];
            const double rsi = si * bi;
            gp((double *)S.x)[q] = 0.0;
            gp((double *)S.r)[q] = bi;
            gp(S.pp)[q] = 0.0;
            gp(S.ps)[q] = 0.0;
            gp(S.rs)[q] = rsi;
            gp(S.zq[1])[q] = rsi * ith;
            v[3 * r] += bi * bi;
            v[3 * r + 1] += bi * (bi * inv);
            v[3 * r + 2] += bi * bi;
        }
    }
    block_sum<3 * kRmax>(v);
    if (threadIdx.x == 0)
        for (int r = 0; r < R; ++r)
            for (int k = 0; k < 3; ++k)
                S.part2[((size_t)t.y * R + r) * 3 + k] = v[3 * r + k];
}

// own-row update of inner step j: z_{j+1} = z_j + c1 (z_j - z_{j-1}) + c2 (rs - A z_j),
// stored unscaled (S z_{j+1}) at the last step
template <int RR>
__device__ __forceinline__ void cheb_row_tail(const BSys &S, int64_t a, int j, int go, gptr<const double> pv,
                                              const double (&acc)[3][RR]) {
    const double c1 = S.cheb[2 * j], c2 = S.cheb[2 * j + 1];
    const bool last = j == S.deg;
    gptr<const double> zm = gp((const double *)S.zq[(j + 2) % 3]);  // z_{j-1}
    gptr<const double> rs = gp((const double *)S.rs);
    gptr<double> zn = gp(S.zq[(j + 1) % 3]);
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        const int64_t d = 3 * a + p;
        const double sd = last ? jscale(S, d) : 1.0;
#pragma unroll
        for (int r = 0; r < RR; ++r) {
            if (!((go >> r) & 1))
                continue;
            const int64_t qd = d * RR + r;
            const double zj = pv[qd];
            const double zmj = j == 1 ? 0.0 : zm[qd];  // z_0 = 0 (not stored)
            const double zv = zj + c1 * (zj - zmj) + c2 * (rs[qd] - acc[p][r]);
            zn[qd] = last ? sd * zv : zv;
        }
    }
}

// A wavefront's rows in the sliced-ELL kernels: 64 consecutive node rows,
// i.e. one contiguous range of 3 * 64 * RR entries of every vector.  Such a
// range is written through LDS with lane-consecutive stores instead of 3 * RR
// strided ones per thread.
struct WaveRows {
    int w0;        // first thread of the wavefront in the block
    int64_t row0;  // its first node row
    int nrow;      // its rows inside the system (0..64)
};

__device__ __forceinline__ WaveRows wave_rows(const BSys &S, int lb, int rows = kBB) {
    WaveRows w;
    w.w0 = (threadIdx.x / 64) * 64;
    w.row0 = (int64_t)lb * rows + w.w0;
    const int64_t left = S.n_nodes - w.row0;
    w.nrow = left <= 0 ? 0 : (left < 64 ? (int)left : 64);
    return w;
}

// dst[row0 * W + e] = tile[w0 * W + e] for the wavefront's entries (W = 3 RR per row)
template <int W>
__device__ __forceinline__ void stage_out(const WaveRows &w, const double *tile, gptr<double> dst) {
    const int lane = threadIdx.x % 64;
    for (int e = lane; e < w.nrow * W; e += 64)
        dst[w.row0 * W + e] = tile[w.w0 * W + e];
}

// The sliced-ELL inner step.  The new iterate goes out through an LDS tile
// as lane-consecutive stores (staging the z_{j-1} and rs reads the same way
// was slower: those reads hit lines the wavefront's other loads bring in).
// Entries of load cases that do not run are written as well (their values
// are never read).
template <int RR, int ROWS>
__device__ __forceinline__ void cheb_sell(const BSys &S, int lb, int j, int go, double *tile) {
    constexpr int W = 3 * RR;  // entries per node row
    const int64_t a = (int64_t)lb * ROWS + threadIdx.x;
    gptr<const double> pv = gp((const double *)S.zq[j % 3]);
    if (a < S.n_nodes) {
        double acc[3][RR];
        zero_acc<RR>(acc);
        sell_apply16<RR>(S, a, pv, acc);
        const double c1 = S.cheb[2 * j], c2 = S.cheb[2 * j + 1];
        const bool last = j == S.deg;
        // the row's 3 RR entries of z_j, z_{j-1} and rs are contiguous: read
        // each as one run (wide loads) before the arithmetic
        double zj[3][RR], zmj[3][RR], rsv[3][RR];
        gather_block<double, double, RR>(pv, (int)a, zj);
        gather_block<double, double, RR>(gp((const double *)S.rs), (int)a, rsv);
        if (j == 1) {  // z_0 = 0 (not stored)
#pragma unroll
            for (int p = 0; p < 3; ++p)
#pragma unroll
                for (int r = 0; r < RR; ++r)
                    zmj[p][r] = 0.0;
        } else {
            gather_block<double, double, RR>(gp((const double *)S.zq[(j + 2) % 3]), (int)a, zmj);
        }
#pragma unroll
        for (int p = 0; p < 3; ++p) {
            const double sd = last ? jscale(S, 3 * a + p) : 1.0;
#pragma unroll
            for (int r = 0; r < RR; ++r) {
                const double zv = zj[p][r] + c1 * (zj[p][r] - zmj[p][r]) + c2 * (rsv[p][r] - acc[p][r]);
                tile[threadIdx.x * W + p * RR + r] = last ? sd * zv : zv;
            }
        }
    }
    __syncthreads();
    stage_out<W>(wave_rows(S, lb, ROWS), tile, gp(S.zq[(j + 1) % 3]));
    (void)go;
}

// ROWS node rows (ROWS / 64 slices) per workgroup; rows past the system's
// end only take part in the barrier
template <int ROWS>
__global__ __launch_bounds__(ROWS) void kp_cheb_sell(BSys *__restrict__ sys, const int2 *__restrict__ tab, int j) {
    const int2 t = tab[blockIdx.x];
    const BSys &S = sys[t.x];
    const int go = S.go;
    if (!go)
        return;
    // one tile for every load-case count (a __shared__ array per template
    // instance would reserve four, and LDS would then hold one block per CU)
    __shared__ double tile[ROWS * 3 * kRmax];
    switch (S.R) {
    case 1: cheb_sell<1, ROWS>(S, t.y, j, go, tile); break;
    case 2: cheb_sell<2, ROWS>(S, t.y, j, go, tile); break;
    case 3: cheb_sell<3, ROWS>(S, t.y, j, go, tile); break;
    default: cheb_sell<4, ROWS>(S, t.y, j, go, tile); break;
    }
}

// The same inner step for the systems with RR load cases only: registers
// and LDS sized for RR (the kernel above carries those of four, and its
// occupancy with them).  The arithmetic is cheb_sell<RR>'s.
template <int RR, int ROWS>
__global__ __launch_bounds__(ROWS) void kp_cheb_sell_r(BSys *__restrict__ sys, const int2 *__restrict__ tab, int j) {
    const int2 t = tab[blockIdx.x];
    const BSys &S = sys[t.x];
    const int go = S.go;
    if (!go)
        return;
    __shared__ double tile[ROWS * 3 * RR];
    cheb_sell<RR, ROWS>(S, t.y, j, go, tile);
}

template <int ROWS>
void launch_cheb_rows(hipStream_t s, BSys *sys, const int2 *tab, int n, const int *n_r, bool split, int j) {
    if (!split) {
        hipLaunchKernelGGL(kp_cheb_sell<ROWS>, dim3(n), dim3(ROWS), 0, s, sys, tab, j);
        return;
    }
    int off = 0;
    for (int r = kRmax - 1; r >= 0; --r) {
        const int c = n_r[r];
        if (!c)
            continue;
        switch (r) {
        case 0: hipLaunchKernelGGL((kp_cheb_sell_r<1, ROWS>), dim3(c), dim3(ROWS), 0, s, sys, tab + off, j); break;
        case 1: hipLaunchKernelGGL((kp_cheb_sell_r<2, ROWS>), dim3(c), dim3(ROWS), 0, s, sys, tab + off, j); break;
        case 2: hipLaunchKernelGGL((kp_cheb_sell_r<3, ROWS>), dim3(c), dim3(ROWS), 0, s, sys, tab + off, j); break;
        default: hipLaunchKernelGGL((kp_cheb_sell_r<4, ROWS>), dim3(c), dim3(ROWS), 0, s, sys, tab + off, j); break;
        }
        off += c;
    }
}

// Inner step j over the sliced-ELL table of `rows`-row groups (kBB or
// kInnerRowsLarge): one launch (split = false) or one per load-case count
// over its segment (n_r[r]: entries with r + 1 load cases; the segments run
// from four load cases down to one).
void launch_cheb_sell(hipStream_t s, BSys *sys, const int2 *tab, int n, const int *n_r, bool split, int rows,
                      int j) {
    if (rows == kInnerRowsLarge)
        launch_cheb_rows<kInnerRowsLarge>(s, sys, tab, n, n_r, split, j);
    else
        launch_cheb_rows<kBB>(s, sys, tab, n, n_r, split, j);
}

template <int G, int RR>
__device__ __forceinline__ void cheb_lanes(const BSys &S, int lb, int j, int go) {
    const int64_t a = (int64_t)lb * (kBB / G) + threadIdx.x / G;
    const int lane = threadIdx.x % G;
    gptr<const double> pv = gp((const double *)S.zq[j % 3]);
    double acc[3][RR];
    zero_acc<RR>(acc);
    lanes_apply16<G, RR>(S, a, lane, pv, acc);
    if (lane == 0 && a < S.n_nodes)
        cheb_row_tail<RR>(S, a, j, go, pv, acc);
}

template <int G>
__global__ __launch_bounds__(kBB) void kp_cheb_lanes(BSys *__restrict__ sys, const int2 *__restrict__ tab, int j) {
    const int2 t = tab[blockIdx.x];
    const BSys &S = sys[t.x];
    const int go = S.go;
    if (!go)
        return;
    switch (S.R) {
    case 1: cheb_lanes<G, 1>(S, t.y, j, go); break;
    case 2: cheb_lanes<G, 2>(S, t.y, j, go); break;
    case 3: cheb_lanes<G, 3>(S, t.y, j, go); break;
    default: cheb_lanes<G, 4>(S, t.y, j, go); break;
    }
}

// FP64 outer SpMV w = K64 z (z = zq[(deg+1) % 3], unscaled): on the
// node-block values with kOuterG lanes per row, or on the sliced-ELL FP64
// operator one thread per row; partials per load case (z.s, z.r, p.r, z.w,
// p.w) with p, s the previous direction and its image -> part1.
// RUNS (the sliced-ELL kernel, every lane a row): the row's 3 RR entries of
// each vector are read as contiguous runs (wide loads), every load case's,
// before the running ones are used; else (the node-block kernel, the tail on
// one lane of kOuterG) entry by entry, running load cases only, which keeps
// that kernel's registers low.
template <int RR, bool RUNS>
__device__ __forceinline__ void outer_row_tail(const BSys &S, int64_t a, int go, gptr<const double> zv,
                                               const double (&acc)[3][RR], double (&v)[5 * kRmax],
                                               double *tile = nullptr) {
    double zr[3][RR], sor[3][RR], por[3][RR], rrr[3][RR];
    if constexpr (RUNS) {
        gather_block<double, double, RR>(zv, (int)a, zr);
        gather_block<double, double, RR>(gp((const double *)S.ps), (int)a, sor);
        gather_block<double, double, RR>(gp((const double *)S.pp), (int)a, por);
        gather_block<double, double, RR>(gp((const double *)S.r), (int)a, rrr);
    }
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        const int64_t d = 3 * a + p;
#pragma unroll
        for (int r = 0; r < RR; ++r) {
            if (!((go >> r) & 1))
                continue;
            const int64_t q = d * RR + r;
            double z, so, po, rr;
            if constexpr (RUNS) {
                z = zr[p][r];
                so = sor[p][r];
                po = por[p][r];
                rr = rrr[p][r];
            } else {
                z = zv[q];
                so = gp((const double *)S.ps)[q];
                po = gp((const double *)S.pp)[q];
                rr = gp((const double *)S.r)[q];
            }
            const double w = acc[p][r];
            if (tile)
                tile[threadIdx.x * 3 * RR + p * RR + r] = w;
            else
                gp(S.pw)[q] = w;
            v[5 * r] += z * so;
            v[5 * r + 1] += z * rr;
            v[5 * r + 2] += po * rr;
            v[5 * r + 3] += z * w;
            v[5 * r + 4] += po * w;
        }
    }
}

template <int RR>
__device__ __forceinline__ void outer_row(const BSys &S, int64_t a, int lane, int go, double (&v)[5 * kRmax]) {
    constexpr int G = kOuterG;
    gptr<const double> zv = gp((const double *)S.zq[(S.deg + 1) % 3]);
    double acc[3][RR];
    zero_acc<RR>(acc);
    if (a < S.n_nodes) {
        gptr<const int32_t> nb_ptr = gp(S.nb_ptr), nb_col = gp(S.nb_col);
        const int b0 = nb_ptr[a], ncol = 3 * (nb_ptr[a + 1] - b0);
        gptr<const double> v0 = gp(S.val64) + 9 * (int64_t)b0;
        for (int j = lane; j < ncol; j += G) {
            const int kb = j / 3;
            const int64_t c = (3 * (int64_t)nb_col[b0 + kb] + (j - 3 * kb)) * RR;
            const double w0 = v0[j], w1 = v0[ncol + j], w2 = v0[2 * ncol + j];
#pragma unroll
            for (int r = 0; r < RR; ++r) {
                const double xj = zv[c + r];
                acc[0][r] += w0 * xj;
                acc[1][r] += w1 * xj;
                acc[2][r] += w2 * xj;
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
    if (lane != 0 || a >= S.n_nodes)
        return;
    outer_row_tail<RR, false>(S, a, go, zv, acc, v);
}

template <int RR, bool SELL>
__device__ __forceinline__ void outer_block(const BSys &S, int lb, int go, double *tile) {
    double v[5 * kRmax];
#pragma unroll
    for (int k = 0; k < 5 * kRmax; ++k)
        v[k] = 0.0;
    if constexpr (SELL) {  // w leaves through LDS as lane-consecutive stores (see cheb_sell)
        const int64_t a = (int64_t)lb * kBB + threadIdx.x;
        if (a < S.n_nodes) {
            gptr<const double> zv = gp((const double *)S.zq[(S.deg + 1) % 3]);
            double acc[3][RR];
            zero_acc<RR>(acc);
            sell_node_rows<double, double, RR>(gp(S.sl_ptr), gp(S.scol), gp(S.sval64), a, zv, acc);
            outer_row_tail<RR, true>(S, a, go, zv, acc, v, tile);
        }
        __syncthreads();
        stage_out<3 * RR>(wave_rows(S, lb), tile, gp(S.pw));
    } else {
        outer_row<RR>(S, (int64_t)lb * (kBB / kOuterG) + threadIdx.x / kOuterG, threadIdx.x % kOuterG, go, v);
    }
    block_sum<5 * kRmax>(v);
    if (threadIdx.x == 0)
        for (int r = 0; r < RR; ++r)
            for (int k = 0; k < 5; ++k)
                S.part1[((size_t)lb * RR + r) * 5 + k] = v[5 * r + k];
}

template <bool SELL>
__global__ __launch_bounds__(kBB) void kp_outer(BSys *__restrict__ sys, const int2 *__restrict__ tab) {
    const int2 t = tab[blockIdx.x];
    const BSys &S = sys[t.x];
    const int go = S.go;
    if (!go)
        return;
    __shared__ double tile[SELL ? kBB * 3 * kRmax : 1];  // one tile for every load-case count
    switch (S.R) {
    case 1: outer_block<1, SELL>(S, t.y, go, tile); break;
    case 2: outer_block<2, SELL>(S, t.y, go, tile); break;
    case 3: outer_block<3, SELL>(S, t.y, go, tile); break;
    default: outer_block<4, SELL>(S, t.y, go, tile); break;
    }
}

// One block per slot: beta (FCG(1): the new direction p = z - beta p_old is
// K-orthogonal to p_old), p.s by expansion, alpha = (p.r)/(p.s); a p.s that
// is not positive (or not finite) is a breakdown.
__global__ __launch_bounds__(kBB) void kp_coef(BSys *__restrict__ sys, const int32_t *__restrict__ list) {
    BSys &S = sys[list[blockIdx.x]];
    const int go = S.go;
    if (!go)
        return;
    __shared__ BRhs H[kRmax];
    stage_rhs(S, H);
    const int R = S.R;
    double red[5 * kRmax];
    bsum_partials<5 * kRmax>(S.part1, S.nparts3, 5 * R, red);
    if (threadIdx.x != 0)
        return;
    int go2 = 0;
#pragma unroll
    for (int r = 0; r < kRmax; ++r) {
        if (r >= R || !((go >> r) & 1))
            continue;
        const BRhs &c = H[r];
        BRhs &h = S.rhs[r];
        const double zs = red[5 * r], zr = red[5 * r + 1], pr = red[5 * r + 2], zw = red[5 * r + 3],
                     pw = red[5 * r + 4];
        const double beta = c.it == 0 ? 0.0 : zs / c.pap;
        const double prn = zr - beta * pr;
        const double psn = ((zw - beta * zs) - beta * pw) + beta * beta * c.pap;
        const double al = prn / psn;
        h.beta = beta;
        h.alpha = al;
        h.pap = psn;
        if (!(psn > 0.0) || !isfinite(al))
            h.status = 2;
        else
            go2 |= 1 << r;
    }
    S.go = go2;
}

// p = z - beta p_old, s = w - beta s_old, x += alpha p, r -= alpha s, and
// the next inner start (rs = S r, z_1 = rs / theta); partials (r.r,
// r.(D^-1 r), 0) -> part2
// RR consecutive doubles from v + i RR
template <int RR>
__device__ __forceinline__ void load_run(gptr<const double> v, int64_t i, double (&o)[RR]) {
#pragma unroll
    for (int r = 0; r < RR; ++r)
        o[r] = v[i * RR + r];
}

template <int RR>
__device__ __forceinline__ void upd_tile(const BSys &S, int lb, int go, double (&v)[3 * kRmax]) {
    const int64_t i = (int64_t)lb * kBB + threadIdx.x;
    if (i >= S.n_dof)
        return;
    const double inv = gp(S.inv)[i], si = jscale(S, i), ith = S.cheb[0];
    // the DOF's RR entries of each vector are contiguous: read as runs (wide
    // loads), every load case's, then update the running ones
    double zr[RR], ppr[RR], pwr[RR], psr[RR], xr[RR], rr0[RR];
    load_run<RR>(gp((const double *)S.zq[(S.deg + 1) % 3]), i, zr);
    load_run<RR>(gp((const double *)S.pp), i, ppr);
    load_run<RR>(gp((const double *)S.pw), i, pwr);
    load_run<RR>(gp((const double *)S.ps), i, psr);
    load_run<RR>(gp((const double *)S.x), i, xr);
    load_run<RR>(gp((const double *)S.r), i, rr0);
#pragma unroll
    for (int r = 0; r < RR; ++r) {
        if (!((go >> r) & 1))
            continue;
        const double al = S.rhs[r].alpha, be = S.rhs[r].beta;
        const int64_t q = i * RR + r;
        const double pn = zr[r] - be * ppr[r];
        const double sn = pwr[r] - be * psr[r];
        const double xn = xr[r] + al * pn;
        const double rn = rr0[r] - al * sn;
        const double rsn = si * rn;
        gp(S.pp)[q] = pn;
        gp(S.ps)[q] = sn;
        gp((double *)S.x)[q] = xn;
        gp((double *)S.r)[q] = rn;
        // the next outer iteration's inner start (unused if this one stops);
        // z_1 may share z's buffer: this thread has read its entry above
        gp(S.rs)[q] = rsn;
        gp(S.zq[1])[q] = rsn * ith;
        v[3 * r] += rn * rn;
        v[3 * r + 1] += rn * (rn * inv);
    }
}

// x of a stopped slot into its pinned host array: the block's DOF tile is
// one contiguous range of kBB * R interleaved entries; the stores are made
// visible to the host before the kernel ends (the next kernel, kp_stop,
// then writes the finish summary the host waits for)
__device__ __forceinline__ void export_tile(const BSys &S, int lb) {
    const int64_t nv = S.n_dof * S.R, q0 = (int64_t)lb * kBB * S.R;
    const int64_t q1 = q0 + (int64_t)kBB * S.R < nv ? q0 + (int64_t)kBB * S.R : nv;
    gptr<const double> x = gp((const double *)S.x);
    for (int64_t q = q0 + threadIdx.x; q < q1; q += kBB)
        S.hx[q] = x[q];
    __threadfence_system();
}

__global__ __launch_bounds__(kBB) void kp_upd(BSys *__restrict__ sys, const int2 *__restrict__ tab) {
    const int2 t = tab[blockIdx.x];
    const BSys &S = sys[t.x];
    const int go = S.go;
    if (!go) {
        if (S.exp && !S.reported)
            export_tile(S, t.y);
        return;
    }
    double v[3 * kRmax];
#pragma unroll
    for (int k = 0; k < 3 * kRmax; ++k)
        v[k] = 0.0;
    switch (S.R) {
    case 1: upd_tile<1>(S, t.y, go, v); break;
    case 2: upd_tile<2>(S, t.y, go, v); break;
    case 3: upd_tile<3>(S, t.y, go, v); break;
    default: upd_tile<4>(S, t.y, go, v); break;
    }
    block_sum<3 * kRmax>(v);
    if (threadIdx.x == 0)
        for (int r = 0; r < S.R; ++r)
            for (int k = 0; k < 3; ++k)
                S.part2[((size_t)t.y * S.R + r) * 3 + k] = v[3 * r + k];
}

// One block per slot, after kp_init (after_update 0) or kp_upd: the stop rule
// per load case (as fx_pcg.hip's: sqrt(r.z) < tol, or ||r||/||b|| <= tol),
// the outer iteration limit, and the mask of load cases the next outer
// iteration works on.  A slot with no load case left running writes its
// finish summary for the host (iterations: inner steps; nupd: outer
// iterations).
__global__ __launch_bounds__(kBB) void kp_stop(BSys *__restrict__ sys, const int32_t *__restrict__ list,
                                               BFin *__restrict__ fin, int after_update) {
    const int slot = list[blockIdx.x];
    BSys &S = sys[slot];
    __shared__ BRhs H[kRmax];
    stage_rhs(S, H);
    const int R = S.R;
    bool any = false;
    for (int r = 0; r < R; ++r)
        any |= H[r].status == 0;
    if (any) {  // block-uniform: every thread read the same staged states
        double red[3 * kRmax];
        bsum_partials<3 * kRmax>(S.part2, S.nparts2, 3 * R, red);
        if (threadIdx.x != 0)
            return;
        int go = 0;
#pragma unroll
        for (int r = 0; r < kRmax; ++r) {
            const BRhs &c = H[r];
            BRhs &h = S.rhs[r];
            if (r >= R || c.status != 0)
                continue;
            const int64_t it = after_update ? c.it + 1 : c.it;
            const double rr = red[3 * r], rzt = red[3 * r + 1];
            const double bb = after_update ? c.bb : red[3 * r + 2];
            double crit;
            bool conv;
            if (c.stop_mode == FX_STOP_ABS_SQRT_RZ) {
                crit = sqrt(rzt);
                conv = crit < c.tol;
            } else {
                crit = bb > 0.0 ? sqrt(rr) / sqrt(bb) : 0.0;
                conv = crit <= c.tol;
            }
            h.it = it;
            h.rr = rr;
            h.crit = crit;
            if (!after_update)
                h.bb = bb;
            if (conv)
                h.status = 1;
            else if (it * (int64_t)S.deg >= c.max_iter || !isfinite(crit))
                h.status = 3;
            else
                go |= 1 << r;
        }
        S.go = go;
        if (go) {  // the inner steps of the next outer iteration (integer-valued sums: exact in any order)
            atomicAdd(S.bytes_acc, (double)S.deg * (S.cheb_bytes0 + (double)__popc(go) * S.cheb_bytes_rhs));
            return;
        }
    } else if (threadIdx.x != 0) {
        return;
    }
    if (!S.reported) {  // thread 0, no load case runs any more: tell the host
        if (S.hx && !S.exp) {  // after the next kp_upd has exported x
            S.exp = 1;
            return;
        }
        BFin &F = fin[slot];
        int64_t it = 0;
        for (int r = 0; r < kRmax; ++r) {
            const BRhs &h = S.rhs[r];  // written by this thread above or in earlier launches
            F.crit[r] = h.crit;
            F.status[r] = h.status;
            if (r < R)
                it = max(it, h.it);
        }
        F.it = it * (int64_t)S.deg;
        F.nupd = it;
        __threadfence_system();
        F.done = 1;
        S.reported = 1;
    }
}

template <typename K>
void launch_inner_tables(const int2 *tab_sell, int n_sell, const int2 *const *tab_g, const int *n_g, K &&launch) {
    if (n_sell)
        launch(0, tab_sell, n_sell);
    for (int g = 0; g < 6; ++g)
        if (n_g[g])
            launch(2 << g, tab_g[g], n_g[g]);
}

}  // namespace

void launch_poly_power_start(hipStream_t s, BSys *sys, const int2 *tab_vec, int n_vec, const int2 *tab_in_sell,
                             int n_in_sell, const int2 *const *tab_in_g, const int *n_in_g, const int32_t *list,
                             int n_list, BFin *fin) {
    hipLaunchKernelGGL(kp_pinit, dim3(n_vec), dim3(kBB), 0, s, sys, tab_vec);
    for (int step = 0; step <= kPowerSteps; ++step) {
        const int last = step == kPowerSteps;
        launch_inner_tables(tab_in_sell, n_in_sell, tab_in_g, n_in_g, [&](int G, const int2 *tab, int n) {
            switch (G) {
            case 0: hipLaunchKernelGGL(kp_power_sell, dim3(n), dim3(kBB), 0, s, sys, tab, step, last); break;
            case 2: hipLaunchKernelGGL(kp_power_lanes<2>, dim3(n), dim3(kBB), 0, s, sys, tab, step, last); break;
            case 4: hipLaunchKernelGGL(kp_power_lanes<4>, dim3(n), dim3(kBB), 0, s, sys, tab, step, last); break;
            case 8: hipLaunchKernelGGL(kp_power_lanes<8>, dim3(n), dim3(kBB), 0, s, sys, tab, step, last); break;
            case 16: hipLaunchKernelGGL(kp_power_lanes<16>, dim3(n), dim3(kBB), 0, s, sys, tab, step, last); break;
            case 32: hipLaunchKernelGGL(kp_power_lanes<32>, dim3(n), dim3(kBB), 0, s, sys, tab, step, last); break;
            default: hipLaunchKernelGGL(kp_power_lanes<64>, dim3(n), dim3(kBB), 0, s, sys, tab, step, last); break;
            }
        });
    }
    hipLaunchKernelGGL(kp_lambda, dim3(n_list), dim3(kBB), 0, s, sys, list);
    hipLaunchKernelGGL(kp_init, dim3(n_vec), dim3(kBB), 0, s, sys, tab_vec);
    hipLaunchKernelGGL(kp_stop, dim3(n_list), dim3(kBB), 0, s, sys, list, fin, 0);
}

void launch_poly_outer(hipStream_t s, BSys *sys, int deg, const int2 *tab_vec, int n_vec, const int2 *tab_in_sell,
                       int n_in_sell, const int *n_in_sell_r, bool split, int inner_rows, const int2 *const *tab_in_g,
                       const int *n_in_g, const int2 *tab_out_sell, int n_out_sell, const int2 *tab_out, int n_out,
                       const int32_t *list, int n_list, BFin *fin, hipEvent_t ev0, hipEvent_t ev1) {
    if (ev0)
        FX_HIP(hipEventRecord(ev0, s));
    for (int j = 1; j <= deg; ++j)
        launch_inner_tables(tab_in_sell, n_in_sell, tab_in_g, n_in_g, [&](int G, const int2 *tab, int n) {
            switch (G) {
            case 0: launch_cheb_sell(s, sys, tab, n, n_in_sell_r, split, inner_rows, j); break;
            case 2: hipLaunchKernelGGL(kp_cheb_lanes<2>, dim3(n), dim3(kBB), 0, s, sys, tab, j); break;
            case 4: hipLaunchKernelGGL(kp_cheb_lanes<4>, dim3(n), dim3(kBB), 0, s, sys, tab, j); break;
            case 8: hipLaunchKernelGGL(kp_cheb_lanes<8>, dim3(n), dim3(kBB), 0, s, sys, tab, j); break;
            case 16: hipLaunchKernelGGL(kp_cheb_lanes<16>, dim3(n), dim3(kBB), 0, s, sys, tab, j); break;
            case 32: hipLaunchKernelGGL(kp_cheb_lanes<32>, dim3(n), dim3(kBB), 0, s, sys, tab, j); break;
            default: hipLaunchKernelGGL(kp_cheb_lanes<64>, dim3(n), dim3(kBB), 0, s, sys, tab, j); break;
            }
        });
    if (ev1)
        FX_HIP(hipEventRecord(ev1, s));
    if (n_out_sell)
        hipLaunchKernelGGL(kp_outer<true>, dim3(n_out_sell), dim3(kBB), 0, s, sys, tab_out_sell);
    if (n_out)
        hipLaunchKernelGGL(kp_outer<false>, dim3(n_out), dim3(kBB), 0, s, sys, tab_out);
    hipLaunchKernelGGL(kp_coef, dim3(n_list), dim3(kBB), 0, s, sys, list);
    hipLaunchKernelGGL(kp_upd, dim3(n_vec), dim3(kBB), 0, s, sys, tab_vec);
    hipLaunchKernelGGL(kp_stop, dim3(n_list), dim3(kBB), 0, s, sys, list, fin, 1);
}

// `reps` launches of the first inner step over the sliced-ELL table, in one
// shape, between two events after one untimed launch; ms per inner step.
// Between outer iterations this has no effect on the solves: the step reads
// the next start z_1 and rs and writes only z_2, which the first inner step
// of the next outer iteration rewrites from the same inputs before anything
// reads it.
double time_poly_inner(hipStream_t s, BSys *sys, const int2 *tab, int n, const int *n_r, bool split, int rows,
                       int reps) {
    hipEvent_t e0, e1;
    FX_HIP(hipEventCreate(&e0));
    FX_HIP(hipEventCreate(&e1));
    float ms = 0.f;
    try {
        launch_cheb_sell(s, sys, tab, n, n_r, split, rows, 1);
        FX_HIP(hipEventRecord(e0, s));
        for (int k = 0; k < reps; ++k)
            launch_cheb_sell(s, sys, tab, n, n_r, split, rows, 1);
        FX_HIP(hipEventRecord(e1, s));
        FX_HIP(hipGetLastError());
        FX_HIP(hipEventSynchronize(e1));
        FX_HIP(hipEventElapsedTime(&ms, e0, e1));
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return (double)ms / reps;
}

}  // namespace fx
