// fx_csr.hip -- Solver::solve on a caller's CSR matrix.
//
// Reference: the solver seam, trait Solver::solve(k_global: &CsMat<f64>,
// f_global, preconditioner, tol, max_iter) (src/solver/solvers/mod.rs:18-44),
// which the step driver calls with the CSR it has just built
// (src/solver/step/static_step.rs:198-204); IterativeSolver::solve
// (src/solver/solvers/iterative.rs:26-121) with the Jacobi preconditioner
// DiagonalPreconditioner::new (src/solver/preconditioner.rs:39-50:
// inv_diag[i] = 1/K_ii if |K_ii| > 1e-9, else 1; a row with no stored
// diagonal keeps 1).
//
// The matrix is a scalar CSR as sprs holds it: row pointers and column
// indices of machine width (usize, int64 here), columns ascending and unique
// per row, both triangles stored.  Nothing about 3x3 node blocks is assumed.
// The iteration is the mesh solve's (fx_pcg_core.h): two kernels per
// iteration, no host round trip inside a chunk of iterations, p formed on
// the fly in the SpMV, block partials reduced in a fixed order.
//
// SpMV: G lanes own one row and walk its entries with stride G (two
// entries per lane in flight); the G partial sums meet in a butterfly.  The
// summation order within a row therefore differs from the reference's
// sequential row sum, which the PCG tolerance absorbs (as the reference's
// rayon reductions do).
//
// Two entry points share the kernels:
//  * fx_solve_csr: host arrays in and out; the matrix is checked on the
//    host, copied with 32-bit column indices (n < 2^31) and solved with
//    work vectors from the context's device pool.
//  * fx_solve_csr_noalloc: device arrays in and out, no device allocation;
//    the work vectors live in `scratch`, a file-scope device array of
//    kScratchBytes per GPU, and the matrix is checked by a kernel before
//    any gather.
#include "fx_pcg_core.h"

#include <chrono>

namespace fx {
namespace {

// ---------------------------------------------------------------------------
// the no-allocation entry point's scratch memory: one array per GPU, held
// until the process ends.  Host code never uses the symbol itself (its host
// name is a stand-in, not a device address): csr_scratch_setup resolves the
// device address once per device with hipGetSymbolAddress.
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = (size_t)64 << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mx;
void *g_scratch_ptr[kMaxDevices] = {};

// Work vectors of one solve carved out of one device block: r, z, p (two
// buffers), Ap, inv_diag (n each), the SpMV block partials (one per row
// group), the update partials (three per 256 rows), the PCG scalars and the
// check flag.  Every section starts on a 256-byte boundary.
struct CsrWork {
    double *r, *z, *p0, *p1, *ap, *inv, *part1, *part2;
    PcgScalars *sc;
    int32_t *flag;
};

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Bytes of the work block for n rows at G lanes per row; with base != null
// also fills *w.
size_t csr_work_layout(int64_t n, int G, char *base, CsrWork *w) {
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char *p = base ? base + off : nullptr;
        off += align256(bytes);
        return p;
    };
    const size_t vec = (size_t)n * sizeof(double);
    double **vecs[6] = {&w->r, &w->z, &w->p0, &w->p1, &w->ap, &w->inv};
    for (double **v : vecs) {
        char *p = take(vec);
        if (base)
            *v = (double *)p;
    }
    char *p1 = take(sizeof(double) * (size_t)nblocks(n, kBlock / G));
    char *p2 = take(sizeof(double) * 3 * (size_t)nblocks(n, kBlock));
    char *sc = take(sizeof(PcgScalars));
    char *fl = take(sizeof(int32_t));
    if (base) {
        w->part1 = (double *)p1;
        w->part2 = (double *)p2;
        w->sc = (PcgScalars *)sc;
        w->flag = (int32_t *)fl;
    }
    return off;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// Structure check of a device CSR before any gather: rp[0] = 0, rp[n] = nnz,
// rows non-decreasing and inside [0, nnz], columns in [0, n) and strictly
// ascending per row.  A thread scans its row only when the row's bounds are
// valid, so every read stays inside rp[0..n] and col[0..nnz).  *flag (zero on
// entry) gets bit 1 for a row-pointer fault and bit 2 for a column fault.
template <typename CI>
__global__ __launch_bounds__(kBlock) void k_csr_check(int64_t n, int64_t nnz, const int64_t *__restrict__ rp,
                                                      const CI *__restrict__ col, int32_t *__restrict__ flag) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i == 0 && (rp[0] != 0 || rp[n] != nnz))
        atomicOr(flag, 1);
    if (i >= n)
        return;
    const int64_t k0 = rp[i], k1 = rp[i + 1];
    if (k0 < 0 || k1 < k0 || k1 > nnz) {
        atomicOr(flag, 1);
        return;
    }
    int64_t prev = -1;
    for (int64_t k = k0; k < k1; ++k) {
        const int64_t c = (int64_t)col[k];
        if (c <= prev || c >= n) {
            atomicOr(flag, 2);
            return;
        }
        prev = c;
    }
}

__global__ void k_csr_clear_flag(int32_t *flag) {
    if (threadIdx.x == 0 && blockIdx.x == 0)
        *flag = 0;
}

// inv[i] = 1/K_ii if |K_ii| > 1e-9, else 1 (no stored diagonal: 1); the
// diagonal is found by binary search in the ascending row.
template <typename CI>
__global__ __launch_bounds__(kBlock) void k_csr_jacobi(int64_t n, const int64_t *__restrict__ rp,
                                                       const CI *__restrict__ col, const double *__restrict__ val,
                                                       double *__restrict__ inv) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n)
        return;
    int64_t lo = rp[i], hi = rp[i + 1];
    double d = 1.0;
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        const int64_t c = (int64_t)col[mid];
        if (c == i) {
            const double v = val[mid];
            if (fabs(v) > kJacobiGuard)
                d = 1.0 / v;
            break;
        }
        if (c < i)
            lo = mid + 1;
        else
            hi = mid;
    }
    inv[i] = d;
}

// The PCG scalars of a new solve (written by a kernel, so no host staging).
__global__ void k_csr_scalars(PcgScalars *sc, double tol, int64_t max_iter, int32_t stop_mode) {
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    PcgScalars init{};
    init.tol = tol;
    init.max_iter = max_iter;
    init.stop_mode = stop_mode;
    *sc = init;
}

// One iteration's SpMV: the head (stop rule, beta), p_i = z + beta p_{i-1}
// for every gathered column, Ap = K p_i, p_i written for the own row, and
// the p_i . Ap block partial (indexed by row group).
template <int G, typename CI>
__global__ __launch_bounds__(kBlock) void k_csr_spmv_p(int64_t n, const int64_t *__restrict__ rp,
                                                       const CI *__restrict__ col, const double *__restrict__ val,
                                                       const double *__restrict__ z, double *__restrict__ pa,
                                                       double *__restrict__ pb, double *__restrict__ ap,
                                                       PcgScalars *__restrict__ sc, const double *__restrict__ part2,
                                                       int nparts2, double *__restrict__ part1) {
    int64_t it;
    double beta, gamma;
    if (!pcg_head(sc, part2, nparts2, it, beta, gamma))
        return;
    const double *pold = (it & 1) ? pb : pa;
    double *pnew = (it & 1) ? pa : pb;
    const unsigned bid = xcd_swizzle(blockIdx.x, gridDim.x);
    const int64_t i = (int64_t)bid * (kBlock / G) + threadIdx.x / G;
    const int lane = threadIdx.x % G;
    double acc = 0.0;
    if (i < n) {
        const int64_t k1 = rp[i + 1];
        for (int64_t k0 = rp[i] + lane; k0 < k1; k0 += kUnroll * G) {
            int64_t c[kUnroll];
            double w[kUnroll], zz[kUnroll], pp[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                const int64_t k = k0 + u * G;
                c[u] = k < k1 ? (int64_t)col[k] : -1;
            }
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                if (c[u] >= 0) {
                    w[u] = val[k0 + u * G];
                    zz[u] = z[c[u]];
                    pp[u] = pold[c[u]];
                }
            }
#pragma unroll
            for (int u = 0; u < kUnroll; ++u)
                if (c[u] >= 0)
                    acc += w[u] * (zz[u] + beta * pp[u]);
        }
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1)
        acc += __shfl_xor(acc, off, G);
    double loc[1] = {0.0};
    if (lane == 0 && i < n) {
        const double q = z[i] + beta * pold[i];
        pnew[i] = q;
        ap[i] = acc;
        loc[0] = q * acc;
    }
    block_sum<1>(loc);
    if (threadIdx.x == 0)
        part1[bid] = loc[0];
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

// Lanes per row: about four entries per lane, 2..64.
int csr_group(int64_t n, int64_t nnz, int requested) {
    if (requested > 0)
        return requested;
    const double per_row = (double)nnz / (double)std::max<int64_t>(1, n);
    int G = 2;
    while (G < 64 && 2 * G <= per_row / 4.0)
        G *= 2;
    return G;
}

template <int G, typename CI>
void launch_csr_iteration(int64_t n, const int64_t *rp, const CI *col, const double *val, double *x,
                          const CsrWork &w, hipStream_t s) {
    const int nparts2 = (int)nblocks(n, kBlock);
    hipLaunchKernelGGL((k_csr_spmv_p<G, CI>), dim3(nblocks(n, kBlock / G)), dim3(kBlock), 0, s, n, rp, col, val,
                       w.z, w.p0, w.p1, w.ap, w.sc, w.part2, nparts2, w.part1);
    hipLaunchKernelGGL(k_update, dim3(nblocks(n, kBlock)), dim3(kBlock), 0, s, n, x, w.r, w.z, w.p0, w.p1, w.ap,
                       w.inv, w.sc, w.part1, (int)nblocks(n, kBlock / G), w.part2);
}

template <typename CI>
void launch_csr_iteration_any(int G, int64_t n, const int64_t *rp, const CI *col, const double *val, double *x,
                              const CsrWork &w, hipStream_t s) {
    switch (G) {
    case 2: launch_csr_iteration<2, CI>(n, rp, col, val, x, w, s); break;
    case 4: launch_csr_iteration<4, CI>(n, rp, col, val, x, w, s); break;
    case 8: launch_csr_iteration<8, CI>(n, rp, col, val, x, w, s); break;
    case 16: launch_csr_iteration<16, CI>(n, rp, col, val, x, w, s); break;
    case 32: launch_csr_iteration<32, CI>(n, rp, col, val, x, w, s); break;
    default: launch_csr_iteration<64, CI>(n, rp, col, val, x, w, s); break;
    }
}

// Jacobi-PCG from x = 0 on device arrays; x receives the solution.
template <typename CI>
fx_status csr_pcg(hipStream_t s, int64_t n, const int64_t *rp, const CI *col, const double *val, const double *b,
                  double *x, const CsrWork &w, const fx_solve_opts &o, int G, fx_solve_report *rep) {
    const auto t0 = std::chrono::steady_clock::now();
    hipLaunchKernelGGL((k_csr_jacobi<CI>), dim3(nblocks(n, kBlock)), dim3(kBlock), 0, s, n, rp, col, val, w.inv);
    hipLaunchKernelGGL(k_csr_scalars, dim3(1), dim3(64), 0, s, w.sc, o.tol, o.max_iter, o.stop_mode);
    hipLaunchKernelGGL(k_init, dim3(nblocks(n, kBlock)), dim3(kBlock), 0, s, n, b, x, w.r, w.z, w.p0, w.inv,
                       w.part2);
    FX_HIP(hipGetLastError());
    PcgScalars h{};
    const int code = iterate_chunks(
        s, w.sc, o.tol, o.max_iter,
        [&](int chunk) {
            for (int k = 0; k < chunk; ++k)
                launch_csr_iteration_any<CI>(G, n, rp, col, val, x, w, s);
            FX_HIP(hipGetLastError());
        },
        &h);
    fx_solve_report R{};
    R.iterations = iterations_done(h);
    R.final_criterion = h.crit;
    R.spmv_group = G;
    R.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rep)
        *rep = R;
    return code == 2 ? FX_ERR_BREAKDOWN : code == 3 ? FX_ERR_NOT_CONVERGED : FX_OK;
}

}  // namespace

// Resolves this device's scratch address once (the device must be
// current); called by fx_ctx_create.
void csr_scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices)
        throw Error{FX_ERR_INVALID, "device index beyond the scratch table"};
    std::lock_guard<std::mutex> lk(g_scratch_mx);
    if (g_scratch_ptr[device])
        return;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        throw Error{FX_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)};
    g_scratch_ptr[device] = ptr;
}

size_t csr_scratch_bytes() { return kScratchBytes; }

fx_status solve_csr_host(fx_ctx *ctx, int64_t n, const int64_t *row_ptr, const int64_t *col_idx,
                         const double *values, const double *rhs, double *x_out, const fx_solve_opts &o,
                         fx_solve_report *rep) {
    hipStream_t s = ctx->stream;
    if (row_ptr[0] != 0)
        throw Error{FX_ERR_INVALID, "fx_solve_csr: row_ptr[0] must be 0"};
    for (int64_t i = 0; i < n; ++i)
        if (row_ptr[i + 1] < row_ptr[i])
            throw Error{FX_ERR_INVALID, "fx_solve_csr: row_ptr decreases at row " + std::to_string(i)};
    const int64_t nnz = row_ptr[n];
    const bool narrow = n < ((int64_t)1 << 31);
    std::vector<int32_t> col32(narrow ? (size_t)nnz : 0);
    for (int64_t i = 0; i < n; ++i) {
        int64_t prev = -1;
        for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k) {
            const int64_t c = col_idx[k];
            if (c <= prev || c >= n)
                throw Error{FX_ERR_INVALID, "fx_solve_csr: row " + std::to_string(i) +
                                                " has a column out of range or not strictly ascending"};
            prev = c;
            if (narrow)
                col32[k] = (int32_t)c;
        }
    }
    const int G = csr_group(n, nnz, o.spmv_group);
    CsrWork w{};
    const size_t work_bytes = csr_work_layout(n, G, nullptr, &w);
    PoolBuf<char> work;
    PoolBuf<int64_t> rp;
    PoolBuf<double> val, b, x;
    PoolBuf<int32_t> c32;
    PoolBuf<int64_t> c64;
    work.alloc(ctx, work_bytes);
    csr_work_layout(n, G, work.p, &w);
    rp.alloc(ctx, n + 1);
    val.alloc(ctx, std::max<int64_t>(1, nnz));
    b.alloc(ctx, n);
    x.alloc(ctx, n);
    FX_HIP(hipMemcpyAsync(rp.p, row_ptr, (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
    FX_HIP(hipMemcpyAsync(val.p, values, nnz * sizeof(double), hipMemcpyHostToDevice, s));
    FX_HIP(hipMemcpyAsync(b.p, rhs, n * sizeof(double), hipMemcpyHostToDevice, s));
    fx_status st;
    if (narrow) {
        c32.alloc(ctx, std::max<int64_t>(1, nnz));
        FX_HIP(hipMemcpyAsync(c32.p, col32.data(), nnz * sizeof(int32_t), hipMemcpyHostToDevice, s));
        st = csr_pcg<int32_t>(s, n, rp.p, c32.p, val.p, b.p, x.p, w, o, G, rep);
    } else {
        c64.alloc(ctx, std::max<int64_t>(1, nnz));
        FX_HIP(hipMemcpyAsync(c64.p, col_idx, nnz * sizeof(int64_t), hipMemcpyHostToDevice, s));
        st = csr_pcg<int64_t>(s, n, rp.p, c64.p, val.p, b.p, x.p, w, o, G, rep);
    }
    FX_HIP(hipMemcpyAsync(x_out, x.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
    FX_HIP(hipStreamSynchronize(s));  // before the pool blocks go back
    return st;
}

fx_status solve_csr_noalloc(fx_ctx *ctx, int64_t n, int64_t nnz, const int64_t *row_ptr_dev,
                            const int64_t *col_dev, const double *val_dev, const double *rhs_dev, double *x_dev,
                            const fx_solve_opts &o, fx_solve_report *rep) {
    hipStream_t s = ctx->stream;
    void *base = ctx->device >= 0 && ctx->device < kMaxDevices ? g_scratch_ptr[ctx->device] : nullptr;
    if (!base)
        throw Error{FX_ERR_INVALID, "fx_solve_csr_noalloc: no scratch memory resolved for this device"};
    const int G = csr_group(n, nnz, o.spmv_group);
    CsrWork w{};
    const size_t need = csr_work_layout(n, G, nullptr, &w);
    if (need > kScratchBytes)
        throw Error{FX_ERR_INVALID, "fx_solve_csr_noalloc: " + std::to_string(n) + " rows need " +
                                        std::to_string(need) + " bytes of work memory, more than the " +
                                        std::to_string(kScratchBytes) + "-byte scratch array holds"};
    csr_work_layout(n, G, (char *)base, &w);
    // structure check before any gather
    hipLaunchKernelGGL(k_csr_clear_flag, dim3(1), dim3(64), 0, s, w.flag);
    hipLaunchKernelGGL((k_csr_check<int64_t>), dim3(nblocks(std::max<int64_t>(n, 1), kBlock)), dim3(kBlock), 0, s,
                       n, nnz, row_ptr_dev, col_dev, w.flag);
    FX_HIP(hipGetLastError());
    int32_t flag = 0;
    FX_HIP(hipMemcpyAsync(&flag, w.flag, sizeof flag, hipMemcpyDeviceToHost, s));
    FX_HIP(hipStreamSynchronize(s));
    if (flag & 1)
        throw Error{FX_ERR_INVALID, "fx_solve_csr_noalloc: row_ptr must start at 0, not decrease and end at nnz"};
    if (flag & 2)
        throw Error{FX_ERR_INVALID,
                    "fx_solve_csr_noalloc: a column index is out of range or not strictly ascending in its row"};
    const fx_status st = csr_pcg<int64_t>(s, n, row_ptr_dev, col_dev, val_dev, rhs_dev, x_dev, w, o, G, rep);
    FX_HIP(hipStreamSynchronize(s));
    return st;
}

}  // namespace fx
