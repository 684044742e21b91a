// fx_assemble.hip -- deterministic global assembly and the step's vector kernels.
//
// Reference:
//   triplet scatter with |v| > 1e-12 and per-element max diagonal
//                                      src/solver/assembler.rs:119-148
//   duplicate summation (TriMat::to_csr) src/solver/step/static_step.rs:198
//   f_int scatter-add                  src/solver/assembler.rs:229-237
//   residual, BC-masked norms          src/solver/step/static_step.rs:138-164
//   penalty on K and on the residual   src/solver/step/static_step.rs:183-195
//   Jacobi inverse diagonal            src/solver/preconditioner.rs:39-47
//
// Assembly is a gather, not a scatter: the host symbolic phase lists, for
// every 3x3 node block of the structural pattern, the element contributions
// in element order.  One lane owns one block: it reads each contribution's
// nine values (72 contiguous bytes) once and sums the kept ones per slot in
// list order, which is the order in which the reference's triplets would be
// summed when elements are visited by ascending ID.  No floating-point
// atomics (the max diagonal is an integer max over non-negative doubles'
// bits).  The structural
// pattern stays the compute pattern; slots whose contributions were all
// dropped hold 0.0 and keep = 0, so the exported CSR matches the filtered
// one while SpMV sums are unchanged (0 * x adds nothing).
#include "fx_reduce.h"
#include "fx_sell.h"

namespace fx {
namespace {

constexpr int kWave = 64;

__global__ __launch_bounds__(256) void k_assemble(int64_t nblk, const int32_t *__restrict__ nb_ptr,
                                                  const int32_t *__restrict__ nb_col,
                                                  const int32_t *__restrict__ blk_row,
                                                  const int32_t *__restrict__ blk_cptr,
                                                  const int32_t *__restrict__ blk_cbase,
                                                  const double *__restrict__ ke, double *__restrict__ kval,
                                                  uint8_t *__restrict__ keep,
                                                  unsigned long long *__restrict__ max_diag_bits,
                                                  const int32_t *__restrict__ skip) {
    if (skip && *skip)
        return;
    __shared__ double s_max[256 / kWave];
    const int64_t blk = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double dmax = 0.0;
    if (blk < nblk) {
        const int a = blk_row[blk];
        const int b0 = nb_ptr[a], deg = nb_ptr[a + 1] - b0;
        const bool diag = nb_col[blk] == a;
        double sum[9];
        bool kept[9];
#pragma unroll
        for (int e = 0; e < 9; ++e) {
            sum[e] = 0.0;
            kept[e] = false;
        }
        const int c1 = blk_cptr[blk + 1];
        for (int c = blk_cptr[blk]; c < c1; ++c) {
            const double *v = ke + blk_cbase[c];  // the element's 3x3 block, row-major
            double w[9];
#pragma unroll
            for (int e = 0; e < 9; ++e)
                w[e] = v[e];
#pragma unroll
            for (int e = 0; e < 9; ++e)
                if (fabs(w[e]) > kDropTol) {
                    sum[e] += w[e];
                    kept[e] = true;
                    if (diag && e % 4 == 0)  // (p, p) of the diagonal block
                        dmax = fmax(dmax, fabs(w[e]));
                }
        }
        const int64_t base = 9 * (int64_t)b0 + 3 * (blk - b0);  // slot (a, p=0, k, q=0)
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                kval[base + (int64_t)p * 3 * deg + q] = sum[3 * p + q];
                keep[base + (int64_t)p * 3 * deg + q] = kept[3 * p + q];
            }
    }
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1)
        dmax = fmax(dmax, __shfl_xor(dmax, off, kWave));
    if (threadIdx.x % kWave == 0)
        s_max[threadIdx.x / kWave] = dmax;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = s_max[0];
        for (int w = 1; w < 256 / kWave; ++w)
            m = fmax(m, s_max[w]);
        if (m > 0.0)
            atomicMax(max_diag_bits, (unsigned long long)__double_as_longlong(m));
    }
}

// f_int[3a+p] = sum over the elements of node a, in element order.
__global__ void k_gather_f(int64_t n_dof, const int32_t *__restrict__ cptr, const int32_t *__restrict__ cbase,
                           const double *__restrict__ fe, double *__restrict__ f) {
    int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= n_dof)
        return;
    const int a = (int)(d / 3), p = (int)(d % 3);
    double s = 0.0;
    for (int c = cptr[a]; c < cptr[a + 1]; ++c)
        s += fe[(int64_t)cbase[c] + p];
    f[d] = s;
}

// Nodal averages of the recovered strain and stress: the (element, node)
// contributions of node a in element order (the f_int contribution lists,
// whose bases count 3 values per contribution against 12 here), divided by
// their count (static_step.rs:362-377).
__global__ void k_gather_recover(int64_t n_nodes, const int32_t *__restrict__ cptr,
                                 const int32_t *__restrict__ cbase, const double *__restrict__ se,
                                 double *__restrict__ S, double *__restrict__ E) {
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n_nodes)
        return;
    double sum[12];
#pragma unroll
    for (int k = 0; k < 12; ++k)
        sum[k] = 0.0;
    const int c0 = cptr[a], c1 = cptr[a + 1];
    for (int c = c0; c < c1; ++c) {
        const double *v = se + 4 * (int64_t)cbase[c];
#pragma unroll
        for (int k = 0; k < 12; ++k)
            sum[k] += v[k];
    }
    const double cnt = (double)(c1 - c0);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        E[6 * a + k] = c1 > c0 ? sum[k] / cnt : 0.0;
        S[6 * a + k] = c1 > c0 ? sum[6 + k] / cnt : 0.0;
    }
}

__global__ void k_jacobi(int64_t n, const int32_t *__restrict__ diag_slot, const double *__restrict__ kval,
                         double *__restrict__ inv, float *__restrict__ inv32) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    double d = kval[diag_slot[i]];
    double v = fabs(d) > kJacobiGuard ? 1.0 / d : 1.0;
    inv[i] = v;
    if (inv32)
        inv32[i] = (float)v;
}

// FP32 copy of K plus, per node, the 3x3 block C_a[p][q] = sum over the
// node row's blocks b of (K_ab - fl32(K_ab))[p][q]: the working operator
// fl32(K) + blockdiag(C) has exactly K's node-row block sums, so the rounding
// error annihilates rigid translations and stays small on the smooth modes
// that PCG resolves last.  One thread per DOF row, columns in slot order.
__global__ void k_convert32(int64_t n_dof, const int32_t *__restrict__ nb_ptr, const double *__restrict__ a,
                            float *__restrict__ b, float *__restrict__ corr) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_dof)
        return;
    const int64_t node = i / 3;
    const int p = (int)(i % 3);
    const int64_t b0 = nb_ptr[node];
    const int deg = nb_ptr[node + 1] - (int)b0;
    const int64_t base = 9 * b0 + (int64_t)p * 3 * deg;
    double c0 = 0.0, c1 = 0.0, c2 = 0.0;
    for (int k = 0; k < deg; ++k) {
        const double v0 = a[base + 3 * k], v1 = a[base + 3 * k + 1], v2 = a[base + 3 * k + 2];
        const float f0 = (float)v0, f1 = (float)v1, f2 = (float)v2;
        b[base + 3 * k] = f0;
        b[base + 3 * k + 1] = f1;
        b[base + 3 * k + 2] = f2;
        c0 += v0 - (double)f0;
        c1 += v1 - (double)f1;
        c2 += v2 - (double)f2;
    }
    corr[9 * node + 3 * p] = (float)c0;
    corr[9 * node + 3 * p + 1] = (float)c1;
    corr[9 * node + 3 * p + 2] = (float)c2;
}

// The Jacobi-scaled operator in FP16: fl16(s_i K_ij s_j) with s = sqrt(inv_diag)
// (1/sqrt(K_ii), or 1 where the Jacobi guard applies), so its diagonal is 1
// and its entries lie in [-1, 1].  Node-block layout (one thread per DOF row)
// for the lane-group operator ...
__global__ void k_scaled16(int64_t n_dof, const int32_t *__restrict__ nb_ptr, const int32_t *__restrict__ nb_col,
                           const double *__restrict__ a, const double *__restrict__ inv, _Float16 *__restrict__ b) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_dof)
        return;
    const int64_t node = i / 3;
    const int p = (int)(i % 3);
    const int64_t b0 = nb_ptr[node];
    const int deg = nb_ptr[node + 1] - (int)b0;
    const int64_t base = 9 * b0 + (int64_t)p * 3 * deg;
    const double si = sqrt(inv[i]);
    for (int k = 0; k < deg; ++k) {
        const int64_t c = 3 * (int64_t)nb_col[b0 + k];
#pragma unroll
        for (int q = 0; q < 3; ++q)
            b[base + 3 * k + q] = (_Float16)(si * a[base + 3 * k + q] * sqrt(inv[c + q]));
    }
}

// ... and in the sliced-ELL arrangement of fx_sell.h with FP16 records: per
// block column kk and lane, entries e < 8 of the 3x3 block as one 16-byte
// record at rec8[(kk * 64 + lane) * 8 + e] and entry 8 at rec1[kk * 64 + lane]
// (two loads per block instead of nine); one 64-thread block per slice,
// padding blocks are zeros.
__global__ void k_sell_scaled16(int64_t n_nodes, const int32_t *__restrict__ nb_ptr,
                                const int32_t *__restrict__ nb_col, const int32_t *__restrict__ sl_ptr,
                                const double *__restrict__ a, const double *__restrict__ inv,
                                _Float16 *__restrict__ rec8, _Float16 *__restrict__ rec1) {
    const int64_t sl = blockIdx.x;
    const int lane = threadIdx.x;
    const int64_t node = sl * kSliceRows + lane;
    const int k0 = sl_ptr[sl], w = sl_ptr[sl + 1] - k0;
    int64_t b0 = 0;
    int deg = 0;
    if (node < n_nodes) {
        b0 = nb_ptr[node];
        deg = nb_ptr[node + 1] - (int)b0;
    }
    for (int k = 0; k < w; ++k) {
        const size_t kk = (size_t)k0 + k;
        const int64_t c = k < deg ? 3 * (int64_t)nb_col[b0 + k] : 0;
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                _Float16 v = (_Float16)0.0f;
                if (k < deg)
                    v = (_Float16)(sqrt(inv[3 * node + p]) * a[9 * b0 + (int64_t)p * 3 * deg + 3 * k + q] *
                                   sqrt(inv[c + q]));
                const int e = 3 * p + q;
                if (e < 8)
                    rec8[(kk * kSliceRows + lane) * 8 + e] = v;
                else
                    rec1[kk * kSliceRows + lane] = v;
            }
    }
}

// Sliced-ELL copy (fx_sell.h) of the node-block values: one 64-thread block
// per slice, one thread per node row; padding is written as zeros.
template <typename ST, typename VT>
__global__ void k_sell_convert(int64_t n_nodes, const int32_t *__restrict__ nb_ptr, const int32_t *__restrict__ sl_ptr,
                               const ST *__restrict__ src, VT *__restrict__ dst) {
    const int64_t sl = blockIdx.x;
    const int lane = threadIdx.x;
    const int64_t a = sl * kSliceRows + lane;
    const int k0 = sl_ptr[sl], w = sl_ptr[sl + 1] - k0;
    int64_t b0 = 0;
    int deg = 0;
    if (a < n_nodes) {
        b0 = nb_ptr[a];
        deg = nb_ptr[a + 1] - (int)b0;
    }
    for (int k = 0; k < w; ++k) {
        const size_t kk = (size_t)k0 + k;
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                VT v = (VT)0;
                if (k < deg)
                    v = (VT)src[9 * b0 + (int64_t)p * 3 * deg + 3 * k + q];
                dst[sell_index<VT>(kk, 3 * p + q, lane)] = v;
            }
    }
}

// Symmetric sliced-ELL FP32 operator (one slice of 64 node rows per block):
// the row's upper blocks (from its diagonal on) rounded to FP32 into uval in
// the sval layout, zeros in the padding; and the row-sum correction of this
// operator, C_a = sum_b (K_ab - Kt_ab), where Kt_ab = fl32(K_ab) for the upper
// blocks and fl32(K_ba)^T -- the owner row's stored record, transposed -- for
// the lower ones.
__global__ void k_ssym_convert(int64_t n_nodes, const int32_t *__restrict__ nb_ptr,
                               const int32_t *__restrict__ ndiag, const int32_t *__restrict__ su_ptr,
                               const int32_t *__restrict__ sll_ptr, const int32_t *__restrict__ lcol,
                               const int32_t *__restrict__ lblk, const double *__restrict__ kval,
                               float *__restrict__ uval, float *__restrict__ kcs) {
    const int64_t q = blockIdx.x;
    const int lane = threadIdx.x;
    const int64_t a = q * kSliceRows + lane;
    int64_t b0 = 0;
    int deg = 0, da = 0;
    if (a < n_nodes) {
        b0 = nb_ptr[a];
        deg = nb_ptr[a + 1] - (int)b0;
        da = ndiag[a];
    }
    double C[9];
    for (int e = 0; e < 9; ++e)
        C[e] = 0.0;
    const int k0 = su_ptr[q], wu = su_ptr[q + 1] - k0;
    for (int k = 0; k < wu; ++k) {
        const size_t kk = (size_t)k0 + k;
        for (int p = 0; p < 3; ++p)
            for (int c = 0; c < 3; ++c) {
                float v = 0.f;
                if (a < n_nodes && da + k < deg) {
                    const double v64 = kval[9 * b0 + (int64_t)p * 3 * deg + 3 * (da + k) + c];
                    v = (float)v64;
                    C[3 * p + c] += v64 - (double)v;
                }
                uval[(kk * 9 + 3 * p + c) * kSliceRows + lane] = v;
            }
    }
    const int l0 = sll_ptr[q], wl = sll_ptr[q + 1] - l0;
    for (int j = 0; j < wl && a < n_nodes && j < da; ++j) {
        const size_t at = ((size_t)l0 + j) * kSliceRows + lane;
        const int cn = lcol[at];
        const int64_t bc0 = nb_ptr[cn];
        const int degc = nb_ptr[cn + 1] - (int)bc0;
        const int t = lblk[at] - (int)bc0;
        for (int p = 0; p < 3; ++p)
            for (int c = 0; c < 3; ++c) {
                const double kac = kval[9 * b0 + (int64_t)p * 3 * deg + 3 * j + c];
                const double kca = kval[9 * bc0 + (int64_t)c * 3 * degc + 3 * t + p];  // K(c, a) entry (c, p)
                C[3 * p + c] += kac - (double)(float)kca;
            }
    }
    if (a < n_nodes)
        for (int e = 0; e < 9; ++e)
            kcs[9 * a + e] = (float)C[e];
}

__global__ void k_residual(int64_t n, const double *__restrict__ fe, const double *__restrict__ fi,
                           double *__restrict__ r) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        r[i] = fe[i] - fi[i];
}

__global__ void k_fill(int64_t n, double *__restrict__ d, double v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        d[i] = v;
}

__global__ void k_scatter(int64_t n, const int32_t *__restrict__ idx, const double *__restrict__ val,
                          double *__restrict__ dst) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        dst[idx[i]] = val[i];
}

// Penalty: one lane per distinct BC DOF; its records are applied in deck
// order: K[d,d] += penalty once per record, rhs[d] += penalty*(t - u[d]).
// md_bits != nullptr: the factor is kPenaltyScale x the max_diag the last
// k_assemble left there (the product the host forms otherwise), and no
// penalty is added when that is not positive.
__global__ void k_penalty(int64_t n_unique, const int32_t *__restrict__ ud_dof, const int32_t *__restrict__ ud_ptr,
                          const int32_t *__restrict__ rec_order, const double *__restrict__ rec_target,
                          const int32_t *__restrict__ diag_slot, double penalty, double *__restrict__ kval,
                          uint8_t *__restrict__ keep, double *__restrict__ rhs, const double *__restrict__ u,
                          const unsigned long long *__restrict__ md_bits, const int32_t *__restrict__ skip) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_unique || (skip && *skip))
        return;
    if (md_bits) {
        const double md = __longlong_as_double((long long)*md_bits);
        if (!(md > 0.0))
            return;
        penalty = md * kPenaltyScale;
    }
    const int d = ud_dof[j];
    const int slot = diag_slot[d];
    double k = kval ? kval[slot] : 0.0;
    double r = rhs ? rhs[d] : 0.0;
    const double ud = u[d];
    for (int c = ud_ptr[j]; c < ud_ptr[j + 1]; ++c) {
        k += penalty;
        r += penalty * (rec_target[rec_order[c]] - ud);
    }
    if (kval) {  // kval == nullptr: one more right-hand side of a K already penalised
        kval[slot] = k;
        keep[slot] = 1;
    }
    if (rhs)
        rhs[d] = r;
}

// Sums of r^2 and f^2 over DOFs with mask == 0; two-level deterministic
// reduction (fixed per-block order, then the last block sums the partials
// in block order).
constexpr int kRedThreads = 256;

__global__ __launch_bounds__(kRedThreads) void k_masked_norms(int64_t n, const double *__restrict__ r,
                                                             const double *__restrict__ f,
                                                             const uint8_t *__restrict__ mask,
                                                             double *__restrict__ part, uint32_t *__restrict__ cnt,
                                                             double *__restrict__ out) {
    double v[2] = {0.0, 0.0};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (!mask[i]) {
            v[0] += r[i] * r[i];
            v[1] += f[i] * f[i];
        }
    }
    double s[2];
    if (grid_sum<2>(v, part, cnt, s)) {
        out[0] = s[0];
        out[1] = s[1];
    }
}

__global__ void k_count_keep(int64_t n, const uint8_t *__restrict__ keep, unsigned long long *__restrict__ out) {
    __shared__ unsigned long long s[4];
    unsigned long long c = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        c += keep[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        c += __shfl_xor(c, off, 64);
    if (threadIdx.x % 64 == 0)
        s[threadIdx.x / 64] = c;
    __syncthreads();
    if (threadIdx.x == 0)
        atomicAdd(out, s[0] + s[1] + s[2] + s[3]);
}

__global__ void k_zero_u64(unsigned long long *__restrict__ p, const int32_t *__restrict__ skip) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && !(skip && *skip))
        *p = 0ull;
}

__global__ void k_reset_status(DevStatus *__restrict__ st) {
    if (threadIdx.x == 0 && blockIdx.x == 0)
        *st = DevStatus{0, 0, INT64_MAX};
}

__global__ void k_axpy(int64_t n, double a, const double *__restrict__ x, double *__restrict__ y) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        y[i] += a * x[i];
}

inline unsigned grid1(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

void launch_assemble(fx_mesh *m) {
    hipStream_t s = m->ctx->stream;
    // a kernel rather than a memset, so that a captured graph can hold it
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, m->max_diag_bits.p, m->k_skip);
    if (m->nblk > 0)
        hipLaunchKernelGGL(k_assemble, dim3(grid1(m->nblk, 256)), dim3(256), 0, s, m->nblk, m->nb_ptr.p,
                           m->nb_col.p, m->blk_row.p, m->blk_cptr.p, m->blk_cbase.p, m->ke.p, m->kval.p, m->keep.p,
                           m->max_diag_bits.p, m->k_skip);
    FX_HIP(hipGetLastError());
}

void launch_gather_f(fx_mesh *m) {
    hipLaunchKernelGGL(k_gather_f, dim3(grid1(m->n_dof, 256)), dim3(256), 0, m->ctx->stream, m->n_dof,
                       m->nod_cptr.p, m->nod_cbase.p, m->fe.p, m->f_int.p);
    FX_HIP(hipGetLastError());
}

void launch_gather_recover(fx_mesh *m, double *S, double *E) {
    hipLaunchKernelGGL(k_gather_recover, dim3(grid1(m->n_nodes, 256)), dim3(256), 0, m->ctx->stream, m->n_nodes,
                       m->nod_cptr.p, m->nod_cbase.p, m->se.p, S, E);
    FX_HIP(hipGetLastError());
}

void launch_jacobi(fx_mesh *m, bool want32, hipStream_t s) {
    hipLaunchKernelGGL(k_jacobi, dim3(grid1(m->n_dof, 256)), dim3(256), 0, s ? s : m->ctx->stream, m->n_dof,
                       m->diag_slot.p, m->kval.p, m->inv_diag.p, want32 ? m->inv_diag32.p : nullptr);
    FX_HIP(hipGetLastError());
}

void launch_convert32(fx_mesh *m, hipStream_t s) {
    m->kval32.alloc(m->nnz_struct);
    m->kcorr32.alloc(9 * m->n_nodes);
    hipLaunchKernelGGL(k_convert32, dim3(grid1(m->n_dof, 256)), dim3(256), 0, s ? s : m->ctx->stream, m->n_dof,
                       m->nb_ptr.p, m->kval.p, m->kval32.p, m->kcorr32.p);
    FX_HIP(hipGetLastError());
}

void launch_sell(fx_mesh *m, bool fp32, hipStream_t s) {
    if (!m->sell)
        return;
    const int64_t ns = (m->n_nodes + kSliceRows - 1) / kSliceRows;
    const size_t slots = (size_t)m->scol.n * 9;  // block columns x 64 lanes x 9 entries
    if (!s)
        s = m->ctx->stream;
    if (fp32) {
        m->sval32.alloc(slots);
        hipLaunchKernelGGL((k_sell_convert<float, float>), dim3((unsigned)ns), dim3(kSliceRows), 0, s, m->n_nodes,
                           m->nb_ptr.p, m->sl_ptr.p, m->kval32.p, m->sval32.p);
    } else {
        m->sval64.alloc(slots);
        hipLaunchKernelGGL((k_sell_convert<double, double>), dim3((unsigned)ns), dim3(kSliceRows), 0, s,
                           m->n_nodes, m->nb_ptr.p, m->sl_ptr.p, m->kval.p, m->sval64.p);
    }
    FX_HIP(hipGetLastError());
}

void launch_scaled16(fx_mesh *m, hipStream_t s) {
    if (!s)
        s = m->ctx->stream;
    if (m->sell) {
        const int64_t ns = (m->n_nodes + kSliceRows - 1) / kSliceRows;
        m->sval16.alloc((size_t)m->scol.n * 9);  // 8-entry records, then the 9th entries
        hipLaunchKernelGGL(k_sell_scaled16, dim3((unsigned)ns), dim3(kSliceRows), 0, s, m->n_nodes, m->nb_ptr.p,
                           m->nb_col.p, m->sl_ptr.p, m->kval.p, m->inv_diag.p, m->sval16.p,
                           m->sval16.p + (size_t)m->scol.n * 8);
    } else {
        m->kval16.alloc(m->nnz_struct);
        hipLaunchKernelGGL(k_scaled16, dim3(grid1(m->n_dof, 256)), dim3(256), 0, s, m->n_dof, m->nb_ptr.p,
                           m->nb_col.p, m->kval.p, m->inv_diag.p, m->kval16.p);
    }
    FX_HIP(hipGetLastError());
}

void launch_ssym(fx_mesh *m, hipStream_t s) {
    if (!m->sell || !m->ssym)
        return;
    const int64_t ns = (m->n_nodes + kSliceRows - 1) / kSliceRows;
    const size_t col = (size_t)9 * kSliceRows;  // floats of one block column
    if (!s)
        s = m->ctx->stream;
    m->uval32.alloc(((size_t)m->su_cols + 1) * col);
    m->kcorr32s.alloc((size_t)9 * m->n_nodes);
    FX_HIP(hipMemsetAsync(m->uval32.p + (size_t)m->su_cols * col, 0, col * sizeof(float), s));  // zero record
    hipLaunchKernelGGL(k_ssym_convert, dim3((unsigned)ns), dim3(kSliceRows), 0, s, m->n_nodes, m->nb_ptr.p,
                       m->ndiag.p, m->su_ptr.p, m->sll_ptr.p, m->lcol.p, m->lblk.p, m->kval.p, m->uval32.p,
                       m->kcorr32s.p);
    FX_HIP(hipGetLastError());
}

void launch_residual(fx_mesh *m, const double *f_ext, const double *f_int, double *r, int64_t n) {
    hipLaunchKernelGGL(k_residual, dim3(grid1(n, 256)), dim3(256), 0, m->ctx->stream, n, f_ext, f_int, r);
    FX_HIP(hipGetLastError());
}

void launch_fill(fx_mesh *m, double *dst, int64_t n, double v, hipStream_t s) {
    if (n <= 0)
        return;
    hipLaunchKernelGGL(k_fill, dim3(grid1(n, 256)), dim3(256), 0, s ? s : m->ctx->stream, n, dst, v);
    FX_HIP(hipGetLastError());
}

void launch_scatter(fx_mesh *m, int64_t n, const int32_t *idx, const double *val, double *dst) {
    if (n <= 0)
        return;
    hipLaunchKernelGGL(k_scatter, dim3(grid1(n, 256)), dim3(256), 0, m->ctx->stream, n, idx, val, dst);
    FX_HIP(hipGetLastError());
}

void launch_penalty(fx_mesh *m, int64_t n_unique, const int32_t *ud_dof, const int32_t *ud_ptr,
                    const double *rec_target, const int32_t *rec_order, double *rhs, const double *u) {
    if (n_unique <= 0)
        return;
    if (std::isnan(m->last_max_diag))  // assembled by fx_batch_add_system: max_diag is on the device
        m->last_max_diag = device_max_diag(m);
    double pen = m->last_max_diag * kPenaltyScale;
    hipLaunchKernelGGL(k_penalty, dim3(grid1(n_unique, 256)), dim3(256), 0, m->ctx->stream, n_unique, ud_dof,
                       ud_ptr, rec_order, rec_target, m->diag_slot.p, pen, m->kval.p, m->keep.p, rhs, u,
                       (const unsigned long long *)nullptr, (const int32_t *)nullptr);
    FX_HIP(hipGetLastError());
}

void launch_penalty_dev(fx_mesh *m, int64_t n_unique, const int32_t *ud_dof, const int32_t *ud_ptr,
                        const double *rec_target, const int32_t *rec_order, double *rhs, const double *u,
                        bool rhs_only) {
    if (n_unique <= 0)
        return;
    hipLaunchKernelGGL(k_penalty, dim3(grid1(n_unique, 256)), dim3(256), 0, m->ctx->stream, n_unique, ud_dof,
                       ud_ptr, rec_order, rec_target, m->diag_slot.p, 0.0, rhs_only ? nullptr : m->kval.p,
                       m->keep.p, rhs, u, (const unsigned long long *)m->max_diag_bits.p, m->k_skip);
    FX_HIP(hipGetLastError());
}

void launch_reset_status(fx_mesh *m) {
    hipLaunchKernelGGL(k_reset_status, dim3(1), dim3(64), 0, m->ctx->stream, m->status.p);
    FX_HIP(hipGetLastError());
}

void launch_masked_norms(fx_mesh *m, const double *r, const double *f, const uint8_t *mask, double *out2) {
    unsigned blocks = (unsigned)std::min<int64_t>(1024, std::max<int64_t>(1, (m->n_dof + 1023) / 1024));
    hipLaunchKernelGGL(k_masked_norms, dim3(blocks), dim3(kRedThreads), 0, m->ctx->stream, m->n_dof, r, f, mask,
                       m->part.p, m->counters.p, out2);
    FX_HIP(hipGetLastError());
}

void axpy_dev(fx_mesh *m, int64_t n, double a, const double *x, double *y) {
    hipLaunchKernelGGL(k_axpy, dim3(grid1(n, 256)), dim3(256), 0, m->ctx->stream, n, a, x, y);
    FX_HIP(hipGetLastError());
}

int64_t count_kept(fx_mesh *m) {
    hipStream_t s = m->ctx->stream;
    unsigned long long *d = reinterpret_cast<unsigned long long *>(m->red2.p);
    FX_HIP(hipMemsetAsync(d, 0, sizeof(unsigned long long), s));
    hipLaunchKernelGGL(k_count_keep, dim3(1024), dim3(256), 0, s, m->nnz_struct, m->keep.p, d);
    FX_HIP(hipGetLastError());
    unsigned long long h = 0;
    FX_HIP(hipMemcpyAsync(&h, d, sizeof h, hipMemcpyDeviceToHost, s));
    FX_HIP(hipStreamSynchronize(s));
    return (int64_t)h;
}

void reset_status(fx_mesh *m) {
    DevStatus st{0, 0, INT64_MAX};
    FX_HIP(hipMemcpyAsync(m->status.p, &st, sizeof st, hipMemcpyHostToDevice, m->ctx->stream));
}

void check_status(fx_mesh *m, const char *what) {
    DevStatus st;
    FX_HIP(hipMemcpyAsync(&st, m->status.p, sizeof st, hipMemcpyDeviceToHost, m->ctx->stream));
    FX_HIP(hipStreamSynchronize(m->ctx->stream));
    ++t_host_syncs;
    if (st.code != 0) {
        std::string msg = std::string(what) + ": ";
        if (st.code == FX_ERR_SINGULAR_JACOBIAN)
            msg += "Singular Jacobian";
        else
            msg += "element has no Young's modulus or Poisson's ratio at its temperature";
        msg += " (element position " + std::to_string(st.bad_element) + ")";
        reset_status(m);
        throw Error{(fx_status)st.code, msg};
    }
}

double device_max_diag(fx_mesh *m) {
    unsigned long long bits = 0;
    FX_HIP(hipMemcpyAsync(&bits, m->max_diag_bits.p, sizeof bits, hipMemcpyDeviceToHost, m->ctx->stream));
    FX_HIP(hipStreamSynchronize(m->ctx->stream));
    ++t_host_syncs;
    double d;
    memcpy(&d, &bits, sizeof d);
    return d;
}

}  // namespace fx
