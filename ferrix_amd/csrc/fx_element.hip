// fx_element.hip -- element stiffness and internal force on gfx950.
//
// Reference: src/solver/mesh_lib/elements/element.rs:217-447 (stiffness and
// internal force per Gauss point), :588-606 (B matrix), material.rs:36-68 and
// :202-230 (LUT and D), c3d4.rs / c3d20.rs (rules and shape functions).
//
// Arithmetic contract (checked bit-for-bit against the CPU oracle):
//  * J = dN_local * X^T summed over nodes in order from 0.0; det by cofactor
//    expansion along row 0; J^-1 = adj(J) / det; dN_global = J^-1 dN_local.
//  * K_el += ((B^T D) B) * w with the dense index order of the contraction;
//    the terms skipped here are the structural zeros of B and D, which never
//    change a sum that starts from +0.0 (a +0.0-seeded IEEE sum is never -0,
//    and adding +/-0 to it is exact).  Likewise eps = B u, sigma = D eps and
//    f_el += (B^T sigma) * w.
//  * built with -ffp-contract=off: no multiply-add is fused.
//
// Layout of the outputs (consumed by the deterministic assembly):
//   ke[block_offset + e*nn*nn*9 + (a*nn + b)*9 + 3p + q] = K_el(3a+p, 3b+q)
//   fe[block_offset + e*3nn + 3a + p]                     = f_el(3a+p)
//
// C3D4: one lane per element (1 Gauss point, 144 outputs).
// C3D20: one 256-lane workgroup per element; lanes 0..26 evaluate J, J^-1,
// dN_global, w and D at their Gauss point into LDS (13 KB), then each lane
// owns node pairs (a, b) and accumulates the 3x3 block over the 27 points.
#include "fx_internal.h"

#include <cstdlib>

namespace fx {
namespace {

__device__ __forceinline__ double lut(int n, const double *t, const double *v, double temp) {
    if (n == 1 || temp <= t[0])
        return v[0];
    if (temp >= t[n - 1])
        return v[n - 1];
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        int mid = lo + (hi - lo) / 2;
        if (t[mid] <= temp)
            lo = mid;
        else
            hi = mid;
    }
    double factor = (temp - t[lo]) / (t[hi] - t[lo]);
    return v[lo] + (v[hi] - v[lo]) * factor;
}

// Isotropic D has three distinct values: d1 = (1-nu) f, d2 = nu f,
// d3 = ((1-2nu)/2) f with f = E / ((1+nu)(1-2nu)).
__device__ __forceinline__ bool d_coeffs(const MatLut &m, double temp, double &d1, double &d2, double &d3) {
    if (m.n_e <= 0 || m.n_nu <= 0)
        return false;
    double e = lut(m.n_e, m.t_e, m.v_e, temp);
    double nu = lut(m.n_nu, m.t_nu, m.v_nu, temp);
    double factor = e / ((1.0 + nu) * (1.0 - 2.0 * nu));
    double c1 = 1.0 - nu, c2 = nu, c3 = (1.0 - 2.0 * nu) / 2.0;
    d1 = c1 * factor;
    d2 = c2 * factor;
    d3 = c3 * factor;
    return true;
}

__device__ __forceinline__ void fail(DevStatus *st, int code, int64_t gpos) {
    st->code = code;  // every failing lane stores the same kind of value
    atomicMin(reinterpret_cast<unsigned long long *>(&st->bad_element), (unsigned long long)gpos);
}

__device__ __forceinline__ bool jinv(const double (&J)[9], double *Ji, double &det);

// J from dN (row-major 3 x nn) and coordinates xc (nn x 3); returns det, Jinv.
template <int NN>
__device__ __forceinline__ bool jac(const double *dN, const double *xc, double *Ji, double &det) {
    double J[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double s = 0.0;
#pragma unroll
            for (int i = 0; i < NN; ++i)
                s += dN[r * NN + i] * xc[3 * i + c];
            J[3 * r + c] = s;
        }
    return jinv(J, Ji, det);
}

// det J and J^-1 = adj(J) / det from J (row-major); false when det == 0.
__device__ __forceinline__ bool jinv(const double (&J)[9], double *Ji, double &det) {
    double m0 = J[4] * J[8] - J[5] * J[7];
    double m1 = J[3] * J[8] - J[5] * J[6];
    double m2 = J[3] * J[7] - J[4] * J[6];
    det = J[0] * m0 - J[1] * m1 + J[2] * m2;
    if (det == 0.0)
        return false;
    Ji[0] = (J[4] * J[8] - J[5] * J[7]) / det;
    Ji[1] = (J[2] * J[7] - J[1] * J[8]) / det;
    Ji[2] = (J[1] * J[5] - J[2] * J[4]) / det;
    Ji[3] = (J[5] * J[6] - J[3] * J[8]) / det;
    Ji[4] = (J[0] * J[8] - J[2] * J[6]) / det;
    Ji[5] = (J[2] * J[3] - J[0] * J[5]) / det;
    Ji[6] = (J[3] * J[7] - J[4] * J[6]) / det;
    Ji[7] = (J[1] * J[6] - J[0] * J[7]) / det;
    Ji[8] = (J[0] * J[4] - J[1] * J[3]) / det;
    return true;
}

// 3x3 block of ((B^T D) B) for node a (dxa,dya,dza) and node b, before
// scaling by w.  T rows are the nonzero entries of row 3a+p of B^T D.
__device__ __forceinline__ void kblock(double dxa, double dya, double dza, double dxb, double dyb, double dzb,
                                       double d1, double d2, double d3, double *M) {
    // p = 0: T = [dx d1, dx d2, dx d2, dy d3, 0, dz d3]
    {
        double t0 = 0.0 + dxa * d1, t1 = 0.0 + dxa * d2, t2 = 0.0 + dxa * d2;
        double t3 = 0.0 + dya * d3, t5 = 0.0 + dza * d3;
        M[0] = ((0.0 + t0 * dxb) + t3 * dyb) + t5 * dzb;
        M[1] = (0.0 + t1 * dyb) + t3 * dxb;
        M[2] = (0.0 + t2 * dzb) + t5 * dxb;
    }
    // p = 1: T = [dy d2, dy d1, dy d2, dx d3, dz d3, 0]
    {
        double t0 = 0.0 + dya * d2, t1 = 0.0 + dya * d1, t2 = 0.0 + dya * d2;
        double t3 = 0.0 + dxa * d3, t4 = 0.0 + dza * d3;
        M[3] = (0.0 + t0 * dxb) + t3 * dyb;
        M[4] = ((0.0 + t1 * dyb) + t3 * dxb) + t4 * dzb;
        M[5] = (0.0 + t2 * dzb) + t4 * dyb;
    }
    // p = 2: T = [dz d2, dz d2, dz d1, 0, dy d3, dx d3]
    {
        double t0 = 0.0 + dza * d2, t1 = 0.0 + dza * d2, t2 = 0.0 + dza * d1;
        double t4 = 0.0 + dya * d3, t5 = 0.0 + dxa * d3;
        M[6] = (0.0 + t0 * dxb) + t5 * dzb;
        M[7] = (0.0 + t1 * dyb) + t4 * dzb;
        M[8] = ((0.0 + t2 * dzb) + t4 * dyb) + t5 * dxb;
    }
}

// ---------------------------------------------------------------------------
// C3D4: one lane per element; the workgroup size (elements per workgroup,
// 64..kC4MaxBlock) is the mesh's c3d4_block, which fx_mesh_autotune_elements
// picks by timing.  An element's arithmetic does not depend on it.
// ---------------------------------------------------------------------------
constexpr int kC4MaxBlock = 512;
template <bool DO_K, bool DO_F>
__global__ __launch_bounds__(kC4MaxBlock) void k_c3d4(int64_t n_el, const int32_t *__restrict__ conn,
                                              const int32_t *__restrict__ mat, const int64_t *__restrict__ gpos,
                                              const double *__restrict__ xyz, const double *__restrict__ temp,
                                              const double *__restrict__ u, int nlgeom,
                                              const MatLut *__restrict__ mats, const double *__restrict__ gN,
                                              const double *__restrict__ gdN, const double *__restrict__ gw,
                                              double *__restrict__ ke, double *__restrict__ fe,
                                              DevStatus *__restrict__ st, const int32_t *__restrict__ skip) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_el || (skip && *skip))
        return;
    double xc[12], ue[12], T[4], dN[12];
    int nd[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        nd[i] = conn[4 * e + i];
        T[i] = temp[nd[i]];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double x = xyz[3 * (int64_t)nd[i] + c];
            double uu = u ? u[3 * (int64_t)nd[i] + c] : 0.0;
            ue[3 * i + c] = uu;
            xc[3 * i + c] = nlgeom ? x + uu : x;
        }
    }
#pragma unroll
    for (int i = 0; i < 12; ++i)
        dN[i] = gdN[i];
    double Ji[9], det;
    if (!jac<4>(dN, xc, Ji, det)) {
        fail(st, FX_ERR_SINGULAR_JACOBIAN, gpos[e]);
        return;
    }
    double dg[12];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < 3; ++k)
                s += Ji[3 * r + k] * dN[4 * k + i];
            dg[4 * r + i] = s;
        }
    double w = fabs(det) * gw[0];
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        t += gN[i] * T[i];
    double d1, d2, d3;
    if (!d_coeffs(mats[mat[e]], t, d1, d2, d3)) {
        fail(st, FX_ERR_INVALID, gpos[e]);
        return;
    }
    if (DO_K) {
        double *out = ke + e * 144;
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                double M[9];
                kblock(dg[a], dg[4 + a], dg[8 + a], dg[b], dg[4 + b], dg[8 + b], d1, d2, d3, M);
#pragma unroll
                for (int k = 0; k < 9; ++k)
                    out[(4 * a + b) * 9 + k] = 0.0 + M[k] * w;
            }
    }
    if (DO_F) {
        double e0 = 0.0, e1 = 0.0, e2 = 0.0, e3 = 0.0, e4 = 0.0, e5 = 0.0;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            double dx = dg[a], dy = dg[4 + a], dz = dg[8 + a];
            double ux = ue[3 * a], uy = ue[3 * a + 1], uz = ue[3 * a + 2];
            e0 += dx * ux;
            e1 += dy * uy;
            e2 += dz * uz;
            e3 += dy * ux;
            e3 += dx * uy;
            e4 += dz * uy;
            e4 += dy * uz;
            e5 += dz * ux;
            e5 += dx * uz;
        }
        double s0 = ((0.0 + d1 * e0) + d2 * e1) + d2 * e2;
        double s1 = ((0.0 + d2 * e0) + d1 * e1) + d2 * e2;
        double s2 = ((0.0 + d2 * e0) + d2 * e1) + d1 * e2;
        double s3 = 0.0 + d3 * e3, s4 = 0.0 + d3 * e4, s5 = 0.0 + d3 * e5;
        double *out = fe + e * 12;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            double dx = dg[a], dy = dg[4 + a], dz = dg[8 + a];
            out[3 * a + 0] = 0.0 + (((0.0 + dx * s0) + dy * s3) + dz * s5) * w;
            out[3 * a + 1] = 0.0 + (((0.0 + dy * s1) + dx * s3) + dz * s4) * w;
            out[3 * a + 2] = 0.0 + (((0.0 + dz * s2) + dy * s4) + dx * s5) * w;
        }
    }
}

// Gauss-point data of one C3D20 element in LDS, shared by the element
// kernels: J (one entry per thread), det, J^-1, w = |det| w_g and the D
// coefficients at the point's temperature (27 threads), and the global
// derivatives dN/dx = J^-1 dN/dxi.  Returns false (for every thread) when
// the element failed (singular Jacobian or no material values).
struct C20Shared {
    double xc[60], u[60], T[20];
    double dg[27][60];  // [g][r*20 + i]
    double w[27], d[27][3], sig[27][6], Ji[27][9], J[27][9];
    int bad;
};

template <int NT>
__device__ __forceinline__ bool c20_gauss_data(C20Shared &sh, int64_t e, const int32_t *__restrict__ conn,
                                               const int32_t *__restrict__ mat, const int64_t *__restrict__ gpos,
                                               const double *__restrict__ xyz, const double *__restrict__ temp,
                                               const double *__restrict__ u, int nlgeom,
                                               const MatLut *__restrict__ mats, const double *__restrict__ gN,
                                               const double *__restrict__ gdN, const double *__restrict__ gw,
                                               DevStatus *__restrict__ st) {
    const int tid = threadIdx.x;
    if (tid == 0)
        sh.bad = 0;
    if (tid < 20) {
        int nd = conn[20 * e + tid];
        sh.T[tid] = temp[nd];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double x = xyz[3 * (int64_t)nd + c];
            double uu = u ? u[3 * (int64_t)nd + c] : 0.0;
            sh.u[3 * tid + c] = uu;
            sh.xc[3 * tid + c] = nlgeom ? x + uu : x;
        }
    }
    __syncthreads();
    // J at every Gauss point, one entry per thread: J(r,c) = sum over the
    // 20 nodes of dN(r,i) x(i,c), in node order (jac's sums)
    if (tid < 27 * 9) {
        const int g = tid / 9, r = (tid % 9) / 3, c = tid % 3;
        const double *dN = gdN + 60 * g;
        double sj = 0.0;
#pragma unroll 4
        for (int i = 0; i < 20; ++i)
            sj += dN[20 * r + i] * sh.xc[3 * i + c];
        sh.J[g][3 * r + c] = sj;
    }
    __syncthreads();
    if (tid < 27) {
        const int g = tid;
        double J[9], Ji[9], det;
#pragma unroll
        for (int k = 0; k < 9; ++k)
            J[k] = sh.J[g][k];
        if (!jinv(J, Ji, det)) {
            sh.bad = 1;
            fail(st, FX_ERR_SINGULAR_JACOBIAN, gpos[e]);
        } else {
#pragma unroll
            for (int k = 0; k < 9; ++k)
                sh.Ji[g][k] = Ji[k];
            sh.w[g] = fabs(det) * gw[g];
            double t = 0.0;
            for (int i = 0; i < 20; ++i)
                t += gN[20 * g + i] * sh.T[i];
            double d1, d2, d3;
            if (!d_coeffs(mats[mat[e]], t, d1, d2, d3)) {
                sh.bad = 1;
                fail(st, FX_ERR_INVALID, gpos[e]);
            }
            sh.d[g][0] = d1;
            sh.d[g][1] = d2;
            sh.d[g][2] = d3;
        }
    }
    __syncthreads();
    if (sh.bad)
        return false;
    // global derivatives dN/dx = J^-1 dN/dxi at every Gauss point
    for (int idx = tid; idx < 27 * 60; idx += NT) {
        const int g = idx / 60, q = idx % 60, r = q / 20, i = q % 20;
        const double *dN = gdN + 60 * g;
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k)
            s += sh.Ji[g][3 * r + k] * dN[20 * k + i];
        sh.dg[g][q] = s;
    }
    __syncthreads();
    return true;
}

// ---------------------------------------------------------------------------
// C3D20: one workgroup per element.  243 threads take one Jacobian entry
// each (27 Gauss points x 9), 27 threads the inverses, every thread a share
// of the 27 x 60 global derivatives, and each of the 400 node pairs its own
// thread for K (one round).  Every value keeps the operation order of the
// oracle's restatement.
// ---------------------------------------------------------------------------
constexpr int kC20Threads = 512;

template <bool DO_K, bool DO_F>
__global__ __launch_bounds__(kC20Threads) void k_c3d20(
    int64_t n_el, const int32_t *__restrict__ conn, const int32_t *__restrict__ mat,
    const int64_t *__restrict__ gpos, const double *__restrict__ xyz, const double *__restrict__ temp,
    const double *__restrict__ u, int nlgeom, const MatLut *__restrict__ mats, const double *__restrict__ gN,
    const double *__restrict__ gdN, const double *__restrict__ gw, double *__restrict__ ke,
    double *__restrict__ fe, DevStatus *__restrict__ st, const int32_t *__restrict__ skip) {
    if (skip && *skip)  // the same flag for the whole grid: no barrier is skipped by part of a block
        return;
    __shared__ C20Shared sh;
    const int64_t e = blockIdx.x;
    const int tid = threadIdx.x;
    if (!c20_gauss_data<kC20Threads>(sh, e, conn, mat, gpos, xyz, temp, u, nlgeom, mats, gN, gdN, gw, st))
        return;
    auto &s_dg = sh.dg;
    auto &s_w = sh.w;
    auto &s_d = sh.d;
    auto &s_sig = sh.sig;
    auto &s_u = sh.u;
    if (DO_K) {
        double *out = ke + e * 3600;
        for (int pair = tid; pair < 400; pair += kC20Threads) {
            const int a = pair / 20, b = pair % 20;
            double acc[9];
#pragma unroll
            for (int k = 0; k < 9; ++k)
                acc[k] = 0.0;
            for (int g = 0; g < 27; ++g) {
                double M[9];
                kblock(s_dg[g][a], s_dg[g][20 + a], s_dg[g][40 + a], s_dg[g][b], s_dg[g][20 + b], s_dg[g][40 + b],
                       s_d[g][0], s_d[g][1], s_d[g][2], M);
                const double w = s_w[g];
#pragma unroll
                for (int k = 0; k < 9; ++k)
                    acc[k] += M[k] * w;
            }
#pragma unroll
            for (int k = 0; k < 9; ++k)
                out[pair * 9 + k] = acc[k];
        }
    }
    if (DO_F) {
        if (tid < 27) {
            const int g = tid;
            double e0 = 0.0, e1 = 0.0, e2 = 0.0, e3 = 0.0, e4 = 0.0, e5 = 0.0;
            for (int a = 0; a < 20; ++a) {
                double dx = s_dg[g][a], dy = s_dg[g][20 + a], dz = s_dg[g][40 + a];
                double ux = s_u[3 * a], uy = s_u[3 * a + 1], uz = s_u[3 * a + 2];
                e0 += dx * ux;
                e1 += dy * uy;
                e2 += dz * uz;
                e3 += dy * ux;
                e3 += dx * uy;
                e4 += dz * uy;
                e4 += dy * uz;
                e5 += dz * ux;
                e5 += dx * uz;
            }
            const double d1 = s_d[g][0], d2 = s_d[g][1], d3 = s_d[g][2];
            s_sig[g][0] = ((0.0 + d1 * e0) + d2 * e1) + d2 * e2;
            s_sig[g][1] = ((0.0 + d2 * e0) + d1 * e1) + d2 * e2;
            s_sig[g][2] = ((0.0 + d2 * e0) + d2 * e1) + d1 * e2;
            s_sig[g][3] = 0.0 + d3 * e3;
            s_sig[g][4] = 0.0 + d3 * e4;
            s_sig[g][5] = 0.0 + d3 * e5;
        }
        __syncthreads();
        if (tid < 60) {
            const int a = tid / 3, p = tid % 3;
            double f = 0.0;
            for (int g = 0; g < 27; ++g) {
                double dx = s_dg[g][a], dy = s_dg[g][20 + a], dz = s_dg[g][40 + a];
                double s;
                if (p == 0)
                    s = ((0.0 + dx * s_sig[g][0]) + dy * s_sig[g][3]) + dz * s_sig[g][5];
                else if (p == 1)
                    s = ((0.0 + dy * s_sig[g][1]) + dx * s_sig[g][3]) + dz * s_sig[g][4];
                else
                    s = ((0.0 + dz * s_sig[g][2]) + dy * s_sig[g][4]) + dx * s_sig[g][5];
                f += s * s_w[g];
            }
            fe[e * 60 + tid] = f;
        }
    }
}

// ---------------------------------------------------------------------------
// C3D20 element K on FP64 MFMA (opt-in: FX_C3D20_MFMA=1).  With G the 162 x 60
// matrix whose rows are the strain rows (Gauss point g, Voigt component c) of
// B_g, and H = blockdiag(w_g D_g) G, K_el = G^T H: a 60 x 60 x 162 product
// on v_mfma_f64_16x16x4_f64 (C[16x16] += A[16x4] B[4x16]).  Only the 10
// upper 16 x 16 tiles of the padded 64 x 64 product are formed, and each
// value with c >= r is stored at both (r, c) and (c, r): K_el comes out
// exactly symmetric.  Operands are generated from the LDS
// Gauss data (B entries are 0 or a global derivative; D is isotropic, so
// (D B) entries are 0 or d_k times a derivative).  The sum over the 162
// strain rows runs in the MFMA's own order, so K_el agrees with the
// oracle's to rounding (tests/test_gpu_mfma.py: 1e-13 relative per
// element), not bit for bit; the assembly's drop rule then sees these
// values.  f_int stays on the VALU order of k_c3d20.
//
// Operand layout of v_mfma_f64_16x16x4_f64 (one wavefront, measured): A[i][k]
// in lane i + 16 k, B[k][j] in lane j + 16 k, and C[i][j] in lane
// j + 16 (i % 4), register i / 4.  Strain rows are taken 4 per instruction, 12 (two Gauss
// points) per round of 3 instructions; a 28th point of zeros pads 27 to 28.
// 5 wavefronts per element, 2 tiles each, every tile split into two
// accumulators (even and odd rounds) so that 4 independent chains
// interleave.
// ---------------------------------------------------------------------------
constexpr int kC20MThreads = 320;

typedef double v4d __attribute__((ext_vector_type(4)));

// B_g[c][3a+p] as (derivative component 0..2, or -1 for a structural zero)
__device__ __forceinline__ int b_comp(int c, int p) {
    // x-DOF: [dx, 0, 0, dy, 0, dz]; y-DOF: [0, dy, 0, dx, dz, 0]; z-DOF: [0, 0, dz, 0, dy, dx]
    constexpr int tab[3][6] = {{0, -1, -1, 1, -1, 2}, {-1, 1, -1, 0, 2, -1}, {-1, -1, 2, -1, 1, 0}};
    return tab[p][c];
}

// (D B)[c][3b+q] = d_k x derivative: packed as comp | (k << 2), or -1
__device__ __forceinline__ int db_code(int c, int q) {
    // d index 0: d1, 1: d2, 2: d3
    constexpr int tab[3][6] = {{0 | 0 << 2, 0 | 1 << 2, 0 | 1 << 2, 1 | 2 << 2, -1, 2 | 2 << 2},
                               {1 | 1 << 2, 1 | 0 << 2, 1 | 1 << 2, 0 | 2 << 2, 2 | 2 << 2, -1},
                               {2 | 1 << 2, 2 | 1 << 2, 2 | 0 << 2, -1, 1 | 2 << 2, 0 | 2 << 2}};
    return tab[q][c];
}

template <bool DO_F>
__global__ __launch_bounds__(kC20MThreads) void k_c3d20_mfma(
    int64_t n_el, const int32_t *__restrict__ conn, const int32_t *__restrict__ mat,
    const int64_t *__restrict__ gpos, const double *__restrict__ xyz, const double *__restrict__ temp,
    const double *__restrict__ u, int nlgeom, const MatLut *__restrict__ mats, const double *__restrict__ gN,
    const double *__restrict__ gdN, const double *__restrict__ gw, double *__restrict__ ke,
    double *__restrict__ fe, DevStatus *__restrict__ st, const int32_t *__restrict__ skip) {
    if (skip && *skip)
        return;
    __shared__ C20Shared sh;
    __shared__ double s_wd[28][3];  // w_g d_k; point 27 is the zero pad
    const int64_t e = blockIdx.x;
    const int tid = threadIdx.x;
    if (!c20_gauss_data<kC20MThreads>(sh, e, conn, mat, gpos, xyz, temp, u, nlgeom, mats, gN, gdN, gw, st))
        return;
    if (tid < 28 * 3) {
        const int g = tid / 3, k = tid % 3;
        s_wd[g][k] = g < 27 ? sh.w[g] * sh.d[g][k] : 0.0;
    }
    __syncthreads();
    // tiles (ti, tj), tj >= ti, of the 4 x 4 grid; wave v takes tiles 2v, 2v + 1
    constexpr int kTi[10] = {0, 0, 0, 0, 1, 1, 1, 2, 2, 3};
    constexpr int kTj[10] = {0, 1, 2, 3, 1, 2, 3, 2, 3, 3};
    const int wave = tid / 64, lane = tid % 64;
    const int li = lane % 16, lk = lane / 16;
    v4d acc[2][2];
#pragma unroll
    for (int t = 0; t < 2; ++t)
        acc[t][0] = acc[t][1] = v4d{0.0, 0.0, 0.0, 0.0};
    // per tile: this lane's A row (DOF of the tile row) and B column
    int arow_a[2], arow_p[2], bcol_b[2], bcol_q[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int r = 16 * kTi[2 * wave + t] + li, c = 16 * kTj[2 * wave + t] + li;
        arow_a[t] = r < 60 ? r / 3 : -1;
        arow_p[t] = r % 3;
        bcol_b[t] = c < 60 ? c / 3 : -1;
        bcol_q[t] = c % 3;
    }
    // strain rows of the 3 instructions of a round: kk = 12 j + 4 s + lk, at
    // Gauss point 2 j + gs[s] and component cs = kk % 6.  Per (s, tile) this
    // lane's operand offsets within a Gauss row of dg (-1: a structural
    // zero) and its D coefficient, fixed for the whole loop.
    int gs[3], ao[3][2], bo[3][2], bk[3][2];
#pragma unroll
    for (int s3 = 0; s3 < 3; ++s3) {
        const int kk = 4 * s3 + lk, c = kk % 6;
        gs[s3] = kk / 6;
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int ca = arow_a[t] >= 0 ? b_comp(c, arow_p[t]) : -1;
            ao[s3][t] = ca >= 0 ? 20 * ca + arow_a[t] : -1;
            const int cb = bcol_b[t] >= 0 ? db_code(c, bcol_q[t]) : -1;
            bo[s3][t] = cb >= 0 ? 20 * (cb & 3) + bcol_b[t] : -1;
            bk[s3][t] = cb >= 0 ? cb >> 2 : 0;
        }
    }
    // rounds j and j + 1 (points 2 j ... 2 j + 3) into accumulator sets 0
    // and 1; every load reads a valid address (point 27, the pad, reads
    // point 26 and selects 0)
    for (int j0 = 0; j0 < 14; j0 += 2) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
#pragma unroll
            for (int s3 = 0; s3 < 3; ++s3) {
                const int g = 2 * (j0 + h) + gs[s3];
                const bool live = g < 27;
                const int gg = live ? g : 26;
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    const double xa = sh.dg[gg][ao[s3][t] >= 0 ? ao[s3][t] : 0];
                    const double xb = sh.dg[gg][bo[s3][t] >= 0 ? bo[s3][t] : 0];
                    const double av = (live && ao[s3][t] >= 0) ? xa : 0.0;
                    const double bv = bo[s3][t] >= 0 ? s_wd[g][bk[s3][t]] * xb : 0.0;
                    acc[t][h] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[t][h], 0, 0, 0);
                }
            }
        }
    }
    double *out = ke + e * 3600;
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int ti = kTi[2 * wave + t], tj = kTj[2 * wave + t];
        const int c = 16 * tj + li;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int r = 16 * ti + lk + 4 * k;
            if (r >= 60 || c >= 60 || c < r)  // the upper triangle, mirrored
                continue;
            const double v = acc[t][0][k] + acc[t][1][k];
            out[((r / 3) * 20 + c / 3) * 9 + 3 * (r % 3) + c % 3] = v;
            if (c != r)
                out[((c / 3) * 20 + r / 3) * 9 + 3 * (c % 3) + r % 3] = v;
        }
    }
    if (DO_F) {  // f_int in k_c3d20's order
        if (tid < 27) {
            const int g = tid;
            double e0 = 0.0, e1 = 0.0, e2 = 0.0, e3 = 0.0, e4 = 0.0, e5 = 0.0;
            for (int a = 0; a < 20; ++a) {
                double dx = sh.dg[g][a], dy = sh.dg[g][20 + a], dz = sh.dg[g][40 + a];
                double ux = sh.u[3 * a], uy = sh.u[3 * a + 1], uz = sh.u[3 * a + 2];
                e0 += dx * ux;
                e1 += dy * uy;
                e2 += dz * uz;
                e3 += dy * ux;
                e3 += dx * uy;
                e4 += dz * uy;
                e4 += dy * uz;
                e5 += dz * ux;
                e5 += dx * uz;
            }
            const double d1 = sh.d[g][0], d2 = sh.d[g][1], d3 = sh.d[g][2];
            sh.sig[g][0] = ((0.0 + d1 * e0) + d2 * e1) + d2 * e2;
            sh.sig[g][1] = ((0.0 + d2 * e0) + d1 * e1) + d2 * e2;
            sh.sig[g][2] = ((0.0 + d2 * e0) + d2 * e1) + d1 * e2;
            sh.sig[g][3] = 0.0 + d3 * e3;
            sh.sig[g][4] = 0.0 + d3 * e4;
            sh.sig[g][5] = 0.0 + d3 * e5;
        }
        __syncthreads();
        if (tid < 60) {
            const int a = tid / 3, p = tid % 3;
            double f = 0.0;
            for (int g = 0; g < 27; ++g) {
                double dx = sh.dg[g][a], dy = sh.dg[g][20 + a], dz = sh.dg[g][40 + a];
                double s;
                if (p == 0)
                    s = ((0.0 + dx * sh.sig[g][0]) + dy * sh.sig[g][3]) + dz * sh.sig[g][5];
                else if (p == 1)
                    s = ((0.0 + dy * sh.sig[g][1]) + dx * sh.sig[g][3]) + dz * sh.sig[g][4];
                else
                    s = ((0.0 + dz * sh.sig[g][2]) + dy * sh.sig[g][4]) + dx * sh.sig[g][5];
                f += s * sh.w[g];
            }
            fe[e * 60 + tid] = f;
        }
    }
}

bool c20_mfma_on() {
    const char *e = std::getenv("FX_C3D20_MFMA");
    return e && e[0] == '1';
}

// types: bit 0 the C3D4 block, bit 1 the C3D20 block
template <bool K, bool F>
void launch_types(fx_mesh *m, const double *u, int nlgeom, const double *const *tab4, const double *const *tab20,
                  bool c20_mfma, int types = 3) {
    hipStream_t s = m->ctx->stream;
    if (m->blk4.n_el && (types & 1)) {
        const ElemBlock &b = m->blk4;
        const int threads = m->c3d4_block;
        int64_t blocks = (b.n_el + threads - 1) / threads;
        hipLaunchKernelGGL((k_c3d4<K, F>), dim3((unsigned)blocks), dim3(threads), 0, s, b.n_el, b.conn.p, b.mat.p,
                           b.gpos.p, m->xyz.p, m->temp.p, u, nlgeom, m->mats.p, tab4[0], tab4[1], tab4[2],
                           m->ke.p + b.ke_offset, m->fe.p + b.fe_offset, m->status.p, K ? m->k_skip : nullptr);
        FX_HIP(hipGetLastError());
    }
    if (!(types & 2)) {
        return;
    } else if (m->blk20.n_el && K && c20_mfma) {
        const ElemBlock &b = m->blk20;
        hipLaunchKernelGGL((k_c3d20_mfma<F>), dim3((unsigned)b.n_el), dim3(kC20MThreads), 0, s, b.n_el, b.conn.p,
                           b.mat.p, b.gpos.p, m->xyz.p, m->temp.p, u, nlgeom, m->mats.p, tab20[0], tab20[1],
                           tab20[2], m->ke.p + b.ke_offset, m->fe.p + b.fe_offset, m->status.p, m->k_skip);
        FX_HIP(hipGetLastError());
    } else if (m->blk20.n_el) {
        const ElemBlock &b = m->blk20;
        hipLaunchKernelGGL((k_c3d20<K, F>), dim3((unsigned)b.n_el), dim3(kC20Threads), 0, s, b.n_el, b.conn.p,
                           b.mat.p, b.gpos.p, m->xyz.p, m->temp.p, u, nlgeom, m->mats.p, tab20[0], tab20[1],
                           tab20[2], m->ke.p + b.ke_offset, m->fe.p + b.fe_offset, m->status.p,
                           K ? m->k_skip : nullptr);
        FX_HIP(hipGetLastError());
    }
}

}  // namespace

// Strain and stress at node i of element e (element.rs:526-564 with the
// recovery of static_step.rs:271-379): B at the node's local coordinates on
// the undeformed geometry, eps = B u_el (the thermal-strain increment is
// identically zero: T never changes), sigma = D(T_node) eps.  One thread per
// (element, node); se[(e NN + i) 12 + k] = eps[k], se[... + 6 + k] = sigma[k].
// Sums run in the oracle's dense index order with B's and D's structural
// zeros skipped (exact, as in the element kernels).
template <int NN>
__global__ __launch_bounds__(128) void k_recover(int64_t n_el, const int32_t *__restrict__ conn,
                                                 const int32_t *__restrict__ mat, const int64_t *__restrict__ gpos,
                                                 const double *__restrict__ xyz, const double *__restrict__ temp,
                                                 const double *__restrict__ u, const MatLut *__restrict__ mats,
                                                 const double *__restrict__ dn_nodes, double *__restrict__ se,
                                                 DevStatus *__restrict__ st) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n_el * NN)
        return;
    const int64_t e = gid / NN;
    const int i = (int)(gid % NN);
    double xc[3 * NN], ue[3 * NN];
    const int32_t *cn = conn + NN * e;
#pragma unroll
    for (int a = 0; a < NN; ++a) {
        const int64_t nd = cn[a];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            xc[3 * a + c] = xyz[3 * nd + c];
            ue[3 * a + c] = u[3 * nd + c];
        }
    }
    // C3D4: dN is the same at every point; C3D20: the table row of node i
    const double *dN = NN == 4 ? dn_nodes : dn_nodes + 3 * NN * i;
    double Ji[9], det;
    if (!jac<NN>(dN, xc, Ji, det)) {
        fail(st, FX_ERR_SINGULAR_JACOBIAN, gpos[e]);
        return;
    }
    double e0 = 0.0, e1 = 0.0, e2 = 0.0, e3 = 0.0, e4 = 0.0, e5 = 0.0;
#pragma unroll
    for (int a = 0; a < NN; ++a) {
        double dg[3];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            double sum = 0.0;
#pragma unroll
            for (int k = 0; k < 3; ++k)
                sum += Ji[3 * r + k] * dN[NN * k + a];
            dg[r] = sum;
        }
        const double dx = dg[0], dy = dg[1], dz = dg[2];
        const double ux = ue[3 * a], uy = ue[3 * a + 1], uz = ue[3 * a + 2];
        e0 += dx * ux;
        e1 += dy * uy;
        e2 += dz * uz;
        e3 += dy * ux;
        e3 += dx * uy;
        e4 += dz * uy;
        e4 += dy * uz;
        e5 += dz * ux;
        e5 += dx * uz;
    }
    double d1, d2, d3;
    if (!d_coeffs(mats[mat[e]], temp[cn[i]], d1, d2, d3)) {
        fail(st, FX_ERR_INVALID, gpos[e]);
        return;
    }
    double *out = se + gid * 12;
    out[0] = e0;
    out[1] = e1;
    out[2] = e2;
    out[3] = e3;
    out[4] = e4;
    out[5] = e5;
    out[6] = ((0.0 + d1 * e0) + d2 * e1) + d2 * e2;
    out[7] = ((0.0 + d2 * e0) + d1 * e1) + d2 * e2;
    out[8] = ((0.0 + d2 * e0) + d2 * e1) + d1 * e2;
    out[9] = 0.0 + d3 * e3;
    out[10] = 0.0 + d3 * e4;
    out[11] = 0.0 + d3 * e5;
}

void launch_recover(fx_mesh *m) {
    DevTables *t = &m->ctx->tables;
    hipStream_t s = m->ctx->stream;
    m->se.alloc(std::max<int64_t>(1, 4 * m->fe_total));
    if (m->blk4.n_el) {
        const int64_t n = 4 * m->blk4.n_el;
        hipLaunchKernelGGL(k_recover<4>, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, s, m->blk4.n_el,
                           m->blk4.conn.p, m->blk4.mat.p, m->blk4.gpos.p, m->xyz.p, m->temp.p, m->u.p, m->mats.p,
                           t->dn4.p, m->se.p + 4 * m->blk4.fe_offset, m->status.p);
    }
    if (m->blk20.n_el) {
        const int64_t n = 20 * m->blk20.n_el;
        hipLaunchKernelGGL(k_recover<20>, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, s, m->blk20.n_el,
                           m->blk20.conn.p, m->blk20.mat.p, m->blk20.gpos.p, m->xyz.p, m->temp.p, m->u.p,
                           m->mats.p, t->dn20n.p, m->se.p + 4 * m->blk20.fe_offset, m->status.p);
    }
    FX_HIP(hipGetLastError());
}

void launch_element(fx_mesh *m, bool want_k, bool want_f, bool nlgeom) {
    DevTables *t = &m->ctx->tables;
    const double *tab4[3] = {t->n4.p, t->dn4.p, t->w4.p};
    const double *tab20[3] = {t->n20.p, t->dn20.p, t->w20.p};
    // the displacement enters the strain (f) always and the coordinates
    // only under NLGEOM; for a linear K it is not read at all
    const double *u = (want_f || nlgeom) ? m->u.p : nullptr;
    int nl = nlgeom ? 1 : 0;
    const bool mf = m->c20_mfma >= 0 ? m->c20_mfma != 0 : c20_mfma_on();
    if (want_k && want_f)
        launch_types<true, true>(m, u, nl, tab4, tab20, mf);
    else if (want_k)
        launch_types<true, false>(m, u, nl, tab4, tab20, mf);
    else if (want_f)
        launch_types<false, true>(m, u, nl, tab4, tab20, mf);
}

// fx_profile_element: `reps` launches of the element-K kernels (reference
// coordinates, K only) between two HIP events after one warm-up launch.
double profile_element(fx_mesh *m, int reps, bool mfma, double *flop, int types) {
    DevTables *t = &m->ctx->tables;
    const double *tab4[3] = {t->n4.p, t->dn4.p, t->w4.p};
    const double *tab20[3] = {t->n20.p, t->dn20.p, t->w20.p};
    hipStream_t s = m->ctx->stream;
    reset_status(m);
    launch_types<true, false>(m, nullptr, 0, tab4, tab20, mfma, types);
    hipEvent_t e0, e1;
    FX_HIP(hipEventCreate(&e0));
    FX_HIP(hipEventCreate(&e1));
    FX_HIP(hipEventRecord(e0, s));
    for (int k = 0; k < reps; ++k)
        launch_types<true, false>(m, nullptr, 0, tab4, tab20, mfma, types);
    FX_HIP(hipEventRecord(e1, s));
    FX_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    FX_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    check_status(m, "fx_profile_element");
    *flop = ((types & 1) ? 2592.0 * (double)m->blk4.n_el : 0.0) +
            ((types & 2) ? (mfma ? 860160.0 : 1283040.0) * (double)m->blk20.n_el : 0.0);
    return (double)ms / reps;
}

// fx_mesh_autotune_elements: the C3D4 workgroup size 64 / 128 / 256 / 512
// and, when allowed, the C3D20 kernel (VALU or MFMA), each the fastest of
// `reps` timed element-K launches of its own element block, kept in the
// mesh and stored for the block's bucket (element type, element count); a
// bucket already timed on the context gives its choice untimed.  Returns the
// chosen kernels' time per launch over both blocks.
double autotune_elements(fx_mesh *m, int reps, bool allow_mfma) {
    double flop = 0.0;
    if (m->blk4.n_el) {
        const auto key = tune_key(2, 4, m->blk4.n_el);
        int best = tune_get(m->ctx, key);
        if (best < 0) {
            const int keep = m->c3d4_block;
            double best_t = 0.0;
            try {
                for (int bs : {64, 128, 256, kC4MaxBlock}) {
                    m->c3d4_block = bs;
                    const double t = profile_element(m, reps, false, &flop, 1);
                    if (best_t == 0.0 || t < best_t) {
                        best_t = t;
                        best = bs;
                    }
                }
            } catch (...) {
                m->c3d4_block = keep;
                throw;
            }
            tune_put(m->ctx, key, best);
        }
        m->c3d4_block = best;
    }
    if (m->blk20.n_el) {
        const auto key = tune_key(3, allow_mfma ? 20 : 0, m->blk20.n_el);
        int mf = tune_get(m->ctx, key);
        if (mf < 0) {
            mf = allow_mfma && profile_element(m, reps, true, &flop, 2) < profile_element(m, reps, false, &flop, 2)
                     ? 1
                     : 0;
            tune_put(m->ctx, key, mf);
        }
        m->c20_mfma = mf;
    }
    return profile_element(m, reps, m->c20_mfma == 1, &flop, 3);
}

}  // namespace fx
