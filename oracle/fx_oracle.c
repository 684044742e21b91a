/*
 * fx_oracle.c -- CPU restatement of the FerriX static-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library in
 * ferrix_amd/: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product never links or calls it.
 *
 * What it restates (reference paths relative to the FerriX tree):
 *   - C3D4 / C3D20 Gauss rules and shape functions
 *       src/solver/mesh_lib/elements/c3d4.rs:31-51
 *       src/solver/mesh_lib/elements/c3d20.rs:31-51 (rule), 90-213 (N, dN)
 *   - temperature LUT and isotropic D matrix
 *       src/solver/material.rs:36-68, 180-230
 *   - B matrix, element stiffness, internal force, thermal force
 *       src/solver/mesh_lib/elements/element.rs:217-318, 352-447, 453-520, 588-606
 *   - triplet assembly with the |v| > 1e-12 drop rule and the per-element
 *     max-diagonal, internal-force scatter
 *       src/solver/assembler.rs:37-151, 157-241
 *   - triplet -> CSR with duplicate summation (sprs TriMat::to_csr,
 *     src/solver/step/static_step.rs:198)
 *   - Jacobi preconditioner and PCG
 *       src/solver/preconditioner.rs:39-58, src/solver/solvers/iterative.rs:26-136
 *
 * Arithmetic order.  The reference uses nalgebra's dynamic GEMM whose
 * operation order cannot be reproduced offline, so this oracle fixes its own
 * order and states it: every sum runs in ascending index order starting from
 * 0.0, products are left-associative exactly as written in the reference
 * expressions, and no multiply-add is fused (Rust never contracts a*b+c, and
 * this file is built with -ffp-contract=off).  J^-1 is the adjugate divided
 * by the determinant (cofactor expansion along the first row).  Elements are
 * visited in the order given (ascending element ID); duplicates of one CSR
 * slot are summed in insertion order.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define FXO_OK 0
#define FXO_ERR_INVALID 1
#define FXO_ERR_OOM 2
#define FXO_ERR_SINGULAR_JACOBIAN 4
#define FXO_ERR_NOT_CONVERGED 5

#define FXO_LUT_MAX 16

/* ------------------------------------------------------------------ */
/* Gauss rules (c3d4.rs:31-36, c3d20.rs:31-51)                         */
/* ------------------------------------------------------------------ */
int fxo_gauss_points(int nn, double *coords, double *w)
{
    if (nn == 4) {
        coords[0] = 0.25; coords[1] = 0.25; coords[2] = 0.25;
        w[0] = 1.0 / 6.0;
        return 1;
    }
    if (nn == 20) {
        const double pts[3] = {-0.774596669241483, 0.0, 0.774596669241483};
        const double wts[3] = {0.5555555555555556, 0.8888888888888888, 0.5555555555555556};
        int g = 0;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) {
                    coords[3 * g + 0] = pts[i];
                    coords[3 * g + 1] = pts[j];
                    coords[3 * g + 2] = pts[k];
                    w[g] = wts[i] * wts[j] * wts[k];
                    ++g;
                }
        return 27;
    }
    return 0;
}

/* c3d4.rs:38-51: N = (1-xi-eta-zeta, xi, eta, zeta); dN constant 3x4 */
void fxo_shape_c3d4(double xi, double eta, double zeta, double *N, double *dN)
{
    N[0] = 1.0 - xi - eta - zeta;
    N[1] = xi;
    N[2] = eta;
    N[3] = zeta;
    static const double d[12] = {-1.0, 1.0, 0.0, 0.0,
                                 -1.0, 0.0, 1.0, 0.0,
                                 -1.0, 0.0, 0.0, 1.0};
    memcpy(dN, d, sizeof d);
}

/* c3d20.rs:90-213: 20-node serendipity brick, CalculiX node order.
 * dN is row-major 3x20 (row 0: d/dxi, row 1: d/deta, row 2: d/dzeta). */
void fxo_shape_c3d20(double xi, double et, double ze, double *N, double *dN)
{
    const double omg = 1.0 - xi, omh = 1.0 - et, omr = 1.0 - ze;
    const double opg = 1.0 + xi, oph = 1.0 + et, opr = 1.0 + ze;

    const double tpgphpr = opg + oph + ze;
    const double tmgphpr = omg + oph + ze;
    const double tmgmhpr = omg + omh + ze;
    const double tpgmhpr = opg + omh + ze;
    const double tpgphmr = opg + oph - ze;
    const double tmgphmr = omg + oph - ze;
    const double tmgmhmr = omg + omh - ze;
    const double tpgmhmr = opg + omh - ze;

    const double omgopg = omg * opg / 4.0;
    const double omhoph = omh * oph / 4.0;
    const double omropr = omr * opr / 4.0;
    const double omgmopg = (omg - opg) / 4.0;
    const double omhmoph = (omh - oph) / 4.0;
    const double omrmopr = (omr - opr) / 4.0;

    N[0] = -omg * omh * omr * tpgphpr / 8.0;
    N[1] = -opg * omh * omr * tmgphpr / 8.0;
    N[2] = -opg * oph * omr * tmgmhpr / 8.0;
    N[3] = -omg * oph * omr * tpgmhpr / 8.0;
    N[4] = -omg * omh * opr * tpgphmr / 8.0;
    N[5] = -opg * omh * opr * tmgphmr / 8.0;
    N[6] = -opg * oph * opr * tmgmhmr / 8.0;
    N[7] = -omg * oph * opr * tpgmhmr / 8.0;
    N[8] = omgopg * omh * omr;
    N[9] = omhoph * opg * omr;
    N[10] = omgopg * oph * omr;
    N[11] = omhoph * omg * omr;
    N[12] = omgopg * omh * opr;
    N[13] = omhoph * opg * opr;
    N[14] = omgopg * oph * opr;
    N[15] = omhoph * omg * opr;
    N[16] = omropr * omg * omh;
    N[17] = omropr * opg * omh;
    N[18] = omropr * opg * oph;
    N[19] = omropr * omg * oph;

    double *dx = dN, *dy = dN + 20, *dz = dN + 40;
    dx[0] = omh * omr * (tpgphpr - omg) / 8.0;
    dx[1] = (opg - tmgphpr) * omh * omr / 8.0;
    dx[2] = (opg - tmgmhpr) * oph * omr / 8.0;
    dx[3] = oph * omr * (tpgmhpr - omg) / 8.0;
    dx[4] = omh * opr * (tpgphmr - omg) / 8.0;
    dx[5] = (opg - tmgphmr) * omh * opr / 8.0;
    dx[6] = (opg - tmgmhmr) * oph * opr / 8.0;
    dx[7] = oph * opr * (tpgmhmr - omg) / 8.0;
    dx[8] = omgmopg * omh * omr;
    dx[9] = omhoph * omr;
    dx[10] = omgmopg * oph * omr;
    dx[11] = -omhoph * omr;
    dx[12] = omgmopg * omh * opr;
    dx[13] = omhoph * opr;
    dx[14] = omgmopg * oph * opr;
    dx[15] = -omhoph * opr;
    dx[16] = -omropr * omh;
    dx[17] = omropr * omh;
    dx[18] = omropr * oph;
    dx[19] = -omropr * oph;

    dy[0] = omg * omr * (tpgphpr - omh) / 8.0;
    dy[1] = opg * omr * (tmgphpr - omh) / 8.0;
    dy[2] = opg * (oph - tmgmhpr) * omr / 8.0;
    dy[3] = omg * (oph - tpgmhpr) * omr / 8.0;
    dy[4] = omg * opr * (tpgphmr - omh) / 8.0;
    dy[5] = opg * opr * (tmgphmr - omh) / 8.0;
    dy[6] = opg * (oph - tmgmhmr) * opr / 8.0;
    dy[7] = omg * (oph - tpgmhmr) * opr / 8.0;
    dy[8] = -omgopg * omr;
    dy[9] = omhmoph * opg * omr;
    dy[10] = omgopg * omr;
    dy[11] = omhmoph * omg * omr;
    dy[12] = -omgopg * opr;
    dy[13] = omhmoph * opg * opr;
    dy[14] = omgopg * opr;
    dy[15] = omhmoph * omg * opr;
    dy[16] = -omropr * omg;
    dy[17] = -omropr * opg;
    dy[18] = omropr * opg;
    dy[19] = omropr * omg;

    dz[0] = omg * omh * (tpgphpr - omr) / 8.0;
    dz[1] = opg * omh * (tmgphpr - omr) / 8.0;
    dz[2] = opg * oph * (tmgmhpr - omr) / 8.0;
    dz[3] = omg * oph * (tpgmhpr - omr) / 8.0;
    dz[4] = omg * omh * (opr - tpgphmr) / 8.0;
    dz[5] = opg * omh * (opr - tmgphmr) / 8.0;
    dz[6] = opg * oph * (opr - tmgmhmr) / 8.0;
    dz[7] = omg * oph * (opr - tpgmhmr) / 8.0;
    dz[8] = -omgopg * omh;
    dz[9] = -omhoph * opg;
    dz[10] = -omgopg * oph;
    dz[11] = -omhoph * omg;
    dz[12] = omgopg * omh;
    dz[13] = omhoph * opg;
    dz[14] = omgopg * oph;
    dz[15] = omhoph * omg;
    dz[16] = omrmopr * omg * omh;
    dz[17] = omrmopr * opg * omh;
    dz[18] = omrmopr * opg * oph;
    dz[19] = omrmopr * omg * oph;
}

static void shape(int nn, double xi, double eta, double zeta, double *N, double *dN)
{
    if (nn == 4)
        fxo_shape_c3d4(xi, eta, zeta, N, dN);
    else
        fxo_shape_c3d20(xi, eta, zeta, N, dN);
}

/* ------------------------------------------------------------------ */
/* Material (material.rs:36-68, 202-230)                               */
/* ------------------------------------------------------------------ */
/* Clamped piecewise-linear LUT; binary search as in material.rs:50-67. */
double fxo_lut_interp(int n, const double *t, const double *v, double temp)
{
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

/* Row-major 6x6 isotropic D in Voigt order [xx,yy,zz,xy,yz,zx]. */
void fxo_d_matrix(double e, double nu, double *D)
{
    double factor = e / ((1.0 + nu) * (1.0 - 2.0 * nu));
    double c1 = 1.0 - nu, c2 = nu, c3 = (1.0 - 2.0 * nu) / 2.0;
    const double m[36] = {c1, c2, c2, 0.0, 0.0, 0.0,
                          c2, c1, c2, 0.0, 0.0, 0.0,
                          c2, c2, c1, 0.0, 0.0, 0.0,
                          0.0, 0.0, 0.0, c3, 0.0, 0.0,
                          0.0, 0.0, 0.0, 0.0, c3, 0.0,
                          0.0, 0.0, 0.0, 0.0, 0.0, c3};
    for (int i = 0; i < 36; ++i)
        D[i] = m[i] * factor;
}

/* Material table: per material a LUT for E and one for nu, FXO_LUT_MAX
 * points each (temperature ascending). */
typedef struct {
    const int32_t *n_e;
    const double *t_e, *v_e;
    const int32_t *n_nu;
    const double *t_nu, *v_nu;
} mat_table;

static void material_d(const mat_table *mt, int m, double temp, double *D)
{
    double e = fxo_lut_interp(mt->n_e[m], mt->t_e + m * FXO_LUT_MAX, mt->v_e + m * FXO_LUT_MAX, temp);
    double nu = fxo_lut_interp(mt->n_nu[m], mt->t_nu + m * FXO_LUT_MAX, mt->v_nu + m * FXO_LUT_MAX, temp);
    fxo_d_matrix(e, nu, D);
}

/* ------------------------------------------------------------------ */
/* Element kernels (element.rs)                                        */
/* ------------------------------------------------------------------ */
/* J = dN_local * X^T, det by cofactor expansion, J^-1 = adj(J)/det.
 * Returns 0 on det == 0 (nalgebra try_inverse fails only then). */
static int jacobian(int nn, const double *dN, const double *xc, double *Jinv, double *det_out)
{
    double J[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            double s = 0.0;
            for (int i = 0; i < nn; ++i)
                s += dN[r * nn + i] * xc[3 * i + c];
            J[3 * r + c] = s;
        }
    double m0 = J[4] * J[8] - J[5] * J[7];
    double m1 = J[3] * J[8] - J[5] * J[6];
    double m2 = J[3] * J[7] - J[4] * J[6];
    double det = J[0] * m0 - J[1] * m1 + J[2] * m2;
    if (det == 0.0)
        return 0;
    Jinv[0] = (J[4] * J[8] - J[5] * J[7]) / det;
    Jinv[1] = (J[2] * J[7] - J[1] * J[8]) / det;
    Jinv[2] = (J[1] * J[5] - J[2] * J[4]) / det;
    Jinv[3] = (J[5] * J[6] - J[3] * J[8]) / det;
    Jinv[4] = (J[0] * J[8] - J[2] * J[6]) / det;
    Jinv[5] = (J[2] * J[3] - J[0] * J[5]) / det;
    Jinv[6] = (J[3] * J[7] - J[4] * J[6]) / det;
    Jinv[7] = (J[1] * J[6] - J[0] * J[7]) / det;
    Jinv[8] = (J[0] * J[4] - J[1] * J[3]) / det;
    *det_out = det;
    return 1;
}

/* element.rs:588-606: B (6 x 3nn, row-major) from dN_global (3 x nn). */
static void build_b(int nn, const double *dg, double *B)
{
    int nd = 3 * nn;
    memset(B, 0, sizeof(double) * 6 * nd);
    for (int i = 0; i < nn; ++i) {
        int c = 3 * i;
        double dx = dg[i], dy = dg[nn + i], dz = dg[2 * nn + i];
        B[0 * nd + c] = dx;
        B[1 * nd + c + 1] = dy;
        B[2 * nd + c + 2] = dz;
        B[3 * nd + c] = dy;
        B[3 * nd + c + 1] = dx;
        B[4 * nd + c + 1] = dz;
        B[4 * nd + c + 2] = dy;
        B[5 * nd + c] = dz;
        B[5 * nd + c + 2] = dx;
    }
}

/*
 * One element: K_el (element.rs:217-318) and/or f_el (element.rs:352-447).
 *   X      nn x 3 reference coordinates
 *   ucfg   nn x 3 displacement added to the coordinates (NLGEOM) or NULL
 *   uel    nn x 3 displacement used for the strain in f, or NULL (zeros)
 *   Tn     nn nodal temperatures (current == initial; thermal strain delta 0)
 *   K      (3nn)^2 row-major output or NULL
 *   f      3nn output or NULL
 */
int fxo_element(int nn, const double *X, const double *ucfg, const double *uel, const double *Tn,
                const mat_table *mt, int m, double *K, double *f)
{
    double gp[81], gw[27], N[20], dN[60], dg[60], Jinv[9], det, D[36];
    double xc[60];
    double B[6 * 60], BtD[60 * 6];
    int nd = 3 * nn;
    int ngp = fxo_gauss_points(nn, gp, gw);
    for (int i = 0; i < 3 * nn; ++i)
        xc[i] = ucfg ? X[i] + ucfg[i] : X[i];
    if (K)
        memset(K, 0, sizeof(double) * nd * nd);
    if (f)
        memset(f, 0, sizeof(double) * nd);
    for (int g = 0; g < ngp; ++g) {
        shape(nn, gp[3 * g], gp[3 * g + 1], gp[3 * g + 2], N, dN);
        if (!jacobian(nn, dN, xc, Jinv, &det))
            return FXO_ERR_SINGULAR_JACOBIAN;
        for (int r = 0; r < 3; ++r)
            for (int i = 0; i < nn; ++i) {
                double s = 0.0;
                for (int k = 0; k < 3; ++k)
                    s += Jinv[3 * r + k] * dN[k * nn + i];
                dg[r * nn + i] = s;
            }
        double weight = fabs(det) * gw[g];
        build_b(nn, dg, B);
        double t = 0.0;
        for (int i = 0; i < nn; ++i)
            t += N[i] * Tn[i];
        material_d(mt, m, t, D);
        if (K) {
            /* ((B^T D) B) * w, then K += */
            for (int i = 0; i < nd; ++i)
                for (int k = 0; k < 6; ++k) {
                    double s = 0.0;
                    for (int j = 0; j < 6; ++j)
                        s += B[j * nd + i] * D[6 * j + k];
                    BtD[6 * i + k] = s;
                }
            for (int i = 0; i < nd; ++i)
                for (int l = 0; l < nd; ++l) {
                    double s = 0.0;
                    for (int k = 0; k < 6; ++k)
                        s += BtD[6 * i + k] * B[k * nd + l];
                    K[i * nd + l] += s * weight;
                }
        }
        if (f) {
            double eps[6], sig[6];
            for (int r = 0; r < 6; ++r) {
                double s = 0.0;
                for (int j = 0; j < nd; ++j)
                    s += B[r * nd + j] * (uel ? uel[j] : 0.0);
                eps[r] = s;
            }
            for (int r = 0; r < 6; ++r) {
                double s = 0.0;
                for (int k = 0; k < 6; ++k)
                    s += D[6 * r + k] * eps[k];
                sig[r] = s;
            }
            for (int i = 0; i < nd; ++i) {
                double s = 0.0;
                for (int r = 0; r < 6; ++r)
                    s += B[r * nd + i] * sig[r];
                f[i] += s * weight;
            }
        }
    }
    return FXO_OK;
}

/* element.rs:453-520 thermal force with separate initial/current nodal
 * temperatures and a per-material expansion LUT (reference temperature
 * t_ref).  Used only to show the term is exactly zero when the two
 * temperature fields are equal, which is always the case in the reference
 * (state.rs:47-78 never changes T). */
int fxo_element_thermal(int nn, const double *X, const double *Ti, const double *Tc, const mat_table *mt,
                        int m, int n_a, const double *t_a, const double *v_a, double t_ref, double *f)
{
    double gp[81], gw[27], N[20], dN[60], dg[60], Jinv[9], det, D[36], B[6 * 60];
    int nd = 3 * nn;
    int ngp = fxo_gauss_points(nn, gp, gw);
    memset(f, 0, sizeof(double) * nd);
    for (int g = 0; g < ngp; ++g) {
        shape(nn, gp[3 * g], gp[3 * g + 1], gp[3 * g + 2], N, dN);
        double tc = 0.0, ti = 0.0;
        for (int i = 0; i < nn; ++i) {
            tc += N[i] * Tc[i];
            ti += N[i] * Ti[i];
        }
        if (n_a <= 0)
            continue;
        double a_c = fxo_lut_interp(n_a, t_a, v_a, tc);
        double th_c = a_c * (tc - t_ref);
        double a_i = fxo_lut_interp(n_a, t_a, v_a, ti);
        double th_i = a_i * (ti - t_ref);
        double dth = th_c - th_i;
        material_d(mt, m, tc, D);
        if (!jacobian(nn, dN, X, Jinv, &det))
            return FXO_ERR_SINGULAR_JACOBIAN;
        for (int r = 0; r < 3; ++r)
            for (int i = 0; i < nn; ++i) {
                double s = 0.0;
                for (int k = 0; k < 3; ++k)
                    s += Jinv[3 * r + k] * dN[k * nn + i];
                dg[r * nn + i] = s;
            }
        double weight = fabs(det) * gw[g];
        build_b(nn, dg, B);
        double eth[6] = {dth, dth, dth, 0.0, 0.0, 0.0}, sig[6];
        for (int r = 0; r < 6; ++r) {
            double s = 0.0;
            for (int k = 0; k < 6; ++k)
                s += D[6 * r + k] * eth[k];
            sig[r] = s;
        }
        for (int i = 0; i < nd; ++i) {
            double s = 0.0;
            for (int r = 0; r < 6; ++r)
                s += B[r * nd + i] * sig[r];
            f[i] += s * weight;
        }
    }
    return FXO_OK;
}

/* Python-friendly wrapper around fxo_element_thermal. */
int fxo_element_thermal_flat(int nn, const double *X, const double *Ti, const double *Tc, const int32_t *n_e,
                             const double *t_e, const double *v_e, const int32_t *n_nu, const double *t_nu,
                             const double *v_nu, int m, int n_a, const double *t_a, const double *v_a, double t_ref,
                             double *f)
{
    mat_table mt = {n_e, t_e, v_e, n_nu, t_nu, v_nu};
    return fxo_element_thermal(nn, X, Ti, Tc, &mt, m, n_a, t_a, v_a, t_ref, f);
}

/* Python-friendly wrapper around fxo_element with flat material arrays. */
int fxo_element_flat(int nn, const double *X, const double *ucfg, const double *uel, const double *Tn,
                     const int32_t *n_e, const double *t_e, const double *v_e, const int32_t *n_nu,
                     const double *t_nu, const double *v_nu, int m, double *K, double *f)
{
    mat_table mt = {n_e, t_e, v_e, n_nu, t_nu, v_nu};
    return fxo_element(nn, X, ucfg, uel, Tn, &mt, m, K, f);
}

/* ------------------------------------------------------------------ */
/* Global assembly (assembler.rs)                                      */
/* ------------------------------------------------------------------ */
/* Mesh as flat arrays: elements in visiting order; el_ptr[e]..el_ptr[e+1]
 * indexes conn (node matrix indices).  nn per element is 4 or 20. */
static void gather3(const double *src, const int32_t *nodes, int nn, double *dst)
{
    for (int i = 0; i < nn; ++i) {
        dst[3 * i + 0] = src[3 * (int64_t)nodes[i] + 0];
        dst[3 * i + 1] = src[3 * (int64_t)nodes[i] + 1];
        dst[3 * i + 2] = src[3 * (int64_t)nodes[i] + 2];
    }
}

/*
 * assembler.rs:37-151.  Emits triplets (row, col, value) for every element
 * entry with |v| > drop_tol in element / i / j / dof_i / dof_j order, and the
 * max |v| over the kept diagonal contributions.  trip_* must hold
 * sum_e (3 nn_e)^2 entries.  u: NULL (linear) or the displacement that moves
 * the coordinates and enters the strain (NLGEOM, static_step.rs:176).
 * On a singular Jacobian, *bad_el receives the failing element position.
 */
int fxo_assemble(int64_t n_el, const int64_t *el_ptr, const int32_t *conn, const int32_t *mat,
                 const double *xyz, const double *Tnode, const double *u, const int32_t *n_e,
                 const double *t_e, const double *v_e, const int32_t *n_nu, const double *t_nu,
                 const double *v_nu, double drop_tol, int64_t *n_trip, int32_t *trip_row,
                 int32_t *trip_col, double *trip_val, double *max_diag, int64_t *bad_el)
{
    mat_table mt = {n_e, t_e, v_e, n_nu, t_nu, v_nu};
    double X[60], ue[60], Tn[20], K[3600];
    int64_t t = 0;
    double md = 0.0;
    for (int64_t e = 0; e < n_el; ++e) {
        const int32_t *nodes = conn + el_ptr[e];
        int nn = (int)(el_ptr[e + 1] - el_ptr[e]);
        if (nn != 4 && nn != 20)
            return FXO_ERR_INVALID;
        gather3(xyz, nodes, nn, X);
        for (int i = 0; i < nn; ++i)
            Tn[i] = Tnode[nodes[i]];
        if (u)
            gather3(u, nodes, nn, ue);
        int rc = fxo_element(nn, X, u ? ue : NULL, u ? ue : NULL, Tn, &mt, mat[e], K, NULL);
        if (rc != FXO_OK) {
            if (bad_el)
                *bad_el = e;
            return rc;
        }
        int nd = 3 * nn;
        for (int i = 0; i < nn; ++i)
            for (int j = 0; j < nn; ++j)
                for (int di = 0; di < 3; ++di)
                    for (int dj = 0; dj < 3; ++dj) {
                        double v = K[(3 * i + di) * nd + 3 * j + dj];
                        if (fabs(v) > drop_tol) {
                            int32_t row = 3 * nodes[i] + di, col = 3 * nodes[j] + dj;
                            trip_row[t] = row;
                            trip_col[t] = col;
                            trip_val[t] = v;
                            ++t;
                            if (row == col)
                                md = fmax(md, fabs(v));
                        }
                    }
    }
    *n_trip = t;
    *max_diag = md;
    return FXO_OK;
}

/* assembler.rs:157-241.  f_int scatter-added in element order.
 * u: total displacement (strain); ucfg: NLGEOM coordinate update or NULL. */
int fxo_internal_force(int64_t n_el, const int64_t *el_ptr, const int32_t *conn, const int32_t *mat,
                       const double *xyz, const double *Tnode, const double *u, const double *ucfg,
                       const int32_t *n_e, const double *t_e, const double *v_e, const int32_t *n_nu,
                       const double *t_nu, const double *v_nu, int64_t n_dof, double *f_int,
                       int64_t *bad_el)
{
    mat_table mt = {n_e, t_e, v_e, n_nu, t_nu, v_nu};
    double X[60], ue[60], uc[60], Tn[20], fe[60];
    for (int64_t i = 0; i < n_dof; ++i)
        f_int[i] = 0.0;
    for (int64_t e = 0; e < n_el; ++e) {
        const int32_t *nodes = conn + el_ptr[e];
        int nn = (int)(el_ptr[e + 1] - el_ptr[e]);
        if (nn != 4 && nn != 20)
            return FXO_ERR_INVALID;
        gather3(xyz, nodes, nn, X);
        gather3(u, nodes, nn, ue);
        if (ucfg)
            gather3(ucfg, nodes, nn, uc);
        for (int i = 0; i < nn; ++i)
            Tn[i] = Tnode[nodes[i]];
        int rc = fxo_element(nn, X, ucfg ? uc : NULL, ue, Tn, &mt, mat[e], NULL, fe);
        if (rc != FXO_OK) {
            if (bad_el)
                *bad_el = e;
            return rc;
        }
        for (int i = 0; i < nn; ++i)
            for (int d = 0; d < 3; ++d)
                f_int[3 * (int64_t)nodes[i] + d] += fe[3 * i + d];
    }
    return FXO_OK;
}

/*
 * Nodal strain and stress recovery (static_step.rs:271-379,
 * element.rs:526-564): for every element, in the order given, and every
 * node i of it, B at the node's local coordinates on the undeformed
 * geometry, strain = B u_el (the thermal-strain increment is identically
 * zero: T never changes, state.rs:47-78), stress = D(T_node) strain; both
 * summed per node and divided by the number of elements containing the
 * node.  Nodes of no element stay zero.  S and E are n_nodes x 6, Voigt
 * order [xx, yy, zz, xy, yz, zx] with engineering shear.
 */
static const double C3D4_NODES[12] = {0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1};  /* c3d4.rs:53-61 */
static const double C3D20_NODES[60] = {                                        /* c3d20.rs:61-84 */
    -1, -1, -1, 1, -1, -1, 1, 1, -1, -1, 1, -1, -1, -1, 1, 1, -1, 1, 1, 1, 1, -1, 1, 1,
    0,  -1, -1, 1, 0,  -1, 0, 1, -1, -1, 0, -1, 0,  -1, 1, 1, 0,  1, 0, 1, 1, -1, 0, 1,
    -1, -1, 0,  1, -1, 0,  1, 1, 0,  -1, 1, 0};

int fxo_stress_strain(int64_t n_nodes, int64_t n_el, const int64_t *el_ptr, const int32_t *conn,
                      const int32_t *mat, const double *xyz, const double *Tnode, const double *u,
                      const int32_t *n_e, const double *t_e, const double *v_e, const int32_t *n_nu,
                      const double *t_nu, const double *v_nu, double *S, double *E, int64_t *bad_el)
{
    mat_table mt = {n_e, t_e, v_e, n_nu, t_nu, v_nu};
    double X[60], ue[60], N[20], dN[60], dg[60], Jinv[9], det, D[36], B[6 * 60];
    int32_t *count = (int32_t *)calloc((size_t)n_nodes, sizeof(int32_t));
    if (!count)
        return FXO_ERR_OOM;
    for (int64_t i = 0; i < 6 * n_nodes; ++i)
        S[i] = E[i] = 0.0;
    for (int64_t e = 0; e < n_el; ++e) {
        const int32_t *nodes = conn + el_ptr[e];
        int nn = (int)(el_ptr[e + 1] - el_ptr[e]);
        if (nn != 4 && nn != 20) {
            free(count);
            return FXO_ERR_INVALID;
        }
        const double *loc = nn == 4 ? C3D4_NODES : C3D20_NODES;
        int nd = 3 * nn;
        gather3(xyz, nodes, nn, X);
        gather3(u, nodes, nn, ue);
        for (int i = 0; i < nn; ++i) {
            shape(nn, loc[3 * i], loc[3 * i + 1], loc[3 * i + 2], N, dN);
            if (!jacobian(nn, dN, X, Jinv, &det)) {
                free(count);
                if (bad_el)
                    *bad_el = e;
                return FXO_ERR_SINGULAR_JACOBIAN;
            }
            for (int r = 0; r < 3; ++r)
                for (int j = 0; j < nn; ++j) {
                    double sum = 0.0;
                    for (int k = 0; k < 3; ++k)
                        sum += Jinv[3 * r + k] * dN[k * nn + j];
                    dg[r * nn + j] = sum;
                }
            build_b(nn, dg, B);
            double eps[6], sig[6];
            for (int r = 0; r < 6; ++r) {
                double sum = 0.0;
                for (int j = 0; j < nd; ++j)
                    sum += B[r * nd + j] * ue[j];
                eps[r] = sum;
            }
            material_d(&mt, mat[e], Tnode[nodes[i]], D);
            for (int r = 0; r < 6; ++r) {
                double sum = 0.0;
                for (int k = 0; k < 6; ++k)
                    sum += D[6 * r + k] * eps[k];
                sig[r] = sum;
            }
            for (int r = 0; r < 6; ++r) {
                S[6 * (int64_t)nodes[i] + r] += sig[r];
                E[6 * (int64_t)nodes[i] + r] += eps[r];
            }
        }
        for (int i = 0; i < nn; ++i)
            count[nodes[i]] += 1;
    }
    for (int64_t n = 0; n < n_nodes; ++n)
        if (count[n] > 0)
            for (int r = 0; r < 6; ++r) {
                S[6 * n + r] /= (double)count[n];
                E[6 * n + r] /= (double)count[n];
            }
    free(count);
    return FXO_OK;
}

/*
 * Triplets -> CSR, summing duplicates in insertion order and sorting the
 * columns of each row ascending (the sprs CsMat invariant).  Counting sort by
 * row, then a stable insertion/merge sort by column within each row.
 * row_ptr has n+1 entries; col/val need n_trip entries; *nnz is the result.
 */
static void sort_row(int32_t *c, double *v, int64_t *ord, int64_t len, int32_t *ctmp, int64_t *otmp)
{
    /* stable merge sort on (col, original order) */
    if (len < 2)
        return;
    for (int64_t width = 1; width < len; width *= 2) {
        for (int64_t lo = 0; lo < len; lo += 2 * width) {
            int64_t mid = lo + width < len ? lo + width : len;
            int64_t hi = lo + 2 * width < len ? lo + 2 * width : len;
            int64_t a = lo, b = mid, k = lo;
            while (a < mid && b < hi) {
                if (c[b] < c[a]) {
                    ctmp[k] = c[b];
                    otmp[k++] = ord[b++];
                } else {
                    ctmp[k] = c[a];
                    otmp[k++] = ord[a++];
                }
            }
            while (a < mid) {
                ctmp[k] = c[a];
                otmp[k++] = ord[a++];
            }
            while (b < hi) {
                ctmp[k] = c[b];
                otmp[k++] = ord[b++];
            }
        }
        memcpy(c, ctmp, sizeof(int32_t) * len);
        memcpy(ord, otmp, sizeof(int64_t) * len);
    }
    (void)v;
}

int fxo_triplets_to_csr(int64_t n, int64_t n_trip, const int32_t *trow, const int32_t *tcol,
                        const double *tval, int64_t *row_ptr, int32_t *col, double *val, int64_t *nnz)
{
    int64_t *cnt = calloc((size_t)n + 1, sizeof(int64_t));
    int64_t *pos = malloc(sizeof(int64_t) * ((size_t)n + 1));
    int32_t *c = malloc(sizeof(int32_t) * (size_t)(n_trip > 0 ? n_trip : 1));
    int64_t *ord = malloc(sizeof(int64_t) * (size_t)(n_trip > 0 ? n_trip : 1));
    int32_t *ctmp = malloc(sizeof(int32_t) * (size_t)(n_trip > 0 ? n_trip : 1));
    int64_t *otmp = malloc(sizeof(int64_t) * (size_t)(n_trip > 0 ? n_trip : 1));
    if (!cnt || !pos || !c || !ord || !ctmp || !otmp) {
        free(cnt); free(pos); free(c); free(ord); free(ctmp); free(otmp);
        return FXO_ERR_OOM;
    }
    for (int64_t t = 0; t < n_trip; ++t) {
        if (trow[t] < 0 || trow[t] >= n || tcol[t] < 0 || tcol[t] >= n) {
            free(cnt); free(pos); free(c); free(ord); free(ctmp); free(otmp);
            return FXO_ERR_INVALID;
        }
        cnt[trow[t] + 1]++;
    }
    for (int64_t i = 0; i < n; ++i)
        cnt[i + 1] += cnt[i];
    memcpy(pos, cnt, sizeof(int64_t) * ((size_t)n + 1));
    for (int64_t t = 0; t < n_trip; ++t) {
        int64_t p = pos[trow[t]]++;
        c[p] = tcol[t];
        ord[p] = t;
    }
    int64_t out = 0;
    row_ptr[0] = 0;
    for (int64_t i = 0; i < n; ++i) {
        int64_t lo = cnt[i], len = cnt[i + 1] - cnt[i];
        sort_row(c + lo, NULL, ord + lo, len, ctmp, otmp);
        int64_t k = lo;
        while (k < lo + len) {
            int32_t cc = c[k];
            double s = tval[ord[k]];
            ++k;
            while (k < lo + len && c[k] == cc) {
                s += tval[ord[k]];
                ++k;
            }
            col[out] = cc;
            val[out] = s;
            ++out;
        }
        row_ptr[i + 1] = out;
    }
    *nnz = out;
    free(cnt); free(pos); free(c); free(ord); free(ctmp); free(otmp);
    return FXO_OK;
}

/* ------------------------------------------------------------------ */
/* Jacobi-PCG (preconditioner.rs:39-58, iterative.rs:26-136)          */
/* ------------------------------------------------------------------ */
static void spmv(int64_t n, const int64_t *rp, const int32_t *ci, const double *v, const double *x, double *y)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double s = 0.0;
        for (int64_t k = rp[i]; k < rp[i + 1]; ++k)
            s += v[k] * x[ci[k]];
        y[i] = s;
    }
}

void fxo_spmv(int64_t n, const int64_t *rp, const int32_t *ci, const double *v, const double *x, double *y,
              int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#endif
    spmv(n, rp, ci, v, x, y);
}

/* inv_diag[i] = 1/K_ii if |K_ii| > 1e-9, else 1.0; a row without a stored
 * diagonal keeps 1.0 (sprs diag() returns only stored entries). */
void fxo_jacobi(int64_t n, const int64_t *rp, const int32_t *ci, const double *v, double *inv_diag)
{
    for (int64_t i = 0; i < n; ++i) {
        inv_diag[i] = 1.0;
        for (int64_t k = rp[i]; k < rp[i + 1]; ++k)
            if (ci[k] == i && fabs(v[k]) > 1e-9)
                inv_diag[i] = 1.0 / v[k];
    }
}

static double dot(int64_t n, const double *a, const double *b)
{
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i)
        s += a[i] * b[i];
    return s;
}

/*
 * PCG exactly as iterative.rs:26-121 with x0 = 0 and the Jacobi
 * preconditioner.  stop_mode 0: converged when sqrt(r.z) < tol at the top of
 * an iteration (the reference rule); stop_mode 1: when ||r||_2 / ||b||_2 <=
 * tol at the same place.  Returns FXO_ERR_NOT_CONVERGED after max_iter
 * iterations, like the reference.  *iters receives the number of SpMVs.
 * Dot products are serial in index order; the SpMV is row-parallel.
 */
int fxo_pcg(int64_t n, const int64_t *rp, const int32_t *ci, const double *v, const double *b, double *x,
            double tol, int64_t max_iter, int stop_mode, int nthreads, int64_t *iters, double *final_res)
{
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#endif
    double *r = malloc(sizeof(double) * (size_t)n), *z = malloc(sizeof(double) * (size_t)n);
    double *p = malloc(sizeof(double) * (size_t)n), *ap = malloc(sizeof(double) * (size_t)n);
    double *md = malloc(sizeof(double) * (size_t)n);
    if (!r || !z || !p || !ap || !md) {
        free(r); free(z); free(p); free(ap); free(md);
        return FXO_ERR_OOM;
    }
    fxo_jacobi(n, rp, ci, v, md);
    for (int64_t i = 0; i < n; ++i) {
        x[i] = 0.0;
        r[i] = b[i];
        z[i] = r[i] * md[i];
        p[i] = z[i];
    }
    double bnorm = sqrt(dot(n, b, b));
    double rs_old = dot(n, r, z);
    int rc = FXO_ERR_NOT_CONVERGED;
    int64_t it;
    for (it = 0; it < max_iter; ++it) {
        double crit = stop_mode == 0 ? sqrt(rs_old) : (bnorm > 0.0 ? sqrt(dot(n, r, r)) / bnorm : 0.0);
        if (stop_mode == 0 ? crit < tol : crit <= tol) {
            rc = FXO_OK;
            if (final_res)
                *final_res = crit;
            break;
        }
        spmv(n, rp, ci, v, p, ap);
        double pkp = dot(n, p, ap);
        double alpha = rs_old / pkp;
        for (int64_t i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * ap[i];
        }
        for (int64_t i = 0; i < n; ++i)
            z[i] = r[i] * md[i];
        double rs_new = dot(n, r, z);
        double beta = rs_new / rs_old;
        for (int64_t i = 0; i < n; ++i)
            p[i] = z[i] + beta * p[i];
        rs_old = rs_new;
    }
    if (rc != FXO_OK && final_res)
        *final_res = stop_mode == 0 ? sqrt(rs_old) : sqrt(dot(n, r, r)) / (bnorm > 0.0 ? bnorm : 1.0);
    *iters = it;
    free(r); free(z); free(p); free(ap); free(md);
    return rc;
}

/*
 * The same PCG with the reference's parallel structure, for the CPU
 * baseline's timing (iterative.rs:26-121 on a rayon pool): the SpMV
 * (iterative.rs:125-136), the three dot products, the x/r and p updates and
 * the Jacobi apply (preconditioner.rs:53-58) are parallel loops, and each
 * iteration allocates K*p zeroed and the preconditioned residual afresh, as
 * the reference's `vec![0.0; n]` and `collect()` do.  Dot products are OpenMP
 * reductions, so their summation order depends on the thread count (rayon's
 * does too); the parity tests use fxo_pcg.
 */
static double dot_par(int64_t n, const double *a, const double *b)
{
    double s = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : s)
    for (int64_t i = 0; i < n; ++i)
        s += a[i] * b[i];
    return s;
}

int fxo_pcg_par(int64_t n, const int64_t *rp, const int32_t *ci, const double *v, const double *b, double *x,
                double tol, int64_t max_iter, int stop_mode, int nthreads, int64_t *iters, double *final_res)
{
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#endif
    double *r = malloc(sizeof(double) * (size_t)n), *z = malloc(sizeof(double) * (size_t)n);
    double *p = malloc(sizeof(double) * (size_t)n), *md = malloc(sizeof(double) * (size_t)n);
    if (!r || !z || !p || !md) {
        free(r); free(z); free(p); free(md);
        return FXO_ERR_OOM;
    }
    fxo_jacobi(n, rp, ci, v, md);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        x[i] = 0.0;
        r[i] = b[i];
        z[i] = r[i] * md[i];
        p[i] = z[i];
    }
    double bnorm = sqrt(dot_par(n, b, b));
    double rs_old = dot_par(n, r, z);
    int rc = FXO_ERR_NOT_CONVERGED;
    int64_t it;
    for (it = 0; it < max_iter; ++it) {
        double crit = stop_mode == 0 ? sqrt(rs_old) : (bnorm > 0.0 ? sqrt(dot_par(n, r, r)) / bnorm : 0.0);
        if (stop_mode == 0 ? crit < tol : crit <= tol) {
            rc = FXO_OK;
            if (final_res)
                *final_res = crit;
            break;
        }
        double *ap = calloc((size_t)n, sizeof(double));
        if (!ap) {
            rc = FXO_ERR_OOM;
            break;
        }
        spmv(n, rp, ci, v, p, ap);
        double alpha = rs_old / dot_par(n, p, ap);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i)
            x[i] += alpha * p[i];
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i)
            r[i] -= alpha * ap[i];
        free(ap);
        free(z);
        z = malloc(sizeof(double) * (size_t)n);
        if (!z) {
            rc = FXO_ERR_OOM;
            break;
        }
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i)
            z[i] = r[i] * md[i];
        double rs_new = dot_par(n, r, z);
        double beta = rs_new / rs_old;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i)
            p[i] = z[i] + beta * p[i];
        rs_old = rs_new;
    }
    if (rc == FXO_ERR_NOT_CONVERGED && final_res)
        *final_res = stop_mode == 0 ? sqrt(rs_old) : sqrt(dot_par(n, r, r)) / (bnorm > 0.0 ? bnorm : 1.0);
    *iters = it;
    free(r); free(z); free(p); free(md);
    return rc;
}

/* Wall seconds of fxo_pcg_par to `tol` (or for `count` iterations when
 * tol is 0): the CPU baseline's solve phase. */
double fxo_pcg_par_time(int64_t n, const int64_t *rp, const int32_t *ci, const double *v, const double *b,
                        double tol, int stop_mode, int64_t count, int nthreads, int64_t *iters)
{
    double *x = malloc(sizeof(double) * (size_t)n);
    double fr, t0 = 0.0, t1 = 0.0;
    *iters = 0;
#ifdef _OPENMP
    t0 = omp_get_wtime();
#endif
    int rc = fxo_pcg_par(n, rp, ci, v, b, x, tol, count, stop_mode, nthreads, iters, &fr);
#ifdef _OPENMP
    t1 = omp_get_wtime();
#endif
    free(x);
    return rc == FXO_OK || (rc == FXO_ERR_NOT_CONVERGED && tol == 0.0) ? t1 - t0 : -1.0;
}

/* Timed pieces for the CPU baseline: seconds for `count` SpMV+vector
 * passes of PCG on a given system (no convergence check). */
double fxo_pcg_time_iters(int64_t n, const int64_t *rp, const int32_t *ci, const double *v, const double *b,
                          int64_t count, int nthreads)
{
    double *x = malloc(sizeof(double) * (size_t)n);
    int64_t it = 0;
    double fr;
    double t0;
#ifdef _OPENMP
    t0 = omp_get_wtime();
#else
    t0 = 0.0;
#endif
    fxo_pcg(n, rp, ci, v, b, x, 0.0, count, 0, nthreads, &it, &fr);
    double t1;
#ifdef _OPENMP
    t1 = omp_get_wtime();
#else
    t1 = 0.0;
#endif
    free(x);
    return t1 - t0;
}

/* FNV-1a (64-bit, offset basis 0xcbf29ce484222325, prime 0x100000001b3) over
 * row_ptr as little-endian int64 followed by col as little-endian int32: the
 * pattern fingerprint of the golden fixtures (test helper, not part of the
 * restated path). */
uint64_t fxo_fnv1a_pattern(int64_t n_rows, const int64_t *row_ptr, int64_t nnz, const int32_t *col)
{
    uint64_t h = 0xcbf29ce484222325ULL;
    const uint64_t prime = 0x100000001b3ULL;
    for (int64_t i = 0; i <= n_rows; ++i) {
        uint64_t v = (uint64_t)row_ptr[i];
        for (int b = 0; b < 8; ++b) {
            h ^= (v >> (8 * b)) & 0xffu;
            h *= prime;
        }
    }
    for (int64_t i = 0; i < nnz; ++i) {
        uint32_t v = (uint32_t)col[i];
        for (int b = 0; b < 4; ++b) {
            h ^= (v >> (8 * b)) & 0xffu;
            h *= prime;
        }
    }
    return h;
}
