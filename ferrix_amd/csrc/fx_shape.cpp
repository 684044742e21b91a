// fx_shape.cpp -- Gauss rules and shape-function tables, built once on the host.
//
// The device kernels read dN_local and N at the Gauss points from these
// tables.  Expressions follow the reference term by term so the tables are
// bit-identical to what the reference's per-call evaluation produces:
//   C3D4  rule and N/dN   src/solver/mesh_lib/elements/c3d4.rs:31-51
//   C3D20 rule            src/solver/mesh_lib/elements/c3d20.rs:31-51
//   C3D20 N/dN            src/solver/mesh_lib/elements/c3d20.rs:90-213
// Built with -ffp-contract=off: no multiply-add is fused.
#include "fx_internal.h"

namespace fx {
namespace {

void c3d20_eval(double xi, double et, double ze, double *N, double *dN) {
    const double omg = 1.0 - xi, omh = 1.0 - et, omr = 1.0 - ze;
    const double opg = 1.0 + xi, oph = 1.0 + et, opr = 1.0 + ze;
    // sums of the corner-node factors (corner n: signs of its local coords)
    const double s_ppp = opg + oph + ze, s_mpp = omg + oph + ze, s_mmp = omg + omh + ze, s_pmp = opg + omh + ze;
    const double s_ppm = opg + oph - ze, s_mpm = omg + oph - ze, s_mmm = omg + omh - ze, s_pmm = opg + omh - ze;
    const double gg = omg * opg / 4.0, hh = omh * oph / 4.0, rr = omr * opr / 4.0;
    const double dg = (omg - opg) / 4.0, dh = (omh - oph) / 4.0, dr = (omr - opr) / 4.0;

    N[0] = -omg * omh * omr * s_ppp / 8.0;
    N[1] = -opg * omh * omr * s_mpp / 8.0;
    N[2] = -opg * oph * omr * s_mmp / 8.0;
    N[3] = -omg * oph * omr * s_pmp / 8.0;
    N[4] = -omg * omh * opr * s_ppm / 8.0;
    N[5] = -opg * omh * opr * s_mpm / 8.0;
    N[6] = -opg * oph * opr * s_mmm / 8.0;
    N[7] = -omg * oph * opr * s_pmm / 8.0;
    N[8] = gg * omh * omr;
    N[9] = hh * opg * omr;
    N[10] = gg * oph * omr;
    N[11] = hh * omg * omr;
    N[12] = gg * omh * opr;
    N[13] = hh * opg * opr;
    N[14] = gg * oph * opr;
    N[15] = hh * omg * opr;
    N[16] = rr * omg * omh;
    N[17] = rr * opg * omh;
    N[18] = rr * opg * oph;
    N[19] = rr * omg * oph;

    double *x = dN, *y = dN + 20, *z = dN + 40;
    x[0] = omh * omr * (s_ppp - omg) / 8.0;
    x[1] = (opg - s_mpp) * omh * omr / 8.0;
    x[2] = (opg - s_mmp) * oph * omr / 8.0;
    x[3] = oph * omr * (s_pmp - omg) / 8.0;
    x[4] = omh * opr * (s_ppm - omg) / 8.0;
    x[5] = (opg - s_mpm) * omh * opr / 8.0;
    x[6] = (opg - s_mmm) * oph * opr / 8.0;
    x[7] = oph * opr * (s_pmm - omg) / 8.0;
    x[8] = dg * omh * omr;
    x[9] = hh * omr;
    x[10] = dg * oph * omr;
    x[11] = -hh * omr;
    x[12] = dg * omh * opr;
    x[13] = hh * opr;
    x[14] = dg * oph * opr;
    x[15] = -hh * opr;
    x[16] = -rr * omh;
    x[17] = rr * omh;
    x[18] = rr * oph;
    x[19] = -rr * oph;

    y[0] = omg * omr * (s_ppp - omh) / 8.0;
    y[1] = opg * omr * (s_mpp - omh) / 8.0;
    y[2] = opg * (oph - s_mmp) * omr / 8.0;
    y[3] = omg * (oph - s_pmp) * omr / 8.0;
    y[4] = omg * opr * (s_ppm - omh) / 8.0;
    y[5] = opg * opr * (s_mpm - omh) / 8.0;
    y[6] = opg * (oph - s_mmm) * opr / 8.0;
    y[7] = omg * (oph - s_pmm) * opr / 8.0;
    y[8] = -gg * omr;
    y[9] = dh * opg * omr;
    y[10] = gg * omr;
    y[11] = dh * omg * omr;
    y[12] = -gg * opr;
    y[13] = dh * opg * opr;
    y[14] = gg * opr;
    y[15] = dh * omg * opr;
    y[16] = -rr * omg;
    y[17] = -rr * opg;
    y[18] = rr * opg;
    y[19] = rr * omg;

    z[0] = omg * omh * (s_ppp - omr) / 8.0;
    z[1] = opg * omh * (s_mpp - omr) / 8.0;
    z[2] = opg * oph * (s_mmp - omr) / 8.0;
    z[3] = omg * oph * (s_pmp - omr) / 8.0;
    z[4] = omg * omh * (opr - s_ppm) / 8.0;
    z[5] = opg * omh * (opr - s_mpm) / 8.0;
    z[6] = opg * oph * (opr - s_mmm) / 8.0;
    z[7] = omg * oph * (opr - s_pmm) / 8.0;
    z[8] = -gg * omh;
    z[9] = -hh * opg;
    z[10] = -gg * oph;
    z[11] = -hh * omg;
    z[12] = gg * omh;
    z[13] = hh * opg;
    z[14] = gg * oph;
    z[15] = hh * omg;
    z[16] = dr * omg * omh;
    z[17] = dr * opg * omh;
    z[18] = dr * opg * oph;
    z[19] = dr * omg * oph;
}

}  // namespace

GaussTable gauss_table(int nn) {
    GaussTable t;
    t.nn = nn;
    if (nn == 4) {
        // one point (1/4, 1/4, 1/4), weight 1/6; N = (1-xi-eta-zeta, xi, eta, zeta)
        t.ngp = 1;
        const double xi = 0.25, eta = 0.25, zeta = 0.25;
        t.N = {1.0 - xi - eta - zeta, xi, eta, zeta};
        t.dN = {-1.0, 1.0, 0.0, 0.0, -1.0, 0.0, 1.0, 0.0, -1.0, 0.0, 0.0, 1.0};
        t.w = {1.0 / 6.0};
        return t;
    }
    // 3x3x3 Gauss-Legendre, xi outermost, zeta innermost
    const double pts[3] = {-0.774596669241483, 0.0, 0.774596669241483};
    const double wts[3] = {0.5555555555555556, 0.8888888888888888, 0.5555555555555556};
    t.ngp = 27;
    t.N.resize(27 * 20);
    t.dN.resize(27 * 60);
    t.w.resize(27);
    int g = 0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) {
                c3d20_eval(pts[i], pts[j], pts[k], &t.N[20 * g], &t.dN[60 * g]);
                t.w[g] = wts[i] * wts[j] * wts[k];
                ++g;
            }
    return t;
}

// dN_local (3 x 20, row-major) at each of the 20 C3D20 node positions
// (c3d20.rs:61-84), for the nodal stress recovery (element.rs:526-564).
std::vector<double> c3d20_node_dn() {
    static const double loc[60] = {-1, -1, -1, 1, -1, -1, 1, 1, -1, -1, 1, -1, -1, -1, 1, 1, -1, 1, 1, 1,
                                   1,  -1, 1,  1, 0, -1, -1, 1, 0,  -1, 0, 1, -1, -1, 0, -1, 0, -1, 1, 1,
                                   0,  1,  0,  1, 1, -1, 0,  1, -1, -1, 0, 1, -1, 0,  1, 1, 0,  -1, 1, 0};
    std::vector<double> out(20 * 60);
    double N[20];
    for (int i = 0; i < 20; ++i)
        c3d20_eval(loc[3 * i], loc[3 * i + 1], loc[3 * i + 2], N, &out[60 * i]);
    return out;
}

}  // namespace fx
