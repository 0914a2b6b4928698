// Overlap, dipole and quadrupole integrals over contracted STO-nG s/p shells,
// and their derivatives with respect to the bra centre.
//
// Cartesian Gaussian moments factorise per direction.  For a primitive pair
// (alpha on A, beta on B) and u = x - P (Gaussian product centre P),
//   E(i, m) = int (x - A)^i (x - B)^m exp(-alpha (x-A)^2 - beta (x-B)^2) dx
// is expanded in the even moments M0, M2, M4 of exp(-gamma u^2).  Multipole
// operators use the ket centre B as origin; moments about A follow from the
// shift t = B - A.  d/dA_x of (x-A)^i exp(-alpha(x-A)^2) is
// 2 alpha (x-A)^(i+1) - i (x-A)^(i-1).
#pragma once
#include "common.h"

namespace qx {

// E(i, m), i = 0..2, m = 0..3 for one direction
QX_DEV void moments_1d(double alpha, double beta, double ax, double bx, double E[3][4])
{
    const double g = alpha + beta;
    const double p = (alpha * ax + beta * bx) / g;
    const double a = p - ax, b = p - bx, d = ax - bx;
    const double m0 = sqrt(M_PI / g) * exp(-alpha * beta / g * d * d);
    const double m2 = m0 / (2.0 * g);
    const double m4 = 1.5 * m2 / g;
    const double a2 = a * a, b2 = b * b, b3 = b2 * b;
    E[0][0] = m0;
    E[0][1] = b * m0;
    E[0][2] = b2 * m0 + m2;
    E[0][3] = b3 * m0 + 3.0 * b * m2;
    E[1][0] = a * m0;
    E[1][1] = a * b * m0 + m2;
    E[1][2] = a * b2 * m0 + (a + 2.0 * b) * m2;
    E[1][3] = a * b3 * m0 + (3.0 * a * b + 3.0 * b2) * m2 + m4;
    E[2][0] = a2 * m0 + m2;
    E[2][1] = a2 * b * m0 + (2.0 * a + b) * m2;
    E[2][2] = a2 * b2 * m0 + (a2 + 4.0 * a * b + b2) * m2 + m4;
    E[2][3] = a2 * b3 * m0 + (b3 + 6.0 * a * b2 + 3.0 * a2 * b) * m2 + (3.0 * b + 2.0 * a) * m4;
}

// quadrupole component k -> cartesian pair
QX_DEV int qu(int k) { return k == 0 ? 0 : (k == 1 ? 0 : (k == 2 ? 1 : (k == 3 ? 0 : (k == 4 ? 1 : 2)))); }
QX_DEV int qw(int k) { return k == 0 ? 0 : (k == 1 ? 1 : (k == 2 ? 1 : (k == 3 ? 2 : (k == 4 ? 2 : 2)))); }

// Values for one AO pair summed over primitives: s, dpb[3] (about B), qpb[6]
// raw second moments about B.  la/lb: AO cartesian exponents.
QX_DEV void ao_pair_values(const ElemParams& ea, int ka, const ElemParams& eb, int kb,
                           const double* ra, const double* rb, const int la[3], const int lb[3],
                           double& s, double dpb[3], double qpb[6])
{
    s = 0.0;
    for (int c = 0; c < 3; ++c) dpb[c] = 0.0;
    for (int k = 0; k < 6; ++k) qpb[k] = 0.0;
    const int nga = ea.nprim[ka], ngb = eb.nprim[kb];
    for (int ip = 0; ip < nga; ++ip) {
        const double al = ea.alpha[ka][ip], ca = ea.coeff[ka][ip];
        for (int jp = 0; jp < ngb; ++jp) {
            const double be = eb.alpha[kb][jp], cc = ca * eb.coeff[kb][jp];
            double E[3][3][4];
            for (int x = 0; x < 3; ++x) moments_1d(al, be, ra[x], rb[x], E[x]);
            double e0[3], e1[3], e2[3];
            for (int x = 0; x < 3; ++x) {
                e0[x] = E[x][la[x]][lb[x]];
                e1[x] = E[x][la[x]][lb[x] + 1];
                e2[x] = E[x][la[x]][lb[x] + 2];
            }
            s += cc * e0[0] * e0[1] * e0[2];
            dpb[0] += cc * e1[0] * e0[1] * e0[2];
            dpb[1] += cc * e0[0] * e1[1] * e0[2];
            dpb[2] += cc * e0[0] * e0[1] * e1[2];
            qpb[0] += cc * e2[0] * e0[1] * e0[2];
            qpb[1] += cc * e1[0] * e1[1] * e0[2];
            qpb[2] += cc * e0[0] * e2[1] * e0[2];
            qpb[3] += cc * e1[0] * e0[1] * e1[2];
            qpb[4] += cc * e0[0] * e1[1] * e1[2];
            qpb[5] += cc * e0[0] * e0[1] * e2[2];
        }
    }
}

// Values and bra-centre derivatives for one AO pair summed over primitives.
// ds[d], ddp[d][c], dqp[d][k]: derivatives of s, dpb, qpb with respect to A_d.
QX_DEV void ao_pair_derivs(const ElemParams& ea, int ka, const ElemParams& eb, int kb,
                           const double* ra, const double* rb, const int la[3], const int lb[3],
                           double& s, double dpb[3], double qpb[6], double ds[3],
                           double ddp[3][3], double dqp[3][6])
{
    s = 0.0;
    for (int c = 0; c < 3; ++c) {
        dpb[c] = 0.0;
        ds[c] = 0.0;
        for (int k = 0; k < 3; ++k) ddp[c][k] = 0.0;
    }
    for (int k = 0; k < 6; ++k) {
        qpb[k] = 0.0;
        for (int d = 0; d < 3; ++d) dqp[d][k] = 0.0;
    }
    const int nga = ea.nprim[ka], ngb = eb.nprim[kb];
    for (int ip = 0; ip < nga; ++ip) {
        const double al = ea.alpha[ka][ip], ca = ea.coeff[ka][ip];
        for (int jp = 0; jp < ngb; ++jp) {
            const double be = eb.alpha[kb][jp], cc = ca * eb.coeff[kb][jp];
            double E[3][3][4];
            for (int x = 0; x < 3; ++x) moments_1d(al, be, ra[x], rb[x], E[x]);
            double e[3][3], de[3][3];  // [dir][order of B power: 0,1,2]
            for (int x = 0; x < 3; ++x) {
                const int i = la[x];
                for (int m = 0; m < 3; ++m) {
                    e[x][m] = E[x][i][lb[x] + m];
                    de[x][m] = 2.0 * al * E[x][i + 1][lb[x] + m] - (i ? E[x][i - 1][lb[x] + m] : 0.0);
                }
            }
            // value products
            const double v000 = e[0][0] * e[1][0] * e[2][0];
            s += cc * v000;
            dpb[0] += cc * e[0][1] * e[1][0] * e[2][0];
            dpb[1] += cc * e[0][0] * e[1][1] * e[2][0];
            dpb[2] += cc * e[0][0] * e[1][0] * e[2][1];
            qpb[0] += cc * e[0][2] * e[1][0] * e[2][0];
            qpb[1] += cc * e[0][1] * e[1][1] * e[2][0];
            qpb[2] += cc * e[0][0] * e[1][2] * e[2][0];
            qpb[3] += cc * e[0][1] * e[1][0] * e[2][1];
            qpb[4] += cc * e[0][0] * e[1][1] * e[2][1];
            qpb[5] += cc * e[0][0] * e[1][0] * e[2][2];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                // factor for direction x: de if x == d else e
                auto f = [&](int x, int m) { return x == d ? de[x][m] : e[x][m]; };
                ds[d] += cc * f(0, 0) * f(1, 0) * f(2, 0);
                ddp[d][0] += cc * f(0, 1) * f(1, 0) * f(2, 0);
                ddp[d][1] += cc * f(0, 0) * f(1, 1) * f(2, 0);
                ddp[d][2] += cc * f(0, 0) * f(1, 0) * f(2, 1);
                dqp[d][0] += cc * f(0, 2) * f(1, 0) * f(2, 0);
                dqp[d][1] += cc * f(0, 1) * f(1, 1) * f(2, 0);
                dqp[d][2] += cc * f(0, 0) * f(1, 2) * f(2, 0);
                dqp[d][3] += cc * f(0, 1) * f(1, 0) * f(2, 1);
                dqp[d][4] += cc * f(0, 0) * f(1, 1) * f(2, 1);
                dqp[d][5] += cc * f(0, 0) * f(1, 0) * f(2, 2);
            }
        }
    }
}

// raw second moments about A from moments about B, t = B - A
QX_DEV void shift_to_a(const double t[3], double s, const double dpb[3], const double qpb[6],
                       double dpa[3], double qpa[6])
{
    for (int c = 0; c < 3; ++c) dpa[c] = dpb[c] + t[c] * s;
    for (int k = 0; k < 6; ++k) {
        const int u = qu(k), w = qw(k);
        qpa[k] = qpb[k] + t[u] * dpb[w] + t[w] * dpb[u] + t[u] * t[w] * s;
    }
}

// traceless quadrupole 3/2 r_a r_b - 1/2 r^2 delta_ab from raw moments
QX_DEV void traceless6(const double raw[6], double out[6])
{
    const double tr = 0.5 * (raw[0] + raw[2] + raw[5]);
    out[0] = 1.5 * raw[0] - tr;
    out[1] = 1.5 * raw[1];
    out[2] = 1.5 * raw[2] - tr;
    out[3] = 1.5 * raw[3];
    out[4] = 1.5 * raw[4];
    out[5] = 1.5 * raw[5] - tr;
}

// transpose of the traceless map: sum_k w[k] traceless(raw)[k] = sum_k wt[k] raw[k]
QX_DEV void traceless6_t(const double w[6], double wt[6])
{
    const double sd = 0.5 * (w[0] + w[2] + w[5]);
    wt[0] = 1.5 * w[0] - sd;
    wt[1] = 1.5 * w[1];
    wt[2] = 1.5 * w[2] - sd;
    wt[3] = 1.5 * w[3];
    wt[4] = 1.5 * w[4];
    wt[5] = 1.5 * w[5] - sd;
}

QX_DEV void ao_cart(int l, int comp, int c[3])
{
    c[0] = c[1] = c[2] = 0;
    if (l) c[comp] = 1;
}

}  // namespace qx
