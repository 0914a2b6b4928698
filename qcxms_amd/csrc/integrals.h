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

// quadrupole component k -> cartesian pair
QX_DEV int qu(int k) { return k == 0 ? 0 : (k == 1 ? 0 : (k == 2 ? 1 : (k == 3 ? 0 : (k == 4 ? 1 : 2)))); }
QX_DEV int qw(int k) { return k == 0 ? 0 : (k == 1 ? 1 : (k == 2 ? 1 : (k == 3 ? 2 : (k == 4 ? 2 : 2)))); }

// One primitive pair (alpha on A, beta on B): per direction the moments
// E[x][i][m], i = 0..2, m = 0..3, without the Gaussian factor
// sqrt(pi/g) exp(-alpha beta/g (A_x - B_x)^2), which for the product of the
// three directions is folded into K = c_a c_b (pi/g)^(3/2) exp(-alpha beta/g |A-B|^2).
struct PrimPair {
    double E[3][3][4];
    double K, al;
};

QX_DEV void fill_1d(double a, double b, double m2, double m4, double E[3][4])
{
    const double a2 = a * a, b2 = b * b, b3 = b2 * b;
    E[0][0] = 1.0;
    E[0][1] = b;
    E[0][2] = b2 + m2;
    E[0][3] = b3 + 3.0 * b * m2;
    E[1][0] = a;
    E[1][1] = a * b + m2;
    E[1][2] = a * b2 + (a + 2.0 * b) * m2;
    E[1][3] = a * b3 + (3.0 * a * b + 3.0 * b2) * m2 + m4;
    E[2][0] = a2 + m2;
    E[2][1] = a2 * b + (2.0 * a + b) * m2;
    E[2][2] = a2 * b2 + (a2 + 4.0 * a * b + b2) * m2 + m4;
    E[2][3] = a2 * b3 + (b3 + 6.0 * a * b2 + 3.0 * a2 * b) * m2 + (3.0 * b + 2.0 * a) * m4;
}

QX_DEV void prim_pair(double al, double be, double cc, const double* ra, const double* rb,
                      PrimPair& q)
{
    const double g = al + be, gi = 1.0 / g;
    const double m2 = 0.5 * gi, m4 = 1.5 * m2 * gi;
    const double px = (al * ra[0] + be * rb[0]) * gi, py = (al * ra[1] + be * rb[1]) * gi,
                 pz = (al * ra[2] + be * rb[2]) * gi;
    fill_1d(px - ra[0], px - rb[0], m2, m4, q.E[0]);
    fill_1d(py - ra[1], py - rb[1], m2, m4, q.E[1]);
    fill_1d(pz - ra[2], pz - rb[2], m2, m4, q.E[2]);
    const double dx = ra[0] - rb[0], dy = ra[1] - rb[1], dz = ra[2] - rb[2];
    const double pg = M_PI * gi;
    q.K = cc * pg * sqrt(pg) * exp(-al * be * gi * (dx * dx + dy * dy + dz * dz));
    q.al = al;
}

// cartesian exponent of component A of a shell with angular momentum L in direction X
template <int L, int A, int X>
struct Cart {
    static constexpr int v = (L == 1 && A == X) ? 1 : 0;
};

// Overlap, dipole and raw second moments about B of one AO pair (bra
// component A of an L_A shell, ket component B of an L_B shell), accumulated
// over one primitive pair.  All table indices are compile-time constants.
template <int LA, int LB, int A, int B>
QX_DEV void ao_values(const PrimPair& q, double& s, double* dpb, double* qpb)
{
    const double* ex = q.E[0][Cart<LA, A, 0>::v] + Cart<LB, B, 0>::v;
    const double* ey = q.E[1][Cart<LA, A, 1>::v] + Cart<LB, B, 1>::v;
    const double* ez = q.E[2][Cart<LA, A, 2>::v] + Cart<LB, B, 2>::v;
    const double K = q.K;
    s += K * ex[0] * ey[0] * ez[0];
    dpb[0] += K * ex[1] * ey[0] * ez[0];
    dpb[1] += K * ex[0] * ey[1] * ez[0];
    dpb[2] += K * ex[0] * ey[0] * ez[1];
    qpb[0] += K * ex[2] * ey[0] * ez[0];
    qpb[1] += K * ex[1] * ey[1] * ez[0];
    qpb[2] += K * ex[0] * ey[2] * ez[0];
    qpb[3] += K * ex[1] * ey[0] * ez[1];
    qpb[4] += K * ex[0] * ey[1] * ez[1];
    qpb[5] += K * ex[0] * ey[0] * ez[2];
}

// Density-weighted bra-centre derivative moments of one shell pair:
//   s = sum P s, dpb[c] = sum P <dp_c>, ds[d] = sum P ds/dA_d,
//   ddp[d][c] = sum P d<dp_c>/dA_d, dqp[d][k] = sum P d<qp_k>/dA_d,
//   wds[d] = sum W ds/dA_d
// (sums over AO pairs and primitive pairs; moments about B).
struct DerivAcc {
    double s, dpb[3], ds[3], ddp[3][3], dqp[3][6], wds[3];
};

template <int IX, int JX>
QX_DEV void dir_tables(const double (&E)[3][4], double al, double e[3], double de[3])
{
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        e[m] = E[IX][JX + m];
        de[m] = 2.0 * al * E[IX + 1][JX + m] - (IX ? E[IX ? IX - 1 : 0][JX + m] : 0.0);
    }
}

template <int LA, int LB, int A, int B>
QX_DEV void ao_derivs(const PrimPair& q, double pm, double wm, DerivAcc& acc)
{
    double e[3][3], de[3][3];
    dir_tables<Cart<LA, A, 0>::v, Cart<LB, B, 0>::v>(q.E[0], q.al, e[0], de[0]);
    dir_tables<Cart<LA, A, 1>::v, Cart<LB, B, 1>::v>(q.E[1], q.al, e[1], de[1]);
    dir_tables<Cart<LA, A, 2>::v, Cart<LB, B, 2>::v>(q.E[2], q.al, e[2], de[2]);
    const double kp = q.K * pm, kw = q.K * wm;
    acc.s += kp * e[0][0] * e[1][0] * e[2][0];
    acc.dpb[0] += kp * e[0][1] * e[1][0] * e[2][0];
    acc.dpb[1] += kp * e[0][0] * e[1][1] * e[2][0];
    acc.dpb[2] += kp * e[0][0] * e[1][0] * e[2][1];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        auto f = [&](int x, int m) { return x == d ? de[x][m] : e[x][m]; };
        const double dsd = f(0, 0) * f(1, 0) * f(2, 0);
        acc.ds[d] += kp * dsd;
        acc.wds[d] += kw * dsd;
        acc.ddp[d][0] += kp * f(0, 1) * f(1, 0) * f(2, 0);
        acc.ddp[d][1] += kp * f(0, 0) * f(1, 1) * f(2, 0);
        acc.ddp[d][2] += kp * f(0, 0) * f(1, 0) * f(2, 1);
        acc.dqp[d][0] += kp * f(0, 2) * f(1, 0) * f(2, 0);
        acc.dqp[d][1] += kp * f(0, 1) * f(1, 1) * f(2, 0);
        acc.dqp[d][2] += kp * f(0, 0) * f(1, 2) * f(2, 0);
        acc.dqp[d][3] += kp * f(0, 1) * f(1, 0) * f(2, 1);
        acc.dqp[d][4] += kp * f(0, 0) * f(1, 1) * f(2, 1);
        acc.dqp[d][5] += kp * f(0, 0) * f(1, 0) * f(2, 2);
    }
}

// all AO pairs of an (L_A, L_B) shell pair for one primitive pair;
// pm/wm: density and energy-weighted density of the AO pairs [a * 3 + b]
template <int LA, int LB, int A = 0, int B = 0>
QX_DEV void shell_derivs(const PrimPair& q, const double* pm, const double* wm, DerivAcc& acc)
{
    ao_derivs<LA, LB, A, B>(q, pm[A * 3 + B], wm[A * 3 + B], acc);
    if constexpr (LB == 1 && B < 2) shell_derivs<LA, LB, A, B + 1>(q, pm, wm, acc);
    else if constexpr (LA == 1 && A < 2) shell_derivs<LA, LB, A + 1, 0>(q, pm, wm, acc);
}

// one bra AO (component A) of an (L_A, L_B) shell pair against all kets:
// overlap, dipole and raw second moments about B, summed over primitives
template <int LA, int LB, int A>
QX_DEV void shell_row_values(const ElemParams& ea, int ka, const ElemParams& eb, int kb,
                             const double* ra, const double* rb, double s[3], double dpb[3][3],
                             double qpb[3][6])
{
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        s[b] = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) dpb[b][c] = 0.0;
#pragma unroll
        for (int k = 0; k < 6; ++k) qpb[b][k] = 0.0;
    }
    const int nga = ea.nprim[ka], ngb = eb.nprim[kb];
    for (int ip = 0; ip < nga; ++ip) {
        const double al = ea.alpha[ka][ip], ca = ea.coeff[ka][ip];
        for (int jp = 0; jp < ngb; ++jp) {
            PrimPair q;
            prim_pair(al, eb.alpha[kb][jp], ca * eb.coeff[kb][jp], ra, rb, q);
            ao_values<LA, LB, A, 0>(q, s[0], dpb[0], qpb[0]);
            if constexpr (LB == 1) {
                ao_values<LA, LB, A, 1>(q, s[1], dpb[1], qpb[1]);
                ao_values<LA, LB, A, 2>(q, s[2], dpb[2], qpb[2]);
            }
        }
    }
}

// raw second moments about A from moments about B, t = B - A
QX_DEV void shift_to_a(const double t[3], double s, const double dpb[3], const double qpb[6],
                       double dpa[3], double qpa[6])
{
    dpa[0] = dpb[0] + t[0] * s;
    dpa[1] = dpb[1] + t[1] * s;
    dpa[2] = dpb[2] + t[2] * s;
    qpa[0] = qpb[0] + 2.0 * t[0] * dpb[0] + t[0] * t[0] * s;
    qpa[1] = qpb[1] + t[0] * dpb[1] + t[1] * dpb[0] + t[0] * t[1] * s;
    qpa[2] = qpb[2] + 2.0 * t[1] * dpb[1] + t[1] * t[1] * s;
    qpa[3] = qpb[3] + t[0] * dpb[2] + t[2] * dpb[0] + t[0] * t[2] * s;
    qpa[4] = qpb[4] + t[1] * dpb[2] + t[2] * dpb[1] + t[1] * t[2] * s;
    qpa[5] = qpb[5] + 2.0 * t[2] * dpb[2] + t[2] * t[2] * s;
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


}  // namespace qx
