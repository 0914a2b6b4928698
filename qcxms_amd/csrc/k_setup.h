// Geometry-dependent setup of one GFN2-xTB evaluation (one workgroup per
// system): coordination numbers, D4 reference weights, repulsion and ATM
// energies, S/H0/dipole/quadrupole integrals, the ES2 gamma matrix, the D4
// charge-response matrix and the Cholesky orthogonaliser X = L^-T of S.
#pragma once
#include "integrals.h"

namespace qx {

QX_DEV double d4_zeta(double a, double c, double qref, double qmod)
{
    if (qmod < 0.0) return exp(a);
    return exp(a * (1.0 - exp(c * (1.0 - qref / qmod))));
}
QX_DEV double d4_dzeta(double a, double c, double qref, double qmod)
{
    if (qmod < 0.0) return 0.0;
    const double e = exp(c * (1.0 - qref / qmod));
    return -a * c * e * exp(a * (1.0 - e)) * qref / (qmod * qmod);
}

// GFN2 coordination-number count and its derivative d(count)/dr
QX_DEV void gfn_count(double rc, double r, double& f, double& df)
{
    const double e1 = exp(-gfn2::kCnKa * (rc / r - 1.0));
    const double e2 = exp(-gfn2::kCnKb * ((rc + gfn2::kCnShift) / r - 1.0));
    const double f1 = 1.0 / (1.0 + e1), f2 = 1.0 / (1.0 + e2);
    const double r2 = r * r;
    const double d1 = -gfn2::kCnKa * rc * e1 * f1 * f1 / r2;
    const double d2 = -gfn2::kCnKb * (rc + gfn2::kCnShift) * e2 * f2 * f2 / r2;
    f = f1 * f2;
    df = d1 * f2 + f1 * d2;
}

// D4 covalent count (dftd4.f90:292-325) and derivative
QX_DEV void d4_count(const ElemParams& ei, const ElemParams& ej, double r, double& f, double& df)
{
    const double rco = ei.rcov_d4 + ej.rcov_d4;
    const double den0 = fabs(ei.en_d4 - ej.en_d4) + gfn2::kD4K5;
    const double den = gfn2::kD4K4 * exp(-den0 * den0 / gfn2::kD4K6);
    const double ex = exp(-gfn2::kD4K1 * (rco / r - 1.0));
    f = den / (1.0 + ex);
    df = (-gfn2::kD4K1 * rco * ex * den) / (r * r * (1.0 + ex) * (1.0 + ex));
}

QX_DEV void mp_radius(const ElemParams& e, double cn, double& r, double& drdcn)
{
    const double arg = cn - e.mpvcn - gfn2::kAesShift;
    const double t = exp(-gfn2::kAesKexp * arg);
    r = e.mprad + (gfn2::kAesRmax - e.mprad) / (1.0 + t);
    drdcn = (gfn2::kAesRmax - e.mprad) * gfn2::kAesKexp * t / ((1.0 + t) * (1.0 + t));
}

// K * X(EN) * Y(zeta) and Pi(R) with dPi/dR for an off-site shell pair
QX_DEV double h0_scale(const ElemParams& ea, int ka, const ElemParams& eb, int kb, double r,
                       double& pir, double& dpir, double& kxy)
{
    const double za = ea.slater[ka], zb = eb.slater[kb];
    const double yz = sqrt(2.0 * sqrt(za * zb) / (za + zb));  // wexp = 1/2
    const double den = ea.en - eb.en;
    const double xen = 1.0 + gfn2::kEnScale * den * den;
    const double rab = ea.rad + eb.rad;
    const double sq = sqrt(r / rab);
    const double fa = 1.0 + ea.shpoly[ka] * sq, fb = 1.0 + eb.shpoly[kb] * sq;
    pir = fa * fb;
    dpir = (r > 0.0) ? (ea.shpoly[ka] * fb + eb.shpoly[kb] * fa) * 0.5 / sqrt(r * rab) : 0.0;
    kxy = 0.5 * (gfn2::kShellK[ka] + gfn2::kShellK[kb]) * xen * yz;
    return kxy * pir;
}

QX_DEV double d4_disp(const ElemParams& ei, const ElemParams& ej, double r2, double* ddisp_dr)
{
    const double r4r2ij = 3.0 * ei.r4r2 * ej.r4r2;
    const double r0 = gfn2::kD4A1 * sqrt(r4r2ij) + gfn2::kD4A2;
    const double r02 = r0 * r0, r06 = r02 * r02 * r02;
    const double oor6 = 1.0 / (r2 * r2 * r2 + r06);
    const double oor8 = 1.0 / (r2 * r2 * r2 * r2 + r06 * r02);
    if (ddisp_dr) {
        const double r = sqrt(r2);
        const double door6 = -6.0 * r2 * r2 * r * oor6 * oor6;
        const double door8 = -8.0 * r2 * r2 * r2 * r * oor8 * oor8;
        *ddisp_dr = gfn2::kD4S6 * door6 + gfn2::kD4S8 * r4r2ij * door8;
    }
    return gfn2::kD4S6 * oor6 + gfn2::kD4S8 * r4r2ij * oor8;
}

struct IntegralOut {
    double *S, *H, *D, *Q;
    int ld;
    size_t ms;
};

// S, H0 (off-atom), D and Q blocks of one bra AO (component A) of a shell
// pair.  S and H0 are symmetric; the moment integrals are stored transposed:
// element (mu, nu), taken about the atom of mu, sits at nu * ld + mu.  A row
// writes element (A, b) at both places; for a shell paired with itself
// (same), row b writes element (b, A) itself, so row A writes only the
// transposed places and every address has one writer.
template <int LA, int LB, int A>
QX_DEV void setup_shell_row(const ShellDesc& sa, const ShellDesc& sb, const ElemParams& ea,
                            const ElemParams& eb, const double* ra, const double* rb,
                            const double t[3], double hs, bool off, bool same, const IntegralOut& io)
{
    double s[3], dpb[3][3], qpb[3][6];
    shell_row_values<LA, LB, A>(ea, sa.k, eb, sb.k, ra, rb, s, dpb, qpb);
    constexpr int NB = LB ? 3 : 1;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        double dpa[3], qpa[6], qta[6], qtb[6];
        shift_to_a(t, s[b], dpb[b], qpb[b], dpa, qpa);
        traceless6(qpa, qta);
        traceless6(qpb[b], qtb);
        const int mu = sa.ao0 + A, nu = sb.ao0 + b;
        const size_t o1 = (size_t)mu * io.ld + nu, o2 = (size_t)nu * io.ld + mu;
        io.S[o2] = s[b];
        if (off) io.H[o2] = hs * s[b];
#pragma unroll
        for (int c = 0; c < 3; ++c) io.D[c * io.ms + o2] = dpa[c];
#pragma unroll
        for (int k = 0; k < 6; ++k) io.Q[k * io.ms + o2] = qta[k];
        if (!same) {
            io.S[o1] = s[b];
            if (off) io.H[o1] = hs * s[b];
#pragma unroll
            for (int c = 0; c < 3; ++c) io.D[c * io.ms + o1] = dpb[b][c];
#pragma unroll
            for (int k = 0; k < 6; ++k) io.Q[k * io.ms + o1] = qtb[k];
        }
    }
}

// ---------------------------------------------------------------------------
// LDS of one system's setup (the Cholesky factor uses the dynamic LDS)
struct SetupLds {
    ShellDesc sh[QX_MAX_NSH];
    int ao_at[QX_MAX_NAO], ao_sh[QX_MAX_NAO];
    double sx[3 * QX_MAX_NAT];
    int sel[QX_MAX_NAT];
    double scn[QX_MAX_NAT], scnd4[QX_MAX_NAT];
    int refoff[QX_MAX_NAT + 1];
    double red[QX_BLOCK / 64];
    // per D4 reference: atom, index within the atom, zeta(q = 0) * gw
    int refat[QX_MAX_REF * QX_MAX_NAT], refii[QX_MAX_REF * QX_MAX_NAT];
    double zg[QX_MAX_REF * QX_MAX_NAT];
};

QX_DEV void setup_system(const BatchPtrs& B, int sys, int with_chol, SetupLds& Ls, double* dyn)
{
    const SysInfo si = B.info[sys];
    const int tid = threadIdx.x;
    const int nat = si.nat, nsh = si.nsh, n = si.nao, ld = si.ld;
    ShellDesc* sh = Ls.sh;
    int* ao_at = Ls.ao_at;
    int* ao_sh = Ls.ao_sh;
    double* sx = Ls.sx;
    int* sel = Ls.sel;
    double* scn = Ls.scn;
    double* scnd4 = Ls.scnd4;
    int* refoff = Ls.refoff;
    double* red = Ls.red;

    const double* gx = B.xyz + 3 * (size_t)si.atom_off;
    for (int i = tid; i < 3 * nat; i += QX_BLOCK) sx[i] = gx[i];
    for (int i = tid; i < nat; i += QX_BLOCK) sel[i] = B.elem[si.atom_off + i];
    build_shell_tables(si, B.elem, sh, ao_at, ao_sh);
    if (tid == 0) {
        int o = 0;
        for (int a = 0; a < nat; ++a) {
            refoff[a] = o;
            o += g_tab.e[sel[a]].nref;
        }
        refoff[nat] = o;
    }
    // 1. coordination numbers: a wavefront per atom, a lane per partner,
    //    tree sums over the lanes
    __syncthreads();
    for (int a = tid >> 6; a < nat; a += QX_BLOCK / 64) {
        const int lane = tid & 63;
        const ElemParams& ea = g_tab.e[sel[a]];
        double cn = 0.0, cn4 = 0.0;
        for (int b = lane; b < nat; b += 64) {
            if (b == a) continue;
            const ElemParams& eb = g_tab.e[sel[b]];
            const double dx = sx[3 * a] - sx[3 * b], dy = sx[3 * a + 1] - sx[3 * b + 1],
                         dz = sx[3 * a + 2] - sx[3 * b + 2];
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double r = sqrt(r2);
            double f, df;
            if (r2 <= gfn2::kCnCutoff * gfn2::kCnCutoff) {
                gfn_count(ea.covrad + eb.covrad, r, f, df);
                cn += f;
            }
            if (r2 <= gfn2::kD4CnThr) {
                d4_count(ea, eb, r, f, df);
                cn4 += f;
            }
        }
        cn = wave_sum(cn);
        cn4 = wave_sum(cn4);
        if (lane != 0) continue;
        scn[a] = cn;
        scnd4[a] = cn4;
        B.cn[si.atom_off + a] = cn;
        B.cnd4[si.atom_off + a] = cn4;
        double mr, dmr;
        mp_radius(ea, cn, mr, dmr);
        B.mrad[si.atom_off + a] = mr;
        B.dmrad[si.atom_off + a] = dmr;
    }
    __syncthreads();
    // 2. self-energies and D4 weights
    for (int s = tid; s < nsh; s += QX_BLOCK) {
        const ElemParams& e = g_tab.e[sh[s].el];
        B.hself[si.shell_off + s] = e.hself0[sh[s].k] - e.kcn[sh[s].k] * scn[sh[s].at];
    }
    for (int a = tid; a < nat; a += QX_BLOCK) {
        const ElemParams& e = g_tab.e[sel[a]];
        const double cn = scnd4[a];
        double norm = 0.0, dnorm = 0.0;
        for (int r = 0; r < e.nref; ++r)
            for (int k = 1; k <= e.refc[r]; ++k) {
                const double twf = k * gfn2::kD4Wf, d = cn - e.refcovcn[r];
                const double t = exp(-twf * d * d);
                norm += t;
                dnorm += 2.0 * twf * (e.refcovcn[r] - cn) * t;
            }
        norm = 1.0 / norm;
        double mx = -1e300;
        for (int r = 0; r < e.nref; ++r) mx = fmax(mx, e.refcovcn[r]);
        for (int r = 0; r < e.nref; ++r) {
            double ex = 0.0, dex = 0.0;
            for (int k = 1; k <= e.refc[r]; ++k) {
                const double twf = k * gfn2::kD4Wf, d = cn - e.refcovcn[r];
                const double t = exp(-twf * d * d);
                ex += t;
                dex += 2.0 * twf * (e.refcovcn[r] - cn) * t;
            }
            double g = ex * norm;
            if (g != g) g = (mx == e.refcovcn[r]) ? 1.0 : 0.0;
            double dg = dex * norm - ex * dnorm * norm * norm;
            if (dg != dg) dg = 0.0;
            B.gw[si.ref_off + refoff[a] + r] = g;
            B.dgw[si.ref_off + refoff[a] + r] = dg;
        }
    }
    // 3. clear the matrix slots (padding stays zero)
    const size_t ms = (size_t)ld * ld;
    double* S = B.S + si.mat_off;
    double* H = B.H0 + si.mat_off;
    double* D = B.D + 3 * si.mat_off;
    double* Q = B.Q + 6 * si.mat_off;
    for (size_t i = tid; i < ms; i += QX_BLOCK) {
        S[i] = 0.0;
        H[i] = 0.0;
    }
    for (size_t i = tid; i < 3 * ms; i += QX_BLOCK) D[i] = 0.0;
    for (size_t i = tid; i < 6 * ms; i += QX_BLOCK) Q[i] = 0.0;
    __syncthreads();
    // 4. integrals, one work item per row of a shell pair (the rows of a p
    //    shell are three items; rows of three AO pairs come first)
    for (int it = tid; it < si.nrowitem; it += QX_BLOCK) {
        const int2 pr = B.spitem[si.rowitem_off + it];
        const int ia = pr.x & 0xffff, row = pr.x >> 16, ib = pr.y;
        const ShellDesc sa = sh[ia], sb = sh[ib];
        const ElemParams& ea = g_tab.e[sa.el];
        const ElemParams& eb = g_tab.e[sb.el];
        const double* ra = sx + 3 * sa.at;
        const double* rb = sx + 3 * sb.at;
        const double t[3] = {rb[0] - ra[0], rb[1] - ra[1], rb[2] - ra[2]};
        double hs = 0.0;
        const bool off = sa.at != sb.at;
        if (off) {
            const double r = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
            double pir, dpir, kxy;
            hs = 0.5 * (B.hself[si.shell_off + ia] + B.hself[si.shell_off + ib]) *
                 h0_scale(ea, sa.k, eb, sb.k, r, pir, dpir, kxy);
        }
        const IntegralOut io{S, H, D, Q, ld, ms};
        const bool same = ia == ib;
        switch ((sa.l * 2 + sb.l) * 3 + row) {
        case 9: setup_shell_row<1, 1, 0>(sa, sb, ea, eb, ra, rb, t, hs, off, same, io); break;
        case 10: setup_shell_row<1, 1, 1>(sa, sb, ea, eb, ra, rb, t, hs, off, same, io); break;
        case 11: setup_shell_row<1, 1, 2>(sa, sb, ea, eb, ra, rb, t, hs, off, same, io); break;
        case 6: setup_shell_row<1, 0, 0>(sa, sb, ea, eb, ra, rb, t, hs, off, same, io); break;
        case 7: setup_shell_row<1, 0, 1>(sa, sb, ea, eb, ra, rb, t, hs, off, same, io); break;
        case 8: setup_shell_row<1, 0, 2>(sa, sb, ea, eb, ra, rb, t, hs, off, same, io); break;
        case 3: setup_shell_row<0, 1, 0>(sa, sb, ea, eb, ra, rb, t, hs, off, same, io); break;
        default: setup_shell_row<0, 0, 0>(sa, sb, ea, eb, ra, rb, t, hs, off, same, io); break;
        }
    }
    __syncthreads();
    for (int s = tid; s < nsh; s += QX_BLOCK)
        for (int a = 0; a < sh[s].nao; ++a) {
            const int mu = sh[s].ao0 + a;
            H[(size_t)mu * ld + mu] = B.hself[si.shell_off + s];
        }
    // 5. ES2 gamma matrix
    double* G = B.gam + si.gam_off;
    for (int p = tid; p < nsh * nsh; p += QX_BLOCK) {
        const int a = p / nsh, b = p % nsh;
        const double* ra = sx + 3 * sh[a].at;
        const double* rb = sx + 3 * sh[b].at;
        const double dx = ra[0] - rb[0], dy = ra[1] - rb[1], dz = ra[2] - rb[2];
        const double eta = 0.5 * (g_tab.e[sh[a].el].hard[sh[a].k] + g_tab.e[sh[b].el].hard[sh[b].k]);
        G[p] = 1.0 / sqrt(dx * dx + dy * dy + dz * dz + 1.0 / (eta * eta));
    }
    // 6. D4 charge-response matrix wd4[k][l] = gw_k gw_l C6ref_kl disp_ij (i != j):
    //    the pair damping disp_ij into LDS (the Cholesky slot is free here),
    //    then one element per thread, coalesced along l
    const int nd = si.ndim;
    double* WD = B.wd4 + si.d4_off;
    const double* gw = B.gw + si.ref_off;
    double* disp = dyn;  // nat x nat
    int* refat = Ls.refat;
    int* refii = Ls.refii;
    double* zg = Ls.zg;
    __syncthreads();
    for (int k = tid; k < nd; k += QX_BLOCK) {
        int a = 0;
        while (refoff[a + 1] <= k) ++a;
        const ElemParams& e = g_tab.e[sel[a]];
        const int r = k - refoff[a];
        refat[k] = a;
        refii[k] = r;
        zg[k] = d4_zeta(gfn2::kD4Ga, e.gam * gfn2::kD4Gc, e.refq[r] + e.zeff, e.zeff) * gw[k];
    }
    for (int p = tid; p < nat * nat; p += QX_BLOCK) {
        const int i = p / nat, j = p % nat;
        const ElemParams& ei = g_tab.e[sel[i]];
        const ElemParams& ej = g_tab.e[sel[j]];
        const double dx = sx[3 * i] - sx[3 * j], dy = sx[3 * i + 1] - sx[3 * j + 1],
                     dz = sx[3 * i + 2] - sx[3 * j + 2];
        const double r2 = dx * dx + dy * dy + dz * dz;
        disp[p] = (i != j && r2 <= gfn2::kD4RThr) ? d4_disp(ei, ej, r2, nullptr) : 0.0;
    }
    __syncthreads();
    for (int p = tid; p < nd * nd; p += QX_BLOCK) {
        const int k = p / nd, l = p - k * nd;
        const int i = refat[k], j = refat[l];
        WD[p] = disp[i * nat + j] * gw[k] * gw[l] * g_tab.c6ref[sel[i]][refii[k]][sel[j]][refii[l]];
    }
    // 6b. damped multipole kernels (AES2) per ordered pair (i, j), rv = R_i - R_j,
    //     stored as planes c * nat^2 + j * nat + i:
    //     h3 rv (3), h5 rv_a rv_b (xx, xy, yy, xz, yz, zz), r^2 h5
    double* ag = B.aesg + (size_t)si.pair_off / QX_PAIR_STRIDE * QX_AES_COMP;
    const int nn = nat * nat;
    for (int p = tid; p < nn; p += QX_BLOCK) {
        const int j = p / nat, i = p % nat;
        double o[QX_AES_COMP];
        for (int k = 0; k < QX_AES_COMP; ++k) o[k] = 0.0;
        if (i != j) {
            const double rv0 = sx[3 * i] - sx[3 * j], rv1 = sx[3 * i + 1] - sx[3 * j + 1],
                         rv2 = sx[3 * i + 2] - sx[3 * j + 2];
            const double r2 = rv0 * rv0 + rv1 * rv1 + rv2 * rv2;
            const double r = sqrt(r2);
            const double r0 = 0.5 * (B.mrad[si.atom_off + i] + B.mrad[si.atom_off + j]);
            const double r03 = r0 * r0 * r0;
            const double h3 = 1.0 / (r2 * r + 6.0 * r03);
            const double h5 = 1.0 / (r * (r2 * r2 + 6.0 * r03 * r0));
            o[0] = h3 * rv0;
            o[1] = h3 * rv1;
            o[2] = h3 * rv2;
            o[3] = h5 * rv0 * rv0;
            o[4] = h5 * rv0 * rv1;
            o[5] = h5 * rv1 * rv1;
            o[6] = h5 * rv0 * rv2;
            o[7] = h5 * rv1 * rv2;
            o[8] = h5 * rv2 * rv2;
            o[9] = r2 * h5;
        }
        for (int k = 0; k < QX_AES_COMP; ++k) ag[(size_t)k * nn + p] = o[k];
    }
    // 7. repulsion energy
    double erep = 0.0;
    for (int p = tid; p < nat * nat; p += QX_BLOCK) {
        const int i = p / nat, j = p % nat;
        if (j >= i) continue;
        const ElemParams& ei = g_tab.e[sel[i]];
        const ElemParams& ej = g_tab.e[sel[j]];
        const double dx = sx[3 * i] - sx[3 * j], dy = sx[3 * i + 1] - sx[3 * j + 1],
                     dz = sx[3 * i + 2] - sx[3 * j + 2];
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        if (r > gfn2::kRepCutoff) continue;
        const double kexp = (ei.z <= 2 && ej.z <= 2) ? gfn2::kRepKexpLight : gfn2::kRepKexp;
        erep += ei.rep_zeff * ej.rep_zeff / r * exp(-sqrt(ei.rep_alpha * ej.rep_alpha) * pow(r, kexp));
    }
    erep = block_sum(erep, red);
    // 8. ATM energy at zero charges: pair C6 into scratch, then triples
    double* c60 = B.pairbuf + si.pair_off;  // nat*nat
    for (int p = tid; p < nat * nat; p += QX_BLOCK) {
        const int i = p / nat, j = p % nat;
        const ElemParams& ei = g_tab.e[sel[i]];
        const ElemParams& ej = g_tab.e[sel[j]];
        double c = 0.0;
        for (int ii = 0; ii < ei.nref; ++ii) {
            const double zi = zg[refoff[i] + ii];
            for (int jj = 0; jj < ej.nref; ++jj) c += zi * zg[refoff[j] + jj] * g_tab.c6ref[sel[i]][ii][sel[j]][jj];
        }
        c60[p] = c;
    }
    __syncthreads();
    // triples i > j > k in contiguous equal chunks per thread (the pair-wise
    // loop left half the threads idle and the rest unevenly loaded)
    double eatm = 0.0;
    {
        const long long ntri = (long long)nat * (nat - 1) * (nat - 2) / 6;
        const long long chunk = (ntri + QX_BLOCK - 1) / QX_BLOCK;
        long long t = (long long)tid * chunk;
        const long long tend = t + chunk < ntri ? t + chunk : ntri;
        // unrank t: i with C(i,3) <= t < C(i+1,3), then j with C(j,2) <= rest
        int i = 2, j = 1, k = 0;
        if (t < tend) {
            while ((long long)(i + 1) * i * (i - 1) / 6 <= t) ++i;
            long long rest = t - (long long)i * (i - 1) * (i - 2) / 6;
            j = 1;
            while ((long long)(j + 1) * j / 2 <= rest) ++j;
            k = (int)(rest - (long long)j * (j - 1) / 2);
        }
        for (; t < tend; ++t) {
            const double dij[3] = {sx[3 * i] - sx[3 * j], sx[3 * i + 1] - sx[3 * j + 1], sx[3 * i + 2] - sx[3 * j + 2]};
            const double u = dij[0] * dij[0] + dij[1] * dij[1] + dij[2] * dij[2];
            const double dik[3] = {sx[3 * i] - sx[3 * k], sx[3 * i + 1] - sx[3 * k + 1], sx[3 * i + 2] - sx[3 * k + 2]};
            const double djk[3] = {sx[3 * j] - sx[3 * k], sx[3 * j + 1] - sx[3 * k + 1], sx[3 * j + 2] - sx[3 * k + 2]};
            const double w = dik[0] * dik[0] + dik[1] * dik[1] + dik[2] * dik[2];
            const double v = djk[0] * djk[0] + djk[1] * djk[1] + djk[2] * djk[2];
            if (u <= gfn2::kD4AtmThr && w <= gfn2::kD4AtmThr && v <= gfn2::kD4AtmThr) {
                const ElemParams& ei = g_tab.e[sel[i]];
                const ElemParams& ej = g_tab.e[sel[j]];
                const ElemParams& ek = g_tab.e[sel[k]];
                const double cij = gfn2::kD4A1 * sqrt(3.0 * ei.r4r2 * ej.r4r2) + gfn2::kD4A2;
                const double cik = gfn2::kD4A1 * sqrt(3.0 * ei.r4r2 * ek.r4r2) + gfn2::kD4A2;
                const double cjk = gfn2::kD4A1 * sqrt(3.0 * ej.r4r2 * ek.r4r2) + gfn2::kD4A2;
                const double uvw = u * v * w, rijk = sqrt(uvw), cijk = cij * cjk * cik;
                const double aa = u + v - w, bb = u + w - v, cc = w + v - u;
                const double atm = (0.375 * aa * bb * cc / uvw + 1.0) / (rijk * rijk * rijk);
                const double x = pow(cijk / rijk, gfn2::kD4Alp / 3.0);
                const double fd = 1.0 / (1.0 + 6.0 * x);
                const double c9 = gfn2::kD4S9 * sqrt(c60[i * nat + j] * c60[j * nat + k] * c60[i * nat + k]);
                eatm += c9 * atm * fd;
            }
            // next triple in (i, j, k) lexicographic order with k < j < i
            if (++k == j) {
                k = 0;
                if (++j == i) {
                    j = 1;
                    ++i;
                }
            }
        }
    }
    eatm = block_sum(eatm, red);
    if (tid == 0) {
        B.eparts[10 * sys + 6] = eatm;
        B.eparts[10 * sys + 7] = erep;
    }
    if (!with_chol) return;
    // 9. Cholesky S = L L^T and X = L^-T
    double* A = (ld <= QX_LDS_MAXM) ? dyn : (B.L + si.mat_off);
    double* Lg = B.L + si.mat_off;
    double* Xg = B.X + si.mat_off;
    __syncthreads();
    for (size_t i = tid; i < ms; i += QX_BLOCK) A[i] = S[i];
    __syncthreads();
    for (int j = 0; j < n; ++j) {
        // every thread takes the pivot's root itself; thread 0 stores it
        // while the others update the trailing block (two barriers a column)
        const double d = sqrt(A[(size_t)j * ld + j]);
        for (int i = j + 1 + tid; i < n; i += QX_BLOCK) A[(size_t)i * ld + j] /= d;
        __syncthreads();
        if (tid == 0) A[(size_t)j * ld + j] = d;
        // trailing lower triangle: row j + 1 + r by lane r, its columns
        // j + 1 + c (c <= r) by the wavefronts in turn
        const int m = n - j - 1;
        for (int r = tid & 63; r < m; r += 64) {
            const int i = j + 1 + r;
            const double lij = A[(size_t)i * ld + j];
            for (int c = tid >> 6; c <= r; c += QX_BLOCK / 64) {
                const int k = j + 1 + c;
                A[(size_t)i * ld + k] -= lij * A[(size_t)k * ld + j];
            }
        }
        __syncthreads();
    }
    // zero the upper triangle, store L
    for (size_t p = tid; p < ms; p += QX_BLOCK) {
        const int i = (int)(p / ld), k = (int)(p % ld);
        const double v = (k <= i && i < n) ? A[p] : 0.0;
        A[p] = v;
        Lg[p] = v;
    }
    __syncthreads();
    // X = (L^-1)^T: column c of L^-1 by forward substitution (thread c),
    // written as row c of X.  With the factor in LDS the columns are built in
    // LDS too (rows of odd stride) and copied out once.
    const bool lx = ld <= QX_LDS_MAXM;
    const int xs = n | 1;
    double* xl = dyn + (size_t)ld * ld;
    if (!lx) {
        for (size_t p = tid; p < ms; p += QX_BLOCK) Xg[p] = 0.0;
        __syncthreads();
    }
    for (int c = tid; c < n; c += QX_BLOCK) {
        double* xr = lx ? xl + (size_t)c * xs : Xg + (size_t)c * ld;  // X[c][i] = Linv[i][c]
        for (int i = c; i < n; ++i) {
            double v = (i == c) ? 1.0 : 0.0;
#pragma unroll 8
            for (int k = c; k < i; ++k) v -= A[(size_t)i * ld + k] * xr[k];
            xr[i] = v / A[(size_t)i * ld + i];
        }
    }
    if (lx) {
        __syncthreads();
        for (size_t p = tid; p < ms; p += QX_BLOCK) {
            const int c = (int)(p / ld), i = (int)(p % ld);
            Xg[p] = (c < n && i >= c && i < n) ? xl[(size_t)c * xs + i] : 0.0;
        }
    }
    // X as stored: row c holds column c of L^-1, i.e. Xg = (L^-1)^T row-major
}

__global__ __launch_bounds__(QX_BLOCK) void k_setup(BatchPtrs B, int with_chol)
{
    const int sys = blockIdx.x;
    if (B.active && !B.active[sys]) return;
    __shared__ SetupLds Ls;
    extern __shared__ double dyn[];
    setup_system(B, sys, with_chol, Ls, dyn);
}

}  // namespace qx
