// Analytic GFN2-xTB gradient (the adjoint of the converged SCC energy), one
// workgroup per system.
//
//   dE/dR = sum_{AB blocks} [2 P dH0 - 2 W dS - P (v_a + v_b) dS
//                            - P (vd . dD + vq : dQ)]      (integral terms)
//         + explicit pair terms (repulsion, ES2 gamma, AES kernels, D4)
//         + CN chain rules (self-energies and multipole radii through the
//           GFN2 CN, D4 weights through the covalent D4 CN).
// Derivative integral blocks are recomputed on the fly by default; with
// B.dblk set, k_dblocks has stored them per AO pair (contracted over the
// primitives) and the gradient only contracts the stored values.
// All reductions are ordered: every partial result has one owner thread, so
// the gradient is bit-reproducible from run to run.
#pragma once
#include "k_scc.h"

namespace qx {

constexpr int kPairStride = QX_PAIR_STRIDE;

// Density-weighted derivative moments of one (L_A, L_B) shell pair
template <int LA, int LB>
QX_DEV void grad_shell_pair(const ElemParams& ea, int ka, const ElemParams& eb, int kb,
                            const double* ra, const double* rb, const double* P, const double* W,
                            int ld, int mu0, int nu0, DerivAcc& acc)
{
    constexpr int NA = LA ? 3 : 1, NB = LB ? 3 : 1;
    double pm[9], wm[9];
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const size_t o = (size_t)(mu0 + a) * ld + nu0 + b;
            pm[a * 3 + b] = P[o];
            wm[a * 3 + b] = W[o];
        }
    const int nga = ea.nprim[ka], ngb = eb.nprim[kb];
    for (int ip = 0; ip < nga; ++ip) {
        const double al = ea.alpha[ka][ip], ca = ea.coeff[ka][ip];
        for (int jp = 0; jp < ngb; ++jp) {
            PrimPair q;
            prim_pair(al, eb.alpha[kb][jp], ca * eb.coeff[kb][jp], ra, rb, q);
            shell_derivs<LA, LB>(q, pm, wm, acc);
        }
    }
}

// Stored derivative blocks: per off-atom shell-pair work item 9 AO-pair slots
// (bra component * 3 + ket component) of kDblk values: s, dpb[3], ds[3],
// ddp[3][3], dqp[3][6] (the DerivAcc members that the density weights, each
// summed over the primitive pairs).
constexpr int kDblk = 34;

template <int LA, int LB, int A = 0, int B = 0>
QX_DEV void store_derivs(const ElemParams& ea, int ka, const ElemParams& eb, int kb, const double* ra,
                         const double* rb, double* out)
{
    DerivAcc acc{};
    const int nga = ea.nprim[ka], ngb = eb.nprim[kb];
    for (int ip = 0; ip < nga; ++ip) {
        const double al = ea.alpha[ka][ip], ca = ea.coeff[ka][ip];
        for (int jp = 0; jp < ngb; ++jp) {
            PrimPair q;
            prim_pair(al, eb.alpha[kb][jp], ca * eb.coeff[kb][jp], ra, rb, q);
            ao_derivs<LA, LB, A, B>(q, 1.0, 0.0, acc);
        }
    }
    double* o = out + (A * 3 + B) * kDblk;
    o[0] = acc.s;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        o[1 + d] = acc.dpb[d];
        o[4 + d] = acc.ds[d];
#pragma unroll
        for (int c = 0; c < 3; ++c) o[7 + 3 * d + c] = acc.ddp[d][c];
#pragma unroll
        for (int k = 0; k < 6; ++k) o[16 + 6 * d + k] = acc.dqp[d][k];
    }
    if constexpr (LB == 1 && B < 2) store_derivs<LA, LB, A, B + 1>(ea, ka, eb, kb, ra, rb, out);
    else if constexpr (LA == 1 && A < 2) store_derivs<LA, LB, A + 1, 0>(ea, ka, eb, kb, ra, rb, out);
}

// The derivative blocks of every off-atom shell pair of every system, stored
// for the gradient (one workgroup per system, one thread per work item).
__global__ __launch_bounds__(QX_BLOCK) void k_dblocks(BatchPtrs B)
{
    const int sys = blockIdx.x;
    if (B.active && !B.active[sys]) return;
    const SysInfo si = B.info[sys];
    __shared__ ShellDesc sh[QX_MAX_NSH];
    __shared__ int ao_at[QX_MAX_NAO], ao_sh[QX_MAX_NAO];
    __shared__ double sx[3 * QX_MAX_NAT];
    build_shell_tables(si, B.elem, sh, ao_at, ao_sh);
    for (int k = threadIdx.x; k < 3 * si.nat; k += QX_BLOCK) sx[k] = B.xyz[3 * (size_t)si.atom_off + k];
    __syncthreads();
    for (int it = threadIdx.x; it < si.nitem_off; it += QX_BLOCK) {
        const int2 pr = B.spitem[si.item_off + it];
        const ShellDesc sa = sh[pr.x], sb = sh[pr.y];
        const ElemParams& ea = g_tab.e[sa.el];
        const ElemParams& eb = g_tab.e[sb.el];
        const double* ra = sx + 3 * sa.at;
        const double* rb = sx + 3 * sb.at;
        double* out = B.dblk + (size_t)(si.item_off + it) * 9 * kDblk;
        switch (sa.l * 2 + sb.l) {
        case 3: store_derivs<1, 1>(ea, sa.k, eb, sb.k, ra, rb, out); break;
        case 2: store_derivs<1, 0>(ea, sa.k, eb, sb.k, ra, rb, out); break;
        case 1: store_derivs<0, 1>(ea, sa.k, eb, sb.k, ra, rb, out); break;
        default: store_derivs<0, 0>(ea, sa.k, eb, sb.k, ra, rb, out); break;
        }
    }
}

// Density-weighted sums of the stored blocks of one work item (the same
// DerivAcc that grad_shell_pair forms from the primitives)
QX_DEV void grad_stored_pair(const double* blk, int na, int nb, const double* P, const double* W, int ld,
                             int mu0, int nu0, DerivAcc& acc)
{
    for (int a = 0; a < na; ++a)
        for (int b = 0; b < nb; ++b) {
            const size_t o = (size_t)(mu0 + a) * ld + nu0 + b;
            const double pm = P[o], wm = W[o];
            const double* v = blk + (a * 3 + b) * kDblk;
            acc.s += pm * v[0];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                acc.dpb[d] += pm * v[1 + d];
                acc.ds[d] += pm * v[4 + d];
                acc.wds[d] += wm * v[4 + d];
#pragma unroll
                for (int c = 0; c < 3; ++c) acc.ddp[d][c] += pm * v[7 + 3 * d + c];
#pragma unroll
                for (int k = 0; k < 6; ++k) acc.dqp[d][k] += pm * v[16 + 6 * d + k];
            }
        }
}

// LDS of one system's gradient
struct GradLds {
    SysShared ss;
    double dedcn[QX_MAX_NAT], dedcn4[QX_MAX_NAT], gacc[3 * QX_MAX_NAT];
};

template <bool STORED = false>
QX_DEV void grad_system(const BatchPtrs& B, int sys, GradLds& Lg)
{
    const SysInfo si = B.info[sys];
    const int tid = threadIdx.x, nat = si.nat, nsh = si.nsh, ld = si.ld;
    SysShared& ss = Lg.ss;
    double* dedcn = Lg.dedcn;
    double* dedcn4 = Lg.dedcn4;
    double* gacc = Lg.gacc;
    load_system(si, B, ss);
    const double* P = B.P + si.mat_off;
    const double* W = B.W + si.mat_off;
    const double* xo = B.xo + si.x_off;
    const double* qsh = xo;
    const double* dp = xo + nsh;
    const double* qp = xo + nsh + 3 * nat;
    const double* vs = B.vs + si.shell_off;
    const double* vd = B.vd + 3 * (size_t)si.atom_off;
    const double* vq = B.vq + 6 * (size_t)si.atom_off;
    const double* hself = B.hself + si.shell_off;
    const double* gw = B.gw + si.ref_off;
    const double* dgw = B.dgw + si.ref_off;
    const double* cn = B.cn + si.atom_off;
    const double* cnd4 = B.cnd4 + si.atom_off;
    const double* dmrad = B.dmrad + si.atom_off;
    double* pb = B.pairbuf + si.pair_off;
    for (int a = tid; a < nat; a += QX_BLOCK) ss.qat[a] = B.qat[si.atom_off + a];
    __syncthreads();

    // D4 charge scaling at the SCC charges (zr, dzr unused here) and at zero
    for (int k = tid; k < si.ndim; k += QX_BLOCK) {
        int a = 0;
        while (ss.refoff[a + 1] <= k) ++a;
        const ElemParams& e = g_tab.e[ss.sel[a]];
        const int r = k - ss.refoff[a];
        const double c = e.gam * gfn2::kD4Gc;
        ss.zr[k] = d4_zeta(gfn2::kD4Ga, c, e.refq[r] + e.zeff, ss.qat[a] + e.zeff);
        ss.yr[k] = d4_zeta(gfn2::kD4Ga, c, e.refq[r] + e.zeff, e.zeff);
    }
    __syncthreads();
    // pre-phase: zero-charge pair C6 and its CN derivative for ATM
    for (int p = tid; p < nat * nat; p += QX_BLOCK) {
        const int i = p / nat, j = p % nat;
        double c = 0.0, dc = 0.0;
        if (i != j) {
            const ElemParams& ei = g_tab.e[ss.sel[i]];
            const ElemParams& ej = g_tab.e[ss.sel[j]];
            for (int ii = 0; ii < ei.nref; ++ii) {
                const int k = ss.refoff[i] + ii;
                for (int jj = 0; jj < ej.nref; ++jj) {
                    const int l = ss.refoff[j] + jj;
                    const double cr = g_tab.c6ref[ss.sel[i]][ii][ss.sel[j]][jj];
                    c += ss.yr[k] * gw[k] * ss.yr[l] * gw[l] * cr;
                    dc += ss.yr[k] * dgw[k] * ss.yr[l] * gw[l] * cr;
                }
            }
        }
        pb[(size_t)p * kPairStride + 10] = c;
        pb[(size_t)p * kPairStride + 11] = dc;
    }
    __syncthreads();

    // Phase A: integral derivative blocks, one work item per off-atom shell
    // pair (atom of ia > atom of ib), grouped by angular-momentum class.
    // The AO-pair density and energy-weighted density enter the primitive
    // sums; the potentials are contracted once per shell pair.
    double* spb = B.spbuf + 5 * si.gam_off;
    for (int it = tid; it < si.nitem_off; it += QX_BLOCK) {
        const int2 pr = B.spitem[si.item_off + it];
        const int ia = pr.x, ib = pr.y;
        const ShellDesc sa = ss.sh[ia], sb = ss.sh[ib];
        const int A = sa.at, Bt = sb.at;
        const ElemParams& ea = g_tab.e[sa.el];
        const ElemParams& eb = g_tab.e[sb.el];
        const double* ra = ss.sx + 3 * A;
        const double* rb = ss.sx + 3 * Bt;
        const double t[3] = {rb[0] - ra[0], rb[1] - ra[1], rb[2] - ra[2]};
        const double r = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
        double pir, dpir, kxy;
        const double scal = h0_scale(ea, sa.k, eb, sb.k, r, pir, dpir, kxy);
        const double hav = 0.5 * (hself[ia] + hself[ib]);
        DerivAcc acc;
        acc.s = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            acc.dpb[d] = acc.ds[d] = acc.wds[d] = 0.0;
#pragma unroll
            for (int c = 0; c < 3; ++c) acc.ddp[d][c] = 0.0;
#pragma unroll
            for (int k = 0; k < 6; ++k) acc.dqp[d][k] = 0.0;
        }
        if constexpr (STORED) {
            grad_stored_pair(B.dblk + (size_t)(si.item_off + it) * 9 * kDblk, sa.nao, sb.nao, P, W, ld,
                             sa.ao0, sb.ao0, acc);
        } else {
            switch (sa.l * 2 + sb.l) {
            case 3: grad_shell_pair<1, 1>(ea, sa.k, eb, sb.k, ra, rb, P, W, ld, sa.ao0, sb.ao0, acc); break;
            case 2: grad_shell_pair<1, 0>(ea, sa.k, eb, sb.k, ra, rb, P, W, ld, sa.ao0, sb.ao0, acc); break;
            case 1: grad_shell_pair<0, 1>(ea, sa.k, eb, sb.k, ra, rb, P, W, ld, sa.ao0, sb.ao0, acc); break;
            default: grad_shell_pair<0, 0>(ea, sa.k, eb, sb.k, ra, rb, P, W, ld, sa.ao0, sb.ao0, acc); break;
            }
        }
        // potential contraction: moments about A follow from those about B by
        // the shift t, which brings in the value terms s and dpb
        double TA[6], TB[6];
        traceless6_t(vq + 6 * A, TA);
        traceless6_t(vq + 6 * Bt, TB);
        const double E00 = 2.0 * TA[0], E01 = TA[1], E11 = 2.0 * TA[2], E02 = TA[3], E12 = TA[4],
                     E22 = 2.0 * TA[5];
        const double Et[3] = {E00 * t[0] + E01 * t[1] + E02 * t[2], E01 * t[0] + E11 * t[1] + E12 * t[2],
                              E02 * t[0] + E12 * t[1] + E22 * t[2]};
        const double* vda = vd + 3 * A;
        const double* vdb = vd + 3 * Bt;
        const double a0 = 2.0 * hav * scal - vs[ia] - vs[ib] -
                          (vda[0] * t[0] + vda[1] * t[1] + vda[2] * t[2]) -
                          0.5 * (Et[0] * t[0] + Et[1] * t[1] + Et[2] * t[2]);
        const double bc[3] = {vda[0] + vdb[0] + Et[0], vda[1] + vdb[1] + Et[1], vda[2] + vdb[2] + Et[2]};
        const double dd[3] = {vda[0] + Et[0], vda[1] + Et[1], vda[2] + Et[2]};
        const double ck[6] = {TA[0] + TB[0], TA[1] + TB[1], TA[2] + TB[2],
                              TA[3] + TB[3], TA[4] + TB[4], TA[5] + TB[5]};
        const double Em[3][3] = {{E00, E01, E02}, {E01, E11, E12}, {E02, E12, E22}};
        const double psum = acc.s;
        const double fpi = 2.0 * psum * hav * kxy * dpir;
        double* o = spb + (size_t)(ia * nsh + ib) * 5;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double g = a0 * acc.ds[d] - 2.0 * acc.wds[d] + dd[d] * acc.s;
#pragma unroll
            for (int c = 0; c < 3; ++c) g += -bc[c] * acc.ddp[d][c] + Em[d][c] * acc.dpb[c];
#pragma unroll
            for (int k = 0; k < 6; ++k) g -= ck[k] * acc.dqp[d][k];
            o[d] = g + fpi * (-t[d] / r);
        }
        o[3] = psum * scal * (-ea.kcn[sa.k]);
        o[4] = psum * scal * (-eb.kcn[sb.k]);
    }

    // Phase B: explicit pair terms, contribution to atom a from ordered pair
    // (a, b).  The three-body loop runs over c > b, so pairs are taken in
    // order of b (heaviest first) and dealt out in snake order over the
    // rounds of the block.
    const int npair = nat * nat;
    for (int r0 = 0; r0 < npair; r0 += QX_BLOCK) {
        const int rank = r0 + (((r0 / QX_BLOCK) & 1) ? QX_BLOCK - 1 - tid : tid);
        if (rank >= npair) continue;
        const int b = rank / nat, a = rank - b * nat;
        const int p = a * nat + b;
        double* o = pb + (size_t)p * kPairStride;
        if (a == b) {
            o[5] = o[6] = o[7] = o[8] = o[9] = 0.0;
            continue;
        }
        const ElemParams& ea = g_tab.e[ss.sel[a]];
        const ElemParams& eb = g_tab.e[ss.sel[b]];
        const double rv[3] = {ss.sx[3 * a] - ss.sx[3 * b], ss.sx[3 * a + 1] - ss.sx[3 * b + 1],
                              ss.sx[3 * a + 2] - ss.sx[3 * b + 2]};
        const double r2 = rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2];
        const double r = sqrt(r2);
        double g[3] = {0, 0, 0};
        double dcn_a = 0.0, dcn4_a = 0.0;
        double sr = 0.0;  // radial factor times 1/r
        // repulsion
        if (r <= gfn2::kRepCutoff) {
            const double kexp = (ea.z <= 2 && eb.z <= 2) ? gfn2::kRepKexpLight : gfn2::kRepKexp;
            const double sa = sqrt(ea.rep_alpha * eb.rep_alpha);
            const double e = ea.rep_zeff * eb.rep_zeff / r * exp(-sa * pow(r, kexp));
            sr += e * (-1.0 / r - sa * kexp * pow(r, kexp - 1.0)) / r;
        }
        // ES2: -q_s q_t gamma^3 R
        for (int s = ss.shoff[a]; s < ss.shoff[a + 1]; ++s) {
            const double ha = g_tab.e[ss.sh[s].el].hard[ss.sh[s].k];
            for (int tt = ss.shoff[b]; tt < ss.shoff[b + 1]; ++tt) {
                const double eta = 0.5 * (ha + g_tab.e[ss.sh[tt].el].hard[ss.sh[tt].k]);
                const double gm = 1.0 / sqrt(r2 + 1.0 / (eta * eta));
                sr += -qsh[s] * qsh[tt] * gm * gm * gm;
            }
        }
        // AES at fixed moments
        {
            const double r0 = 0.5 * (ss.mrad[a] + ss.mrad[b]);
            const double r03 = r0 * r0 * r0, r04 = r03 * r0;
            const double h3 = 1.0 / (r2 * r + 6.0 * r03);
            const double h5 = 1.0 / (r * (r2 * r2 + 6.0 * r04));
            const double dh3 = -3.0 * r2 * h3 * h3;
            const double dh5 = -(5.0 * r2 * r2 + 6.0 * r04) * h5 * h5;
            const double h5b = r2 * h5, dh5b = 2.0 * r * h5 + r2 * dh5;
            const double dh3r0 = -18.0 * r0 * r0 * h3 * h3;
            const double dh5r0 = -24.0 * r03 * r * h5 * h5;
            const double dh5br0 = r2 * dh5r0;
            const double* mi = dp + 3 * a;
            const double* mj = dp + 3 * b;
            const double* ti = qp + 6 * a;
            const double* tj = qp + 6 * b;
            const double mjr = mj[0] * rv[0] + mj[1] * rv[1] + mj[2] * rv[2];
            const double mir = mi[0] * rv[0] + mi[1] * rv[1] + mi[2] * rv[2];
            const double mimj = mi[0] * mj[0] + mi[1] * mj[1] + mi[2] * mj[2];
            const double tjr[3] = {tj[0] * rv[0] + tj[1] * rv[1] + tj[3] * rv[2],
                                   tj[1] * rv[0] + tj[2] * rv[1] + tj[4] * rv[2],
                                   tj[3] * rv[0] + tj[4] * rv[1] + tj[5] * rv[2]};
            const double tir[3] = {ti[0] * rv[0] + ti[1] * rv[1] + ti[3] * rv[2],
                                   ti[1] * rv[0] + ti[2] * rv[1] + ti[4] * rv[2],
                                   ti[3] * rv[0] + ti[4] * rv[1] + ti[5] * rv[2]};
            const double tjrr = tjr[0] * rv[0] + tjr[1] * rv[1] + tjr[2] * rv[2];
            const double tirr = tir[0] * rv[0] + tir[1] * rv[1] + tir[2] * rv[2];
            const double qi = ss.qat[a], qj = ss.qat[b];
            const double srad = (qi * mjr - qj * mir) * dh3 + mimj * dh5b - 3.0 * mir * mjr * dh5 +
                                (qi * tjrr + qj * tirr) * dh5;
            sr += srad / r;
            for (int c = 0; c < 3; ++c)
                g[c] += (qi * mj[c] - qj * mi[c]) * h3 - 3.0 * (mi[c] * mjr + mj[c] * mir) * h5 +
                        2.0 * (qi * tjr[c] + qj * tir[c]) * h5;
            const double der0 = (qi * mjr - qj * mir) * dh3r0 + mimj * dh5br0 -
                                3.0 * mir * mjr * dh5r0 + (qi * tjrr + qj * tirr) * dh5r0;
            dcn_a += 0.5 * der0 * dmrad[a];
        }
        // D4 two-body at the SCC charges
        if (r2 <= gfn2::kD4RThr) {
            double c6 = 0.0, dc = 0.0;
            for (int ii = 0; ii < ea.nref; ++ii) {
                const int k = ss.refoff[a] + ii;
                for (int jj = 0; jj < eb.nref; ++jj) {
                    const int l = ss.refoff[b] + jj;
                    const double cr = g_tab.c6ref[ss.sel[a]][ii][ss.sel[b]][jj];
                    c6 += ss.zr[k] * gw[k] * ss.zr[l] * gw[l] * cr;
                    dc += ss.zr[k] * dgw[k] * ss.zr[l] * gw[l] * cr;
                }
            }
            double dd;
            const double disp = d4_disp(ea, eb, r2, &dd);
            sr += -c6 * dd / r;
            dcn4_a += -dc * disp;
        }
        // ATM at zero charges: triples {a, b, c} with c > b, c != a
        const double u = r2;
        if (u <= gfn2::kD4AtmThr) {
            const double cab = gfn2::kD4A1 * sqrt(3.0 * ea.r4r2 * eb.r4r2) + gfn2::kD4A2;
            const double c6ab = pb[(size_t)(a * nat + b) * kPairStride + 10];
            const double dc6ab = pb[(size_t)(a * nat + b) * kPairStride + 11];
            for (int c = b + 1; c < nat; ++c) {
                if (c == a) continue;
                const double dac[3] = {ss.sx[3 * a] - ss.sx[3 * c], ss.sx[3 * a + 1] - ss.sx[3 * c + 1],
                                       ss.sx[3 * a + 2] - ss.sx[3 * c + 2]};
                const double dbc[3] = {ss.sx[3 * b] - ss.sx[3 * c], ss.sx[3 * b + 1] - ss.sx[3 * c + 1],
                                       ss.sx[3 * b + 2] - ss.sx[3 * c + 2]};
                const double w = dac[0] * dac[0] + dac[1] * dac[1] + dac[2] * dac[2];
                const double v = dbc[0] * dbc[0] + dbc[1] * dbc[1] + dbc[2] * dbc[2];
                if (w > gfn2::kD4AtmThr || v > gfn2::kD4AtmThr) continue;
                const ElemParams& ec = g_tab.e[ss.sel[c]];
                const double cac = gfn2::kD4A1 * sqrt(3.0 * ea.r4r2 * ec.r4r2) + gfn2::kD4A2;
                const double cbc = gfn2::kD4A1 * sqrt(3.0 * eb.r4r2 * ec.r4r2) + gfn2::kD4A2;
                const double uvw = u * v * w, rijk = sqrt(uvw), cijk = cab * cbc * cac;
                const double A_ = u + v - w, B_ = u + w - v, C_ = w + v - u;
                const double tt = A_ * B_ * C_ / uvw;
                const double r3 = 1.0 / (uvw * rijk);
                const double atm = (0.375 * tt + 1.0) * r3;
                const double alp3 = gfn2::kD4Alp / 3.0;
                const double x = pow(cijk / rijk, alp3);
                const double fd = 1.0 / (1.0 + 6.0 * x);
                const double c6ac = pb[(size_t)(a * nat + c) * kPairStride + 10];
                const double c6bc = pb[(size_t)(b * nat + c) * kPairStride + 10];
                const double c9 = gfn2::kD4S9 * sqrt(c6ab * c6bc * c6ac);
                const double eabc = c9 * atm * fd;
                const double dtt_du = (B_ * C_ + A_ * C_ - A_ * B_) / uvw - tt / u;
                const double dtt_dw = (-B_ * C_ + A_ * C_ + A_ * B_) / uvw - tt / w;
                const double datm_du = 0.375 * dtt_du * r3 - 1.5 * (0.375 * tt + 1.0) * r3 / u;
                const double datm_dw = 0.375 * dtt_dw * r3 - 1.5 * (0.375 * tt + 1.0) * r3 / w;
                const double df = fd * fd * 6.0 * alp3 * x * 0.5;
                const double de_du = c9 * (datm_du * fd + atm * df / u);
                const double de_dw = c9 * (datm_dw * fd + atm * df / w);
                for (int q = 0; q < 3; ++q) g[q] += 2.0 * (de_du * rv[q] + de_dw * dac[q]);
                const double dc6ac = pb[(size_t)(a * nat + c) * kPairStride + 11];
                dcn4_a += 0.5 * eabc * (dc6ab / c6ab + dc6ac / c6ac);
            }
        }
        o[5] = g[0] + sr * rv[0];
        o[6] = g[1] + sr * rv[1];
        o[7] = g[2] + sr * rv[2];
        o[8] = dcn_a;
        o[9] = dcn4_a;
    }
    __syncthreads();
    // Phase C: per-atom reduction of forces and CN derivatives: a wavefront
    // per atom, a lane per partner, tree sums over the lanes
    {
        const int lane = tid & 63;
        for (int a = tid >> 6; a < nat; a += QX_BLOCK / 64) {
            double g0 = 0.0, g1 = 0.0, g2 = 0.0, dc = 0.0, dc4 = 0.0;
            const int b = lane;
            if (b < nat && b != a) {
                const double* o = pb + (size_t)(a * nat + b) * kPairStride;
                g0 = o[5];
                g1 = o[6];
                g2 = o[7];
                dc = o[8];
                dc4 = o[9];
                if (b < a) {
                    for (int ia = ss.shoff[a]; ia < ss.shoff[a + 1]; ++ia)
                        for (int ib = ss.shoff[b]; ib < ss.shoff[b + 1]; ++ib) {
                            const double* sp = spb + (size_t)(ia * nsh + ib) * 5;
                            g0 += sp[0];
                            g1 += sp[1];
                            g2 += sp[2];
                            dc += sp[3];
                        }
                } else {
                    for (int ia = ss.shoff[b]; ia < ss.shoff[b + 1]; ++ia)
                        for (int ib = ss.shoff[a]; ib < ss.shoff[a + 1]; ++ib) {
                            const double* sp = spb + (size_t)(ia * nsh + ib) * 5;
                            g0 -= sp[0];
                            g1 -= sp[1];
                            g2 -= sp[2];
                            dc += sp[4];
                        }
                }
            }
            g0 = wave_sum(g0);
            g1 = wave_sum(g1);
            g2 = wave_sum(g2);
            dc = wave_sum(dc);
            dc4 = wave_sum(dc4);
            if (lane == 0) {
                // on-site self-energies
                for (int s = ss.shoff[a]; s < ss.shoff[a + 1]; ++s) {
                    double pd = 0.0;
                    for (int k = 0; k < ss.sh[s].nao; ++k) {
                        const int mu = ss.sh[s].ao0 + k;
                        pd += P[(size_t)mu * ld + mu];
                    }
                    dc += pd * (-g_tab.e[ss.sh[s].el].kcn[ss.sh[s].k]);
                }
                gacc[3 * a] = g0;
                gacc[3 * a + 1] = g1;
                gacc[3 * a + 2] = g2;
                dedcn[a] = dc;
                dedcn4[a] = dc4;
            }
        }
    }
    __syncthreads();
    // Phase D: CN chain rules, ordered pairs
    for (int p = tid; p < nat * nat; p += QX_BLOCK) {
        const int a = p / nat, b = p % nat;
        double* o = pb + (size_t)p * kPairStride;
        if (a == b) {
            o[12] = o[13] = o[14] = 0.0;
            continue;
        }
        const ElemParams& ea = g_tab.e[ss.sel[a]];
        const ElemParams& eb = g_tab.e[ss.sel[b]];
        const double rv[3] = {ss.sx[3 * a] - ss.sx[3 * b], ss.sx[3 * a + 1] - ss.sx[3 * b + 1],
                              ss.sx[3 * a + 2] - ss.sx[3 * b + 2]};
        const double r2 = rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2];
        const double r = sqrt(r2);
        double f = 0.0, df = 0.0, fac = 0.0;
        if (r2 <= gfn2::kCnCutoff * gfn2::kCnCutoff) {
            gfn_count(ea.covrad + eb.covrad, r, f, df);
            fac += (dedcn[a] + dedcn[b]) * df / r;
        }
        if (r2 <= gfn2::kD4CnThr) {
            d4_count(ea, eb, r, f, df);
            fac += (dedcn4[a] + dedcn4[b]) * df / r;
        }
        o[12] = fac * rv[0];
        o[13] = fac * rv[1];
        o[14] = fac * rv[2];
    }
    __syncthreads();
    for (int a = tid; a < nat; a += QX_BLOCK) {
        double g[3] = {gacc[3 * a], gacc[3 * a + 1], gacc[3 * a + 2]};
        for (int b = 0; b < nat; ++b) {
            if (b == a) continue;
            const double* o = pb + (size_t)(a * nat + b) * kPairStride;
            g[0] += o[12];
            g[1] += o[13];
            g[2] += o[14];
        }
        for (int c = 0; c < 3; ++c) B.grad[3 * (size_t)(si.atom_off + a) + c] = g[c];
    }
}

__global__ __launch_bounds__(QX_BLOCK) void k_grad(BatchPtrs B)
{
    const int sys = blockIdx.x;
    if (B.active && !B.active[sys]) return;
    __shared__ GradLds Lg;
    if (B.dblk) grad_system<true>(B, sys, Lg);
    else grad_system<false>(B, sys, Lg);
}

}  // namespace qx
