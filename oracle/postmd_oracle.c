/* Test infrastructure only (the checker, never the product): CPU restatement
 * of QCxMS's post-MD fragment analysis for finished EI trajectories.
 *
 *   orc_eself            iniqm.f90:430-451 (+ get_core_e 17-41, electrons_amount 45-63)
 *   orc_boltz            utility.f90:469-498 (units = 2, eV)
 *   orc_res_record       write_fragments.f90:402-441 record, format
 *                        (F10.7,i3,2i5,2i2,2x,i3,2x,20(i4,i3))
 *   orc_analyse_ips      analyse.f90:14-448 with eqm (iniqm.f90:241-427), prog 8
 *                        (GFN2, etemp 300 K) over the oracle's GFN2 energies
 *   orc_manage_fragments write_fragments.f90:12-457 (EI, method /= 3)
 *
 * Arrays follow the Fortran layout: fragip(nfrag, nchrg) column-major with
 * element (i, j) at [(i-1) + (j-1)*nfrag].
 *
 * Pinning: eself's hardness tables and the record format against the
 * compiled reference and a Fortran writer of the same format
 * (tests/golden/gen_post_md_golden.py); boltz and the charge bookkeeping
 * against closed forms and limits (tests/test_post_md_oracle.py); the IPs
 * are differences of the oracle's GFN2 energies, whose absolute values stay
 * parity unpinned against tblite (see gfn2_oracle.c).
 *
 * Two reference lines read outside their arrays and are restated by their
 * evident intent: write_fragments.f90:119-121 zero fragip(i,1) and then use
 * fragip(i,0); here ip_diff(i,1) = fragip(i,1) and
 * ip_diff(i,j) = fragip(i,j) - fragip(i,j-1) for j > 1 (the neutral as the
 * zeroth charge state).  write_fragments.f90:164 masks with
 * ip_ranking(count_ip-1) at count_ip = 1; here that bound is -huge, the
 * value the array is filled with at write_fragments.f90:148. */
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define NFMAX 10
#define NCMAX 3

static int core_e(int z)
{
    static const int lim[10] = {2, 10, 18, 29, 36, 47, 54, 71, 79, 86};
    static const int nc[10] = {0, 2, 10, 18, 28, 36, 46, 54, 68, 78};
    for (int k = 0; k < 10; ++k)
        if (z <= lim[k]) return nc[k];
    return 0;
}

int orc_nel(int n, const int32_t *iat, int chrg)
{
    double zs = 0.0;
    for (int i = 0; i < n; ++i) zs += (double)(iat[i] - core_e(iat[i]));
    return (int)zs - chrg;
}

/* returns 0, or 5 for an element without tabulated hardness */
int orc_eself(int n, const int32_t *iat, double *e)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) {
        float g, g3;
        if (iat[i] == 1) { g = 0.47259288f; g3 = -0.02448f; }
        else if (iat[i] == 6) { g = 0.42195412f; g3 = 0.167160f; }
        else if (iat[i] == 7) { g = 0.50438193f; g3 = 0.156306f; }
        else if (iat[i] == 8) { g = 0.58691863f; g3 = 0.161466f; }
        else return 5;
        const double z = (double)(iat[i] - core_e(iat[i]));
        s = s + 0.5 * pow(z, 2) * (double)g + pow(z, 3) * (double)g3 / 3.0;
    }
    *e = s;
    return 0;
}

void orc_boltz(int nfrag, int mchrg, double temp, const double *ip, double *fragchrg2)
{
    const double f = temp * ORC_KB * ORC_AUTOEV;
    double esum = 0;
    for (int i = 1; i <= nfrag; ++i)
        for (int j = 1; j <= mchrg; ++j) esum = esum + exp(-ip[(i - 1) + (j - 1) * nfrag] / f);
    for (int i = 1; i <= nfrag; ++i)
        for (int j = 1; j <= mchrg; ++j)
            fragchrg2[(i - 1) + (j - 1) * nfrag] = exp(-ip[(i - 1) + (j - 1) * nfrag] / f) / esum;
}

/* Fortran Iw / Fw.d output: right-justified, asterisks when too wide */
static void fmt_i(char **p, int v, int w)
{
    char b[40];
    int n = snprintf(b, sizeof b, "%d", v);
    if (n > w) {
        for (int k = 0; k < w; ++k) *(*p)++ = '*';
        return;
    }
    for (int k = 0; k < w - n; ++k) *(*p)++ = ' ';
    memcpy(*p, b, (size_t)n);
    *p += n;
}

static void fmt_f(char **p, double v, int w, int d)
{
    char b[80];
    int n = snprintf(b, sizeof b, "%.*f", d, v);
    if (n > w || !isfinite(v)) {
        for (int k = 0; k < w; ++k) *(*p)++ = '*';
        return;
    }
    for (int k = 0; k < w - n; ++k) *(*p)++ = ' ';
    memcpy(*p, b, (size_t)n);
    *p += n;
}

/* icoll < 0: EI item list (no collision number).  out: at least 160 bytes. */
void orc_res_record(double q, int mchrg, int itrj, int icoll, int isec, int j, int l,
                    const int32_t *idum2, const int32_t *idum1, char *out)
{
    int items[8 + 2 * 200];
    int n = 0;
    items[n++] = mchrg;
    items[n++] = itrj;
    if (icoll >= 0) items[n++] = icoll;
    items[n++] = isec;
    items[n++] = j;
    items[n++] = l;
    for (int m = 0; m < l; ++m) {
        items[n++] = idum2[m];
        items[n++] = idum1[m];
    }
    /* edit descriptors after F10.7, in order; -2 = 2x */
    int desc[8 + 40];
    int nd = 0;
    desc[nd++] = 3; desc[nd++] = 5; desc[nd++] = 5; desc[nd++] = 2; desc[nd++] = 2;
    desc[nd++] = -2; desc[nd++] = 3; desc[nd++] = -2;
    for (int r = 0; r < 20; ++r) { desc[nd++] = 4; desc[nd++] = 3; }
    char *p = out;
    fmt_f(&p, q, 10, 7);
    int k = 0;
    for (int d = 0; d < nd; ++d) {
        if (desc[d] == -2) { *p++ = ' '; *p++ = ' '; continue; }
        if (k == n) break;
        fmt_i(&p, items[k++], desc[d]);
    }
    *p = '\0';
}

/* analyse for one trajectory: fragments 1..nfrag of list at coordinates axyz;
 * fragip(nfrag, |mchrg|) out.  Returns 0, or -1 where the reference stops
 * (SCC failure / no electrons to remove). */
int orc_analyse_ips(int nuc, const int32_t *iat, const double *axyz, const int32_t *list,
                    int nfrag, int mchrg, const orc_scc_opts *opts, double *fragip, int *ipok)
{
    const int nch = abs(mchrg);
    int rc = 0;
    *ipok = 1;
    for (int k = 0; k < nfrag * nch; ++k) fragip[k] = 0.0;
    if (nfrag <= 1) return 0;
    int32_t *fz = (int32_t *)malloc(sizeof(int32_t) * (size_t)nuc);
    double *fx = (double *)malloc(sizeof(double) * 3 * (size_t)nuc);
    for (int i = 1; i <= nfrag; ++i) {
        /* avg_frag_struc: the fragment's atoms in input order */
        int na = 0;
        for (int a = 0; a < nuc; ++a)
            if (list[a] == i) {
                fz[na] = iat[a];
                memcpy(fx + 3 * na, axyz + 3 * a, 3 * sizeof(double));
                ++na;
            }
        double e_neut = 0.0, e_ion = 0.0;
        for (int lp = 0; lp <= nch; ++lp) {
            const int chrg = (mchrg < 0) ? -lp : lp;
            const int nel = orc_nel(na, fz, chrg);
            double e = 0.0;
            int st = 0;
            if (nel == 0) st = orc_eself(na, fz, &e);
            else if (nel < 0) st = -1;
            else st = orc_gfn2_energy(na, fz, fx, chrg, 300.0, opts, &e, NULL);
            if (st != 0) rc = -1;
            if (fabs(e) < 1.0e-10 || st != 0) *ipok = 0;
            if (lp == 0) {
                e_neut = e;
                continue;
            }
            e_ion = e;
            if (e_ion != 0 && e_neut != 0) {
                const double ip = (e_ion - e_neut) * ORC_AUTOEV;
                fragip[(i - 1) + (lp - 1) * nfrag] = ip;
                if (mchrg < 0) {
                    if (ip > 40.0 || ip < -35.0) *ipok = 0;
                } else if (mchrg == 1) {
                    if (ip < 0.0 || ip > 50.0) *ipok = 0;
                } else if (mchrg > 1) {
                    if (ip < 0.0 || ip > 100.0) *ipok = 0;
                }
            }
        }
    }
    if (!*ipok) /* attempts 2 (GFN2 again) and 3 (IPEA1, absent) fail the same way */
        for (int i = 1; i <= nfrag; ++i)
            for (int j = 1; j <= mchrg; ++j) fragip[(i - 1) + (j - 1) * nfrag] = 0.0;
    free(fz);
    free(fx);
    return rc;
}

/* manage_fragments, EI.  In/out: *mchrg, *chrgcont, *nfrag_ok.  Out: list,
 * *nfrag, *tcont, fragchrg3[10], fragip[10*3] (column-major nfrag x |mchrg|),
 * *ipok, records: asave (continued fragment) and res[nres] (each 160 bytes). */
int orc_manage_fragments(int nuc, const int32_t *iat, const double *xyz, const double *axyz,
                         const double *qat, const double *mass, const int32_t *imass,
                         double atlast, int itrj, int isec, int maxsec, double btf,
                         const orc_scc_opts *opts, int *mchrg, double *chrgcont, int *nfrag_ok,
                         int32_t *list, int *nfrag_out, int *tcont_out, double *fragchrg3,
                         double *fragchrg, double *fragip, int *ipok, char *asave, char *res,
                         int *nres)
{
    int32_t fragat[200 * 10];
    double fragm[10];
    int32_t nfrag;
    int rc = 0;
    orc_fragment_structure(nuc, iat, xyz, 3.0, 1, 0, list);
    orc_fragmass(nuc, iat, list, mass, imass, &nfrag, fragm, fragat);
    *nfrag_out = nfrag;
    *tcont_out = 0;
    *nres = 0;
    *ipok = 1;
    asave[0] = '\0';
    for (int k = 0; k < 10; ++k) fragchrg3[k] = fragchrg[k] = 0.0;
    for (int k = 0; k < 30; ++k) fragip[k] = 0.0;
    if (nfrag > 5) *nfrag_ok = 0;
    if (!*nfrag_ok) return 0;
    nfrag = 0;
    for (int a = 0; a < nuc; ++a)
        if (list[a] > nfrag) nfrag = list[a];
    const int nch = abs(*mchrg);
    /* natf(i) and iatf(natf(i), i), the element of the fragment's last atom */
    int natf[NFMAX] = {0}, zlast[NFMAX] = {0};
    for (int a = 0; a < nuc; ++a) {
        const int f = list[a] - 1;
        natf[f] += 1;
        zlast[f] = iat[a];
    }
    double fip[NFMAX * NCMAX];
    rc = orc_analyse_ips(nuc, iat, axyz, list, nfrag, *mchrg, opts, fip, ipok);
    for (int k = 0; k < nfrag * nch; ++k) fragip[k] = fip[k];

    double q2[NFMAX * NCMAX], q3[NFMAX];
    for (int k = 0; k < NFMAX * NCMAX; ++k) q2[k] = 0.0;
    for (int k = 0; k < NFMAX; ++k) q3[k] = 0.0;
    if (nfrag > 1) {
        double ipd[NFMAX * NCMAX];
        for (int i = 1; i <= nfrag; ++i)
            for (int j = 1; j <= nch; ++j) {
                const double lower = (j == 1) ? 0.0 : fip[(i - 1) + (j - 2) * nfrag];
                ipd[(i - 1) + (j - 1) * nfrag] = fip[(i - 1) + (j - 1) * nfrag] - lower;
                if (natf[i - 1] == 1 && zlast[i - 1] == 1 && j == 1)
                    ipd[(i - 1) + (j - 1) * nfrag] = fip[(i - 1) + (j - 1) * nfrag];
                else if (natf[i - 1] == 1 && zlast[i - 1] == 1 && j > 1)
                    ipd[(i - 1) + (j - 1) * nfrag] = DBL_MAX;
            }
        const int msize = nfrag * nch;
        if (nch > 1) {
            double rank[1 + NFMAX * NCMAX];
            int sfrag[1 + NFMAX * NCMAX], schrg[1 + NFMAX * NCMAX];
            double ipd2[NFMAX * NCMAX];
            for (int k = 0; k < msize; ++k) ipd2[k] = DBL_MAX;
            rank[0] = -DBL_MAX;
            int found = 0;
            for (int c = 1; c <= msize; ++c) {
                /* minloc(ip_diff, mask = ip_diff > rank(c-1)), column-major */
                int bi = 0, bj = 0;
                double bv = 0.0;
                for (int j = 1; j <= nch; ++j)
                    for (int i = 1; i <= nfrag; ++i) {
                        const double v = ipd[(i - 1) + (j - 1) * nfrag];
                        if (v > rank[c - 1] && (bi == 0 || v < bv)) { bi = i; bj = j; bv = v; }
                    }
                if (bi == 0) break;
                rank[c] = bv;
                sfrag[c] = bi;
                schrg[c] = bj;
                found = c;
            }
            if (found >= nch)
                for (int c = nch; c <= found; ++c)
                    ipd2[(sfrag[c] - 1) + (schrg[c] - 1) * nfrag] =
                        ipd[(sfrag[c] - 1) + (schrg[c] - 1) * nfrag] -
                        ipd[(sfrag[nch - 1] - 1) + (schrg[nch - 1] - 1) * nfrag];
            for (int c = 1; c <= nch - 1 && c <= found; ++c)
                ipd2[(sfrag[c] - 1) + (schrg[c] - 1) * nfrag] = DBL_MAX;
            orc_boltz(nfrag, nch, atlast * btf, ipd2, q2);
            for (int c = 1; c <= nch - 1 && c <= found; ++c)
                q2[(sfrag[c] - 1) + (schrg[c] - 1) * nfrag] = 1.0;
        } else {
            orc_boltz(nfrag, nch, atlast * btf, ipd, q2);
        }
    }
    if (*mchrg < 0)
        for (int k = 0; k < NFMAX * NCMAX; ++k) q2[k] = -q2[k];
    if (nfrag > 1) {
        for (int i = 1; i <= nfrag; ++i)
            for (int j = 1; j <= nch; ++j) q3[i - 1] = q3[i - 1] + q2[(i - 1) + (j - 1) * nfrag];
    } else {
        for (int i = 0; i < NFMAX; ++i) q3[i] = *chrgcont;
    }
    int tcont = 0;
    if (*mchrg > 0) {
        double largest = -1;
        for (int i = 1; i <= nfrag; ++i) {
            const double fa = (double)(float)natf[i - 1];
            if (fa * q3[i - 1] > largest) { largest = fa * q3[i - 1]; tcont = i; }
        }
    } else {
        double largest = 1;
        for (int i = 1; i <= nfrag; ++i) {
            const double fa = (double)(float)natf[i - 1];
            if (fa * q3[i - 1] < largest) { largest = fa * q3[i - 1]; tcont = i; }
        }
    }
    if (nfrag == 1 || isec == maxsec + 1) {
        tcont = 0;
    } else {
        for (int i = 1; i <= nfrag; ++i) q3[i - 1] = q3[i - 1] * fabs(*chrgcont) / abs(*mchrg);
        *chrgcont = q3[tcont - 1];
        const double v = q3[tcont - 1];
        const int nearest = (int)(v < 0 ? -floor(-v + 0.5) : floor(v + 0.5));
        if (nearest > 0 && *mchrg > 0) *mchrg = nearest;
        if (nearest == 0 && *mchrg > 0) *mchrg = 1;
        if (nearest < 0 && *mchrg < 0) *mchrg = nearest;
        if (nearest == 0 && *mchrg < 0) *mchrg = -1;
    }
    for (int a = 0; a < nuc; ++a) fragchrg[list[a] - 1] = fragchrg[list[a] - 1] + qat[a];
    for (int i = 0; i < NFMAX; ++i) fragchrg3[i] = q3[i];
    *tcont_out = tcont;
    for (int j = 1; j <= nfrag; ++j) {
        int32_t id1[200], id2[200];
        int l = 0;
        for (int k = 1; k <= 200; ++k)
            if (fragat[(k - 1) + (j - 1) * 200] != 0) {
                id1[l] = fragat[(k - 1) + (j - 1) * 200];
                id2[l] = k;
                ++l;
            }
        char *dst = (tcont > 0 && tcont == j) ? asave : res + 160 * (*nres)++;
        orc_res_record(q3[j - 1], *mchrg, itrj, -1, isec, j, l, id2, id1, dst);
    }
    return rc;
}
