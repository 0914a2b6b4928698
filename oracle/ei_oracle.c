/* CPU oracle: fragmentation generations of one EI trajectory.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Restates the production loop of
 * src/main.F90:2208-2345 (EI branch) on top of the oracle's MD
 * (orc_md_run_avg, md.f90) and manage_fragments (orc_manage_fragments,
 * write_fragments.f90): MD of the current structure; if the run counts
 * (mdok), manage_fragments; if it picks a fragment tcont and maxsec > 0,
 * the next generation continues with that fragment's atoms recentred on
 * sum(Z x)/sum(Z) (main.F90:2303-2325), velocities kept, velof = 0,
 * tadd = 0, eimp = 0, nmax shrunk (main.F90:2296-2320), isec + 1 (which
 * makes nfragexit 2, main.F90:2255).  The QC initialisation at the top of
 * each generation (iniqm, main.F90:2240-2248) has no effect on the MD and is
 * left out. */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>

int orc_ei_run(int nat, const int32_t *iat, const int32_t *imass, const double *mass,
               const double *velof, const double *xyz, const double *velo, int chrg, double eimp,
               double tadd, double tstep, int nmax0, int nfragexit, int maxsec, double btf,
               int itrj, const orc_scc_opts *opts, orc_ei_gen *gens, int maxgen)
{
    int n = nat, mchrg = chrg, nfrag_ok = 1, nmax = nmax0, isec = 0, ng = 0;
    double chrgcont = (double)mchrg, ttime = 0.0;
    int32_t *z = (int32_t *)malloc(sizeof(int32_t) * nat), *im = (int32_t *)malloc(sizeof(int32_t) * nat);
    int32_t *atoms = (int32_t *)malloc(sizeof(int32_t) * nat), *list = (int32_t *)malloc(sizeof(int32_t) * nat);
    double *m = (double *)malloc(sizeof(double) * nat), *vf = (double *)malloc(sizeof(double) * nat);
    double *x = (double *)malloc(sizeof(double) * 3 * nat), *v = (double *)malloc(sizeof(double) * 3 * nat);
    double *g = (double *)malloc(sizeof(double) * 3 * nat), *q = (double *)malloc(sizeof(double) * nat);
    double *ax = (double *)malloc(sizeof(double) * 3 * nat), *aq = (double *)malloc(sizeof(double) * nat);
    for (int i = 0; i < nat; ++i) {
        z[i] = iat[i];
        im[i] = imass ? imass[i] : -1;
        m[i] = mass[i];
        vf[i] = velof[i];
        atoms[i] = i;
    }
    memcpy(x, xyz, sizeof(double) * 3 * nat);
    memcpy(v, velo, sizeof(double) * 3 * nat);
    int last = -1;
    for (;;) {
        isec = isec + 1;
        if (ng >= maxgen) break;
        orc_md_state st;
        memset(&st, 0, sizeof(st));
        st.nat = n;
        st.chrg = mchrg;
        st.nfragexit = nfragexit;
        st.max_steps = nmax;
        st.isec = isec;
        st.tstep = tstep;
        st.tadd = tadd;
        st.eimp = eimp;
        st.etempin = -1.0;
        st.ttime = ttime;
        orc_md_run_avg(&st, z, im, m, vf, x, v, g, q, list, orc_gfn2_md_egrad, (void *)opts, 0, ax, aq);
        orc_ei_gen *gen = &gens[ng++];
        memset(gen, 0, sizeof(*gen));
        gen->isec = isec;
        gen->nat = n;
        gen->nstep = st.nstep;
        gen->exit_reason = st.exit_reason;
        gen->mdok = st.mdok;
        gen->fragstate = st.fragstate;
        gen->nfrag = st.nfrag;
        gen->max_steps = nmax;
        gen->mchrg = mchrg;
        gen->nfrag_ok = nfrag_ok;
        gen->chrgcont = chrgcont;
        ttime = st.ttime;
        gen->ttime = ttime;
        gen->epot = st.epot;
        for (int i = 0; i < n; ++i) gen->atoms[i] = atoms[i];
        if (!st.mdok) {
            /* main.F90:2340-2345: the continued fragment's record is written */
            if (last >= 0 && gens[last].asave[0]) gens[last].asave_written = 1;
            break;
        }
        int nf = 0, tc = 0, ipok = 0, nres = 0;
        double q3[10], fq[10], fip[30];
        char res[10 * 160];
        orc_manage_fragments(n, z, x, ax, aq, m, im, st.atlast, itrj, isec, maxsec, btf, opts, &mchrg,
                             &chrgcont, &nfrag_ok, list, &nf, &tc, q3, fq, fip, &ipok, gen->asave, res,
                             &nres);
        gen->tcont = tc;
        gen->nfrag_ok = nfrag_ok;
        gen->chrgcont = chrgcont;
        gen->nres = nres;
        for (int k = 0; k < nres && k < 10; ++k) memcpy(gen->res[k], res + 160 * k, 160);
        for (int i = 0; i < n; ++i) gen->list[i] = list[i];
        last = ng - 1;
        if (!nfrag_ok) break;
        if (!(tc > 0 && maxsec > 0)) break;
        nmax = nmax - nmax0 / 5;
        if (nmax < nmax0 / 2) nmax = nmax0 / 2;
        if (st.fragstate == 2) nmax = nmax0 / 3;
        int k = 0;
        double dum = 0.0, cema[3] = {0.0, 0.0, 0.0};
        for (int i = 0; i < n; ++i) {
            if (list[i] != tc) continue;
            z[k] = z[i];
            im[k] = im[i];
            m[k] = m[i];
            atoms[k] = atoms[i];
            for (int c = 0; c < 3; ++c) {
                x[3 * k + c] = x[3 * i + c];
                v[3 * k + c] = v[3 * i + c];
            }
            dum = dum + (double)z[k];
            for (int c = 0; c < 3; ++c) cema[c] = cema[c] + x[3 * k + c] * (double)z[k];
            ++k;
        }
        for (int c = 0; c < 3; ++c) cema[c] = cema[c] / dum;
        n = k;
        if (n <= 10) nmax = nmax0 / 10;
        if (n <= 3) {
            gen->asave_written = gen->asave[0] ? 1 : 0;
            break;
        }
        gen->continued = 1;
        for (int i = 0; i < n; ++i) {
            for (int c = 0; c < 3; ++c) x[3 * i + c] = x[3 * i + c] - cema[c];
            vf[i] = 0.0;
        }
        tadd = 0.0;
        eimp = 0.0;
    }
    free(z); free(im); free(atoms); free(list); free(m); free(vf); free(x); free(v); free(g); free(q);
    free(ax); free(aq);
    return ng;
}
