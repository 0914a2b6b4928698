/* CPU oracle: MD bookkeeping of the QCxMS EI production loop.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Each function restates the cited
 * reference lines with the same operation order, so that with FMA contraction
 * off (oracle/Makefile) the results are bit-identical to the Fortran.
 * Default-real literals of the reference are REAL(4); they are reproduced as
 * float constants widened to double where they meet double operands.
 */
#include "oracle.h"
#include <math.h>
#include <string.h>
#include <stdlib.h>

/* xtb_mctc_convert.f90:24-63 */
static const double c_aatoau = 1.0 / 0.52917726;
static const double c_amutokg = 1.660539040e-27;
static const double c_metokg = 9.10938356e-31;

double orc_aatoau(void) { return c_aatoau; }
double orc_amutoau(void) { return c_amutokg * (1.0 / c_metokg); }
double orc_autoamu(void) { return (1.0 / c_amutokg) * c_metokg; }

/* covalent_radii.f90:89-114: "Radius used in QCxMS", Angstrom */
static const double rad_aa[118] = {
    0.32, 0.37,
    1.30, 0.99, 0.84, 0.75, 0.71, 0.64, 0.60, 0.62,
    1.60, 1.40, 1.24, 1.14, 1.09, 1.04, 1.00, 1.01,
    2.00, 1.74,
    1.59, 1.48, 1.44, 1.30, 1.29, 1.24, 1.18, 1.17, 1.22, 1.20,
    1.23, 1.20, 1.20, 1.18, 1.17, 1.16,
    2.15, 1.90,
    1.76, 1.64, 1.56, 1.46, 1.38, 1.36, 1.34, 1.30, 1.36, 1.40,
    1.42, 1.40, 1.40, 1.37, 1.36, 1.36,
    2.38, 2.06,
    1.94, 1.84, 1.90, 1.88, 1.86, 1.85, 1.83,
    1.82, 1.81, 1.80, 1.79, 1.77, 1.77, 1.78,
    1.74, 1.64, 1.58, 1.50, 1.41, 1.36, 1.32, 1.30, 1.30, 1.32,
    1.44, 1.45, 1.50, 1.42, 1.48, 1.46,
    2.42, 2.11,
    2.01, 1.90, 1.84, 1.83, 1.80, 1.80, 1.49,
    1.49, 1.51, 1.51, 1.48, 1.50, 1.56, 1.58,
    1.45, 1.41, 1.34, 1.29, 1.27, 1.21, 1.16, 1.15, 1.09, 1.22,
    1.36, 1.43, 1.46, 1.58, 1.48, 1.57};

/* covalent_radii.f90:27-50 (Pyykko & Atsumi 2009), used x 4/3 for CN */
static const double covrad2009_aa[118] = {
    0.32, 0.46,
    1.20, 0.94, 0.77, 0.75, 0.71, 0.63, 0.64, 0.67,
    1.40, 1.25, 1.13, 1.04, 1.10, 1.02, 0.99, 0.96,
    1.76, 1.54,
    1.33, 1.22, 1.21, 1.10, 1.07, 1.04, 1.00, 0.99, 1.01, 1.09,
    1.12, 1.09, 1.15, 1.10, 1.14, 1.17,
    1.89, 1.67,
    1.47, 1.39, 1.32, 1.24, 1.15, 1.13, 1.13, 1.08, 1.15, 1.23,
    1.28, 1.26, 1.26, 1.23, 1.32, 1.31,
    2.09, 1.76,
    1.62, 1.47, 1.58, 1.57, 1.56, 1.55, 1.51,
    1.52, 1.51, 1.50, 1.49, 1.49, 1.48, 1.53,
    1.46, 1.37, 1.31, 1.23, 1.18, 1.16, 1.11, 1.12, 1.13, 1.32,
    1.30, 1.30, 1.36, 1.31, 1.38, 1.42,
    2.01, 1.81,
    1.67, 1.58, 1.52, 1.53, 1.54, 1.55, 1.49,
    1.49, 1.51, 1.51, 1.48, 1.50, 1.56, 1.58,
    1.45, 1.41, 1.34, 1.29, 1.27, 1.21, 1.16, 1.15, 1.09, 1.22,
    1.36, 1.43, 1.46, 1.58, 1.48, 1.57};

/* atomic_masses.f90:24-53, amu */
static const double mass_amu[118] = {
    1.00794075, 4.00260193, 6.94003660, 9.01218307, 10.81102805, 12.01073590, 14.00670321,
    15.99940492, 18.99840316, 20.18004638, 22.98976928, 24.30505162, 26.98153853,
    28.08549871, 30.97376200, 32.06478741, 35.45293758, 39.94779856, 39.09830091,
    40.07802251, 44.95590828, 47.86674496, 50.94146504, 51.99613176, 54.93804391,
    55.84514443, 58.93319429, 58.69334711, 63.54603995, 65.37778253, 69.72306607,
    72.62755016, 74.92159457, 78.95938856, 79.90352778, 83.79800000, 85.46766360,
    87.61664447, 88.90584030, 91.22364160, 92.90637300, 95.95978854, 97.90721240,
    101.06494014, 102.90549800, 106.41532751, 107.86814963, 112.41155782, 114.81808663,
    118.71011259, 121.75978367, 127.60312648, 126.90447190, 131.29276145, 132.90545196,
    137.32689163, 138.90546887, 140.11573074, 140.90765760, 144.24159603, 144.91275590,
    150.36635571, 151.96437813, 157.25213065, 158.92535470, 162.49947282, 164.93032880,
    167.25908265, 168.93421790, 173.05415017, 174.96681496, 178.48497872, 180.94787564,
    183.84177755, 186.20670455, 190.22485963, 192.21605165, 195.08445686, 196.96656879,
    200.59916703, 204.38341284, 207.21690806, 208.98039910, 208.98243080, 209.98714790,
    222.01757820, 223.01973600, 226.02541030, 227.02775230, 232.03805580, 231.03588420,
    238.02891046, 237.04817360, 244.06420530, 243.06138130, 247.07035410, 247.07030730,
    251.07958860, 252.08298000, 257.09510610, 258.09843150, 259.10103000, 262.10961000,
    267.12179000, 269.12791000, 271.13393000, 270.13336000, 276.14846000, 276.15159000,
    280.16131000, 282.16912000, 284.17416000, 284.17873000, 289.19042000, 288.19274000,
    293.20449000, 292.20746000, 294.21392000};

double orc_rad(int z) { return (z >= 1 && z <= 118) ? c_aatoau * rad_aa[z - 1] : 0.0; }
double orc_covrad_d3(int z)
{
    return (z >= 1 && z <= 118) ? 4.0 / 3.0 * (c_aatoau * covrad2009_aa[z - 1]) : 0.0;
}
double orc_atomic_mass(int z)
{
    return (z >= 1 && z <= 118) ? orc_amutoau() * mass_amu[z - 1] : 0.0;
}

/* md.f90:749-773 */
void orc_leapfrog(int nat, const double *grad, const double *mass, double tstp, double *xyz,
                  double *vel, double *ke)
{
    double k_ = 0.0;
    for (int k = 0; k < nat; ++k) {
        double m = mass[k];
        for (int j = 0; j < 3; ++j) {
            double vold = vel[3 * k + j];
            vel[3 * k + j] = vel[3 * k + j] - (tstp * grad[3 * k + j] / m);
            double vavg = 0.5 * (vold + vel[3 * k + j]);
            xyz[3 * k + j] = xyz[3 * k + j] + (tstp * vel[3 * k + j]);
            k_ = k_ + 0.5 * (m * vavg * vavg);
        }
    }
    *ke = k_;
}

/* mdinit.f90:57-76 */
void orc_ekinet(int nat, const double *velo, const double *mass, double *e, double *t)
{
    double ek = 0.0;
    for (int i = 0; i < nat; ++i) {
        const double *v = velo + 3 * i;
        ek = ek + mass[i] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
    ek = ek * 0.5;
    *e = ek;
    *t = ek / (0.5 * 3.0 * nat * ORC_KB);
}

/* impact.f90:12-55.  Returns the iteration count k; k >= 20000 is the
 * reference's "stop 'error in impactscale'" (velocities are still scaled). */
int orc_impactscale(int nuc, double *velo, const double *mass, const double *velof,
                    double eimp, double ff, double e0)
{
    const double inc = (double)0.0002f, tol = (double)0.001f;
    double esoll = eimp * ff + e0;
    double scal = 0.0, e, t;
    int k = 0;
    double *v = (double *)malloc(sizeof(double) * 3 * (size_t)(nuc > 0 ? nuc : 1));
    for (;;) {
        k = k + 1;
        for (int i = 0; i < nuc; ++i)
            for (int j = 0; j < 3; ++j)
                v[3 * i + j] = velo[3 * i + j] * (1.0 + velof[i] * scal);
        orc_ekinet(nuc, v, mass, &e, &t);
        scal = scal + inc;
        if (esoll - e > tol && k < 20000)
            continue;
        break;
    }
    for (int i = 0; i < nuc; ++i)
        for (int j = 0; j < 3; ++j)
            velo[3 * i + j] = velo[3 * i + j] * (1.0 + velof[i] * scal);
    free(v);
    return k;
}

/* utility.f90:69-86 */
double orc_setetemp(int nfrag, double eimp, double ax, double ieetemp)
{
    double etemp = (double)5000.0f + (double)20000.0f * ax;
    if (eimp > 0 && nfrag <= 1) {
        double tmp = eimp > 0.0 ? eimp : 0.0;
        etemp = etemp + tmp * ieetemp;
    }
    return etemp;
}

/* utility.f90:449-464 */
int orc_getspin(int nat, const int32_t *ic, int chrg)
{
    int j = 0;
    for (int i = 0; i < nat; ++i) j += ic[i];
    j = j - abs(chrg);
    int isp = 1 + j % 2;
    if (j < 1) isp = -1;
    return isp;
}

/* iniqm.f90:718-732, including the y-component counted twice */
double orc_gnorm(int nuc, const double *grad)
{
    double gn = 0.0;
    for (int i = 0; i < nuc; ++i)
        gn = gn + grad[3 * i] * grad[3 * i] + grad[3 * i + 1] * grad[3 * i + 1] +
             grad[3 * i + 1] * grad[3 * i + 1];
    return sqrt(gn);
}

/* iniqm.f90:684-715.  Returns ok; may zero *e exactly as the reference. */
int orc_checkqc(int nuc, double *e, const double *grad, const double *qat, int mchrg)
{
    if (fabs(*e) < 1.0e-8) return 0;
    double gn = orc_gnorm(nuc, grad);
    if (gn < 1.0e-8 || gn > (double)20.0f) {
        *e = 0.0;
        return 0;
    }
    if (mchrg > 0) {
        double mx = -HUGE_VAL;
        for (int i = 0; i < nuc; ++i)
            if (qat[i] > mx) mx = qat[i];
        if (fabs(mx) < 1.0e-5) {
            *e = 0.0;
            return 0;
        }
    }
    return 1;
}

/* fragments.f90:93-182, same loops and visiting order (1-based labels) */
void orc_fragment_structure(int nat, const int32_t *oz, const double *xyz, double rcut,
                            int at1, int at2, int32_t *frag)
{
    unsigned char *connect = (unsigned char *)calloc((size_t)nat * (size_t)nat, 1);
    for (int i = 0; i < nat - 1; ++i)
        for (int j = i + 1; j < nat; ++j) {
            double dx = xyz[3 * i] - xyz[3 * j], dy = xyz[3 * i + 1] - xyz[3 * j + 1],
                   dz = xyz[3 * i + 2] - xyz[3 * j + 2];
            double r = sqrt(dx * dx + dy * dy + dz * dz);
            double rcov = rcut * 0.5 * (orc_rad(oz[i]) + orc_rad(oz[j]));
            if (r < rcov) {
                connect[(size_t)i * nat + j] = 1;
                connect[(size_t)j * nat + i] = 1;
            }
        }
    if (at1 == 0 && at2 == 0) {
        for (int i = 0; i < nat; ++i) frag[i] = 1;
        free(connect);
        return;
    }
    for (int i = 0; i < nat; ++i) frag[i] = 0;
    frag[at1 - 1] = 1;
    int attotal = 1;
    if (at2 != 0) {
        connect[(size_t)(at1 - 1) * nat + (at2 - 1)] = 0;
        connect[(size_t)(at2 - 1) * nat + (at1 - 1)] = 0;
    }
    int finish = 0, currentfrag = 0;
    while (attotal != nat) {
        currentfrag = currentfrag + 1;
        while (!finish) {
            finish = 1;
            for (int i = 0; i < nat; ++i) {
                if (frag[i] != currentfrag) continue;
                for (int j = 0; j < nat; ++j) {
                    if (connect[(size_t)i * nat + j] && frag[j] == 0) {
                        frag[j] = currentfrag;
                        attotal = attotal + 1;
                        finish = 0;
                    }
                }
            }
        }
        for (int i = 0; i < nat; ++i)
            if (frag[i] == 0) {
                frag[i] = currentfrag + 1;
                attotal = attotal + 1;
                break;
            }
        finish = 0;
    }
    free(connect);
}

/* fragments.f90:10-84 (formula strings omitted: they are built on the host
 * from fragat).  fragm is returned in amu as fragx in the reference.
 * Labels above 10 index past fragm(10)/fragel(:,10) in the reference
 * (undefined behaviour); here such atoms are left out of the sums. */
void orc_fragmass(int nat, const int32_t *iat, const int32_t *list, const double *mass,
                  const int32_t *imass, int32_t *nfrag, double *fragm_out, int32_t *fragat)
{
    double fragm[10] = {0};
    int32_t fragel[10][200];
    memset(fragel, 0, sizeof(fragel));
    for (int i = 0; i < nat; ++i) {
        int l = list[i];
        if (l < 1 || l > 10) continue;
        fragm[l - 1] = fragm[l - 1] + mass[i];
        int j = iat[i];
        if (imass[i] > 0) j = 100 + imass[i];
        if (j >= 1 && j <= 200) fragel[l - 1][j - 1] += 1;
    }
    int nf = 0;
    for (int i = 0; i < 10; ++i) fragm_out[i] = 0.0;
    memset(fragat, 0, sizeof(int32_t) * 200 * 10);
    const double autoamu = orc_autoamu();
    for (int i = 0; i < 10; ++i) {
        if (fragm[i] > 0) {
            fragm_out[nf] = fragm[i] * autoamu;
            memcpy(fragat + 200 * nf, fragel[i], sizeof(int32_t) * 200);
            nf = nf + 1;
        }
    }
    *nfrag = nf;
}

/* md.f90:715-741 */
void orc_intenergy(int nuc, const int32_t *list, const double *mass, const double *velo,
                   int nfrag, double *t)
{
    double eint[10] = {0};
    int n[10] = {0};
    for (int i = 0; i < nuc; ++i) {
        int j = list[i];
        if (j < 1 || j > 10) continue;
        const double *v = velo + 3 * i;
        eint[j - 1] = eint[j - 1] + 0.5 * mass[i] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        n[j - 1] = n[j - 1] + 1;
    }
    for (int i = 0; i < nfrag && i < 10; ++i)
        t[i] = eint[i] / ((double)(1.5f * (float)n[i]) * ORC_KB);
}

/* Optional per-step record for the parity tests (per thread): after the
 * connectivity of every loop iteration, ORC_MD_TRACE doubles: nstep, the
 * electronic temperature set for the next egrad, epot, ekin, nfrag,
 * fragT(1), fragT(2), eerror.  cap bounds the number of records. */
static _Thread_local double *md_trace;
static _Thread_local int md_trace_cap;
void orc_md_set_trace(double *buf, int cap)
{
    md_trace = buf;
    md_trace_cap = buf ? cap : 0;
}

/* Optional per-step coordinates (per thread): 3 * nat doubles after the
 * leap-frog of every loop iteration, at most cap records. */
static _Thread_local double *md_xyz_trace;
static _Thread_local int md_xyz_trace_cap;
void orc_md_set_xyz_trace(double *buf, int cap)
{
    md_xyz_trace = buf;
    md_xyz_trace_cap = buf ? cap : 0;
}

/* md.f90:34-708 restricted to the production branch (it>0, no restart) of
 * EI (method<3, icoll=0) and of CID after a collision (st->cid: method=3,
 * icoll>=1).  egrad follows iniqm.f90:457-679 for prog=8: the
 * callback returns E, grad, qat; checkqc is applied here.  steps_limit
 * bounds the number of loop iterations run by this call (<=0: no bound). */
int orc_md_run(orc_md_state *st, const int32_t *iat, const int32_t *imass, const double *mass,
               const double *velof, double *xyz, double *velo, double *grad, double *achrg,
               int32_t *list, orc_egrad_fn egrad, void *ctx, int steps_limit)
{
    return orc_md_run_avg(st, iat, imass, mass, velof, xyz, velo, grad, achrg, list, egrad, ctx,
                          steps_limit, NULL, NULL);
}

/* As orc_md_run; also returns the averages md hands back to main
 * (md.f90:355-363, 691-698, EI branch): axyz = avxyz / kdump and
 * avg_chrg = avchrg / kdump (either may be NULL). */
int orc_md_run_avg(orc_md_state *st, const int32_t *iat, const int32_t *imass, const double *mass,
                   const double *velof, double *xyz, double *velo, double *grad, double *achrg,
                   int32_t *list, orc_egrad_fn egrad, void *ctx, int steps_limit, double *axyz,
                   double *avg_chrg)
{
    const int nuc = st->nat;
    const int more = 250, avdump = 50, cnt_steps = 50;
    const int cid = st->cid;
    /* later generations exit at two fragments (main.F90:2255) */
    const int nfragexit = st->isec > 1 ? 2 : st->nfragexit;
    double ekin, t, epot;
    double *avchrg = (double *)calloc((size_t)nuc, sizeof(double));
    double *avxyz = (double *)calloc(3 * (size_t)nuc, sizeof(double));
    double *avxyz2 = (double *)calloc(3 * (size_t)nuc, sizeof(double));
    double *store = (double *)calloc(3 * (size_t)nuc, sizeof(double));
    st->mdok = 0;
    st->nfrag = 1;
    st->fragstate = 0;
    st->exit_reason = ORC_MD_RUNNING;
    orc_ekinet(nuc, velo, mass, &ekin, &t);
    double ekinstart = ekin;
    double etemp = (st->etempin < 0) ? orc_setetemp(1, st->eimp, st->ax, st->ieetemp)
                                     : st->etempin;
    for (int i = 0; i < 3 * nuc; ++i) grad[i] = 0.0;
    for (int i = 0; i < nuc; ++i) achrg[i] = 0.0;
    epot = 0.0;
    egrad(ctx, nuc, iat, xyz, st->chrg, etemp, &epot, grad, achrg);
    orc_checkqc(nuc, &epot, grad, achrg, st->chrg);
    st->epot = epot;
    if (epot == 0.0) {
        st->exit_reason = ORC_MD_ERR_SCF;
        free(avchrg);
        free(avxyz);
        free(avxyz2);
        free(store);
        return 0;
    }
    /* the ion's centre-of-mass motion (md.f90:247-255): with CID its kinetic
     * energy is taken out of Ekin, at the start and after every step */
    double old_cm[3] = {0, 0, 0}, summass = 0.0, tm = 0.0;
    for (int i = 0; i < nuc; ++i) {
        for (int c = 0; c < 3; ++c) old_cm[c] = old_cm[c] + mass[i] * xyz[3 * i + c];
        tm = tm + mass[i];
    }
    for (int c = 0; c < 3; ++c) old_cm[c] = old_cm[c] / tm;
    for (int i = 0; i < nuc; ++i) summass = summass + mass[i];
    double new_velo = st->new_velo;
    const double vv0 = new_velo * ORC_MSTOAU;
    const double e_kin0 = 0.5 * summass * (vv0 * vv0);
    const double e_kin_diff0 = ekin - e_kin0;
    double new_temp = (2.0 * e_kin_diff0) / (3.0 * ORC_KB * nuc);
    if (cid) ekin = e_kin_diff0;
    int check_fragmented = 1, count_average = st->count_average, cnt = 0, save_natf[10] = {0};
    int max_steps = st->max_steps;
    const int add_steps = orc_cid_add_steps(nuc);
    /* ttime carries over from an earlier generation (main.F90 passes it through md) */
    double tav = 0, epav = 0, ekav = 0, edum = 0, eerror = 0, atlast = 0, dtime = 0, ttime = st->ttime;
    int nstep = 0, morestep = 0, fconst = 0, kdump = avdump, nfrag = 1;
    double fragT[10] = {0}, fragm[10] = {0};
    int32_t fragat[2000];
    const double tstep = st->tstep;
    int nadd = (int)((st->tadd + tstep) / tstep - 1);
    double fadd = tstep / (st->tadd + tstep);
    int run = 0;
    for (;;) {
        if (steps_limit > 0 && run >= steps_limit) break;
        run++;
        nstep = nstep + 1;
        t = ekin / ((double)(1.5f * (float)nuc) * ORC_KB);
        tav = tav + t;
        epav = epav + epot;
        ekav = ekav + ekin;
        double eav;
        if (nstep > nadd) {
            edum = edum + epot + ekin;
            eav = edum / (double)(float)(nstep - nadd);
        } else {
            eav = epot + ekin;
        }
        eerror = eav - epot - ekin;
        if (nfrag == 1) fragT[0] = tav / nstep;
        int err1 = (epot == 0.0), err2 = (fabs(eerror) > (cid ? (double)0.2f : (double)0.1f));
        if (err1 || err2) {
            st->mdok = (nfrag > 1 && nfrag <= 4) || st->isec > 1;
            st->exit_reason = err1 ? ORC_MD_ERR_SCF : ORC_MD_ERR_DRIFT;
            break;
        }
        if (kdump > avdump - 1) {
            kdump = 0;
            for (int i = 0; i < nuc; ++i) avchrg[i] = 0.0;
            for (int i = 0; i < 3 * nuc; ++i) avxyz[i] = 0.0;
            atlast = 0;
        }
        for (int i = 0; i < nuc; ++i) avchrg[i] = avchrg[i] + achrg[i];
        for (int i = 0; i < 3 * nuc; ++i) avxyz[i] = avxyz[i] + xyz[i];
        atlast = atlast + (cid ? new_temp : t);

        orc_leapfrog(nuc, grad, mass, tstep, xyz, velo, &ekin);
        ttime = ttime + tstep / ORC_FSTOAU;
        for (int i = 0; i < 3 * nuc; ++i) grad[i] = 0.0;
        for (int i = 0; i < nuc; ++i) achrg[i] = 0.0;
        epot = 0.0;
        egrad(ctx, nuc, iat, xyz, st->chrg, etemp, &epot, grad, achrg);
        orc_checkqc(nuc, &epot, grad, achrg, st->chrg);
        kdump = kdump + 1;
        if (nfrag == 1) morestep = 0;
        if (nfrag > 1 && dtime < 1.0e-6) dtime = ttime / (double)1000.0f;

        if (!cid) {
            /* heating and the electronic temperature: EI only (md.f90:438-445) */
            if (nstep <= nadd && nfrag == 1) {
                int k = orc_impactscale(nuc, velo, mass, velof, st->eimp, fadd * nstep, ekinstart);
                if (k >= 20000) {
                    st->exit_reason = ORC_MD_HEAT_FAIL;
                    break;
                }
            }
            double dum = st->eimp - st->eimp * (double)(float)nstep / (double)(float)nadd;
            etemp = orc_setetemp(nfrag, dum, st->ax, st->ieetemp);
        }

        orc_fragment_structure(nuc, iat, xyz, 3.0, 1, 0, list);
        int32_t nf;
        orc_fragmass(nuc, iat, list, mass, imass, &nf, fragm, fragat);
        nfrag = nf;
        if (nfrag > 1) orc_intenergy(nuc, list, mass, velo, nfrag, fragT);
        if (md_xyz_trace && nstep <= md_xyz_trace_cap)
            memcpy(md_xyz_trace + 3 * (size_t)nuc * (size_t)(nstep - 1), xyz, sizeof(double) * 3 * (size_t)nuc);
        if (md_trace && nstep <= md_trace_cap) {
            double *r = md_trace + ORC_MD_TRACE * (size_t)(nstep - 1);
            r[0] = nstep;
            r[1] = etemp;
            r[2] = epot;
            r[3] = ekin;
            r[4] = nfrag;
            r[5] = fragT[0];
            r[6] = fragT[1];
            r[7] = eerror;
        }
        if (nfrag > 6) {
            st->exit_reason = ORC_MD_NFRAG6;
            break;
        }
        if (cid) {
            /* the ion's speed, Ekin without it (md.f90:466-478) */
            double cm[3] = {0, 0, 0};
            double tmm = 0.0;
            for (int i = 0; i < nuc; ++i) {
                for (int c = 0; c < 3; ++c) cm[c] = cm[c] + mass[i] * xyz[3 * i + c];
                tmm = tmm + mass[i];
            }
            for (int c = 0; c < 3; ++c) cm[c] = cm[c] / tmm;
            const double d0 = cm[0] - old_cm[0], d1 = cm[1] - old_cm[1], d2 = cm[2] - old_cm[2];
            const double cm_out = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
            new_velo = (cm_out / tstep) / ORC_MSTOAU;
            for (int c = 0; c < 3; ++c) old_cm[c] = cm[c];
            const double vv = new_velo * ORC_MSTOAU;
            const double e_kin = 0.5 * summass * (vv * vv);
            const double e_kin_diff = ekin - e_kin;
            new_temp = (2.0 * e_kin_diff) / (3.0 * ORC_KB * nuc);
            ekin = e_kin_diff;
            /* the averaged fragment structure and the extension (md.f90:496-627) */
            if (nfrag > check_fragmented) {
                count_average = 1;
                check_fragmented = nfrag;
                max_steps = nstep + add_steps;
            }
            if (nfrag < check_fragmented && count_average) {
                cnt = 0;
                memset(avxyz2, 0, sizeof(double) * 3 * (size_t)nuc);
                memset(store, 0, sizeof(double) * 3 * (size_t)nuc);
                count_average = 0;
                check_fragmented = 1;
            }
            if (count_average) {
                cnt = cnt + 1;
                for (int k = 0; k < 3 * nuc; ++k) {
                    avxyz2[k] = avxyz2[k] + xyz[k];
                    store[k] = avxyz2[k] / cnt;
                }
                int natf[10] = {0};
                for (int i = 0; i < nuc; ++i)
                    if (list[i] >= 1 && list[i] <= 10) natf[list[i] - 1] += 1;
                for (int f = 0; f < nfrag && f < 10; ++f) {
                    if (cnt == 1) save_natf[f] = natf[f];
                    if (natf[f] != save_natf[f]) {
                        cnt = 0;
                        memset(store, 0, sizeof(double) * 3 * (size_t)nuc);
                        memset(avxyz2, 0, sizeof(double) * 3 * (size_t)nuc);
                        break;
                    }
                }
                if (cnt == cnt_steps) {
                    memset(avxyz2, 0, sizeof(double) * 3 * (size_t)nuc);
                    cnt = 0;
                    count_average = 0;
                }
            }
            if (nstep >= max_steps) {
                st->fragstate = 1;
                st->mdok = 1;
                st->exit_reason = ORC_MD_MAXSTEP;
                break;
            }
            continue;
        }
        if (nfrag > nfragexit) {
            st->fragstate = 1;
            st->mdok = 1;
            st->exit_reason = ORC_MD_NFRAGEXIT;
            break;
        }
        if (nfrag >= 2) fconst = fconst + 1;
        else fconst = 0;
        if (fconst > 1000) {
            st->fragstate = 2;
            st->mdok = 1;
            st->exit_reason = ORC_MD_FCONST;
            break;
        }
        if (nfrag >= nfragexit) {
            morestep = morestep + 1;
            if (morestep > more) {
                st->fragstate = 1;
                st->mdok = 1;
                st->exit_reason = ORC_MD_MORESTEP;
                break;
            }
        }
        if (nstep >= st->max_steps) {
            st->fragstate = 1;
            st->mdok = 1;
            st->exit_reason = ORC_MD_MAXSTEP;
            break;
        }
    }
    st->nstep = nstep;
    st->nfrag = nfrag;
    st->tav = nstep ? tav / nstep : 0;
    st->epav = nstep ? epav / nstep : 0;
    st->ekav = nstep ? ekav / nstep : 0;
    st->ttime = ttime;
    st->dtime = dtime;
    st->atlast = kdump ? atlast / kdump : 0;
    st->epot = epot;
    st->ekin = ekin;
    st->eerror = eerror;
    st->etemp = etemp;
    for (int i = 0; i < 10; ++i) {
        st->fragm[i] = fragm[i];
        st->fragT[i] = fragT[i];
    }
    for (int i = 0; i < nuc; ++i)
        if (avg_chrg) avg_chrg[i] = (nstep > 0 && kdump) ? avchrg[i] / kdump : 0.0;
    for (int i = 0; i < 3 * nuc; ++i)
        if (axyz)
            axyz[i] = check_fragmented > 1 ? store[i] : ((nstep > 0 && kdump) ? avxyz[i] / kdump : 0.0);
    st->check_fragmented = check_fragmented;
    st->count_average = count_average;
    st->new_velo = new_velo;
    free(avchrg);
    free(avxyz);
    free(avxyz2);
    free(store);
    return nstep;
}
