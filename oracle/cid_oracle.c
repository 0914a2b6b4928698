/* CPU oracle: the collision MD of QCxMS's CID mode for one collision
 * (cid.f90:724-1111, icoll = 1 run of cid() after its set-up).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  The system holds the ion's nuc
 * atoms followed by the collision gas (one atom, or two for N2); the last
 * atom is the one whose distance to the ion's centre of mass is tracked
 * (distArCOM, cid.f90:1124-1139).  Every step: leap-frog over all atoms,
 * egrad of the whole system with checkqc (iniqm.f90:641-655), the ion's
 * centre-of-mass speed, its internal temperature, connectivity and fragment
 * masses of the ion, the averaged fragment structure, the distance test every
 * dumpdist steps (a collision after five receding tests extends the run by
 * nint(800 * 2 dt / fs) steps), the fragmentation extension and the stop
 * test.  Screen and file dumps are left out.  Operation order follows the
 * Fortran (sequential sums, REAL(4) literals where the source has them).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define KB 3.166808578545117e-06               /* xtb_mctc_constants.f90:35 */
#define FSTOAU 41.3413733365614                /* xtb_mctc_convert.f90:54 */
#define MSTOAU (1.0 / 2.18769126364e+06)       /* xtb_mctc_convert.f90:63 */
#define AUTOEV 27.21138505                     /* xtb_mctc_convert.f90:27 */

/* utility.f90:541-562 */
static void center_of_mass(int nuc, const double *mass, const double *xyz, double cm[3])
{
    double tm = 0.0;
    cm[0] = cm[1] = cm[2] = 0.0;
    for (int i = 0; i < nuc; ++i) {
        cm[0] = cm[0] + mass[i] * xyz[3 * i];
        cm[1] = cm[1] + mass[i] * xyz[3 * i + 1];
        cm[2] = cm[2] + mass[i] * xyz[3 * i + 2];
        tm = tm + mass[i];
    }
    cm[0] = cm[0] / tm;
    cm[1] = cm[1] / tm;
    cm[2] = cm[2] / tm;
}

/* cid.f90:1124-1139: distance of the last atom to cm */
static double dist_to_com(int nuc0, const double *xyz0, const double cm[3])
{
    const double d0 = xyz0[3 * (nuc0 - 1)] - cm[0], d1 = xyz0[3 * (nuc0 - 1) + 1] - cm[1],
                 d2 = xyz0[3 * (nuc0 - 1) + 2] - cm[2];
    return sqrt(d0 * d0 + d1 * d1 + d2 * d2);
}

/* cid.f90:232-234 */
int orc_cid_add_steps(int nuc)
{
    int a = 0;
    if (nuc > 10) a = (nuc / 10) * 500;
    if (nuc >= 40) a = (nuc / 10) * 1000;
    return a;
}

/* md.f90:715-741 with its E_int output */
static void intenergy_e(int nuc, const int32_t *list, const double *mass, const double *velo,
                        int nfrag, double *t, double *eint)
{
    int n[10] = {0};
    for (int f = 0; f < 10; ++f) eint[f] = 0.0;
    for (int i = 0; i < nuc; ++i) {
        const int j = list[i];
        if (j < 1 || j > 10) continue;
        const double *v = velo + 3 * i;
        eint[j - 1] = eint[j - 1] + 0.5 * mass[i] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        n[j - 1] = n[j - 1] + 1;
    }
    for (int f = 0; f < nfrag && f < 10; ++f) t[f] = eint[f] / ((double)(1.5f * (float)n[f]) * KB);
}

int orc_cid_run(orc_cid_state *st, const int32_t *iat0, const double *mass0, double *xyz0,
                double *velo0, double *grad0, double *achrg0, int32_t *list, double *axyz,
                orc_egrad_fn egrad, void *ctx, int steps_limit)
{
    const int nuc = st->nuc, nuc0 = st->nuc0;
    const int dumpdist = 10, dumpavg = 50, cnt_steps = 50;
    const double dt = st->tstep;
    const int add_steps = orc_cid_add_steps(nuc);
    double summass = 0.0;
    for (int i = 0; i < nuc; ++i) summass = summass + mass0[i];
    double *avxyz = calloc(3 * (size_t)nuc, sizeof(double));
    double *avxyz2 = calloc(3 * (size_t)nuc, sizeof(double));
    double *store = calloc(3 * (size_t)nuc, sizeof(double));
    int32_t *imass = malloc(sizeof(int32_t) * (size_t)nuc);
    int32_t *fragat = malloc(sizeof(int32_t) * 200 * 10);
    for (int i = 0; i < nuc; ++i) imass[i] = -1;

    /* the initial forces (cid.f90:703) */
    double e = 0.0;
    int rc = egrad(ctx, nuc0, iat0, xyz0, st->chrg, st->etemp, &e, grad0, achrg0);
    int ok = (rc == 0);
    if (ok) ok = orc_checkqc(nuc0, &e, grad0, achrg0, st->chrg);
    st->epot = e;

    double cm[3], old_cm[3];
    center_of_mass(nuc, mass0, xyz0, cm);
    old_cm[0] = cm[0];
    old_cm[1] = cm[1];
    old_cm[2] = cm[2];
    double new_dist = dist_to_com(nuc0, xyz0, cm);
    double lowest = new_dist, highest = new_dist;
    int nstep = 0, m = 0, xyzavg_dump = 0, average_dump = 0, distance_dump = 0, step_counter = 0;
    int total_steps = st->ntot, check_fragmented = 1, cnt = 0, nfrag = 1;
    int collided = st->collided, count_average = st->count_average, fragmented = 0;
    int save_natf[10] = {0};
    double tav = 0.0, avgt = 0.0, atlast = 0.0, new_velo = 0.0, ecom = 0.0;
    int gradfail = !ok, steps_run = 0;
    st->collision_step = 0;

    while (!gradfail) {
        if (steps_limit > 0 && steps_run >= steps_limit) break;
        ++steps_run;
        nstep = nstep + 1;
        if (xyzavg_dump == dumpavg) {
            xyzavg_dump = 0;
            memset(avxyz, 0, sizeof(double) * 3 * (size_t)nuc);
        }
        st->ttime = st->ttime + dt * (1.0 / FSTOAU);
        distance_dump = distance_dump + 1;
        xyzavg_dump = xyzavg_dump + 1;
        average_dump = average_dump + 1;

        double ke;
        orc_leapfrog(nuc0, grad0, mass0, dt, xyz0, velo0, &ke);
        st->ekin = ke;
        rc = egrad(ctx, nuc0, iat0, xyz0, st->chrg, st->etemp, &e, grad0, achrg0);
        ok = (rc == 0);
        if (ok) ok = orc_checkqc(nuc0, &e, grad0, achrg0, st->chrg);
        st->epot = e;
        if (!ok) {
            gradfail = 1;
            break;
        }
        center_of_mass(nuc, mass0, xyz0, cm);
        const double d0 = cm[0] - old_cm[0], d1 = cm[1] - old_cm[1], d2 = cm[2] - old_cm[2];
        const double cm_out = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
        old_cm[0] = cm[0];
        old_cm[1] = cm[1];
        old_cm[2] = cm[2];
        new_velo = (nstep != 1) ? (cm_out / dt) / MSTOAU : 0.0;

        double ekin, tdum;
        orc_ekinet(nuc, velo0, mass0, &ekin, &tdum);
        const double vv = new_velo * MSTOAU;
        const double e_velo = 0.5 * summass * (vv * vv);
        const double e_kin_diff = ekin - e_velo;
        double new_temp = (2.0 * e_kin_diff) / (3.0 * KB * nuc);
        if (nstep == 1) new_temp = st->tinit;
        tav = tav + new_temp;
        m = m + 1;
        avgt = tav / m;
        ecom = (st->beta * e_velo) * AUTOEV;

        /* connectivity and fragment masses of the ion (cid.f90:839-840) */
        double fragm[10];
        orc_fragment_structure(nuc, iat0, xyz0, 3.0, 1, 0, list);
        orc_fragmass(nuc, iat0, list, mass0, imass, &nfrag, fragm, fragat);
        for (int k = 0; k < 3 * nuc; ++k) avxyz[k] = avxyz[k] + xyz0[k];

        /* the averaged fragment structure (cid.f90:847-958) */
        if (nfrag > check_fragmented) {
            count_average = 1;
            check_fragmented = nfrag;
        }
        if (nfrag < check_fragmented) {
            if (count_average) {
                cnt = 0;
                memset(avxyz2, 0, sizeof(double) * 3 * (size_t)nuc);
                memset(store, 0, sizeof(double) * 3 * (size_t)nuc);
                count_average = 0;
                check_fragmented = 1;
            }
        }
        if (count_average) {
            cnt = cnt + 1;
            for (int k = 0; k < 3 * nuc; ++k) {
                avxyz2[k] = avxyz2[k] + xyz0[k];
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
                for (int k = 0; k < 3 * nuc; ++k) store[k] = avxyz2[k] / cnt;
                memset(avxyz2, 0, sizeof(double) * 3 * (size_t)nuc);
                cnt = 0;
                count_average = 0;
            }
        }
        if (average_dump == dumpavg) {
            average_dump = 0;
            atlast = 0.0;
        }
        atlast = avgt;

        /* distance test every dumpdist steps (cid.f90:997-1026) */
        if (distance_dump == dumpdist) {
            distance_dump = 0;
            center_of_mass(nuc, mass0, xyz0, cm);
            new_dist = dist_to_com(nuc0, xyz0, cm);
            if (new_dist > highest) highest = new_dist;
            if (new_dist < lowest) lowest = new_dist;
            if (lowest < new_dist) step_counter = step_counter + 1;
            else step_counter = 0;
            if (step_counter == 5) {
                total_steps = nstep + (int)lround(800.0 * (2.0 * dt * (1.0 / FSTOAU)));
                collided = 1;
                st->collision_step = nstep;
                tav = 0.0;
                m = 0;
            }
        }
        if (nfrag > 1 && collided && !fragmented) {
            total_steps = nstep + add_steps;
            fragmented = 1;
        }
        if (nstep >= total_steps) {
            st->velo_diff = st->velo_cm - new_velo;
            break;
        }
    }

    /* after the loop (cid.f90:1050-1107) */
    intenergy_e(nuc, list, mass0, velo0, nfrag, st->fragT, st->eint);
    st->nstep = nstep;
    st->nfrag = nfrag;
    st->gradfail = gradfail;
    st->collided = collided;
    st->fragmented = fragmented;
    st->count_average = count_average;
    st->total_steps = total_steps;
    st->step_counter = step_counter;
    st->check_fragmented = check_fragmented;
    st->lowest_com = lowest;
    st->new_velo = new_velo;
    st->avgt = avgt;
    st->atlast = atlast;
    st->ecom = ecom;
    st->done = gradfail || nstep >= total_steps;
    if (axyz) {
        if (check_fragmented > 1) memcpy(axyz, store, sizeof(double) * 3 * (size_t)nuc);
        else
            for (int k = 0; k < 3 * nuc; ++k) axyz[k] = avxyz[k] / xyzavg_dump;
    }
    free(avxyz);
    free(avxyz2);
    free(store);
    free(imass);
    free(fragat);
    return 0;
}
