/* CPU oracle for the QCxMS GFN2-xTB molecular-dynamics hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm, used by tests/ as the parity checker, by __graft_entry__.smoke()
 * as the checker, and by bench.py's cpu_baseline leg as the CPU baseline.
 * The product (qcxms_amd/) never includes, links or calls anything here.
 *
 * Units follow the reference: Bohr, Hartree, atomic time units, masses in
 * electron masses (xtb_mctc_convert.f90:24-63, mass.f90:3-22).
 */
#ifndef QCXMS_ORACLE_H
#define QCXMS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- constants (xtb_mctc_constants.f90:35, xtb_mctc_convert.f90:24-63) ---- */
#define ORC_KB      3.166808578545117e-06
#define ORC_AUTOAA  0.52917726
#define ORC_AUTOEV  27.21138505
#define ORC_FSTOAU  41.3413733365614
#define ORC_MSTOAU  (1.0 / 2.18769126364e+06) /* xtb_mctc_convert.f90:63 */

double orc_aatoau(void);
double orc_amutoau(void);
double orc_autoamu(void);
double orc_rad(int z);            /* covalent_radii.f90:89-114 (Bohr) */
double orc_covrad_d3(int z);      /* covalent_radii.f90:83-84 (Bohr) */
double orc_atomic_mass(int z);    /* atomic_masses.f90:24-67 (m_e) */

/* ---- MD bookkeeping (bit-exact restatements) ---- */
void orc_leapfrog(int nat, const double *grad, const double *mass, double tstp,
                  double *xyz, double *vel, double *ke);                 /* md.f90:749-773 */
void orc_ekinet(int nat, const double *velo, const double *mass,
                double *e, double *t);                                   /* mdinit.f90:57-76 */
int  orc_impactscale(int nuc, double *velo, const double *mass, const double *velof,
                     double eimp, double ff, double e0);                 /* impact.f90:12-55 */
double orc_setetemp(int nfrag, double eimp, double ax, double ieetemp);  /* utility.f90:69-86 */
int  orc_getspin(int nat, const int32_t *ic, int chrg);                  /* utility.f90:449-464 */
double orc_gnorm(int nuc, const double *grad);                           /* iniqm.f90:718-732 */
int  orc_checkqc(int nuc, double *e, const double *grad, const double *qat,
                 int mchrg);                                             /* iniqm.f90:684-715 */
void orc_fragment_structure(int nat, const int32_t *oz, const double *xyz, double rcut,
                            int at1, int at2, int32_t *frag);            /* fragments.f90:93-182 */
void orc_fragmass(int nat, const int32_t *iat, const int32_t *list, const double *mass,
                  const int32_t *imass, int32_t *nfrag, double *fragm,
                  int32_t *fragat);                                      /* fragments.f90:10-84 */
void orc_intenergy(int nuc, const int32_t *list, const double *mass, const double *velo,
                   int nfrag, double *t);                                /* md.f90:715-741 */

/* ---- MD driver: the EI production loop of md.f90:285-682 (it>0, method<3) ---- */
typedef int (*orc_egrad_fn)(void *ctx, int nat, const int32_t *iat, const double *xyz,
                            int chrg, double etemp, double *e, double *grad,
                            double *qat);

typedef struct {
    /* inputs; isec: fragmentation generation (main.F90:2228), 0 or 1 for the first */
    int nat, chrg, nfragexit, max_steps, isec;
    double tstep, tadd, eimp, ax, ieetemp, etempin;
    /* outputs (md.f90:684-700); ttime is also an input: the time of earlier
     * generations, to which this run adds */
    int nstep, nfrag, fragstate, mdok, exit_reason;
    double tav, epav, ekav, ttime, dtime, atlast, epot, ekin, eerror, etemp;
    double fragm[10], fragT[10];
    /* CID mode (md.f90 with method = 3, icoll >= 1: the MD after a
     * collision): cid != 0 selects it; new_velo (m/s) is the ion's
     * centre-of-mass speed handed over by cid() and returned at the end;
     * count_average is kept by the reference between calls (SAVE);
     * check_fragmented is returned */
    int cid, count_average, check_fragmented;
    double new_velo;
} orc_md_state;

/* exit_reason codes shared with the product's MD engine */
enum { ORC_MD_RUNNING = 0, ORC_MD_ERR_SCF = 1, ORC_MD_ERR_DRIFT = 2, ORC_MD_NFRAG6 = 3,
       ORC_MD_NFRAGEXIT = 4, ORC_MD_FCONST = 5, ORC_MD_MORESTEP = 6, ORC_MD_MAXSTEP = 7,
       ORC_MD_HEAT_FAIL = 8 };

int orc_md_run(orc_md_state *st, const int32_t *iat, const int32_t *imass,
               const double *mass, const double *velof, double *xyz, double *velo,
               double *grad, double *achrg, int32_t *list, orc_egrad_fn egrad, void *ctx,
               int steps_limit);
/* orc_md_run plus the averages md returns to main: axyz[3*nat] and
 * avg_chrg[nat] over the last kdump steps (md.f90:691-698); NULL skips. */
int orc_md_run_avg(orc_md_state *st, const int32_t *iat, const int32_t *imass,
                   const double *mass, const double *velof, double *xyz, double *velo,
                   double *grad, double *achrg, int32_t *list, orc_egrad_fn egrad, void *ctx,
                   int steps_limit, double *axyz, double *avg_chrg);

/* per-step records of orc_md_run on this thread (buf NULL: off) */
#define ORC_MD_TRACE 8
void orc_md_set_trace(double *buf, int cap);
/* per-step coordinates (3 * nat doubles a step) of orc_md_run on this thread */
void orc_md_set_xyz_trace(double *buf, int cap);

/* ---- CID: the collision MD of cid.f90:724-1111 (one collision, icoll = 1) ----
 * The system holds the ion's nuc atoms, then the gas atoms (nuc0 in all; the
 * last atom is the tracked one).  collided and count_average are inputs as
 * well: the reference keeps them (SAVE) from one collision to the next. */
typedef struct {
    /* inputs */
    int nuc, nuc0, chrg, ntot, collided, count_average;
    double tstep, etemp, tinit, beta, velo_cm, ttime;
    /* outputs */
    int nstep, nfrag, gradfail, fragmented, total_steps, step_counter, check_fragmented,
        collision_step, done;
    double epot, ekin, lowest_com, new_velo, avgt, atlast, ecom, velo_diff;
    double fragT[10], eint[10];
} orc_cid_state;

int orc_cid_add_steps(int nuc);                                          /* cid.f90:232-234 */
/* xyz0/velo0/grad0 [3*nuc0] (grad0 and achrg0 [nuc0] are set by the initial
 * egrad), list [nuc], axyz [3*nuc] (the averaged structure cid returns; may be
 * NULL); steps_limit > 0 bounds the steps of this call. */
int orc_cid_run(orc_cid_state *st, const int32_t *iat0, const double *mass0, double *xyz0,
                double *velo0, double *grad0, double *achrg0, int32_t *list, double *axyz,
                orc_egrad_fn egrad, void *ctx, int steps_limit);

/* ---- GFN2-xTB single point with analytic gradient ---- */
typedef struct {
    double econv, pconv;     /* SCC convergence: |dE| and RMS charge/moment change */
    int maxiter;
    double broyden_damp;     /* damping of the modified Broyden mixer */
    int cold_start;          /* 1: start from zero charges (reference behaviour) */
} orc_scc_opts;

typedef struct {
    double energy, e_rep, e_disp, e_es2, e_es3, e_aes, e_axc, e_band, e_ts;
    int iterations, converged, nao, nsh, nel_int;
} orc_gfn2_info;

void orc_scc_default_opts(orc_scc_opts *o);
/* returns 0 on success, 5 unsupported element, -1 SCC not converged */
int orc_gfn2_egrad(int nat, const int32_t *num, const double *xyz, int charge, double etemp,
                   const orc_scc_opts *opts, double *energy, double *grad, double *qat,
                   double *dipole_at, double *qpole_at, orc_gfn2_info *info);
/* energy only (no gradient), same model */
int orc_gfn2_energy(int nat, const int32_t *num, const double *xyz, int charge,
                    double etemp, const orc_scc_opts *opts, double *energy, double *qat);

/* orbitals of the converged SCC (getspec's single point): eps[nao] ascending
 * (Eh), focc[nao] electrons per orbital, coeff[nao*nao] row-major with column
 * k = orbital k, overlap[nao*nao], aoat[nao] 0-based atom of each AO; any
 * output may be NULL */
int orc_gfn2_orbitals(int nat, const int32_t *num, const double *xyz, int charge, double etemp,
                      const orc_scc_opts *opts, double *energy, double *eps, double *focc,
                      double *coeff, double *overlap, int32_t *aoat);
/* MO populations of write_qmo (mo_energ.f90:36-52): qmo[nao*nat] */
void orc_mo_populations(int nat, int n, const double *coeff, const double *overlap,
                        const int32_t *aoat, double *qmo);

/* building blocks exposed for unit pinning */
int  orc_gfn2_supported(int z);
void orc_gfn2_cn(int nat, const int32_t *num, const double *xyz, double *cn, double *dcn);
void orc_d4_cn(int nat, const int32_t *num, const double *xyz, double *cn, double *dcn);
int  orc_sto_ng(int ng, int n, int l, double zeta, double *alpha, double *coeff);
void orc_overlap_shellpair(int la, int ga, const double *aa, const double *ca,
                           int lb, int gb, const double *ab, const double *cb,
                           const double *ra, const double *rb, double *s, double *dpb,
                           double *qpb);
int  orc_gfn2_matrices(int nat, const int32_t *num, const double *xyz, double *s,
                       double *h0, double *dpint, double *qpint, int *nao_out);
void orc_sygv(int n, double *f, double *s, double *eval, double *work);
int  orc_gfn2_gamma(int nat, const int32_t *num, const double *xyz, double *gam, int *nsh);
int  orc_gfn2_interaction(int nat, const int32_t *num, const double *xyz, const double *x,
                          double *e, double *vs, double *vd, double *vq);
double orc_fermi_fill(int n, const double *e, double kt, double nalpha, double *f);
double orc_d4_twobody(int nat, const int32_t *num, const double *xyz, const double *q,
                      double *grad, double *dedq);
double orc_d4_atm(int nat, const int32_t *num, const double *xyz, double *grad);

/* batched CPU baseline: nthreads workers, each owning whole systems */
int orc_gfn2_egrad_batch(int nsys, const int32_t *offsets, const int32_t *num,
                         const double *xyz, const int32_t *charge, const double *etemp,
                         const orc_scc_opts *opts, double *energy, double *grad,
                         double *qat, int32_t *stat, int nthreads);

/* MD driven by the GFN2 oracle (ctx: orc_scc_opts*), and its threaded batch */
int orc_gfn2_md_egrad(void *ctx, int nat, const int32_t *iat, const double *xyz, int chrg,
                      double etemp, double *e, double *g, double *q);
int orc_md_run_batch(int ntraj, const int32_t *offsets, const int32_t *num, const int32_t *imass,
                     const double *mass, const double *velof, double *xyz, double *velo,
                     double *grad, double *qat, int32_t *list, orc_md_state *states,
                     const orc_scc_opts *opts, int steps_limit, int nthreads);

/* ---- post-MD fragment analysis (postmd_oracle.c) ---- */
int orc_nel(int n, const int32_t *iat, int chrg);
int orc_eself(int n, const int32_t *iat, double *e);
void orc_boltz(int nfrag, int mchrg, double temp, const double *ip, double *fragchrg2);
void orc_res_record(double q, int mchrg, int itrj, int icoll, int isec, int j, int l,
                    const int32_t *idum2, const int32_t *idum1, char *out);
int orc_analyse_ips(int nuc, const int32_t *iat, const double *axyz, const int32_t *list,
                    int nfrag, int mchrg, const orc_scc_opts *opts, double *fragip, int *ipok);
int orc_manage_fragments(int nuc, const int32_t *iat, const double *xyz, const double *axyz,
                         const double *qat, const double *mass, const int32_t *imass,
                         double atlast, int itrj, int isec, int maxsec, double btf,
                         const orc_scc_opts *opts, int *mchrg, double *chrgcont, int *nfrag_ok,
                         int32_t *list, int *nfrag_out, int *tcont_out, double *fragchrg3,
                         double *fragchrg, double *fragip, int *ipok, char *asave, char *res,
                         int *nres);

/* ---- fragmentation generations of one EI trajectory (ei_oracle.c) ---- */
typedef struct {
    int isec, nat, nstep, exit_reason, mdok, fragstate, nfrag, max_steps, tcont, mchrg, nfrag_ok,
        nres, continued, asave_written;
    double chrgcont, ttime, epot;
    int32_t atoms[64], list[64];
    char asave[160];
    char res[10][160];
} orc_ei_gen;
/* main.F90:2208-2345 for one trajectory; returns the number of generations
 * written to gens (at most maxgen) */
int orc_ei_run(int nat, const int32_t *iat, const int32_t *imass, const double *mass,
               const double *velof, const double *xyz, const double *velo, int chrg, double eimp,
               double tadd, double tstep, int nmax0, int nfragexit, int maxsec, double btf,
               int itrj, const orc_scc_opts *opts, orc_ei_gen *gens, int maxgen);

#ifdef __cplusplus
}
#endif
#endif
