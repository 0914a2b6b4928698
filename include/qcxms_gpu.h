/* C ABI of the MI355X GFN2-xTB molecular-dynamics engine (libqcxms_gpu.so).
 *
 * Plain pointers and sizes only.  Units follow the reference: coordinates in
 * Bohr, energies in Hartree, gradients in Hartree/Bohr, temperatures in K,
 * masses in electron masses, time in atomic units.  Arrays are row-major
 * [atom][xyz], which is the memory layout of the Fortran xyz(3,nat).
 * Status codes follow src/tblite.f90:55-58: 0 success, 5 unknown or
 * unsupported method/element, -1 calculation failed (SCC not converged).
 * Every function returns an int status; -2 reports an invalid argument or a
 * HIP runtime error, with the message available from qcxms_gpu_last_error().
 */
#ifndef QCXMS_GPU_H
#define QCXMS_GPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QCXMS_OK 0
#define QCXMS_UNKNOWN_METHOD 5
#define QCXMS_FATAL (-1)
#define QCXMS_INVALID (-2)

/* method selectors, src/tblite.f90:34-40 */
#define QCXMS_GFN1_XTB 1
#define QCXMS_GFN2_XTB 2
#define QCXMS_IPEA1_XTB 11

/* Select the GPU and build the parameter tables.  Called implicitly by the
 * other entry points with device 0; no counterpart in the reference. */
int qcxms_gpu_init(int device);
/* Message of the last failure on this thread ("" if none). */
const char *qcxms_gpu_last_error(void);

/* Drop-in for get_xtb_egrad (src/tblite.f90:65-93): one structure, fresh
 * wavefunction per call.  multiplicity enters as uhf = min(mult-1, 0) = 0
 * (src/tblite.f90:111), i.e. a spin-restricted Fermi-smeared solution.
 * etemp <= 0 (getspec's call, src/mo_spec.f90:63 via src/iniqm.f90:409)
 * asks for integer occupations and is evaluated at 1e-3 K.
 * qat[nat], gradient[3*nat] are written; *stat receives the reference's
 * status code (0, 5 or -1) and is also the return value.  The device arrays
 * of the structure are kept for the next call with the same atoms and
 * charge (one structure per MD step), until the process ends. */
int qcxms_get_xtb_egrad(int32_t nat, const int32_t *num, const double *xyz, int32_t charge,
                        int32_t multiplicity, int32_t method, double etemp, double *qat,
                        double *energy, double *gradient, int32_t *stat);
/* spec_calc of get_xtb_egrad (src/tblite.f90:153-166 feeding write_qmo,
 * src/mo_energ.f90:7-80): the MO spectrum of the structure of the last
 * qcxms_get_xtb_egrad call in this process, as qcxms_batch_mo_spectrum. */
int qcxms_get_xtb_mo_spectrum(int32_t cap, int32_t *nao, int32_t *ihomo, double *emo,
                              double *focc, double *qmo);

/* Batched single points over ragged systems (the per-trajectory egrad calls
 * of iniqm.f90:641-655, many at once).  System s owns atoms
 * offsets[s] .. offsets[s+1]-1 of num/xyz/qat/gradient.  stat[nsys]. */
int qcxms_egrad_batch(int32_t nsys, const int32_t *offsets, const int32_t *num,
                      const double *xyz, const int32_t *charge, const int32_t *multiplicity,
                      int32_t method, const double *etemp, double *qat, double *energy,
                      double *gradient, int32_t *stat);

/* Allocation-free variant of qcxms_egrad_batch: same arguments, results and
 * status codes.  No counterpart in the reference; it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Every device array of the call
 * lives in one static device array of the library, 64 MiB per GPU, held
 * until the process ends (about 1.5 MB per system of 26 atoms, 5.8 MB per
 * system of 60 atoms: 44 and 11 systems);
 * a batch that needs more is rejected with QCXMS_INVALID and a message.  It
 * calls no hipMalloc, hipFree or stream creation and runs on the library's
 * stream.  Calls in one process are serialised by the library (they share
 * the array, which keeps the last call's contents); calls that overlapped on
 * one GPU would race. */
int qcxms_egrad_batch_noalloc(int32_t nsys, const int32_t *offsets, const int32_t *num,
                              const double *xyz, const int32_t *charge,
                              const int32_t *multiplicity, int32_t method, const double *etemp,
                              double *qat, double *energy, double *gradient, int32_t *stat);

/* ---- persistent batches: device-resident state for repeated evaluations --- */
typedef struct qcxms_batch qcxms_batch;

/* Create a batch of nsys systems (layout as qcxms_egrad_batch).  Returns NULL
 * on failure (see qcxms_gpu_last_error). */
qcxms_batch *qcxms_batch_create(int32_t nsys, const int32_t *offsets, const int32_t *num,
                                const int32_t *charge, int32_t method);
void qcxms_batch_destroy(qcxms_batch *b);
/* SCC controls: energy and RMS moment-change thresholds, iteration cap,
 * Broyden damping, Jacobi off-diagonal tolerance.  Defaults: 1e-9, 1e-8, 100,
 * 0.4, 1e-11. */
int qcxms_batch_set_scc(qcxms_batch *b, double econv, double pconv, int32_t maxiter,
                        double damping, double jacobi_tol);
/* warm != 0: start each SCC from the previous call's converged charges,
 * moments and orbitals (same fixed point, fewer iterations).  Default 0
 * (fresh wavefunction per call, as src/tblite.f90:133). */
int qcxms_batch_set_warm_start(qcxms_batch *b, int32_t warm);
/* SCC launch plan of single-point batches, chosen per basis size by the
 * autotuner (qcxms_amd/autotune.py; its MI355X table is compiled in): bin k
 * takes the systems with nao <= max_nao[k] (the first bin that fits, the
 * last one whatever is larger); eig[k] = 1 runs the MFMA eigen refinement
 * before Jacobi, lds[k] = 1 keeps SCC matrices in LDS where they fit (up to
 * 56 basis functions).  A batch whose systems fall into one bin launches
 * with that bin's choice; a batch spanning several bins launches once with
 * (mixed_eig, mixed_lds), or, with mixed_eig = -1, once per contiguous run of
 * systems of one bin, LDS sized for that run, on streams forked from and
 * joined to its own.  The table is read when a batch is created; nbins = 0
 * makes later batches use one launch with their own options.  MD ensembles
 * always use one launch.  No counterpart in the reference. */
int qcxms_set_tuning(int32_t nbins, const int32_t *max_nao, const int32_t *eig, const int32_t *lds,
                     int32_t mixed_eig, int32_t mixed_lds);
int qcxms_get_tuning(int32_t cap, int32_t *nbins, int32_t *max_nao, int32_t *eig, int32_t *lds,
                     int32_t *mixed_eig, int32_t *mixed_lds);
/* on != 0: the gradient contracts derivative integral blocks that a kernel
 * (k_dblocks) stores per AO pair of every off-atom shell pair (34 doubles per
 * AO pair, allocated at the first request and kept); 0 (default): the
 * gradient recomputes them from the primitives and stores nothing.  The
 * results agree to rounding; the two forms trade device memory for work.
 * No counterpart in the reference (tblite forms them inside its gradient). */
int qcxms_batch_set_stored_derivatives(qcxms_batch *b, int32_t on);
/* Device memory the batch holds (bytes), for the memory comparison above. */
int qcxms_batch_device_bytes(qcxms_batch *b, uint64_t *bytes);
/* Energy, gradient and Mulliken charges for all systems at xyz (host
 * arrays).  etemp[nsys] in K.  Outputs may be NULL when not needed. */
int qcxms_batch_egrad(qcxms_batch *b, const double *xyz, const double *etemp, double *energy,
                      double *gradient, double *qat, int32_t *stat);
/* Per-system SCC iteration counts and energy parts
 * (band, es2, es3, aes, axc, disp2, atm, rep, -TS, electronic) [nsys*10]. */
int qcxms_batch_info(qcxms_batch *b, int32_t *iterations, double *eparts);
/* SCC iterations per system summed over all evaluations so far [nsys]. */
int qcxms_batch_scc_iterations_total(qcxms_batch *b, uint64_t *total);
/* Atomic dipoles [3*natoms] and traceless quadrupoles [6*natoms, xx xy yy xz yz zz]. */
int qcxms_batch_multipoles(qcxms_batch *b, double *dipoles, double *quadrupoles);
/* Orbitals of system sys after the batch's last evaluation, in ascending
 * energy, as tblite's wavefunction holds them for the MO spectrum of
 * getspec (tblite.f90:153-166, mo_energ.f90:7-80): emo[nao] (Eh),
 * focc[nao] (0..2), coeff[nao*nao] with orbital k in coeff[k*nao .. k*nao+nao-1]
 * (column-major over the AOs), overlap[nao*nao], aoat[nao] (1-based atom of
 * each AO).  cap is the capacity of the per-AO buffers; *nao receives the
 * basis size.  Any output pointer except nao may be NULL. */
int qcxms_batch_orbitals(qcxms_batch *b, int32_t sys, int32_t cap, int32_t *nao, double *emo,
                         double *focc, double *coeff, double *overlap, int32_t *aoat);

/* MO spectrum of system sys after the batch's last evaluation, computed on
 * the GPU: emo[nao] orbital energies ascending (Eh), focc[nao] electrons per
 * orbital, qmo[nao*nat] (orbital-major) the Mulliken population of each
 * orbital on each atom plus 1e-10, normalised per orbital (write_qmo,
 * src/mo_energ.f90:36-52), *ihomo the 1-based index of the highest orbital
 * holding more than 1e-6 electrons (wfn%homo, src/tblite.f90:159).  cap is
 * the capacity of the per-orbital buffers (qmo holds cap*nat). */
int qcxms_batch_mo_spectrum(qcxms_batch *b, int32_t sys, int32_t cap, int32_t *nao, int32_t *ihomo,
                            double *emo, double *focc, double *qmo);

/* ---- timing hooks (HIP events on the library's stream) ---- */
/* enable != 0 records events around every kernel launch of this batch. */
int qcxms_batch_set_profiling(qcxms_batch *b, int32_t enable);
/* Accumulated kernel time (ms) and launch count for kernel names
 * "setup", "scc", "grad", "dblocks", "md_pre", "md_post", "fused"; counters
 * reset on read. */
int qcxms_batch_kernel_time(qcxms_batch *b, const char *kernel, double *ms, int32_t *launches);


/* ---- molecular dynamics: the EI production loop of src/md.f90:285-682 ----
 * All trajectories of an ensemble advance together; each step is one
 * leap-frog update, one GFN2 energy+gradient (warm-started SCC), heating by
 * impactscale, setetemp, fragment connectivity and the exit tests.  A
 * trajectory that meets an exit condition stops while the others continue. */
typedef struct qcxms_md qcxms_md;

typedef struct {
    double tstep;       /* time step in atomic units (0.5 fs = 0.5*41.3413733365614) */
    double etempin;     /* electronic temperature (K); < 0: setetemp (utility.f90:69-86) */
    double ax;          /* common1 ax (0 for GFN2 runs, input.f90) */
    double ieetemp;     /* IEE-dependent etemp scaling (input.f90:199 default 0) */
    int32_t nfragexit;  /* exit when more fragments than this appear (input.f90:216: 3) */
    int32_t max_steps;  /* nmax = tmax / tstep */
    int32_t use_graph;  /* != 0: replay the five kernels of a step as one hipGraph */
    int32_t extrapolate; /* SCC start of each step from the previous steps' converged
                            moments: polynomial extrapolation of order 1 (linear) to 4
                            (quartic), -1 none (previous step's moments), 0 library
                            default (4) */
    double econv, pconv;  /* SCC thresholds (defaults 1e-9, 1e-8 when 0) */
    int32_t broyden_keep; /* > 0: keep up to this many Broyden secant pairs from step to
                             step (same fixed point); 0: fresh mixing history per step */
    int32_t fused;        /* != 0: one persistent launch runs whole MD iterations per
                             trajectory (bookkeeping, integrals, SCC, gradient) back to
                             back; 0: one launch per kernel and step */
    int32_t eigensolver;  /* SCC eigenvectors of systems with at most 50 basis functions:
                             1 simultaneous rotations by FP64 MFMA products while every
                             angle is below 0.1, then parallel cyclic Jacobi; 2 Jacobi
                             only; 0 library default (1).  Same fixed point and
                             tolerances either way. */
} qcxms_md_params;

/* Final or current state of one trajectory (md.f90:684-700 outputs). */
typedef struct {
    int32_t nstep, nfrag, fragstate, mdok, exit_reason, running, scc_iterations, pad;
    double tav, epav, ekav;  /* averages over the steps run */
    double ttime, dtime, atlast, epot, ekin, eerror, etemp;
    double fragm[10];        /* fragment masses (amu), compacted as fragmass */
    double fragT[10];        /* fragment temperatures (K) */
    /* CID mode (qcxms_md_set_cid; 0 otherwise): the highest fragment count
     * since the last reset (check_fragmented), the averaging flag md keeps
     * for the next call (count_average, SAVE in the reference) and the ion's
     * last centre-of-mass speed in m/s (new_velo, handed back to cid()) */
    int32_t check_fragmented, count_average;
    double new_velo;
} qcxms_md_result;

/* exit_reason codes: 0 running, 1 SCF/checkqc failure (E = 0), 2 energy drift
 * > 0.1 Eh, 3 more than 6 fragments, 4 nfrag > nfragexit, 5 nothing happens
 * (fconst > 1000), 6 nfrag = nfragexit for 250 more steps, 7 max_steps
 * reached, 8 impactscale found no scaling within 20000 iterations. */

/* ntraj trajectories (offsets/num/charge as qcxms_batch_create); per atom:
 * mass (m_e), imass (isotope mass number or -1), xyz (Bohr), velo (au),
 * velof (heating factors, main.F90:843-853); per trajectory: eimp (Eh) and
 * tadd (au).  Runs the initial egrad (md.f90:182). */
qcxms_md *qcxms_md_create(int32_t ntraj, const int32_t *offsets, const int32_t *num,
                          const int32_t *charge, const double *mass, const int32_t *imass,
                          const double *xyz, const double *velo, const double *velof,
                          const double *eimp, const double *tadd, const qcxms_md_params *params);
void qcxms_md_destroy(qcxms_md *md);
/* Advance every running trajectory by up to nsteps loop iterations;
 * *running receives the number still running afterwards (may be NULL). */
int qcxms_md_run(qcxms_md *md, int32_t nsteps, int32_t *running);
/* Advance n ensembles together, each on its own stream: step s of every
 * ensemble is enqueued before step s + 1 of any, so that the device overlaps
 * their kernels and fills the gaps that an ensemble's slowest trajectory
 * leaves in each step.  running[n] (may be NULL) receives each ensemble's
 * count of running trajectories.  No counterpart in the reference (each
 * trajectory is a separate process there, bin/pqcxms:98). */
int qcxms_md_run_group(qcxms_md **mds, int32_t n, int32_t nsteps, int32_t *running);
/* Copy state back: any pointer may be NULL.  list = fragment labels. */
int qcxms_md_read(qcxms_md *md, double *xyz, double *velo, double *grad, double *qat,
                  int32_t *list, qcxms_md_result *results);
/* Averages over the last avdump (50) steps that md returns to main
 * (md.f90:355-363, 691-698): axyz[3*natoms] = average geometry (Bohr), the
 * structure the fragment IPs are computed at (in the CID mode after a new
 * fragment count: the structure averaged over the 50 steps after it,
 * store_avxyz), and achrg[natoms] = average Mulliken charges.  Either
 * pointer may be NULL. */
int qcxms_md_read_averages(qcxms_md *md, double *axyz, double *achrg);
/* Timing hooks of the ensemble's underlying batch (see qcxms_batch_kernel_time). */
qcxms_batch *qcxms_md_batch(qcxms_md *md);
/* Fragmentation generation isec[ntraj] (>= 1) of each trajectory, the isec
 * argument of md (src/md.f90:34, src/main.F90:2228): from the second
 * generation on a trajectory exits at more than two fragments
 * (src/main.F90:2255) and an SCF or drift failure keeps the run
 * (src/md.f90:329).  Default 1. */
int qcxms_md_set_generation(qcxms_md *md, const int32_t *isec);
/* The CID mode of the production MD (md.f90 with method = 3 after a
 * collision, icoll >= 1; the call main.F90 makes after each cid()):
 * trajectories with cid[s] != 0 take the ion's flight out of the kinetic
 * energy (velo_cm[s]: the ion's centre-of-mass speed in m/s, new_velo of
 * cid(); Ekin and the temperatures are those of the motion inside the ion,
 * md.f90:247-255, 464-487), get no heating and no setetemp, a drift limit of
 * 0.2 Eh, and no fragment exits: a new fragment count starts the averaged
 * structure over the next 50 steps and moves the end to add_steps after it
 * (cid.f90:232-234, md.f90:496-627); the run ends at max_steps.
 * count_average[s] is the averaging flag the reference keeps between calls
 * (may be NULL: 0).  Set once per trajectory, before its first step. */
int qcxms_md_set_cid(qcxms_md *md, const int32_t *cid, const double *velo_cm,
                     const int32_t *count_average);

/* ---- fragmentation generations: the EI loop of src/main.F90:2208-2345 ----
 * Each trajectory runs its MD (md.f90); when it ends with mdok,
 * manage_fragments (write_fragments.f90:12-457) assigns fragment charges and
 * picks the fragment tcont that the next generation continues with:
 * its atoms recentred on their nuclear-charge-weighted centre (sum Z x / sum Z),
 * velocities kept, no heating (velof = 0, tadd = 0, eimp = 0), isec + 1,
 * the new total charge, and nmax shrunk as main.F90:2296-2320 does.  The
 * ensemble advances in chunks of repack_every steps; after a chunk in which
 * trajectories ended, the running trajectories and the new generations are
 * laid out again on the device (state gathered in place, largest basis
 * first). */
typedef struct qcxms_ei qcxms_ei;

/* One generation of one trajectory. */
typedef struct {
    int32_t traj;        /* trajectory index (creation order, from 0) */
    int32_t isec;        /* generation, 1 for the first MD */
    int32_t nat;         /* atoms of this generation */
    int32_t nstep, exit_reason, mdok, fragstate, nfrag, max_steps;
    int32_t tcont;       /* fragment continued by the next generation, 0: none */
    int32_t mchrg;       /* total charge of this generation's MD */
    int32_t nfrag_ok;    /* manage_fragments' verdict (0: more than 5 fragments) */
    int32_t nres;        /* records in res */
    int32_t continued;   /* 1 when a next generation was started */
    int32_t asave_written; /* 1 when main writes asave to qcxms.res (the next
                              generation failed, or the fragment had <= 3 atoms) */
    int32_t pad;
    double chrgcont;     /* statistical charge after manage_fragments */
    double ttime, epot;
    int32_t atoms[64];   /* the generation's atoms as indices into the first generation */
    int32_t list[64];    /* their fragment labels at the end of the generation */
    char asave[128];     /* record of the continued fragment */
    char res[10][128];   /* qcxms.res records of this generation */
} qcxms_ei_generation;

/* ntraj trajectories as qcxms_md_create; maxsec: generation limit
 * (input.f90:214, default 7); btf: Boltzmann temperature factor
 * (input.f90:204, default 1); repack_every: MD steps between re-layouts
 * (>= 1).  Trajectory t is numbered itrj = t + 1 in the records. */
qcxms_ei *qcxms_ei_create(int32_t ntraj, const int32_t *offsets, const int32_t *num,
                          const int32_t *charge, const double *mass, const int32_t *imass,
                          const double *xyz, const double *velo, const double *velof,
                          const double *eimp, const double *tadd, const qcxms_md_params *params,
                          int32_t maxsec, double btf, int32_t repack_every);
void qcxms_ei_destroy(qcxms_ei *ei);
/* Advance by up to nsteps MD steps (every running generation); *running
 * receives the number of trajectories still running (may be NULL). */
int qcxms_ei_run(qcxms_ei *ei, int32_t nsteps, int32_t *running);
/* Generations finished so far, in the order they finished; *count receives
 * their number (at most cap are copied). */
int qcxms_ei_generations(qcxms_ei *ei, int32_t cap, qcxms_ei_generation *out, int32_t *count);
/* Re-layouts done so far. */
int qcxms_ei_repacks(qcxms_ei *ei, int32_t *count);
/* Trajectory-steps executed so far over all generations. */
int qcxms_ei_steps(qcxms_ei *ei, uint64_t *total);
/* The current layout's ensemble (timing hooks: qcxms_md_batch). */
qcxms_md *qcxms_ei_md(qcxms_ei *ei);

/* ---- batched drop-ins for single reference routines ---- */
/* fragment_structure(nat,iat,xyz,3.0,1,0,list) + fragmass masses
 * (src/fragments.f90:10-182, as called at src/md.f90:450-451); systems with
 * up to 64 atoms.  nfrag[nsys], fragm[10*nsys] (amu). */
int qcxms_fragments_batch(int32_t nsys, const int32_t *offsets, const int32_t *num,
                          const double *xyz, const double *mass, int32_t *list, int32_t *nfrag,
                          double *fragm);
/* leapfrog (src/md.f90:749-773); ekin[nsys]. */
int qcxms_leapfrog_batch(int32_t nsys, const int32_t *offsets, const double *grad,
                         const double *mass, double tstep, double *xyz, double *velo,
                         double *ekin);
/* impactscale (src/impact.f90:12-55); per system eimp, ff, e0; iterations[nsys]
 * (>= 20000: the reference stops with 'error in impactscale'). */
int qcxms_impactscale_batch(int32_t nsys, const int32_t *offsets, double *velo,
                            const double *mass, const double *velof, const double *eimp,
                            const double *ff, const double *e0, int32_t *iterations);

/* ---- post-MD fragment analysis (once per finished trajectory) ---- */
/* analyse (src/analyse.f90:14-448; GFN2 IPs, iprog 8: etemp 300 K) for many
 * trajectories at once: the neutral and ion single points of every fragment
 * run as one ragged batch on the GPU.  Trajectory t owns atoms
 * offsets[t] .. offsets[t+1]-1 with averaged coordinates axyz (Bohr) and
 * fragment labels list (1..nfrag[t], nfrag <= 10); mchrg[t] is its total
 * charge (1 <= |mchrg| <= 3).  Out: fragip[30*t + 10*(j-1) + (i-1)] = IP
 * (EA when mchrg < 0) of fragment i for charge j, eV, zero where not computed
 * (one fragment: analyse.f90:125-130); ipok[t] = 0 when an energy failed or an
 * IP left the bounds of analyse.f90:383-400 (IPs then zeroed as after the
 * last attempt, analyse.f90:406-414; the reference's third attempt would use
 * IPEA1, which is not built).  A fragment left without electrons takes eself
 * (iniqm.f90:430-451); a charge that needs a negative electron count or an SCC
 * failure gives stat[t] = -1 (the reference stops there, iniqm.f90:411-413). */
int qcxms_fragment_ips_batch(int32_t ntraj, const int32_t *offsets, const int32_t *num,
                             const double *axyz, const int32_t *list, const int32_t *nfrag,
                             const int32_t *mchrg, double *fragip, int32_t *ipok, int32_t *stat);

/* State of one trajectory for manage_fragments. */
typedef struct {
    int32_t mchrg;      /* in/out: total charge (write_fragments.f90:302-310) */
    int32_t nfrag_ok;   /* in/out: run still counted (0 after more than 5 fragments) */
    double chrgcont;    /* in/out: charge carried by the continued fragment */
    int32_t nfrag, tcont, ipok, nres, stat;  /* out; tcont 0: no continuation */
    int32_t icoll;      /* in: collision number of a CID run (> 0: the records carry
                           it, write_fragments.f90:411-420); 0: EI records */
    double fragchrg3[10]; /* IP-Boltzmann fragment charges (fractional) */
    double fragchrg[10];  /* summed Mulliken charges qat per fragment */
    double fragm[10];     /* fragment masses (amu) */
    double fragip[30];    /* IPs as qcxms_fragment_ips_batch */
    char asave[128];      /* record of the continued fragment ("" when tcont = 0) */
    char res[10][128];    /* records for qcxms.res, nres of them */
} qcxms_fragment_report;

/* manage_fragments (src/write_fragments.f90:12-457; EI records, or CID
 * records with the collision number when report[t].icoll > 0) for many
 * finished trajectories: fragment_structure + fragmass on the final
 * coordinates xyz (list[natoms] out), IPs of the fragments at the averaged
 * coordinates axyz (qcxms_md_read_averages), Boltzmann charges at
 * atlast[t] * btf (boltz, utility.f90:469-498), the continued fragment, the new
 * total charge and charge contribution, and the qcxms.res records with the
 * reference's format (F10.7,i3,2i5,2i2,2x,i3,2x,20(i4,i3)).  qat: the run's
 * averaged Mulliken charges; mass in m_e; imass isotope numbers or <= 0
 * (NULL: none); itrj[t] trajectory number; isec / maxsec the fragmentation
 * generation and its limit; at most 64 atoms per trajectory.  The first IP
 * difference of each fragment is its first IP, relative to the neutral
 * (write_fragments.f90:119-121 zero fragip(i,1) and then read fragip(i,0), an
 * element before the array, so the reference's value there depends on heap
 * contents). */
int qcxms_manage_fragments_batch(int32_t ntraj, const int32_t *offsets, const int32_t *num,
                                 const double *xyz, const double *axyz, const double *qat,
                                 const double *mass, const int32_t *imass, const double *atlast,
                                 const int32_t *itrj, int32_t isec, int32_t maxsec, double btf,
                                 int32_t *list, qcxms_fragment_report *report);

/* ---- CID: the collision MD of src/cid.f90:724-1111 (one collision, icoll = 1)
 * for many trajectories.  System s holds the ion's nion[s] atoms followed by
 * the collision gas (one atom, or two for N2; only gases whose elements have
 * GFN2 parameters here can run, i.e. N2); the last atom is the one whose
 * distance to the ion's centre of mass is tracked.  xyz and velo are the
 * state after cid()'s set-up (cid.f90:276-676: the ion's velocities include
 * the flight towards the gas); tinit[s] is the ion's initial temperature,
 * beta[s] = m_gas / (m_gas + m_ion), velo_cm[s] the centre-of-mass speed of
 * the previous collision (m/s, 0 for the first).  Each step: leap-frog of
 * ion and gas, GFN2 energy and gradient of the whole system with checkqc, the
 * ion's speed and temperature, its connectivity, the averaged fragment
 * structure, the distance test every 10 steps (five receding tests are the
 * collision: 800 more steps of 1 fs), the fragmentation extension, the stop
 * test. */
typedef struct qcxms_cid qcxms_cid;

typedef struct {
    double tstep;   /* time step (au) */
    double etemp;   /* electronic temperature (K), constant in CID */
    double econv;   /* SCC energy criterion (0: library default) */
    double pconv;   /* SCC moment criterion (0: library default) */
    int32_t ntot;   /* maximum number of steps (cid.f90:217: 15000) */
    int32_t reserved;
} qcxms_cid_params;

typedef struct {
    int32_t nstep, nfrag, done, gradfail, collided, fragmented, total_steps, collision_step,
        check_fragmented, count_average;
    double ttime, epot, ekin, new_velo, avgt, atlast, ecom, velo_diff, lowest_com;
    double fragT[10], eint[10];  /* fragment temperatures and kinetic energies at the end */
} qcxms_cid_result;

qcxms_cid *qcxms_cid_create(int32_t ntraj, const int32_t *offsets, const int32_t *nion,
                            const int32_t *num, const int32_t *charge, const double *mass,
                            const int32_t *imass, const double *xyz, const double *velo,
                            const double *tinit, const double *beta, const double *velo_cm,
                            const qcxms_cid_params *params);
void qcxms_cid_destroy(qcxms_cid *cd);
/* The two flags cid() keeps from one call to the next of a trajectory
 * (cid.f90:125,147: locals with an initial value, hence SAVE): collided[s]
 * (a collision was seen in an earlier call, so a fragmentation extends the
 * run at once) and count_average[s] (the averaged-structure count still
 * running).  Both start at 0; set them before the first step of a later
 * collision (icoll > 1).  Either pointer may be NULL (left unchanged). */
int qcxms_cid_set_saved(qcxms_cid *cd, const int32_t *collided, const int32_t *count_average);
/* advance every running trajectory by up to nsteps steps; *running = those
 * still going */
int qcxms_cid_run(qcxms_cid *cd, int32_t nsteps, int32_t *running);
/* state of all atoms (ion and gas), the ion's fragment labels, the averaged
 * structure cid returns (cid.f90:1097-1102; gas rows 0) and one record per
 * trajectory; any output may be NULL */
int qcxms_cid_read(qcxms_cid *cd, double *xyz, double *velo, double *grad, double *qat,
                   int32_t *list, double *axyz, qcxms_cid_result *results);
/* the ensemble's underlying batch, for the timing hooks */
qcxms_batch *qcxms_cid_batch(qcxms_cid *cd);

#ifdef __cplusplus
}
#endif
#endif
