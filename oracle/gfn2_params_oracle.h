/* GFN2-xTB parameters for the CPU oracle (H, C, N, O).
 *
 * TEST INFRASTRUCTURE ONLY.  The reference obtains these from tblite v0.2.1
 * (new_gfn2_calculator, called at src/tblite.f90:123), which is not in this
 * container.  The values below are the published GFN2-xTB parameters
 * (Bannwarth, Ehlert, Grimme, JCTC 15, 1652 (2019)) as recalled; they cannot
 * be checked against tblite here.  The HIP library carries its own copy.
 */
#ifndef QCXMS_GFN2_PARAMS_ORACLE_H
#define QCXMS_GFN2_PARAMS_ORACLE_H

typedef struct {
    int z, nsh;
    int l[2], n[2], ng[2];
    double zeta[2];        /* Slater exponents */
    double selfenergy[2];  /* eV */
    double kcn[2];         /* eV per unit CN */
    double shpoly[2];      /* shell polynomial coefficients */
    double refocc[2];      /* reference shell occupations */
    double hubbard;        /* atomic chemical hardness (Eh) */
    double shell_hubbard[2];
    double hubbard_deriv;  /* third-order parameter (Eh) */
    double rep_alpha, rep_zeff, en, atomic_rad_aa;
    double dkernel, qkernel, mprad, mpvcn;
    double nval;           /* valence electrons */
} orc_gfn2_elem;

static const orc_gfn2_elem orc_gfn2_elems[4] = {
    /* H */
    {1, 1, {0, 0}, {1, 0}, {3, 0}, {1.230000, 0.0}, {-10.707211, 0.0}, {-0.0500000, 0.0},
     {-0.00953618, 0.0}, {1.0, 0.0}, 0.405771, {0.0, 0.0}, 0.0800000,
     2.213717, 1.105388, 2.20, 0.32, 0.05563889, 0.00027431, 1.4, 1.0, 1.0},
    /* C */
    {6, 2, {0, 1}, {2, 2}, {4, 4}, {2.096432, 1.800000}, {-13.970922, -10.063292},
     {-0.0102144, 0.0161657}, {-0.02294321, -0.00271102}, {1.0, 3.0}, 0.538015,
     {0.0, 0.1056358}, 0.1500000, 1.247655, 4.231078, 2.55, 0.75, -0.00411674, 0.00213583,
     3.0, 3.0, 4.0},
    /* N */
    {7, 2, {0, 1}, {2, 2}, {4, 4}, {2.339881, 2.014332}, {-16.686243, -12.523956},
     {-0.0198407, 0.0230461}, {-0.08506003, -0.02504201}, {1.5, 3.5}, 0.461493,
     {0.0, 0.1164892}, -0.0639780, 1.682689, 5.242592, 3.04, 0.71, 0.03521273, 0.02026786,
     1.9, 3.0, 5.0},
    /* O */
    {8, 2, {0, 1}, {2, 2}, {4, 4}, {2.439742, 2.137023}, {-20.229985, -15.503117},
     {-0.0173888, 0.0279349}, {-0.14955291, -0.03350819}, {2.0, 4.0}, 0.451896,
     {0.0, 0.1497020}, -0.0517134, 2.165712, 5.784415, 3.44, 0.64, -0.04935670, -0.00310828,
     1.8, 2.0, 6.0},
};

/* global GFN2 parameters */
#define ORC_GFN2_KSHELL_S 1.85
#define ORC_GFN2_KSHELL_P 2.23
#define ORC_GFN2_ENSCALE 0.02
#define ORC_GFN2_WEXP 0.5
#define ORC_GFN2_GEXP 2.0
#define ORC_GFN2_THIRD_S 1.0
#define ORC_GFN2_THIRD_P 0.5
#define ORC_GFN2_REP_KEXP 1.5
#define ORC_GFN2_REP_KEXP_LIGHT 1.0
#define ORC_GFN2_AES_DMP3 3.0
#define ORC_GFN2_AES_DMP5 4.0
#define ORC_GFN2_AES_SHIFT 1.2
#define ORC_GFN2_AES_KEXP 4.0
#define ORC_GFN2_AES_RMAX 5.0
#define ORC_GFN2_CN_KA 10.0
#define ORC_GFN2_CN_KB 20.0
#define ORC_GFN2_CN_RSHIFT 2.0
#define ORC_GFN2_CN_CUTOFF 25.0
#define ORC_GFN2_REP_CUTOFF 25.0
/* D4 for GFN2 */
#define ORC_D4_S6 1.0
#define ORC_D4_S8 2.7
#define ORC_D4_A1 0.52
#define ORC_D4_A2 5.0
#define ORC_D4_S9 5.0
#define ORC_D4_ALP 16
#define ORC_D4_WF 6.0
#define ORC_D4_GA 3.0
#define ORC_D4_GC 2.0
/* D4 covalent CN, dftd4.f90:10-19 */
#define ORC_D4_K1 16.0
#define ORC_D4_K4 4.10451
#define ORC_D4_K5 19.08857
#define ORC_D4_K6 (2 * 11.28174 * 11.28174)
#define ORC_D4_CN_THR 1600.0
#define ORC_D4_R_THR 5000.0
#define ORC_D4_ATM_THR 1600.0

#endif
