// Shared device-side definitions of the MI355X GFN2-xTB engine.
//
// One workgroup owns one system (trajectory).  Per-system data lives in flat
// device arrays addressed through SysInfo offsets; per-system n x n matrices
// are stored row-major with a leading dimension ld = round_up(nao, 16) and
// zero padding, so that 16 x 16 FP64 MFMA tiles cover them exactly.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gfn2_params.h"

#define QX_BLOCK 256       // threads per system in the setup and gradient kernels
#define QX_SCC_BLOCK 256   // threads per system in the SCC kernel
#define QX_MAX_NAT 64
#define QX_MAX_NSH 128
#define QX_MAX_NAO 256
#define QX_MAX_REF 7
#define QX_LDS_MAXM 64  // Jacobi matrices up to this order live in LDS
#define QX_PAIR_STRIDE 16  // doubles per ordered atom pair in the gradient pair buffer
#define QX_AES_COMP 10     // multipole-kernel factors per ordered atom pair

#ifndef QX_DEFAULT_EXTRAPOLATE
#define QX_DEFAULT_EXTRAPOLATE 4  // MD start of the SCC when the caller passes 0
#endif

#define QX_DEV __device__ __forceinline__

struct ElemParams {
    int z, nsh, nprim[2], pqn[2];
    double alpha[2][6], coeff[2][6];  // scaled exponents, normalised contraction coefficients
    double slater[2], hself0[2], kcn[2], shpoly[2], refocc[2];  // self-energies/kcn in Eh
    double hard[2], third[2];                                     // shell hardness, Gamma_l
    double rep_alpha, rep_zeff, en, rad, dkernel, qkernel, mprad, mpvcn, nval, covrad;
    // D4
    int nref, refc[QX_MAX_REF];
    double refq[QX_MAX_REF], refcovcn[QX_MAX_REF], zeff, gam, r4r2, rcov_d4, en_d4;
};

struct DevTables {
    ElemParams e[gfn2::kElem];
    double c6ref[gfn2::kElem][QX_MAX_REF][gfn2::kElem][QX_MAX_REF];
};

extern __constant__ DevTables g_tab;
extern __constant__ double g_rad[118];  // QCxMS bond radii (covalent_radii.f90:89-114), Bohr

struct SysInfo {
    int nat, nsh, nao, ld, ndim, nx;
    int atom_off, shell_off, ref_off, x_off;
    long long mat_off;   // one ld*ld slot
    long long gam_off;   // nsh*nsh
    long long d4_off;    // ndim*ndim
    long long hist_off;  // Broyden history: 2*mem*nx + mem*mem
    long long pair_off;  // nat*nat*QX_PAIR_STRIDE gradient scratch
    int item_off;        // shell-pair work items: off-atom pairs first (grouped by
    int nitem_off;       // angular-momentum class), then same-atom pairs
    int nitem;
    int rowitem_off;     // the same pairs split into rows of the first shell's AOs
    int nrowitem;        // (rows of three AO pairs first), for the setup integrals
    int charge;
    double nel, kt;
};

struct SccOpts {
    double econv, pconv, damp, jacobi_tol;
    int maxiter, mem;
    int extrapolate;  // MD start: 1 linear 2 x(n-1) - x(n-2), 2 quadratic 3 (x(n-1) - x(n-2)) + x(n-3),
                      // 3 cubic 4 (x(n-1) + x(n-3)) - 6 x(n-2) - x(n-4),
                      // 4 quartic 5 (x(n-1) - x(n-4)) - 10 (x(n-2) - x(n-3)) + x(n-5)
    int keep;         // > 0: Broyden secant pairs kept across warm-started calls (ring of keep)
    int lds_doubles;  // dynamic LDS of the launch in doubles (set per launch by the host)
    int eig_mfma;     // 1: MFMA simultaneous-rotation eigen refinement before Jacobi
    int lds_max_m;    // systems of at most this (even) basis size keep their SCC matrices in
                      // LDS, the others in global memory (set per launch, <= qx::kFastM)
    int first_sys;    // system of workgroup 0 (a launch over a range of the batch's systems)
};

// Device-resident batch state.  Every pointer is a flat array over all systems.
struct BatchPtrs {
    int nsys;
    const SysInfo* info;
    const int* elem;  // per atom: element slot
    double* xyz;      // 3 per atom
    double* grad;     // 3 per atom
    double* qat;      // per atom
    double* cn;       // per atom (GFN2 CN)
    double* cnd4;     // per atom (D4 covalent CN)
    double* mrad;     // per atom
    double* dmrad;    // per atom
    double* hself;    // per shell
    double* gw;       // per D4 reference
    double* dgw;      // per D4 reference
    double* S;        // ld^2 per system
    double* H0;
    double* D;        // 3 ld^2 per system: dipole integral (mu, nu) about mu's atom at nu*ld+mu
    double* Q;        // 6 ld^2 per system: traceless quadrupole, same layout
    double* L;        // Cholesky factor of S (lower)
    double* X;        // L^-T (upper): X^T S X = 1
    double* C;        // eigenvectors (AO basis)
    double* T;        // scratch
    double* P;        // density
    double* W;        // energy-weighted density
    double* U;        // eigenvectors in the orthogonalised basis (warm start)
    double* E;        // scratch matrix of the MFMA eigen refinement (systems whose SCC
                      // matrices are in global memory)
    double* gam;      // nsh^2 per system
    double* wd4;      // ndim^2 per system
    double* x;        // mixing vector [qsh, dp, qp] (converged, warm start)
    double* xo;       // output moments of the last iteration
    double* vs;       // per shell potential at the output moments
    double* vd;       // 3 per atom
    double* vq;       // 6 per atom
    double* eps;      // orbital energies, ld per system
    double* focc;     // occupations, ld per system
    double* hist;     // Broyden history
    double* pairbuf;  // gradient scratch
    const int2* spitem;  // shell-pair work items (ia, ib), ia > ib or same shell
    double* spbuf;    // per shell pair (ia, ib): gradient on A (3), dE/dCN_A, dE/dCN_B
    double* aesg;     // multipole-kernel factors, QX_AES_COMP planes of nat x nat per system
    double* dblk;     // stored derivative blocks per off-atom shell-pair item (null: recomputed)
    double* xprev;    // mixing vector of the step before (extrapolated MD start)
    double* xprev2;   // and of the step before that
    double* xprev3;   // and of the one before that (cubic extrapolation)
    double* xprev4;   // and of the one before that (quartic extrapolation)
    double* energy;   // per system: total
    double* eparts;   // per system: 10 parts (band, es2, es3, aes, axc, d4, atm, rep, ts, scc)
    int* iters;       // per system
    unsigned long long* iters_total;  // per system: SCC iterations summed over calls
    int* status;      // per system: 0 ok, -1 not converged
    int* warm;        // per system: 1 if U/x hold a previous solution
    const int* active;  // per system: 0 skips the system (finished trajectories), may be null
};

QX_DEV int pad16(int n) { return (n + 15) & ~15; }

// block-wide sum (all threads must call); smem needs QX_BLOCK/64 doubles
QX_DEV double block_sum(double v, double* smem)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) smem[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < QX_BLOCK / 64; ++i) s += smem[i];
    __syncthreads();
    return s;
}

QX_DEV double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Shell descriptor built per system in LDS
struct ShellDesc {
    int at, l, el, k, ao0, nao;  // k: shell index within the element
};

// Fill shell and AO tables of one system (called by all threads).
QX_DEV void build_shell_tables(const SysInfo& si, const int* elem, ShellDesc* sh, int* ao_at,
                               int* ao_sh)
{
    if (threadIdx.x == 0) {
        int is = 0, ia = 0;
        for (int a = 0; a < si.nat; ++a) {
            int el = elem[si.atom_off + a];
            for (int k = 0; k < g_tab.e[el].nsh; ++k) {
                ShellDesc d;
                d.at = a;
                d.l = k;
                d.el = el;
                d.k = k;
                d.ao0 = ia;
                d.nao = k ? 3 : 1;
                sh[is] = d;
                for (int c = 0; c < d.nao; ++c) {
                    ao_at[ia] = a;
                    ao_sh[ia] = is;
                    ++ia;
                }
                ++is;
            }
        }
    }
    __syncthreads();
}
